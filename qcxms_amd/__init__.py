"""MI355X-native GFN2-xTB molecular-dynamics hot path of QCxMS.

The compute runs in libqcxms_gpu.so (HIP, gfx950) through the C ABI declared
in include/qcxms_gpu.h; this package is its Python host layer, mirroring the
reference's Fortran interface (src/tblite.f90 get_xtb_egrad, src/md.f90).
"""
from .gpu import (EXIT_REASONS, GFN1_XTB, GFN2_XTB, IPEA1_XTB, Batch, FragmentReport,
                  EiEnsemble, LibraryMissing, MdEnsemble, egrad_batch, egrad_many, fragment_ips_batch,
                  fragments_batch, get_tuning, get_xtb_egrad, get_xtb_mo_spectrum,
                  impactscale_batch, leapfrog_batch, library, library_path, manage_fragments_batch,
                  set_tuning)
from .cid import CidEnsemble, cid_setup, cid_setup_next
from .cid_run import CidRun

__all__ = ["EXIT_REASONS", "GFN1_XTB", "GFN2_XTB", "IPEA1_XTB", "Batch", "FragmentReport",
           "EiEnsemble", "LibraryMissing", "MdEnsemble", "egrad_batch", "egrad_many", "fragment_ips_batch",
           "fragments_batch", "get_tuning", "get_xtb_egrad", "get_xtb_mo_spectrum", "impactscale_batch", "leapfrog_batch", "library",
           "library_path", "manage_fragments_batch", "set_tuning", "CidEnsemble", "cid_setup", "cid_setup_next", "CidRun"]
