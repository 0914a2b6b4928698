"""The C ABI boundary without a GPU: libqcxms_gpu.so loads and exports every
function that include/qcxms_gpu.h declares, and the Fortran drop-in module
(qcxms_amd/fortran/qcxms_tblite.f90, replacing src/tblite.f90) compiled and
linked against it.  No compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qcxms_gpu.h")
LIB = os.path.join(ROOT, "qcxms_amd", "libqcxms_gpu.so")
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "egrad_driver")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(qcxms_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("qcxms_get_xtb_egrad", "qcxms_egrad_batch", "qcxms_md_create", "qcxms_md_run",
                 "qcxms_md_read", "qcxms_fragments_batch", "qcxms_leapfrog_batch",
                 "qcxms_impactscale_batch"):
        assert must in names


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_are_plain_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in declared_functions():
        assert n in exported, n


def test_fortran_drop_in_module_linked():
    if not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("no Fortran compiler in this image")
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    out = subprocess.run(["nm", DRIVER], capture_output=True, text=True, check=True).stdout
    assert "get_xtb_egrad" in out and "qcxms_get_xtb_egrad" in out
    mod = os.path.join(ROOT, "qcxms_amd", "fortran", "_build", "qcxms_tblite.mod")
    assert os.path.exists(mod)


def test_tuning_table_is_validated_before_any_device_call():
    """qcxms_set_tuning rejects malformed tables on the host (no GPU needed)."""
    lib = C.CDLL(LIB)
    lib.qcxms_set_tuning.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.c_int32, C.c_int32]
    lib.qcxms_gpu_last_error.restype = C.c_char_p
    arr = lambda *v: (C.c_int32 * len(v))(*v)
    assert lib.qcxms_set_tuning(2, arr(40, 24), arr(0, 0), arr(1, 1), -1, 0) == -2
    assert b"increasing" in lib.qcxms_gpu_last_error()
    assert lib.qcxms_set_tuning(1, arr(40), arr(2), arr(1), -1, 0) == -2
    assert b"0 or 1" in lib.qcxms_gpu_last_error()
    assert lib.qcxms_set_tuning(33, arr(*range(1, 34)), arr(*[0] * 33), arr(*[0] * 33), -1, 0) == -2
    assert lib.qcxms_set_tuning(1, arr(40), arr(1), arr(1), 2, 0) == -2
    assert b"mixed" in lib.qcxms_gpu_last_error()
