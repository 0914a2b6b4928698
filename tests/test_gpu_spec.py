"""The MO spectrum of a batch on the GPU (qcxms_batch_mo_spectrum: write_qmo's
Mulliken MO populations, mo_energ.f90:36-52) against the oracle's orbitals
and populations of the same SCC."""
import numpy as np
import pytest

import oracle_bind as ob
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu
OPTS = ob.default_opts(econv=1e-12, pconv=1e-10)


@pytest.mark.parametrize("case", ["octane_neutral_aufbau", "octane_cation_5000K", "c12_stretched"])
def test_mo_spectrum_matches_oracle(case):
    import qcxms_amd
    if case == "octane_neutral_aufbau":
        num, xyz = W.octane(noise=0.02, seed=4)
        charge, etemp = 0, 1e-3
    elif case == "octane_cation_5000K":
        num, xyz = W.octane(noise=0.02, seed=5)
        charge, etemp = 1, 5000.0
    else:
        # 88 basis functions: the SCC matrices live in global memory
        num, xyz = W.chain_c12h24o2n2(seed=1)
        charge, etemp = 1, 5000.0
    nat = len(num)
    b = qcxms_amd.Batch([(num, charge)])
    b.set_scc(econv=1e-11, pconv=1e-9)
    e, g, q, st = b.egrad(xyz, etemp)
    emo, focc, qmo, ihomo = b.mo_spectrum(0)
    b.close()
    o = ob.gfn2_orbitals(num, xyz, charge, etemp, OPTS)
    ref = ob.mo_populations(o["coeff"], o["S"], o["aoat"], nat)
    print(f"{case}: nat {nat} nao {len(emo)} ihomo {ihomo} max|d eps| "
          f"{np.abs(emo - o['eps']).max():.2e}", flush=True)
    assert st[0] == 0 and o["stat"] == 0
    assert len(emo) == len(o["eps"])
    np.testing.assert_allclose(emo, o["eps"], atol=1e-7)
    np.testing.assert_allclose(focc, o["focc"], atol=1e-6)
    occ = np.nonzero(o["focc"] > 1e-6)[0]
    assert ihomo == occ.max() + 1
    gap = np.minimum(np.diff(o["eps"], prepend=-np.inf), np.diff(o["eps"], append=np.inf))
    keep = gap > 1e-3
    assert keep.sum() >= len(emo) * 3 // 4
    np.testing.assert_allclose(qmo[keep], ref[keep], atol=1e-6)
    np.testing.assert_allclose(qmo.sum(axis=1), 1.0, atol=1e-12)
