"""The oracle's MO spectrum: the orbitals of the converged SCC
(orc_gfn2_orbitals) and write_qmo's Mulliken MO populations
(orc_mo_populations, mo_energ.f90:36-52), pinned by closed forms, by the
symmetry of H2 and by the Mulliken charges of the same SCC, which the
populations weighted with the occupations must rebuild."""
import numpy as np

import oracle_bind as ob
from qcxms_amd import workloads as W

OPTS = ob.default_opts(econv=1e-12, pconv=1e-10)
ZVAL = {1: 1.0, 6: 4.0, 7: 5.0, 8: 6.0}


def test_populations_single_centre_and_two_centre_closed_forms():
    # orbitals on single atoms with S = 1: each orbital lives on its own atom
    aoat = np.array([0, 0, 1, 2], np.int32)
    q = ob.mo_populations(np.eye(4), np.eye(4), aoat, 3)
    np.testing.assert_allclose(q.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    assert q[0, 0] > 1 - 1e-9 and q[1, 0] > 1 - 1e-9 and q[2, 1] > 1 - 1e-9 and q[3, 2] > 1 - 1e-9
    # a bonding combination over two atoms with overlap s splits evenly; an
    # uneven one (c1, c2) gives c1^2 + c1 c2 s on the first atom
    s = 0.3
    S = np.array([[1.0, s], [s, 1.0]])
    c = np.array([[1.0], [1.0]]) / np.sqrt(2 + 2 * s)
    q = ob.mo_populations(np.hstack([c, c]), S, np.array([0, 1], np.int32), 2)
    np.testing.assert_allclose(q[0], [0.5, 0.5], atol=1e-12)
    c1, c2 = 0.8, 0.35
    nrm = np.sqrt(c1 * c1 + c2 * c2 + 2 * c1 * c2 * s)
    cc = np.array([[c1, 0.0], [c2, 1.0]]) / np.array([nrm, 1.0])
    q = ob.mo_populations(cc, S, np.array([0, 1], np.int32), 2)
    assert abs(q[0, 0] - (c1 * c1 + c1 * c2 * s) / nrm ** 2) < 1e-9


def test_h2_orbitals_split_evenly():
    num = np.array([1, 1], np.int32)
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])
    o = ob.gfn2_orbitals(num, xyz, 0, 1e-3, OPTS)
    assert o["stat"] == 0
    q = ob.mo_populations(o["coeff"], o["S"], o["aoat"], 2)
    np.testing.assert_allclose(q, 0.5, atol=1e-12)
    np.testing.assert_allclose(o["focc"], [2.0, 0.0], atol=1e-12)


def test_orbitals_orthonormal_and_same_fixed_point_as_egrad():
    num, xyz = W.octane(noise=0.02, seed=4)
    o = ob.gfn2_orbitals(num, xyz, 1, 300.0, OPTS)
    r = ob.gfn2_egrad(num, xyz, 1, 300.0, opts=OPTS)
    assert o["stat"] == 0 and o["energy"] == r["energy"]
    n = len(o["eps"])
    C, S = o["coeff"], o["S"]
    np.testing.assert_allclose(C.T @ S @ C, np.eye(n), atol=1e-10)
    assert np.all(np.diff(o["eps"]) >= 0)
    assert abs(o["focc"].sum() - 49.0) < 1e-9


def test_populations_rebuild_the_mulliken_charges():
    """sum_k focc_k pop_k(a) = Z_a - q_a, with pop_k(a) the unnormalised
    population (qmo times its orbital's norm 1 + nat 1e-10, minus 1e-10)."""
    num, xyz = W.octane(noise=0.02, seed=4)
    nat = len(num)
    o = ob.gfn2_orbitals(num, xyz, 1, 5000.0, OPTS)
    r = ob.gfn2_egrad(num, xyz, 1, 5000.0, opts=OPTS)
    q = ob.mo_populations(o["coeff"], o["S"], o["aoat"], nat)
    np.testing.assert_allclose(q.sum(axis=1), 1.0, atol=1e-14)
    norm = np.diag(o["coeff"].T @ o["S"] @ o["coeff"]) + nat * 1e-10
    raw = q * norm[:, None] - 1e-10
    gross = o["focc"] @ raw
    zval = np.array([ZVAL[int(z)] for z in num])
    np.testing.assert_allclose(zval - gross, r["qat"], atol=1e-9)
