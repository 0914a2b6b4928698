"""Pins for the CPU oracle's GFN2-xTB restatement: closed forms, invariants,
finite differences, a library eigensolver and the reference's in-tree D4.
(tblite itself is not available here: absolute parity with it is unpinned.)"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.linalg

import oracle_bind as ob
from qcxms_amd import workloads as W

AA = W.AATOAU


def water():
    return np.array([8, 1, 1], np.int32), np.array(
        [[0, 0, 0.1173], [0, 0.7572, -0.4692], [0, -0.7572, -0.4692]]) * AA


def fd_gradient(num, xyz, charge, etemp, h=1e-4, opts=None):
    fd = np.zeros_like(xyz)
    for a in range(len(num)):
        for c in range(3):
            xp = xyz.copy(); xp[a, c] += h
            xm = xyz.copy(); xm[a, c] -= h
            ep = ob.gfn2_egrad(num, xp, charge, etemp, opts=opts, grad=False)["energy"]
            em = ob.gfn2_egrad(num, xm, charge, etemp, opts=opts, grad=False)["energy"]
            fd[a, c] = (ep - em) / (2 * h)
    return fd


TIGHT = dict(econv=1e-12, pconv=1e-11)


@pytest.mark.parametrize("case", ["water", "cnoc_cation", "c4h10_cation"])
def test_gradient_matches_finite_differences(case):
    if case == "water":
        num, xyz = water(); charge, etemp = 0, 300.0
    elif case == "cnoc_cation":
        num, xyz = W.build_chain([6, 7, 8, 6], noise=0.05, seed=5); charge, etemp = 1, 5000.0
    else:
        num, xyz = W.build_chain([6, 6, 6, 6], noise=0.08, seed=9); charge, etemp = 1, 5000.0
    opts = ob.default_opts(**TIGHT)
    r = ob.gfn2_egrad(num, xyz, charge, etemp, opts=opts)
    assert r["stat"] == 0
    fd = fd_gradient(num, xyz, charge, etemp, opts=opts)
    assert np.abs(fd - r["grad"]).max() < 2e-8


def test_invariants_octane_cation():
    num, xyz = W.octane(seed=0)
    r = ob.gfn2_egrad(num, xyz, 1, 5000.0)
    assert r["stat"] == 0 and r["info"].converged
    assert abs(r["qat"].sum() - 1.0) < 1e-10
    g = r["grad"]
    assert np.abs(g.sum(0)).max() < 1e-10                      # translation
    assert np.abs(np.cross(xyz, g).sum(0)).max() < 1e-9        # rotation
    # rigid rotation + translation leaves E unchanged
    th = 0.7
    rot = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    r2 = ob.gfn2_egrad(num, xyz @ rot.T + np.array([1.0, -2.0, 0.5]), 1, 5000.0)
    assert abs(r2["energy"] - r["energy"]) < 1e-9


def test_size_consistency_two_methanes():
    num1, x1 = W.build_chain([6], noise=0.0)
    num = np.concatenate([num1, num1])
    xyz = np.concatenate([x1, x1 + np.array([60.0, 0, 0])])
    e1 = ob.gfn2_egrad(num1, x1, 0, 300.0, grad=False)["energy"]
    e2 = ob.gfn2_egrad(num, xyz, 0, 300.0, grad=False)["energy"]
    assert abs(e2 - 2 * e1) < 1e-7


def test_overlap_closed_forms():
    num, xyz = W.build_chain([6, 8], noise=0.1, seed=2)
    s, h0, d, q = ob.gfn2_matrices(num, xyz)
    assert np.allclose(np.diag(s), 1.0, atol=1e-14)
    assert np.allclose(s, s.T, atol=1e-15)
    assert np.linalg.eigvalsh(s).min() > 0
    assert np.allclose(h0, h0.T, atol=1e-15)
    # dipole integrals about the row atom vs the column atom: D^A - D^B = (B - A) S
    # check on the C-O pair block: atoms C,H,H,H,O,H -> AOs C 0..3, O 7..10
    assert list(num) == [6, 1, 1, 1, 8, 1]
    ra, rb = xyz[0], xyz[4]
    for c in range(3):
        da = d[c][0:4, 7:11]         # <C| (r - A) |O>
        db = d[c][7:11, 0:4].T       # <C| (r - B) |O>
        np.testing.assert_allclose(da - db, (rb[c] - ra[c]) * s[0:4, 7:11], atol=1e-13)
    # quadrupole integrals are traceless
    tr = q[0] + q[2] + q[5]
    assert np.abs(tr).max() < 1e-13


def test_ss_overlap_primitive_closed_form():
    L = ob.oracle()
    a, b = np.array([0.7]), np.array([1.9])
    ca = np.array([(2 * 0.7 / math.pi) ** 0.75]); cb = np.array([(2 * 1.9 / math.pi) ** 0.75])
    ra = np.array([0.1, -0.3, 0.2]); rb = np.array([1.0, 0.4, -0.6])
    s = np.zeros(9); dpb = np.zeros(27); qpb = np.zeros(54)
    L.orc_overlap_shellpair(0, 1, ob.ptr(a), ob.ptr(ca), 0, 1, ob.ptr(b), ob.ptr(cb), ob.ptr(ra),
                            ob.ptr(rb), ob.ptr(s), ob.ptr(dpb), ob.ptr(qpb))
    r2 = np.sum((ra - rb) ** 2)
    want = ca[0] * cb[0] * (math.pi / 2.6) ** 1.5 * math.exp(-0.7 * 1.9 / 2.6 * r2)
    assert abs(s[0] - want) < 1e-15
    # dipole about B: <a|(r-B)|b> = S (P - B), P the Gaussian product centre
    p = (0.7 * ra + 1.9 * rb) / 2.6
    for c in range(3):
        assert abs(dpb[c * 9] - want * (p[c] - rb[c])) < 1e-15


def test_generalized_eigensolver_against_scipy():
    L = ob.oracle()
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 50):
        a = rng.normal(size=(n, n)); f = a + a.T
        b = rng.normal(size=(n, n)); s = b @ b.T + n * np.eye(n)
        fc, sc = f.copy(), s.copy()
        e = np.zeros(n); work = np.zeros(3 * n * n + 2 * n)
        L.orc_sygv(n, ob.ptr(fc), ob.ptr(sc), ob.ptr(e), ob.ptr(work))
        ew = scipy.linalg.eigh(f, s, eigvals_only=True)
        np.testing.assert_allclose(e, ew, rtol=1e-12, atol=1e-12)
        cmat = fc
        np.testing.assert_allclose(cmat.T @ s @ cmat, np.eye(n), atol=1e-11)
        np.testing.assert_allclose(f @ cmat, s @ cmat * e[None, :], atol=1e-10)


def test_d4_reference_tables_match_reference(golden):
    L = ob.oracle()
    L.orc_d4_literal_mode(1)
    try:
        for k, z in enumerate((1, 6, 7, 8)):
            nref = C.c_int(); refal = np.zeros((7, 23)); refc = np.zeros(7, np.int32)
            rc = np.zeros(7); rq = np.zeros(7)
            L.orc_d4_reference(z, C.byref(nref), ob.ptr(refal), ob.ptr(refc, ob.ip), ob.ptr(rc),
                               ob.ptr(rq))
            n = nref.value
            assert n == golden["d4_nref"][k]
            assert np.array_equal(refc[:n], golden["d4_refc"][k][:n])
            np.testing.assert_allclose(refal[:n], golden["d4_refal"][k][:n], rtol=1e-14, atol=1e-14)
            np.testing.assert_array_equal(rc[:n], golden["d4_refcovcn"][k][:n])
    finally:
        L.orc_d4_literal_mode(0)


def test_d4_pieces_match_reference(golden):
    L = ob.oracle()
    for a, c, qref, qmod, z, dz in golden["d4_zeta"]:
        assert L.orc_d4_zeta_fn(a, c, qref, qmod) == pytest.approx(z, rel=1e-15, abs=1e-300)
        assert L.orc_d4_dzeta_fn(a, c, qref, qmod) == pytest.approx(dz, rel=1e-14, abs=1e-300)
    L.orc_d4_trapzd.argtypes = [ob.dp]
    for p, t in zip(golden["d4_trapzd_in"], golden["d4_trapzd_out"]):
        assert L.orc_d4_trapzd(ob.ptr(np.ascontiguousarray(p))) == pytest.approx(t, rel=1e-15)


def test_d4_energy_gradient_match_reference_dispgrad(golden):
    """Two-body at charges q plus ATM at q=0, energy and gradient, against the
    reference's dispgrad (mbd=3) on the same inputs."""
    L = ob.oracle()
    L.orc_d4_twobody.argtypes = [C.c_int, ob.ip, ob.dp, ob.dp, ob.dp, ob.dp]
    L.orc_d4_atm.argtypes = [C.c_int, ob.ip, ob.dp, ob.dp]
    L.orc_d4_literal_mode(1)
    try:
        off = golden["disp_off"]
        for k in range(len(off) - 1):
            oz = golden["disp_oz"][off[k]:off[k + 1]].astype(np.int32)
            xyz = np.ascontiguousarray(golden["disp_xyz"][off[k]:off[k + 1]])
            q = np.ascontiguousarray(golden["disp_q"][off[k]:off[k + 1]])
            nat = len(oz)
            g2 = np.zeros((nat, 3)); g3 = np.zeros((nat, 3)); dq = np.zeros(nat)
            e2 = L.orc_d4_twobody(nat, ob.ptr(oz, ob.ip), ob.ptr(xyz), ob.ptr(q), ob.ptr(g2), ob.ptr(dq))
            e3 = L.orc_d4_atm(nat, ob.ptr(oz, ob.ip), ob.ptr(xyz), ob.ptr(g3))
            assert e2 + e3 == pytest.approx(golden["disp_e"][k], rel=1e-12)
            np.testing.assert_allclose(g2 + g3, golden["disp_g"][off[k]:off[k + 1]], rtol=1e-9,
                                       atol=1e-13)
    finally:
        L.orc_d4_literal_mode(0)


def test_d4_charge_potential_by_finite_difference():
    L = ob.oracle()
    L.orc_d4_twobody.argtypes = [C.c_int, ob.ip, ob.dp, ob.dp, ob.dp, ob.dp]
    num, xyz = W.build_chain([6, 7, 8, 6], noise=0.05, seed=5)
    rng = np.random.default_rng(1)
    q = rng.normal(0, 0.3, len(num))
    dq = np.zeros(len(num))
    L.orc_d4_twobody(len(num), ob.ptr(num, ob.ip), ob.ptr(xyz), ob.ptr(q), None, ob.ptr(dq))
    h = 1e-5
    for i in range(len(num)):
        qp = q.copy(); qp[i] += h; qm = q.copy(); qm[i] -= h
        ep = L.orc_d4_twobody(len(num), ob.ptr(num, ob.ip), ob.ptr(xyz), ob.ptr(qp), None, None)
        em = L.orc_d4_twobody(len(num), ob.ptr(num, ob.ip), ob.ptr(xyz), ob.ptr(qm), None, None)
        assert abs((ep - em) / (2 * h) - dq[i]) < 1e-9
