"""Pins of the oracle's GFN2 pieces by what mathematics and physics fix,
independent of the code under test (tblite v0.2.1, which defines the model,
is absent; tblite.f90:123-136 is the call it would answer):

* overlap, dipole and quadrupole integrals of s and p Gaussian shells
  against brute-force quadrature on a 3D grid;
* the ES2 kernel: on-site value = shell hardness, R -> 0 limit between atoms
  = arithmetic mean of the two hardnesses, R -> infinity = 1/R, and
  1/gamma^2 - R^2 constant (the exponent-2 Klopman-Ohno form);
* the anisotropic (AES2) energy of two distant atoms = the classical
  charge-dipole, dipole-dipole and charge-quadrupole energies of point
  charges carrying the same multipoles;
* the SCC potentials = derivatives of the interaction energy;
* the Fermi entropy term -TS = kT dF/dkT of the free energy at fixed
  orbital energies and electron count.
Each check fails for a plausible slip (a sign, a factor 3/2 or 2, a
transposed index, harmonic instead of arithmetic mean)."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob

dp, ip = ob.dp, ob.ip


def L():
    lib = ob.oracle()
    lib.orc_overlap_shellpair.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp,
                                          dp, dp, dp, dp, dp]
    lib.orc_sto_ng.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, dp, dp]
    lib.orc_gfn2_gamma.argtypes = [C.c_int, ip, dp, dp, C.POINTER(C.c_int)]
    lib.orc_gfn2_interaction.argtypes = [C.c_int, ip, dp, dp, dp, dp, dp, dp]
    lib.orc_fermi_fill.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp]
    lib.orc_fermi_fill.restype = C.c_double
    return lib


# ---------------------------------------------------------------- integrals
def _shell(ng, n, l, zeta):
    a, c = np.zeros(6), np.zeros(6)
    assert L().orc_sto_ng(ng, n, l, zeta, ob.ptr(a), ob.ptr(c)) == 0
    return a[:ng].copy(), c[:ng].copy()


def _prim(l, a):
    """One normalised primitive (coefficient includes the normalisation)."""
    norm = (2 * a / np.pi) ** 0.75 * (np.sqrt(4 * a) if l else 1.0)
    return np.array([a]), np.array([norm])


def _values(l, alpha, coeff, centre, x, y, z):
    """AO values of a shell on grid points: s, or p as x, y, z."""
    dx, dy, dz = x - centre[0], y - centre[1], z - centre[2]
    r2 = dx * dx + dy * dy + dz * dz
    rad = sum(c * np.exp(-a * r2) for a, c in zip(alpha, coeff))
    return [rad] if l == 0 else [dx * rad, dy * rad, dz * rad]


def _quadrature(la, sa, lb, sb, ra, rb, h, half):
    """S, D (about B), raw Q (about B) by the trapezoid rule on a cube of
    side 2*half around the midpoint (the integrands vanish at its faces)."""
    mid = 0.5 * (ra + rb)
    ax = np.arange(-half, half + 1e-12, h)
    na, nb = (1 if la == 0 else 3), (1 if lb == 0 else 3)
    S = np.zeros((na, nb))
    D = np.zeros((3, na, nb))
    Q = np.zeros((6, na, nb))
    qidx = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]
    X, Y = np.meshgrid(ax + mid[0], ax + mid[1], indexing="ij")
    for zv in ax + mid[2]:
        Z = np.full_like(X, zv)
        fa = _values(la, *sa, ra, X, Y, Z)
        fb = _values(lb, *sb, rb, X, Y, Z)
        rel = [X - rb[0], Y - rb[1], Z - rb[2]]
        for i in range(na):
            for j in range(nb):
                w = fa[i] * fb[j]
                S[i, j] += w.sum()
                for c in range(3):
                    D[c, i, j] += (w * rel[c]).sum()
                for k, (u, v) in enumerate(qidx):
                    Q[k, i, j] += (w * rel[u] * rel[v]).sum()
    v = h ** 3
    return S * v, D * v, Q * v


def _oracle_pair(la, sa, lb, sb, ra, rb):
    s, d, q = np.zeros(9), np.zeros(27), np.zeros(54)
    aa, ca = (np.ascontiguousarray(t) for t in sa)
    ab, cb = (np.ascontiguousarray(t) for t in sb)
    L().orc_overlap_shellpair(la, len(aa), ob.ptr(aa), ob.ptr(ca), lb, len(ab), ob.ptr(ab),
                              ob.ptr(cb), ob.ptr(np.ascontiguousarray(ra)),
                              ob.ptr(np.ascontiguousarray(rb)), ob.ptr(s), ob.ptr(d), ob.ptr(q))
    na, nb = (1 if la == 0 else 3), (1 if lb == 0 else 3)
    return (s.reshape(3, 3)[:na, :nb], d.reshape(3, 3, 3)[:, :na, :nb],
            q.reshape(6, 3, 3)[:, :na, :nb])


PAIRS = {
    "s-s": (0, _prim(0, 1.3), 0, _prim(0, 0.7)),
    "s-p": (0, _prim(0, 0.9), 1, _prim(1, 1.6)),
    "p-s": (1, _prim(1, 2.2), 0, _prim(0, 0.5)),
    "p-p": (1, _prim(1, 1.1), 1, _prim(1, 0.8)),
    "H1s-C2p": (0, _shell(3, 1, 0, 1.23), 1, _shell(4, 2, 1, 1.8)),
}


@pytest.mark.parametrize("name", sorted(PAIRS))
def test_shell_pair_integrals_match_grid_quadrature(name):
    la, sa, lb, sb = PAIRS[name]
    ra = np.array([0.3, -0.2, 0.1])
    rb = ra + np.array([1.1, 0.7, -0.9])  # 1.58 bohr, no symmetry axis
    s, d, q = _oracle_pair(la, sa, lb, sb, ra, rb)
    S, D, Q = _quadrature(la, sa, lb, sb, ra, rb, h=0.09, half=9.0)
    np.testing.assert_allclose(s, S, atol=2e-10)
    np.testing.assert_allclose(d, D, atol=2e-10)
    np.testing.assert_allclose(q, Q, atol=2e-10)
    assert np.abs(S).max() > 0.05  # the pair overlaps noticeably


# ---------------------------------------------------------------- ES2 kernel
def _gamma(num, xyz):
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nsh = C.c_int()
    L().orc_gfn2_gamma(len(num), ob.ptr(num, ip), ob.ptr(xyz), None, C.byref(nsh))
    g = np.zeros(nsh.value ** 2)
    L().orc_gfn2_gamma(len(num), ob.ptr(num, ip), ob.ptr(xyz), ob.ptr(g), C.byref(nsh))
    return g.reshape(nsh.value, nsh.value)


def test_es2_kernel_limits_and_form():
    num = [1, 6]  # shells: H s | C s, C p
    u = np.array([0.36, -0.48, 0.8])
    onsite = np.diag(_gamma(num, [[0, 0, 0], [50.0, 0, 0]]))
    # R -> 0: arithmetic mean of the two shells' hardnesses
    g0 = _gamma(num, [[0, 0, 0], 1e-7 * u])
    for a in range(3):
        for b in range(3):
            assert abs(g0[a, b] - 0.5 * (onsite[a] + onsite[b])) < 1e-12
    # R -> infinity: 1/R
    R = 1e5
    gi = _gamma(num, [[0, 0, 0], R * u])
    assert abs(gi[0, 1] * R - 1.0) < 1e-9 and abs(gi[0, 2] * R - 1.0) < 1e-9
    # in between: 1/gamma^2 - R^2 does not depend on R
    vals = []
    for R in (0.5, 1.7, 3.2, 8.0):
        g = _gamma(num, [[0, 0, 0], R * u])
        vals.append(1.0 / g[0, 2] ** 2 - R * R)
    assert np.ptp(vals) < 1e-11 * max(vals)
    assert abs(vals[0] - (2.0 / (onsite[0] + onsite[2])) ** 2) < 1e-12


# ----------------------------------------------------- AES2 and potentials
def _interaction(num, xyz, x):
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nat = len(num)
    nsh = len(x) - 9 * nat
    e = np.zeros(5)
    vs, vd, vq = np.zeros(nsh), np.zeros(3 * nat), np.zeros(6 * nat)
    L().orc_gfn2_interaction(nat, ob.ptr(num, ip), ob.ptr(xyz), ob.ptr(x), ob.ptr(e), ob.ptr(vs),
                             ob.ptr(vd), ob.ptr(vq))
    return e, np.concatenate([vs, vd, vq])


def _point_charges(q, mu, theta, centre, d=0.5):
    """Point charges with total charge q, dipole mu and traceless (Buckingham)
    quadrupole theta = 1/2 sum q_k (3 r r - r^2 1) about centre."""
    pts = [(q, np.array(centre, float))]
    mu = np.asarray(mu, float)
    if np.linalg.norm(mu) > 0:
        u = mu / np.linalg.norm(mu)
        qq = np.linalg.norm(mu) / (2 * d)
        pts += [(qq, centre + d * u), (-qq, centre - d * u)]
    lam, vec = np.linalg.eigh(theta)
    for k in range(3):
        qk = lam[k] / (3 * d * d)
        v = vec[:, k]
        pts += [(qk, centre + d * v), (qk, centre - d * v), (-2 * qk, np.array(centre, float))]
    return pts


def _coulomb(pa, pb):
    return sum(qa * qb / np.linalg.norm(ra - rb) for qa, ra in pa for qb, rb in pb)


def _theta6(t):
    """6-vector (xx, xy, yy, xz, yz, zz) of a symmetric 3x3 matrix."""
    return np.array([t[0, 0], t[0, 1], t[1, 1], t[0, 2], t[1, 2], t[2, 2]])


def test_aes_energy_of_distant_atoms_is_classical_multipole_energy():
    """Each AES2 term on its own: charge-dipole, dipole-dipole and
    charge-quadrupole energies of C and O 20000 bohr apart against point
    charges that carry the same multipoles (the damping is 6 r0^3 / R^3
    ~ 1e-11 relative there)."""
    rng = np.random.default_rng(3)
    num = [6, 8]
    u = np.array([0.48, 0.6, -0.64])
    R = 2.0e4
    ra, rb = np.zeros(3), R * u
    q = [0.31, -0.47]
    mu = [rng.normal(size=3) * 0.4, rng.normal(size=3) * 0.3]
    th = []
    for _ in range(2):
        a = rng.normal(size=(3, 3)) * 0.3
        a = 0.5 * (a + a.T)
        th.append(a - np.trace(a) / 3 * np.eye(3))
    zero, nil = np.zeros((3, 3)), np.zeros(3)

    def aes(qa, qb, ma, mb, ta, tb):
        # shell charges: C s, C p | O s, O p carrying the atomic charges
        x = np.concatenate([[0.1, qa - 0.1, -0.2, qb + 0.2], ma, mb, _theta6(ta), _theta6(tb)])
        return _interaction(num, [ra, rb], x)[0][2]

    def classical(qa, qb, ma, mb, ta, tb):
        return _coulomb(_point_charges(qa, ma, ta, ra), _point_charges(qb, mb, tb, rb))

    e_qq = classical(q[0], q[1], nil, nil, zero, zero)
    want = {
        "charge-dipole": classical(q[0], q[1], mu[0], mu[1], zero, zero) - e_qq
        - (classical(0, 0, mu[0], mu[1], zero, zero) - classical(0, 0, nil, nil, zero, zero)),
        "dipole-dipole": classical(0, 0, mu[0], mu[1], zero, zero),
        "charge-quadrupole": classical(q[0], q[1], nil, nil, th[0], th[1]) - e_qq
        - classical(0, 0, nil, nil, th[0], th[1]),
    }
    got = {
        "charge-dipole": aes(q[0], q[1], mu[0], mu[1], zero, zero)
        - aes(0, 0, mu[0], mu[1], zero, zero),
        "dipole-dipole": aes(0, 0, mu[0], mu[1], zero, zero),
        "charge-quadrupole": aes(q[0], q[1], nil, nil, th[0], th[1]),
    }
    for k in want:
        assert abs(got[k] - want[k]) < 1e-6 * abs(want[k]), (k, got[k], want[k])
        assert abs(want[k]) > 1e-15


def test_scc_potentials_are_interaction_energy_derivatives():
    from qcxms_amd import workloads as W
    num, xyz = W.build_chain([6, 8], noise=0.05, seed=4)
    nat = len(num)
    nsh = int(sum(1 if z == 1 else 2 for z in num))
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 0.2, nsh), rng.normal(0, 0.1, 3 * nat),
                        rng.normal(0, 0.1, 6 * nat)])
    _, v = _interaction(num, xyz, x)
    h = 1e-5
    for k in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd = (_interaction(num, xyz, xp)[0].sum() - _interaction(num, xyz, xm)[0].sum()) / (2 * h)
        assert abs(fd - v[k]) < 2e-8 * max(1.0, abs(v[k])), (k, fd, v[k])


# ------------------------------------------------------------ Fermi entropy
def _fermi(eps, kt, nalpha):
    f = np.zeros(len(eps))
    mts = L().orc_fermi_fill(len(eps), ob.ptr(np.ascontiguousarray(eps)), kt, nalpha, ob.ptr(f))
    return f, mts


def test_fermi_entropy_term_is_kt_times_free_energy_derivative():
    rng = np.random.default_rng(7)
    eps = np.sort(rng.normal(-0.3, 0.25, 24))
    nalpha = 6.5  # an odd electron count, half an electron per spin
    for kt in (5000 * 3.166808578545117e-6, 300 * 3.166808578545117e-6 * 10):
        f, mts = _fermi(eps, kt, nalpha)
        assert abs(f.sum() - nalpha) < 1e-12
        free = lambda t: 2.0 * float(_fermi(eps, t, nalpha)[0] @ eps) + _fermi(eps, t, nalpha)[1]
        h = kt * 1e-4
        dfdt = (free(kt + h) - free(kt - h)) / (2 * h)
        # F = E - T S at fixed N: dF/dT = -S, so kT dF/dkT = -T S
        assert abs(kt * dfdt - mts) < 1e-7 * abs(mts), (kt * dfdt, mts)
        assert mts < 0.0
