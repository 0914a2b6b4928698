"""CPU oracle MD bookkeeping pinned bit-for-bit to the reference's Fortran
(golden vectors from oracle/_ref; live cross-check when it is built)."""
import numpy as np
import pytest

import oracle_bind as ob


def split(a, off, k):
    return a[off[k]:off[k + 1]]


def test_tables_match_reference(golden):
    L = ob.oracle()
    for z in range(1, 119):
        assert L.orc_rad(z) == golden["rad"][z - 1]
        assert L.orc_atomic_mass(z) == golden["mass"][z - 1]
        assert L.orc_covrad_d3(z) == golden["covrad_d3"][z - 1]


def test_fragment_structure_golden(golden):
    L = ob.oracle()
    off = golden["frag_off"]
    for k in range(len(off) - 1):
        oz = split(golden["frag_oz"], off, k).astype(np.int32)
        xyz = golden["frag_xyz"][off[k]:off[k + 1]]
        want = split(golden["frag_out"], off, k)
        got = ob.fragment_structure(L, "orc_", oz, xyz)
        assert np.array_equal(got, want), k


def test_fragment_labels_are_canonical_components(golden):
    """Labels equal connected components numbered by lowest atom index."""
    L = ob.oracle()
    rng = np.random.default_rng(7)
    rad = golden["rad"]
    for _ in range(50):
        nat = int(rng.integers(1, 40))
        oz = rng.choice([1, 6, 7, 8], size=nat).astype(np.int32)
        xyz = rng.normal(0, 5.0, size=(nat, 3))
        got = ob.fragment_structure(L, "orc_", oz, xyz)
        d = np.sqrt(((xyz[:, None] - xyz[None]) ** 2).sum(-1))
        adj = d < 3.0 * 0.5 * (rad[oz - 1][:, None] + rad[oz - 1][None])
        comp = -np.ones(nat, int)
        lab = 0
        for i in range(nat):
            if comp[i] >= 0:
                continue
            lab += 1
            stack = [i]
            comp[i] = lab
            while stack:
                a = stack.pop()
                for b in np.nonzero(adj[a])[0]:
                    if comp[b] < 0:
                        comp[b] = lab
                        stack.append(b)
        assert np.array_equal(got, comp)


def test_fragmass_golden(golden):
    L = ob.oracle()
    off = golden["frag_off"]
    for k in range(len(off) - 1):
        if golden["fm_nfrag"][k] < 0:
            continue
        oz = split(golden["frag_oz"], off, k).astype(np.int32)
        frag = split(golden["frag_out"], off, k).astype(np.int32)
        mass = golden["mass"][oz - 1]
        nf, fm, fa = ob.fragmass(L, "orc_", oz, frag, mass)
        assert nf == golden["fm_nfrag"][k]
        assert np.array_equal(fm, golden["fm_fragm"][k])
        assert np.array_equal(fa, golden["fm_fragat"][k])


def test_ekinet_impactscale_golden(golden):
    L = ob.oracle()
    off = golden["ek_off"]
    for k in range(len(off) - 1):
        v = golden["ek_velo"][off[k]:off[k + 1]]
        m = golden["ek_mass"][off[k]:off[k + 1]]
        e, t = ob.ekinet(L, "orc_", v, m)
        assert e == golden["ek_out"][k][0] and t == golden["ek_out"][k][1]
    off = golden["is_off"]
    for k in range(len(off) - 1):
        v = golden["is_velo"][off[k]:off[k + 1]]
        m = golden["is_mass"][off[k]:off[k + 1]]
        f = golden["is_velof"][off[k]:off[k + 1]]
        eimp, ff, e0 = golden["is_par"][k]
        got, kk = ob.impactscale(L, "orc_", v, m, f, eimp, ff, e0)
        assert kk < 20000
        assert np.array_equal(got, golden["is_out"][off[k]:off[k + 1]])


def test_leapfrog_closed_forms():
    """Zero force: linear motion; constant force: exact quadratic positions."""
    import ctypes as C
    L = ob.oracle()
    rng = np.random.default_rng(3)
    nat = 7
    mass = rng.uniform(1800, 30000, nat)
    x0 = rng.normal(size=(nat, 3))
    v0 = rng.normal(scale=1e-3, size=(nat, 3))
    g = rng.normal(scale=1e-2, size=(nat, 3))
    dt = 0.5 * 41.3413733365614
    x, v, ke = x0.copy(), v0.copy(), C.c_double()
    nsteps = 50
    for _ in range(nsteps):
        L.orc_leapfrog(nat, ob.ptr(g), ob.ptr(mass), C.c_double(dt), ob.ptr(x), ob.ptr(v), C.byref(ke))
    a = -g / mass[:, None]
    # v_n = v0 + n dt a ; x_n = x0 + dt sum_{k=1..n} v_k = x0 + n dt v0 + dt^2 a n(n+1)/2
    np.testing.assert_allclose(v, v0 + nsteps * dt * a, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(x, x0 + nsteps * dt * v0 + dt * dt * a * nsteps * (nsteps + 1) / 2,
                               rtol=1e-11, atol=1e-12)
    vm = v - 0.5 * dt * a
    assert abs(ke.value - 0.5 * np.sum(mass[:, None] * vm * vm)) < 1e-12 * abs(ke.value)


def test_checkqc_and_getspin():
    import ctypes as C
    L = ob.oracle()
    g = np.zeros((3, 3)); g[0, 1] = 1.0; g[1, 2] = 30.0  # z never enters gnorm
    L.orc_gnorm.argtypes = [C.c_int, ob.dp]
    assert L.orc_gnorm(3, ob.ptr(g)) == np.sqrt(2.0)
    num = np.array([6, 1, 1, 1, 1], np.int32)
    assert L.orc_getspin(5, ob.ptr(num, ob.ip), 0) == 1
    assert L.orc_getspin(5, ob.ptr(num, ob.ip), 1) == 2
    h = np.array([1], np.int32)
    assert L.orc_getspin(1, ob.ptr(h, ob.ip), 1) == -1
    e = C.c_double(-5.0)
    q = np.array([0.5, 0.5, 0.0])
    g2 = np.ones((3, 3)) * 0.01
    assert L.orc_checkqc(3, C.byref(e), ob.ptr(g2), ob.ptr(q), 1) == 1
    g3 = np.ones((3, 3)) * 20.0
    assert L.orc_checkqc(3, C.byref(e), ob.ptr(g3), ob.ptr(q), 1) == 0 and e.value == 0.0
    assert L.orc_setetemp(1, 0.3, 0.0, 0.0) == 5000.0


@pytest.mark.skipif(ob.ref() is None, reason="reference Fortran not built")
def test_live_against_reference_fortran():
    L, R = ob.oracle(), ob.ref()
    rng = np.random.default_rng(11)
    rad = np.array([R.ref_rad(z) for z in range(1, 119)])
    for _ in range(100):
        nat = int(rng.integers(1, 70))
        oz = rng.integers(1, 87, size=nat).astype(np.int32)
        xyz = rng.normal(0, float(rng.choice([2.0, 5.0, 10.0])), size=(nat, 3))
        assert np.array_equal(ob.fragment_structure(L, "orc_", oz, xyz),
                              ob.fragment_structure(R, "ref_", oz, xyz))


def test_md_averages_window_closed_form():
    """md.f90:352-363, 691-698: the running sums restart when kdump exceeds
    avdump - 1 = 49 (at the first step and every 50 steps), collect the
    geometry and charges before each leap-frog step, and are divided by
    kdump at the end.  Under a constant force the geometries are the exact
    quadratic x_n = x0 + n dt v0 + dt^2 a n(n+1)/2, so after 56 steps the
    average is the mean of x_50 .. x_55 (six terms)."""
    import ctypes as C
    L = ob.oracle()
    num = np.array([6, 1], dtype=np.int32)
    mass = np.array([12.0, 1.008]) * 1822.888486
    x0 = np.array([[0.0, 0.0, 0.0], [2.05, 0.0, 0.0]])
    v0 = np.array([[1e-5, -2e-5, 0.5e-5], [1e-5, -2e-5, 0.5e-5]])
    g = np.array([[1e-4, 0.0, -1e-4], [1e-4 * 1.008 / 12.0, 0.0, -1e-4 * 1.008 / 12.0]])
    q = np.array([0.3, 0.7])

    @ob.EGRAD_FN
    def egrad(ctx, nat, iat, xyz, chrg, etemp, e, grad, qat):
        e[0] = -5.0
        for k in range(3 * nat):
            grad[k] = g.reshape(-1)[k]
        for k in range(nat):
            qat[k] = q[k]
        return 0

    st = ob.MdState()
    st.nat, st.chrg, st.nfragexit, st.max_steps = 2, 1, 3, 56
    dt = 0.5 * 41.3413733365614
    st.tstep, st.tadd, st.eimp, st.ax, st.ieetemp, st.etempin = dt, 0.0, 0.0, 0.0, 0.0, 300.0
    x, v = x0.copy(), v0.copy()
    grad, qat = np.zeros((2, 3)), np.zeros(2)
    lst = np.zeros(2, dtype=np.int32)
    axyz, achrg = np.zeros((2, 3)), np.zeros(2)
    n = L.orc_md_run_avg(C.byref(st), ob.ptr(num, ob.ip), ob.ptr(np.full(2, -1, np.int32), ob.ip),
                         ob.ptr(mass), ob.ptr(np.ones(2)), ob.ptr(x), ob.ptr(v), ob.ptr(grad),
                         ob.ptr(qat), ob.ptr(lst, ob.ip), egrad, None, 0, ob.ptr(axyz), ob.ptr(achrg))
    assert n == 56 and st.exit_reason == 7
    a = -g / mass[:, None]
    xs = [x0 + k * dt * v0 + dt * dt * a * k * (k + 1) / 2 for k in range(50, 56)]
    np.testing.assert_allclose(axyz, np.mean(xs, axis=0), rtol=0, atol=1e-12)
    np.testing.assert_allclose(achrg, q, rtol=0, atol=1e-15)


def test_impactscale_increments_are_exact():
    """impactscale (impact.f90:12-55) adds the REAL(4) literal 0.0002 to a
    REAL(8) scal once per iteration from 0.  The increment is a multiple of
    2^-36, so every partial sum below 4 (20000 iterations) is exact and equal
    to (k - 1) * 0.0002f: the property the batched search evaluates the
    iterations independently by."""
    c = float(np.float32(0.0002))
    s = 0.0
    for k in range(1, 20001):
        assert s == (k - 1) * c
        s = s + c


def test_intenergy_is_ekinet_of_each_fragment():
    """md.f90:715-741: the temperature of fragment j is the ekinet temperature
    (mdinit.f90:57-76, pinned to the compiled reference) of its atoms alone."""
    import ctypes as C
    L = ob.oracle()
    rng = np.random.default_rng(8)
    nat = 23
    lst = rng.integers(1, 5, nat).astype(np.int32)
    lst[:4] = [1, 2, 3, 4]
    mass = rng.uniform(1800.0, 30000.0, nat)
    velo = rng.normal(0.0, 3e-4, (nat, 3))
    t = np.zeros(10)
    L.orc_intenergy.argtypes = [C.c_int, ob.ip, ob.dp, ob.dp, C.c_int, ob.dp]
    L.orc_intenergy(nat, ob.ptr(lst, ob.ip), ob.ptr(mass), ob.ptr(np.ascontiguousarray(velo)), 4,
                    ob.ptr(t))
    for j in range(1, 5):
        sel = lst == j
        e, temp = C.c_double(), C.c_double()
        L.orc_ekinet(int(sel.sum()), ob.ptr(np.ascontiguousarray(velo[sel])),
                     ob.ptr(np.ascontiguousarray(mass[sel])), C.byref(e), C.byref(temp))
        assert abs(t[j - 1] - temp.value) < 1e-12 * temp.value, (j, t[j - 1], temp.value)
