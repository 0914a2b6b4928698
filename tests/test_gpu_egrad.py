"""GPU GFN2-xTB energy/gradient/charges against the CPU oracle.

Tolerances (the hot path's acceptance bar): |dE| <= 1e-8 Eh,
max |dg| <= 1e-6 Eh/bohr, max |dq| <= 1e-6 e.  Both sides converge the SCC
well below these (GPU pconv 1e-9, oracle pconv 1e-10), so warm- and
cold-started iteration paths land on the same fixed point.
"""
import numpy as np
import pytest

import oracle_bind as ob
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu

E_TOL, G_TOL, Q_TOL = 1e-8, 1e-6, 1e-6
ORC_OPTS = dict(econv=1e-12, pconv=1e-10)


def gpu():
    import qcxms_amd
    return qcxms_amd


def water():
    return np.array([8, 1, 1], np.int32), np.array(
        [[0, 0, 0.1173], [0, 0.7572, -0.4692], [0, -0.7572, -0.4692]]) * W.AATOAU


def oracle(num, xyz, charge, etemp):
    r = ob.gfn2_egrad(num, xyz, charge, etemp, opts=ob.default_opts(**ORC_OPTS))
    assert r["stat"] == 0
    return r


def check(num, xyz, charge, etemp, e, g, q, label=""):
    r = oracle(num, xyz, charge, etemp)
    de = abs(e - r["energy"])
    dg = np.abs(g - r["grad"]).max()
    dq = np.abs(q - r["qat"]).max()
    print(f"{label} nat={len(num)} dE={de:.2e} dg={dg:.2e} dq={dq:.2e}", flush=True)
    assert de <= E_TOL, (label, de)
    assert dg <= G_TOL, (label, dg)
    assert dq <= Q_TOL, (label, dq)


def test_get_xtb_egrad_water():
    q = gpu()
    num, xyz = water()
    qat, e, g, st = q.get_xtb_egrad(num, xyz, 0, 1, q.GFN2_XTB, 300.0)
    assert st == 0
    check(num, xyz, 0, 300.0, e, g, qat, "water")


def test_octane_cation_5000K():
    q = gpu()
    num, xyz = W.octane(seed=0)
    b = q.Batch([(num, 1)])
    b.set_scc(econv=1e-11, pconv=1e-9)
    e, g, qat, st = b.egrad(xyz, 5000.0)
    assert st[0] == 0
    check(num, xyz, 1, 5000.0, e[0], g, qat, "octane+")


def test_unknown_method_and_element():
    q = gpu()
    num, xyz = water()
    _, _, _, st = q.get_xtb_egrad(num, xyz, 0, 1, q.GFN1_XTB, 300.0)
    assert st == 5
    num2 = np.array([17, 1], np.int32)
    _, _, _, st = q.get_xtb_egrad(num2, xyz[:2], 0, 1, q.GFN2_XTB, 300.0)
    assert st == 5


def test_ragged_batch_random_chains():
    q = gpu()
    systems, coords, meta = [], [], []
    for s in range(24):
        num, xyz, charge = W.random_chain(1000 + s, nheavy_max=12)
        systems.append((num, charge))
        coords.append(xyz)
        meta.append((num, xyz, charge))
    b = q.Batch(systems)
    b.set_scc(econv=1e-11, pconv=1e-9)
    e, g, qat, st = b.egrad(np.concatenate(coords), 5000.0)
    assert np.all(st == 0)
    off = b.offsets
    for i, (num, xyz, charge) in enumerate(meta):
        check(num, xyz, charge, 5000.0, e[i], g[off[i]:off[i + 1]], qat[off[i]:off[i + 1]],
              f"chain{i}")


def test_large_basis_outside_lds_path():
    """C20H36O2N2 cation: 132 basis functions, matrices beyond the LDS path."""
    q = gpu()
    num, xyz = W.chain_c20h36o2n2(seed=0)
    b = q.Batch([(num, 1)])
    b.set_scc(econv=1e-11, pconv=1e-9)
    e, g, qat, st = b.egrad(xyz, 5000.0)
    assert st[0] == 0
    check(num, xyz, 1, 5000.0, e[0], g, qat, "c20")


def test_warm_start_same_fixed_point():
    q = gpu()
    num, xyz = W.octane(seed=3)
    b = q.Batch([(num, 1)])
    b.set_scc(econv=1e-11, pconv=1e-9)
    b.set_warm_start(True)
    e0, g0, q0, _ = b.egrad(xyz, 5000.0)
    it0 = b.info()[0][0]
    # a new geometry (every atom displaced independently, as after an MD step)
    xyz2 = xyz + np.random.default_rng(11).normal(scale=0.01, size=xyz.shape)
    e1, g1, q1, st = b.egrad(xyz2, 5000.0)
    it1 = b.info()[0][0]
    assert st[0] == 0
    assert it1 < it0
    check(num, xyz2, 1, 5000.0, e1[0], g1, q1, "warm")


@pytest.mark.parametrize("case", ["octane", "c12_stretched_global", "ragged"])
def test_stored_derivative_blocks_give_the_same_gradient(case):
    """The gradient from stored derivative blocks (k_dblocks) equals the
    recomputed one to rounding, on LDS and global-memory systems and a ragged
    batch, and matches the oracle; the stored form holds more device memory."""
    q = gpu()
    if case == "octane":
        systems = [W.octane(seed=3)]
        charges = [1]
    elif case == "c12_stretched_global":
        systems = [W.chain_c12h24o2n2(seed=2)]
        charges = [1]
    else:
        systems, charges = [], []
        for s in range(6):
            num, xyz, chg = W.random_chain(300 + s)
            systems.append((num, xyz))
            charges.append(chg)
    b = q.Batch([(n, c) for (n, _), c in zip(systems, charges)])
    b.set_scc(econv=1e-11, pconv=1e-9)
    xyz = np.concatenate([x for _, x in systems])
    e0, g0, q0, st0 = b.egrad(xyz, 5000.0)
    mem0 = b.device_bytes()
    b.set_stored_derivatives(True)
    e1, g1, q1, st1 = b.egrad(xyz, 5000.0)
    mem1 = b.device_bytes()
    b.set_stored_derivatives(False)
    e2, g2, _, _ = b.egrad(xyz, 5000.0)
    b.close()
    dg = np.abs(g1 - g0).max()
    print(f"{case}: max|g_stored - g_recomputed| {dg:.2e}; device memory {mem0 / 2**20:.1f} -> "
          f"{mem1 / 2**20:.1f} MiB", flush=True)
    assert np.all(st0 == 0) and np.all(st1 == 0)
    assert np.array_equal(e0, e1) and np.array_equal(q0, q1)
    assert dg <= 1e-10
    assert np.array_equal(g2, g0)
    assert mem1 > mem0
    off = 0
    for i, ((num, x), c) in enumerate(zip(systems, charges)):
        k = slice(off, off + len(num))
        check(num, x, c, 5000.0, e1[i], g1[k], q1[k], f"{case} system {i}")
        off += len(num)
