"""The allocation-free entry point qcxms_egrad_batch_noalloc against
qcxms_egrad_batch, both through the C ABI: two calls in a row (the second
with fewer, smaller systems, so that state left in the shared device array
by the first call must not leak into it), each to the parity bar, and a
batch too large for the array rejected by the host."""
import ctypes as C

import numpy as np
import pytest

from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu
dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)


def _lib():
    import qcxms_amd
    L = qcxms_amd.library()
    for name in ("qcxms_egrad_batch", "qcxms_egrad_batch_noalloc"):
        getattr(L, name).argtypes = [C.c_int32, ip, ip, dp, ip, ip, C.c_int32, dp, dp, dp, dp, ip]
        getattr(L, name).restype = C.c_int
    return L


def _call(name, systems):
    L = _lib()
    off = np.cumsum([0] + [len(n) for n, _, _ in systems]).astype(np.int32)
    num = np.concatenate([n for n, _, _ in systems]).astype(np.int32)
    xyz = np.ascontiguousarray(np.concatenate([x for _, x, _ in systems]), dtype=np.float64)
    chg = np.array([c for _, _, c in systems], dtype=np.int32)
    mult = np.array([W.getspin(n, c) for n, _, c in systems], dtype=np.int32)
    et = np.full(len(systems), 5000.0)
    nat, ns = int(off[-1]), len(systems)
    q, e, g, st = np.zeros(nat), np.zeros(ns), np.zeros((nat, 3)), np.zeros(ns, dtype=np.int32)
    p = lambda a, t=dp: a.ctypes.data_as(t)
    rc = getattr(L, name)(ns, p(off, ip), p(num, ip), p(xyz), p(chg, ip), p(mult, ip), 2, p(et), p(q),
                          p(e), p(g), p(st, ip))
    return rc, e, g, q, st


def _systems(seeds, big=()):
    out = []
    for s in seeds:
        num, xyz, chg = W.random_chain(900 + s, nheavy_max=12)
        out.append((num, xyz, chg))
    for s in big:
        num, xyz = W.chain_c12h24o2n2(seed=s)  # 88 basis functions: global-memory SCC path
        out.append((num, xyz, 1))
    num, xyz = W.octane(seed=3)
    out.append((num, xyz, 1))
    return out


def _check(systems, label):
    rc0, e0, g0, q0, s0 = _call("qcxms_egrad_batch", systems)
    rc1, e1, g1, q1, s1 = _call("qcxms_egrad_batch_noalloc", systems)
    print(f"{label}: {len(systems)} systems, rc {rc0}/{rc1}, max|dE| {np.abs(e1 - e0).max():.2e} "
          f"max|dg| {np.abs(g1 - g0).max():.2e} max|dq| {np.abs(q1 - q0).max():.2e} "
          f"identical {np.array_equal(e0, e1) and np.array_equal(g0, g1)}", flush=True)
    assert rc0 == rc1 == 0 and np.all(s0 == 0) and np.array_equal(s0, s1)
    assert np.abs(e1 - e0).max() <= 1e-8
    assert np.abs(g1 - g0).max() <= 1e-6
    assert np.abs(q1 - q0).max() <= 1e-6


def test_noalloc_matches_egrad_batch_twice_then_smaller():
    _check(_systems(range(10), big=(1,)), "first call")
    _check(_systems(range(20, 23)), "second call (smaller)")


def test_noalloc_rejects_a_batch_beyond_its_array():
    num, xyz = W.octane(seed=1)
    systems = [(num, xyz, 1)] * 200
    rc, e, g, q, st = _call("qcxms_egrad_batch_noalloc", systems)
    from qcxms_amd.gpu import last_error
    msg = last_error()
    print(f"200 octanes: rc {rc}: {msg}", flush=True)
    assert rc == -2 and "scratch" in msg
    # and the entry point still works afterwards
    _check(_systems(range(30, 32)), "after the rejection")
