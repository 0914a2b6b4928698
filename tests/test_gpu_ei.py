"""Fragmentation generations (main.F90:2208-2345) on the GPU against the
oracle's restatement, and the device re-layout of running trajectories.

Every generation must agree with the oracle: MD length, exit, mdok,
fragment count, the fragment continued (tcont), the new total charge, the
atoms of the generation and their fragment labels (bit-exact), and the
qcxms.res records (integers exactly, charges to 2e-7)."""
import numpy as np
import pytest

import ei_cases as EC
import oracle_bind as ob

pytestmark = pytest.mark.gpu
NMAX = 60
GPU_SCC = dict(econv=1e-11, pconv=1e-9)
ORC = ob.default_opts(econv=1e-12, pconv=1e-10)
KEYS = ("isec", "nat", "nstep", "exit_reason", "mdok", "fragstate", "nfrag", "max_steps", "tcont",
        "mchrg", "nfrag_ok", "nres", "continued", "asave_written")


def q():
    import qcxms_amd
    return qcxms_amd


def _records_equal(a, b):
    fa, fb = a.split(), b.split()
    assert len(fa) == len(fb), (a, b)
    assert abs(float(fa[0]) - float(fb[0])) <= 2e-7, (a, b)
    assert fa[1:] == fb[1:], (a, b)


def _compare(g, o):
    print(f"  traj {g['traj']} isec {g['isec']}: " + " ".join(f"{k} {g[k]}/{o[k]}" for k in KEYS[2:9]),
          flush=True)
    for k in KEYS:
        assert g[k] == o[k], (k, g[k], o[k])
    assert np.array_equal(g["atoms"], o["atoms"])
    if g["mdok"]:
        assert np.array_equal(g["list"], o["list"])
    assert abs(g["chrgcont"] - o["chrgcont"]) < 1e-6
    assert abs(g["ttime"] - o["ttime"]) < 1e-9
    assert abs(g["epot"] - o["epot"]) < 1e-6
    for ra, rb in zip(g["res"], o["res"]):
        _records_equal(ra, rb)
    assert bool(g["asave"]) == bool(o["asave"].strip())
    if g["asave"]:
        _records_equal(g["asave"], o["asave"])


def test_generations_match_oracle_through_relayouts():
    cs = EC.cases()
    names = ["split2", "split4", "intact"]
    ei = q().EiEnsemble([cs[n] for n in names], max_steps=NMAX, use_graph=False, repack=20, **GPU_SCC)
    running = ei.run(160)
    gens = ei.generations()
    repacks = ei.repacks()
    ei.close()
    print(f"{len(gens)} generations, {repacks} re-layouts, {running} running", flush=True)
    assert running == 0 and repacks >= 2
    for t, name in enumerate(names):
        ref = ob.ei_run(cs[name], nmax=NMAX, itrj=t + 1, opts=ORC)
        mine = [g for g in gens if g["traj"] == t]
        print(f"{name}: {len(mine)} generations (oracle {len(ref)})", flush=True)
        assert len(mine) == len(ref)
        for g, o in zip(mine, ref):
            _compare(g, o)


@pytest.mark.parametrize("fused", [False, True], ids=["per_step", "persistent"])
def test_relayout_keeps_a_running_trajectory_bit_identical(fused):
    """The intact trajectory runs on while the split one ends at its first
    step: re-laid out after 10 steps (with the split trajectory's next
    generation beside it) it must give the same bits as when it stays in the
    first layout for all its steps."""
    cs = EC.cases()
    trajs = [cs["split4"], cs["intact"]]
    out = []
    for repack in (10, 1000):
        ei = q().EiEnsemble(trajs, max_steps=NMAX, use_graph=False, fused=fused, repack=repack,
                            maxsec=1)
        ei.run(NMAX)
        out.append([g for g in ei.generations() if g["traj"] == 1][0])
        n = ei.repacks()
        ei.close()
        print(f"repack {repack}: {n} re-layouts, intact epot {out[-1]['epot']!r}", flush=True)
        if repack == 10:
            assert n >= 1
    a, b = out
    assert a["nstep"] == b["nstep"] == NMAX
    assert a["epot"] == b["epot"] and a["ttime"] == b["ttime"]
    assert a["res"] == b["res"]
