"""Whole CID runs on the GPU (qcxms_amd.CidRun: the collision loop of
main.F90:1586-2110 with its mean-free-path MD loop, main.F90:1774-1974) for
two butane ions and a pentane ion hit by N2 at 60 eV, up to three collisions.

Every stage the run logs is replayed on the oracle from the same inputs:
each collision (cid_oracle.c, with the saved collided / count_average flags
of the trajectory's earlier collisions) and each mean-free-path MD
(md_oracle.c in its CID mode, with the speed and averaging flag handed over)
ends at the same step with the same fragment count, labels and flags on both
sides; the speed and the averaged structure agree to the bars below (the
stages run 800-1700 steps of hot ions).  The driver's own rules are checked
on the log: the stage order, nmax of each MD, the continued fragment
recentred on its nuclear-charge-weighted centre with its velocities kept,
the collision number in the qcxms.res records, and each run's end."""
import concurrent.futures as cf

import numpy as np
import pytest

import cid_cases as CC
import oracle_bind as ob

pytestmark = pytest.mark.gpu
ORC = ob.default_opts(econv=1e-12, pconv=1e-10)
ETEMP = 5000.0


def q():
    import qcxms_amd
    return qcxms_amd


_RUN = {}


def whole_run():
    if "run" not in _RUN:
        run = q().CidRun(CC.run_ions(), charge=1, elab_ev=60.0, collisions=3, rng=np.random.default_rng(11),
                         log=True, econv=1e-11, pconv=1e-9, etemp=ETEMP)
        run.run()
        _RUN["run"] = run
    return _RUN["run"]


def replay(e):
    if e["kind"] == "collision":
        return ob.cid_run(e["system"], charge=e["charge"], etemp=ETEMP, opts=ORC, collided=e["collided"],
                          count_average=e["count_average"])
    tr = e["traj"]
    return ob.md_run(tr, 0, max_steps=e["nmax"], etempin=ETEMP, opts=ORC, charge=tr["charge"], isec=tr["isec"],
                     cid=True, new_velo=tr["new_velo"], count_average=tr["count_average"])


def test_every_stage_matches_the_oracle():
    run = whole_run()
    stages = [e for e in run.log if e["kind"] in ("collision", "mfp")]
    assert sum(e["kind"] == "collision" for e in stages) >= 4 and sum(e["kind"] == "mfp" for e in stages) >= 2
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        refs = list(ex.map(replay, stages))
    for e, ref in zip(stages, refs):
        st = ref["state"]
        nuc = len(ref["list"])
        print(f"{e['kind']} traj {e['itrj']} icoll {e['icoll']}: nstep {e['nstep']}/{st.nstep} nfrag "
              f"{e['nfrag']}/{st.nfrag} v {e['new_velo']:.2f}/{st.new_velo:.2f} "
              f"daxyz {np.abs(e['axyz'][:nuc] - ref['axyz']).max():.1e}", flush=True)
        assert e["nstep"] == st.nstep and e["nfrag"] == st.nfrag
        assert np.array_equal(e["list"][:nuc], ref["list"])
        assert e["count_average_out"] == st.count_average
        assert abs(e["new_velo"] - st.new_velo) <= 1e-3 * max(st.new_velo, 1.0)
        assert np.abs(e["axyz"][:nuc] - ref["axyz"]).max() <= 1e-4
        if e["kind"] == "collision":
            assert e["gradfail"] == st.gradfail == 0 and e["collided_out"] == st.collided
        else:
            assert e["exit_reason"] == st.exit_reason and e["mdok"] == st.mdok
            assert e["check_fragmented"] == st.check_fragmented


def test_driver_rules_on_the_log():
    from qcxms_amd.cid_run import nmax_mfp
    run = whole_run()
    for t in run.trajs:
        log = [e for e in run.log if e["itrj"] == t.itrj]
        assert log[0]["kind"] == "collision" and log[0]["icoll"] == 1
        prev = None
        for e in log:
            if e["kind"] == "fragments":
                assert prev["kind"] in ("collision", "mfp") and e["icoll"] == prev["icoll"]
                assert e["isec"] == (1 if prev["kind"] == "collision" else prev["isec"])
            elif e["kind"] == "mfp":
                assert prev["kind"] == "fragments" and e["isec"] == prev["isec"] + 1
                tr = e["traj"]
                assert e["nmax"] == nmax_mfp(len(tr["num"]), e["isec"])
                # the continued fragment of the stage before, recentred on sum Z x / sum Z
                src = [x for x in log if x["kind"] in ("collision", "mfp")
                       and (x["icoll"], x.get("isec", 1)) == (prev["icoll"], prev["isec"])][0]
                nuc_src = len(prev["list"])
                sel = np.arange(nuc_src) if prev["tcont"] == 0 else np.nonzero(prev["list"] == prev["tcont"])[0]
                assert len(tr["num"]) == len(sel)
                assert np.array_equal(tr["velo"], src["velo"][sel])
                z = np.asarray(tr["num"], dtype=float)
                assert np.abs((z[:, None] * tr["xyz"]).sum(0)).max() < 1e-10
                assert np.abs(tr["xyz"] - src["xyz"][sel] - (tr["xyz"] - src["xyz"][sel])[0]).max() < 1e-10
                assert tr["cid"] and tr["new_velo"] == src["new_velo"]
            elif e["kind"] == "collision" and e["icoll"] > 1:
                assert prev["kind"] == "fragments" and prev["tcont"] == 0
                assert e["system"]["velo_cm"] > 800.0
            prev = e
        assert t.stop in ("small", "little mass", "low velocity", "collisions done", "too many fragments")
        if t.stop == "small":
            assert len(t.num) <= 7
        if t.stop == "collisions done":
            assert t.icoll == 3
        # every qcxms.res record carries the collision number (i5 after itrj)
        assert t.records and all(int(r[13:18]) == t.itrj and 1 <= int(r[18:23]) <= t.icoll for r in t.records)
