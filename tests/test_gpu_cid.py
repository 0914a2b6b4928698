"""CID collision MD on the GPU (qcxms_cid_*: the loop of cid.f90:724-1111 for
many trajectories) against the oracle's restatement (cid_oracle.c) on ions
hit by N2.

Over the first 60 steps both follow the same trajectory (|dx| <= 1e-8 bohr,
|dv| <= 1e-10, |dE| <= 1e-8 Eh).  Over the whole run the bookkeeping agrees
exactly: the step of the collision (five receding distance tests), the
extended step count, the fragmentation, the step the run ends and the
fragment labels; the continuous quantities to the bars below."""
import numpy as np
import pytest

import cid_cases as CC
import oracle_bind as ob

pytestmark = pytest.mark.gpu
ORC = ob.default_opts(econv=1e-12, pconv=1e-10)
GPU = dict(econv=1e-11, pconv=1e-9)


def q():
    from qcxms_amd import cid
    return cid


def test_first_steps_follow_the_oracle():
    cs = CC.cases()
    names = list(cs)
    e = q().CidEnsemble([cs[n] for n in names], charge=1, **GPU)
    e.run(60)
    out = e.read()
    e.close()
    off = np.cumsum([0] + [len(cs[n]["num"]) for n in names])
    for k, n in enumerate(names):
        ref = ob.cid_run(cs[n], charge=1, opts=ORC, steps_limit=60)
        r, st = out["results"][k], ref["state"]
        sl = slice(off[k], off[k + 1])
        dx = np.abs(out["xyz"][sl] - ref["xyz"]).max()
        dv = np.abs(out["velo"][sl] - ref["velo"]).max()
        print(f"{n}: nstep {r.nstep}/{st.nstep} dE {abs(r.epot - st.epot):.2e} dx {dx:.2e} dv {dv:.2e}",
              flush=True)
        assert r.nstep == st.nstep == 60
        assert abs(r.epot - st.epot) <= 1e-8
        assert dx <= 1e-8 and dv <= 1e-10
        assert abs(r.ttime - st.ttime) <= 1e-12
        assert abs(r.avgt - st.avgt) <= 1e-6 * abs(st.avgt)
        assert abs(r.new_velo - st.new_velo) <= 1e-6 * max(abs(st.new_velo), 1.0)
        assert abs(r.ecom - st.ecom) <= 1e-8 * max(abs(st.ecom), 1.0)


def test_whole_collisions_match_the_oracle():
    cs = CC.cases()
    names = list(cs)
    e = q().CidEnsemble([cs[n] for n in names], charge=1, **GPU)
    running = e.run(3000)
    out = e.read()
    e.close()
    assert running == 0
    off = np.cumsum([0] + [len(cs[n]["num"]) for n in names])
    for k, n in enumerate(names):
        ref = ob.cid_run(cs[n], charge=1, opts=ORC)
        r, st = out["results"][k], ref["state"]
        nuc = cs[n]["nion"]
        sl = slice(off[k], off[k] + nuc)
        print(f"{n}: nstep {r.nstep}/{st.nstep} collision {r.collision_step}/{st.collision_step} "
              f"total {r.total_steps}/{st.total_steps} nfrag {r.nfrag}/{st.nfrag} "
              f"dE {abs(r.epot - st.epot):.2e} fragT {list(np.round(r.fragT[:r.nfrag]))}", flush=True)
        for key in ("nstep", "collision_step", "total_steps", "nfrag", "fragmented", "gradfail",
                    "check_fragmented", "count_average"):
            assert getattr(r, key) == getattr(st, key), key
        assert r.collided == 1 and r.done == 1
        assert np.array_equal(out["list"][sl], ref["list"])
        assert abs(r.epot - st.epot) <= 1e-6
        for f in range(st.nfrag):
            assert abs(r.fragT[f] - st.fragT[f]) <= 1e-4 * st.fragT[f]
        assert np.abs(out["axyz"][sl] - ref["axyz"]).max() <= 1e-5


def test_gradient_failure_ends_the_run():
    """The gas placed 0.2 bohr from a hydrogen: the initial gradient norm is
    above checkqc's limit (iniqm.f90:684-715), cid() stops before the loop on
    both sides (stopcid)."""
    sysd = dict(CC.cases()["c2_60eV"])
    sysd["xyz"] = sysd["xyz"].copy()
    sysd["xyz"][-1] = sysd["xyz"][sysd["nion"] - 1] + np.array([0.2, 0.0, 0.0])
    e = q().CidEnsemble([sysd], charge=1, **GPU)
    running = e.run(10)
    r = e.read()["results"][0]
    e.close()
    st = ob.cid_run(sysd, charge=1, opts=ORC)["state"]
    assert running == 0
    assert (r.gradfail, r.nstep, r.done) == (st.gradfail, st.nstep, 1) == (1, 0, 1)


def test_octane_with_n2_global_memory_path_follows_the_oracle():
    """C8H18+ + N2 (28 atoms, 58 basis functions: SCC matrices in global
    memory, MFMA eigen refinement), the system of the CID benchmark, over 60
    steps of the approach."""
    from qcxms_amd import workloads as W
    t = W.md_ensemble(1, base_seed=1, temp=500.0, tadd_fs=100.0)[0]
    sysd = q().cid_setup(t["num"], t["xyz"], t["velo"], t["mass"], CC.TSTEP, gas="N2", elab_ev=40.0,
                         rand=(0.4, 0.6, 0.2, 0.8))
    assert sum(1 if z == 1 else 4 for z in sysd["num"]) == 58
    e = q().CidEnsemble([sysd], charge=1, **GPU)
    e.run(60)
    out = e.read()
    e.close()
    ref = ob.cid_run(sysd, charge=1, opts=ORC, steps_limit=60)
    r, st = out["results"][0], ref["state"]
    dx = np.abs(out["xyz"] - ref["xyz"]).max()
    print(f"octane + N2: dE {abs(r.epot - st.epot):.2e} dx {dx:.2e}", flush=True)
    assert r.nstep == st.nstep == 60
    assert abs(r.epot - st.epot) <= 1e-8 and dx <= 1e-8
    assert np.abs(out["velo"] - ref["velo"]).max() <= 1e-10
