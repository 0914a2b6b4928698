"""GPU MD bookkeeping: bit-exact against the reference Fortran (golden
vectors) and the oracle; batched MD trajectories against the oracle MD."""
import numpy as np
import pytest

import oracle_bind as ob
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu
TSTEP = 0.5 * W.FSTOAU


def q():
    import qcxms_amd
    return qcxms_amd


def test_fragments_bitexact_against_reference(golden):
    off = golden["frag_off"]
    nums, xyzs, masses, idx = [], [], [], []
    for k in range(len(off) - 1):
        oz = golden["frag_oz"][off[k]:off[k + 1]].astype(np.int32)
        if len(oz) > 64:
            continue
        nums.append(oz)
        xyzs.append(golden["frag_xyz"][off[k]:off[k + 1]])
        masses.append(golden["mass"][oz - 1])
        idx.append(k)
    lists, nf, fm = q().fragments_batch(nums, xyzs, masses)
    for j, k in enumerate(idx):
        want = golden["frag_out"][off[k]:off[k + 1]]
        assert np.array_equal(lists[j], want), k
        if golden["fm_nfrag"][k] >= 0:
            assert nf[j] == golden["fm_nfrag"][k]
            assert np.array_equal(fm[j], golden["fm_fragm"][k]), k


def test_impactscale_bitexact_against_reference(golden):
    off = golden["is_off"]
    n = len(off) - 1
    velos = [golden["is_velo"][off[k]:off[k + 1]] for k in range(n)]
    masses = [golden["is_mass"][off[k]:off[k + 1]] for k in range(n)]
    velofs = [golden["is_velof"][off[k]:off[k + 1]] for k in range(n)]
    par = golden["is_par"]
    v, it = q().impactscale_batch(velos, masses, velofs, par[:, 0], par[:, 1], par[:, 2])
    assert np.all(it < 20000)
    assert np.array_equal(v, golden["is_out"])


def test_leapfrog_bitexact_against_oracle():
    import ctypes as C
    L = ob.oracle()
    rng = np.random.default_rng(42)
    grads, masses, xyzs, velos, want_x, want_v, want_e = [], [], [], [], [], [], []
    for s in range(30):
        nat = int(rng.integers(1, 65))
        m = W.atomic_mass(6) * rng.uniform(0.08, 1.4, nat)
        g = rng.normal(0, 0.05, (nat, 3)); x = rng.normal(0, 5, (nat, 3)); v = rng.normal(0, 1e-3, (nat, 3))
        grads.append(g); masses.append(m); xyzs.append(x); velos.append(v)
        x2, v2, ke = x.copy(), v.copy(), C.c_double()
        L.orc_leapfrog(nat, ob.ptr(g), ob.ptr(m), C.c_double(TSTEP), ob.ptr(x2), ob.ptr(v2), C.byref(ke))
        want_x.append(x2); want_v.append(v2); want_e.append(ke.value)
    x, v, ek = q().leapfrog_batch(grads, masses, TSTEP, xyzs, velos)
    assert np.array_equal(x, np.concatenate(want_x))
    assert np.array_equal(v, np.concatenate(want_v))
    assert np.array_equal(ek, np.array(want_e))


def ensemble(n, seed0=1, tadd_fs=5.0):
    return W.md_ensemble(n, base_seed=seed0, tadd_fs=tadd_fs)


def test_md_ensemble_matches_oracle_md():
    trajs = ensemble(3)
    nsteps = 56  # crosses the avdump = 50 reset of the running averages
    md = q().MdEnsemble(trajs, max_steps=nsteps, use_graph=False, econv=1e-11, pconv=1e-9)
    running = md.run(nsteps + 5)
    out = md.read()
    axyz, achrg = md.read_averages()
    assert running == 0
    opts = ob.default_opts(econv=1e-12, pconv=1e-10)
    for t, tr in enumerate(trajs):
        ref = ob.md_run(tr, steps_limit=0, max_steps=nsteps, opts=opts)
        st, res = ref["state"], out["results"][t]
        sl = slice(md.offsets[t], md.offsets[t + 1])
        print(f"traj {t}: nstep {res.nstep}/{st.nstep} exit {res.exit_reason}/{st.exit_reason} "
              f"dEpot {abs(res.epot - st.epot):.2e} dEkin {abs(res.ekin - st.ekin):.2e} "
              f"dx {np.abs(out['xyz'][sl] - ref['xyz']).max():.2e}", flush=True)
        assert res.nstep == st.nstep == nsteps
        assert res.exit_reason == st.exit_reason == 7
        assert res.nfrag == st.nfrag
        assert np.array_equal(out["list"][sl], ref["list"])
        assert abs(res.epot - st.epot) < 1e-8
        assert abs(res.ekin - st.ekin) < 1e-9
        assert abs(res.tav - st.tav) < 1e-4
        assert np.abs(out["xyz"][sl] - ref["xyz"]).max() < 1e-8
        assert np.abs(out["velo"][sl] - ref["velo"]).max() < 1e-10
        # averages handed to the fragment analysis (md.f90:691-698)
        assert np.abs(axyz[sl] - ref["axyz"]).max() < 1e-8
        assert np.abs(achrg[sl] - ref["achrg"]).max() < 1e-6
        assert abs(res.atlast - st.atlast) < 1e-4


def test_md_graph_replay_equals_eager():
    trajs = ensemble(2, seed0=11)
    a = q().MdEnsemble(trajs, max_steps=100, use_graph=False)
    b = q().MdEnsemble(trajs, max_steps=100, use_graph=True)
    for _ in range(3):
        a.run(4)
        b.run(4)
    ra, rb = a.read(), b.read()
    assert np.array_equal(ra["xyz"], rb["xyz"])
    assert np.array_equal(ra["velo"], rb["velo"])
    assert [r.nstep for r in ra["results"]] == [r.nstep for r in rb["results"]] == [12, 12]


def test_md_fragment_exit_matches_oracle():
    """Two methyl fragments 8 A apart: nfrag = 2 > nfragexit = 1 ends the run
    after the first step with fragstate 1 and mdok."""
    n1, x1 = W.build_chain([6], noise=0.0)
    num = np.concatenate([n1[:4], n1[:4]])
    xyz = np.concatenate([x1[:4], x1[:4] + np.array([8.0 * W.AATOAU, 0, 0])])
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    tr = dict(num=num, xyz=xyz, velo=W.maxwell_boltzmann(mass, 300.0, 5), mass=mass,
              eimp=0.05, tadd=0.0)
    md = q().MdEnsemble([tr], max_steps=50, nfragexit=1, use_graph=False)
    md.run(5)
    res = md.read()["results"][0]
    ref = ob.md_run(tr, steps_limit=0, max_steps=50, nfragexit=1)["state"]
    assert (res.nstep, res.nfrag, res.exit_reason, res.mdok, res.fragstate) == \
        (ref.nstep, ref.nfrag, ref.exit_reason, ref.mdok, ref.fragstate) == (1, 2, 4, 1, 1)


def test_bench_workload_at_default_thresholds_meets_parity():
    """The batched MD exactly as bench.py runs it by default (persistent MD
    kernel, default SCC thresholds and eigensolver, warm-started and
    extrapolated SCC) and as the per-step kernels replayed as hipGraphs run
    it, both within the parity bar of the tightly converged oracle MD:
    per-step energies to 1e-8 Eh, forces to 1e-6 Eh/bohr, charges to 1e-6 e."""
    import bench
    trajs = bench.workload(4)
    nsteps = 12
    outs = []
    for kw in (dict(fused=True), dict(use_graph=True)):
        md = q().MdEnsemble(trajs, max_steps=nsteps, **kw)
        md.run(nsteps)
        outs.append(md.read())
        md.close()
    opts = ob.default_opts(econv=1e-12, pconv=1e-10)
    for t, tr in enumerate(trajs):
        ref = ob.md_run(tr, steps_limit=0, max_steps=nsteps, opts=opts)
        for mode, out in zip(("fused", "graph"), outs):
            st, res = ref["state"], out["results"][t]
            sl = slice(md.offsets[t], md.offsets[t + 1])
            de = abs(res.epot - st.epot)
            dg = np.abs(out["grad"][sl] - ref["grad"]).max()
            dq = np.abs(out["qat"][sl] - ref["qat"]).max()
            dx = np.abs(out["xyz"][sl] - ref["xyz"]).max()
            print(f"bench traj {t} {mode}: dE={de:.2e} dg={dg:.2e} dq={dq:.2e} dx={dx:.2e}", flush=True)
            assert res.nstep == st.nstep == nsteps
            assert de <= 1e-8 and dg <= 1e-6 and dq <= 1e-6


def test_md_lanes_equal_single_ensemble():
    """Trajectories split into lanes (ensembles on separate streams advanced
    together by qcxms_md_run_group) reproduce the single ensemble bit for
    bit: every trajectory's arithmetic is independent of its neighbours."""
    trajs = ensemble(5, seed0=17)
    a = q().MdEnsemble(trajs, max_steps=30, use_graph=True)
    b = q().MdEnsemble(trajs, max_steps=30, use_graph=True, lanes=3)
    ra_n = a.run(6)
    rb_n = b.run(6)
    ra, rb = a.read(), b.read()
    xa, qa_ = a.read_averages()
    xb, qb_ = b.read_averages()
    a.close()
    b.close()
    assert ra_n == rb_n
    assert np.array_equal(ra["xyz"], rb["xyz"]) and np.array_equal(ra["velo"], rb["velo"])
    assert np.array_equal(ra["qat"], rb["qat"]) and np.array_equal(xa, xb)
    assert [r.nstep for r in ra["results"]] == [r.nstep for r in rb["results"]]
    assert [r.epot for r in ra["results"]] == [r.epot for r in rb["results"]]


def test_fused_md_matches_oracle_md():
    """The persistent MD kernel (whole iterations per launch) against the
    oracle MD over 30 steps of heating trajectories, as
    test_md_ensemble_matches_oracle_md does for the per-kernel path."""
    trajs = ensemble(4, seed0=41)
    nsteps = 30
    md = q().MdEnsemble(trajs, max_steps=nsteps, fused=True, econv=1e-11, pconv=1e-9)
    running = md.run(nsteps + 5)
    out = md.read()
    axyz, achrg = md.read_averages()
    md.close()
    assert running == 0
    opts = ob.default_opts(econv=1e-12, pconv=1e-10)
    for t, tr in enumerate(trajs):
        ref = ob.md_run(tr, steps_limit=0, max_steps=nsteps, opts=opts)
        st, res = ref["state"], out["results"][t]
        sl = slice(md.offsets[t], md.offsets[t + 1])
        print(f"fused traj {t}: nstep {res.nstep}/{st.nstep} dEpot {abs(res.epot - st.epot):.2e} "
              f"dx {np.abs(out['xyz'][sl] - ref['xyz']).max():.2e}", flush=True)
        assert res.nstep == st.nstep and res.exit_reason == st.exit_reason
        assert np.array_equal(out["list"][sl], ref["list"])
        assert abs(res.epot - st.epot) < 1e-8
        assert np.abs(out["xyz"][sl] - ref["xyz"]).max() < 1e-8
        assert np.abs(out["velo"][sl] - ref["velo"]).max() < 1e-10
        assert np.abs(axyz[sl] - ref["axyz"]).max() < 1e-8


def test_fused_md_chunks_equal_one_launch():
    """Steps split over several persistent launches give the state of one
    launch over all of them, and match the per-kernel path to the parity bar."""
    trajs = ensemble(3, seed0=43)
    a = q().MdEnsemble(trajs, max_steps=40, fused=True)
    b = q().MdEnsemble(trajs, max_steps=40, fused=True)
    c = q().MdEnsemble(trajs, max_steps=40, use_graph=True)
    a.run(12)
    for _ in range(3):
        b.run(4)
    c.run(12)
    ra, rb, rc = a.read(), b.read(), c.read()
    for m in (a, b, c):
        m.close()
    assert np.array_equal(ra["xyz"], rb["xyz"]) and np.array_equal(ra["velo"], rb["velo"])
    assert [r.nstep for r in ra["results"]] == [r.nstep for r in rc["results"]] == [12, 12, 12]
    assert np.abs(ra["xyz"] - rc["xyz"]).max() < 1e-8
    assert max(abs(x.epot - y.epot) for x, y in zip(ra["results"], rc["results"])) < 1e-8


def test_jacobi_eigensolver_md_matches_oracle_md():
    """MD with the SCC eigenvectors from parallel cyclic Jacobi alone
    (eigensolver 2; the default refines them by simultaneous rotations on the
    FP64 matrix cores first, which every other MD test exercises) against the
    oracle MD over 30 heating steps; an unknown eigensolver is rejected."""
    trajs = ensemble(4, seed0=57)
    nsteps = 30
    with pytest.raises(RuntimeError):
        q().MdEnsemble(trajs, max_steps=nsteps, eigensolver=3)
    md = q().MdEnsemble(trajs, max_steps=nsteps, eigensolver=2, use_graph=False, econv=1e-11, pconv=1e-9)
    running = md.run(nsteps + 5)
    out = md.read()
    md.close()
    assert running == 0
    opts = ob.default_opts(econv=1e-12, pconv=1e-10)
    for t, tr in enumerate(trajs):
        ref = ob.md_run(tr, steps_limit=0, max_steps=nsteps, opts=opts)
        st, res = ref["state"], out["results"][t]
        sl = slice(md.offsets[t], md.offsets[t + 1])
        print(f"jacobi traj {t}: nstep {res.nstep}/{st.nstep} dEpot {abs(res.epot - st.epot):.2e} "
              f"dx {np.abs(out['xyz'][sl] - ref['xyz']).max():.2e}", flush=True)
        assert res.nstep == st.nstep and res.exit_reason == st.exit_reason
        assert np.array_equal(out["list"][sl], ref["list"])
        assert abs(res.epot - st.epot) < 1e-8
        assert np.abs(out["xyz"][sl] - ref["xyz"]).max() < 1e-8
        assert np.abs(out["velo"][sl] - ref["velo"]).max() < 1e-10


def test_md_mixed_sizes_lds_fallback_matches_oracle():
    """One ensemble of C8H18+ (50 basis functions: LDS matrices with the MFMA
    eigen refinement) and two C/H/N/O chain cations of 53 and 56 basis
    functions (LDS matrices without room for the refinement's fifth matrix:
    Jacobi), default MD settings, 10 heating steps against the oracle; the
    per-step kernels and the persistent kernel give the same bits."""
    trajs = ensemble(1, seed0=73)
    for seed in (5155, 5180):
        num, xyz, charge = W.random_chain(seed)
        mass = np.array([W.atomic_mass(int(z)) for z in num])
        trajs.append(dict(num=num, xyz=xyz, velo=W.maxwell_boltzmann(mass, 500.0, seed + 7919), mass=mass,
                          eimp=W.iee_poisson(len(num), 0.6, 1.5, seed + 104729), tadd=5.0 * W.FSTOAU,
                          charge=charge))
    nsteps = 10
    outs = []
    for fused in (False, True):
        md = q().MdEnsemble(trajs, max_steps=nsteps, use_graph=False, fused=fused, econv=1e-11, pconv=1e-9)
        md.run(nsteps + 2)
        outs.append(md.read())
        md.close()
    # the persistent kernel runs the same code paths per trajectory
    assert np.array_equal(outs[0]["xyz"], outs[1]["xyz"])
    out = outs[0]
    opts = ob.default_opts(econv=1e-12, pconv=1e-10)
    for t, tr in enumerate(trajs):
        ref = ob.md_run(tr, steps_limit=0, max_steps=nsteps, opts=opts, charge=tr.get("charge", 1))
        st, res = ref["state"], out["results"][t]
        sl = slice(md.offsets[t], md.offsets[t + 1])
        print(f"mixed traj {t} ({len(tr['num'])} atoms): nstep {res.nstep}/{st.nstep} "
              f"dEpot {abs(res.epot - st.epot):.2e} dx {np.abs(out['xyz'][sl] - ref['xyz']).max():.2e}", flush=True)
        assert res.nstep == st.nstep and res.exit_reason == st.exit_reason
        assert abs(res.epot - st.epot) < 1e-8
        assert np.abs(out["xyz"][sl] - ref["xyz"]).max() < 1e-8
