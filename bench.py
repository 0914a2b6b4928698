"""Throughput of the batched GFN2-xTB EI production MD on one MI355X.

Workload (the batched configuration the metric is quoted on): 256 independent
trajectories of the n-octane radical cation C8H18+ (26 atoms, 50 basis
functions), each with its own seeded geometry noise (0.02 A), Maxwell-Boltzmann
velocities at 500 K and an internal excess energy drawn from a Poisson-shaped
distribution (mean 0.6 eV/atom, truncated at 1.5 eV/atom) added by impactscale
over tadd = 100 fs; FP64 GFN2-xTB forces at etemp 5000 K (setetemp), 0.5 fs
leap-frog steps, trajectory length 5000 steps.  A step is one pass of the MD
loop (src/md.f90:285-682) over every running trajectory.

value = trajectory-steps executed / wall time of the timed steps, with all
state resident in HBM.  By default the steps run in the persistent MD kernel
(k_md_fused: whole MD iterations per trajectory, many steps per launch) with
the SCC eigenvectors refined by FP64 MFMA rotations; --fused 0 runs one launch
per kernel and step.  The roofline object describes the dominant kernel
(k_md_fused, or the SCC kernel k_scc with --fused 0), timed by HIP events on
the library's stream.  Its algorithmic FP64 flops: per trajectory-step
F_setup + (SCC iterations) F_iter + F_int + F_grad + F_pair, with
F_setup = 4/3 n^3, F_iter = 15 n^3 + 30 n^2 (n = basis functions) and the
integral, integral-derivative and pair-term costs per primitive pair, AO pair,
atom pair and atom triple counted in the oracle's FP64 operations
(profiles/flop_model.json); k_scc alone is priced with F_iter only.  cpu_baseline: the C oracle's restatement of the same MD loop with the
same forces (not QCxMS+tblite, which cannot be built here), one trajectory per
host thread as bin/pqcxms runs it, on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "trajectory-MD-steps/sec on 1 GPU (batched vs 1-per-launch) at |dE|≤1e-8 Eh"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def scc_flops(n):
    """Analytic FP64 flops of one SCC iteration of an n-function system."""
    return 15.0 * n ** 3 + 30.0 * n ** 2


FLOP_MODEL = os.path.join(ROOT, "profiles", "flop_model.json")


def fixed_step_flops(num):
    """FP64 flops of one energy + gradient of the structure num outside the
    SCC iterations: {setup, int, grad, pair} from the frozen per-unit costs
    (shell pairs ia >= ib with their primitive and AO pairs for the integrals,
    pairs on different atoms for their derivatives, atom pairs and triples for
    the pair terms)."""
    m = json.load(open(FLOP_MODEL))
    nprim = {int(k): v for k, v in m["nprim"].items()}
    shells = [(a, l, g) for a, z in enumerate(num) for l, g in enumerate(nprim[int(z)])]
    c = dict(pp_all=0.0, pp_ao_all=0.0, sp_all=0.0, pp_off=0.0, pp_ao_off=0.0, sp_off=0.0)
    for i, (ai, li, gi) in enumerate(shells):
        for aj, lj, gj in shells[:i + 1]:
            npp, nao = gi * gj, (2 * li + 1) * (2 * lj + 1)
            c["pp_all"] += npp
            c["pp_ao_all"] += npp * nao
            c["sp_all"] += 1
            if ai != aj:
                c["pp_off"] += npp
                c["pp_ao_off"] += npp * nao
                c["sp_off"] += 1
    n = sum(2 * l + 1 for _, l, _ in shells)
    nat, nsh = len(num), len(shells)
    fi, fg, fp = m["F_int"], m["F_grad"], m["F_pair"]
    return {"setup": 4.0 / 3.0 * n ** 3,
            "int": fi["per_prim_pair"] * c["pp_all"] + fi["per_prim_pair_ao_pair"] * c["pp_ao_all"]
            + fi["per_shell_pair"] * c["sp_all"],
            "grad": 2.0 * n ** 3 + fg["per_prim_pair"] * c["pp_off"]
            + fg["per_prim_pair_ao_pair"] * c["pp_ao_off"] + fg["per_shell_pair"] * c["sp_off"],
            "pair": fp["per_atom"] * nat + fp["per_atom_pair"] * nat * (nat - 1) / 2.0
            + fp["per_atom_triple"] * nat * (nat - 1) * (nat - 2) / 6.0
            + fp["per_shell_pair_sq"] * nsh * nsh,
            "n": n}


def workload(ntraj):
    from qcxms_amd import workloads as W
    return W.md_ensemble(ntraj, base_seed=1, mean_ev=0.6, max_ev=1.5, temp=500.0, noise=0.02,
                         tadd_fs=100.0)


TRAFFIC_FILES = {False: os.path.join(ROOT, "profiles", "k_scc_hbm_traffic_c8h18_256.txt"),
                 True: os.path.join(ROOT, "profiles", "k_md_fused_hbm_traffic_c8h18_256.txt")}


def traffic_from_profiles(fused=False, steps_per_launch=1.0):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain-copy
    calibration), or None.  The persistent kernel's summary gives bytes per
    MD step, scaled here by the steps one launch runs."""
    path = TRAFFIC_FILES[fused]
    if not os.path.exists(path):
        return None
    for line in open(path):
        for key, scale in (("bytes_per_launch", 1.0), ("bytes_per_step", steps_per_launch)):
            if line.startswith(key):
                try:
                    return float(line.split("=")[1].split()[0]) * scale
                except (IndexError, ValueError):
                    return None
    return None


def run_gpu(ntraj, steps, warmup, use_graph_check=True, lanes=1, fused=False, eigensolver=0):
    import qcxms_amd
    trajs = workload(ntraj)
    md = qcxms_amd.MdEnsemble(trajs, max_steps=5000, use_graph=False, lanes=lanes, fused=fused,
                              eigensolver=eigensolver)
    md.run(warmup)
    md.set_profiling(True)
    names = ("fused",) if fused else ("setup", "scc", "grad", "md_pre", "md_post")
    for k in names:
        md.kernel_time(k)
    it0 = md.scc_iterations_total().astype(np.float64).sum()
    n0 = sum(r.nstep for r in md.read()["results"])
    t0 = time.perf_counter()
    md.run(steps)
    t1 = time.perf_counter()
    res = md.read()["results"]
    n1 = sum(r.nstep for r in res)
    it1 = md.scc_iterations_total().astype(np.float64).sum()
    kt = {k: md.kernel_time(k) for k in names}
    md.set_profiling(False)
    running = sum(r.running for r in res)
    md.close()
    wall = t1 - t0
    traj_steps = n1 - n0
    # the dominant kernel: the SCC, or the persistent kernel that contains it
    # (and with it setup, gradient and MD bookkeeping of every step)
    scc_ms, scc_launches = kt["fused" if fused else "scc"]
    fx = fixed_step_flops(trajs[0]["num"])
    scc_part = (it1 - it0) * scc_flops(fx["n"])
    fixed_part = traj_steps * (fx["setup"] + fx["int"] + fx["grad"] + fx["pair"])
    flops = scc_part + (fixed_part if fused else 0.0)
    achieved = flops / (scc_ms * 1e-3) / 1e12 if scc_ms > 0 else 0.0
    per_step = {k: fx[k] / 1e6 for k in ("setup", "int", "grad", "pair")}
    per_step["scc"] = scc_part / max(traj_steps, 1) / 1e6
    out = dict(wall=wall, traj_steps=traj_steps, value=traj_steps / wall,
               ms_per_step=1e3 * wall / steps, running=running,
               scc_iters_per_traj_step=(it1 - it0) / max(traj_steps, 1),
               dominant_launches=scc_launches,
               kernels={k: dict(ms_total=v[0], launches=v[1],
                                ms_avg=(v[0] / v[1] if v[1] else 0.0)) for k, v in kt.items()},
               achieved_tflops=achieved, scc_avg_ms=scc_ms / max(scc_launches, 1),
               scc_flops_per_launch=flops / max(scc_launches, 1),
               achieved_scc_only=scc_part / (scc_ms * 1e-3) / 1e12 if scc_ms > 0 else 0.0,
               mflop_per_traj_step=per_step)
    # the same steps through the per-step kernels, each step replayed as one
    # captured hipGraph (no per-kernel events)
    if use_graph_check:
        md = qcxms_amd.MdEnsemble(trajs, max_steps=5000, use_graph=True, lanes=lanes, fused=False,
                                  eigensolver=eigensolver)
        md.run(warmup)
        n0 = sum(r.nstep for r in md.read()["results"])
        t0 = time.perf_counter()
        md.run(steps)
        t1 = time.perf_counter()
        n1 = sum(r.nstep for r in md.read()["results"])
        md.close()
        out["per_step_hipgraph_value"] = (n1 - n0) / (t1 - t0)
    return out


def run_singlepoints(name, reps):
    """Batched single points of the uneven (1024 random C/H/N/O chains, 3-62
    atoms) or stretched (64 x 40-atom C12H24O2N2+) configuration: cold-started
    SCC on every call, systems ordered largest basis first."""
    import qcxms_amd
    from qcxms_amd import workloads as W
    if name == "uneven1024":
        sysl = []
        for s in range(1024):
            num, xyz, charge = W.random_chain(5000 + s)
            sysl.append((num, charge, xyz))
        desc = "1024 single points: random saturated C/H/N/O chains, 1-20 heavy atoms (3-62 atoms)"
    else:
        sysl = []
        for s in range(64):
            num, xyz = W.chain_c12h24o2n2(seed=s)
            sysl.append((num, 1, xyz))
        desc = "64 single points: C12H24O2N2+ (40 atoms, 88 basis functions), bonds stretched <= +30 %"
    nbf = lambda num: int(sum(1 if z == 1 else 4 for z in num))
    sysl.sort(key=lambda t: -nbf(t[0]))
    xyz = np.concatenate([t[2] for t in sysl])
    kernels = ("setup", "scc", "dblocks", "grad")

    def timed_batch(stored):
        b = qcxms_amd.Batch([(t[0], t[1]) for t in sysl])
        b.set_warm_start(0)
        b.set_stored_derivatives(stored)
        b.egrad(xyz, 5000.0)
        b.set_profiling(True)
        for k in kernels:
            b.kernel_time(k)
        t0 = time.perf_counter()
        for _ in range(reps):
            e, g, q, st = b.egrad(xyz, 5000.0)
        wall = time.perf_counter() - t0
        kt = {k: b.kernel_time(k) for k in kernels}
        iters = np.asarray(b.info()[0], dtype=np.float64)
        mem = b.device_bytes()
        b.close()
        assert np.all(st == 0)
        return wall, kt, iters, mem, g

    wall, kt, iters, mem, g_rec = timed_batch(False)
    kt = {k: v for k, v in kt.items() if v[1]}
    # the same batch with the launch plan before tuning (one launch, Jacobi,
    # LDS matrices up to 56 basis functions)
    table, mixed = qcxms_amd.get_tuning()
    qcxms_amd.set_tuning([(256, 0, 1)], (0, 1))
    try:
        w_un, kt_un, _, _, _ = timed_batch(False)
    finally:
        qcxms_amd.set_tuning(table, mixed)
    autotuned = {"table": table, "mixed": mixed,
                 "untuned_value": len(sysl) * reps / w_un,
                 "untuned_scc_ms_avg": kt_un["scc"][0] / max(kt_un["scc"][1], 1),
                 "source": "qcxms_amd/autotune.py, profiles/tuning_mi355x.json"}
    # algorithmic flops of one launch of the three kernels together (cold SCC:
    # the same iterations on every repetition)
    flops = 0.0
    for (num, _, _), it in zip(sysl, iters):
        fx = fixed_step_flops(num)
        flops += fx["setup"] + fx["int"] + fx["grad"] + fx["pair"] + it * scc_flops(fx["n"])
    kernel_ms = sum(v[0] for v in kt.values()) / max(reps, 1)
    achieved = flops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    # the same single points with the derivative integral blocks stored by
    # k_dblocks and contracted by the gradient, against recomputing them
    w2, kt2, _, mem2, g_sto = timed_batch(True)
    form = lambda w, k, m: {"value": len(sysl) * reps / w, "unit": "single-points/s",
                            "device_mib": m / 2 ** 20,
                            "kernels_ms_avg": {n: (v[0] / v[1]) for n, v in k.items() if v[1]}}
    derivative_blocks = {"recomputed": form(wall, kt, mem), "stored": form(w2, kt2, mem2),
                         "max_abs_grad_difference": float(np.abs(g_sto - g_rec).max())}
    return {"metric": "GFN2-xTB energy+gradient single points/sec on 1 GPU (batched)",
            "value": len(sysl) * reps / wall, "unit": "single-points/s", "n_gpus": 1,
            "reps": reps, "dtype": "f64", "data": "synthetic",
            "config": {"workload": desc, "nsys": len(sysl),
                       "scc_iterations_mean": float(iters.mean())},
            "kernels_ms_avg": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kt.items()},
            "roofline": {"bound": "mfma", "kernel": "k_setup + k_scc + k_grad",
                         "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "flops_per_launch": flops,
                         "source": "profiles/flop_model.json"},
            "derivative_blocks": derivative_blocks,
            "autotuned_launch_plan": autotuned}


def run_c20_md(ntraj, steps, warmup, repack=100):
    """Batched fragmentation MD of C20H36O2N2+ (60 atoms, 132 basis
    functions, matrices in global memory), IEE Poisson mean 0.8 eV/atom,
    through the fragmentation-generation driver: a trajectory that ends
    continues with the fragment manage_fragments picks (main.F90:2208-2345),
    and after every `repack` steps in which trajectories ended the running
    ones are laid out again on the device."""
    import qcxms_amd
    from qcxms_amd import workloads as W
    trajs = W.md_ensemble(ntraj, base_seed=1, mean_ev=0.8, max_ev=2.0, temp=500.0,
                          builder=W.chain_c20h36o2n2, tadd_fs=100.0)
    ei = qcxms_amd.EiEnsemble(trajs, max_steps=10000, use_graph=True, repack=repack)
    ei.run(warmup)
    n0, r0, g0 = ei.steps(), ei.repacks(), len(ei.generations())
    t0 = time.perf_counter()
    running = ei.run(steps)
    wall = time.perf_counter() - t0
    n1, r1 = ei.steps(), ei.repacks()
    gens = ei.generations()
    ei.close()
    return {"metric": "trajectory-MD-steps/sec on 1 GPU (batched)", "value": (n1 - n0) / wall,
            "unit": "trajectory-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * wall / max(steps, 1), "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"batched fragmentation MD: {ntraj} x C20H36O2N2+ (60 atoms, "
                                   "132 basis functions), IEE Poisson 0.8 eV/atom (<= 2.0), "
                                   "tadd 100 fs, fragment continuation (main.F90:2208-2345)",
                       "ntraj": ntraj, "repack_every": repack, "relayouts": r1 - r0,
                       "generations_ended": len(gens) - g0, "running_after": running,
                       "trajectory_steps": n1 - n0,
                       "exit_reasons": np.bincount([g["exit_reason"] for g in gens],
                                                   minlength=9).tolist() if gens else []}}


def run_cid(ntraj, steps, warmup):
    """Batched CID collision MD (cid.f90:724-1111): ntraj x C8H18+ (26 atoms)
    hit by N2 at E_LAB = 40 eV, set up as first collisions with seeded random
    offsets; the ion and the gas (28 atoms, 58 basis functions) in every
    GFN2 evaluation."""
    import qcxms_amd
    from qcxms_amd import workloads as W
    trajs = W.md_ensemble(ntraj, base_seed=1, temp=500.0, tadd_fs=100.0)
    rng = np.random.default_rng(4242)
    systems = [qcxms_amd.cid_setup(t["num"], t["xyz"], t["velo"], t["mass"], 0.5 * W.FSTOAU, gas="N2",
                                   elab_ev=40.0, rand=tuple(rng.random(4))) for t in trajs]
    e = qcxms_amd.CidEnsemble(systems, charge=1)
    e.run(warmup)
    n0 = sum(r.nstep for r in e.read()["results"])
    t0 = time.perf_counter()
    running = e.run(steps)
    wall = time.perf_counter() - t0
    res = e.read()["results"]
    e.close()
    n1 = sum(r.nstep for r in res)
    return {"metric": "CID trajectory-MD-steps/sec on 1 GPU (batched)", "value": (n1 - n0) / wall,
            "unit": "trajectory-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * wall / max(steps, 1), "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"batched CID collision MD: {ntraj} x C8H18+ + N2 (28 atoms, 58 basis "
                                   "functions), E_LAB 40 eV, GFN2-xTB FP64, etemp 5000 K, 0.5 fs",
                       "ntraj": ntraj, "trajectory_steps": n1 - n0, "running_after": running,
                       "collided": sum(r.collided for r in res),
                       "fragmented": sum(r.fragmented for r in res)}}


# algorithmic HBM bytes of the MD bookkeeping scans per atom and per
# trajectory record (reads + writes of the arrays each kernel touches once)
SCAN_BYTES = {
    # qat, avchrg, xyz, avxyz, grad, velo, mass read; avchrg, avxyz, velo, xyz written
    "md_pre": (8 + 8 + 24 + 24 + 24 + 24 + 8) + (8 + 24 + 24 + 24),
    # grad, qat (checkqc), velo, velof, mass (heating), num, xyz read; velo, list written
    "md_post": (24 + 8 + 24 + 8 + 8 + 4 + 24) + (24 + 4),
}
# the EI fields of the trajectory record (328 of MdTraj's 440 bytes; the
# CID-mode fields are read as one flag), read and written
TRAJ_RECORD_BYTES = 2 * 328


def scan_kernels(ntraj, steps=20, warmup=5):
    """HBM rate of the integrator/bookkeeping scans (k_md_pre: leap-frog and
    averages; k_md_post: checkqc, heating, connectivity), timed with HIP
    events in a short run of the per-step kernels of the same workload."""
    import qcxms_amd
    trajs = workload(ntraj)
    natoms = sum(len(t["num"]) for t in trajs)
    md = qcxms_amd.MdEnsemble(trajs, max_steps=5000, use_graph=False, fused=False)
    md.run(warmup)
    md.set_profiling(True)
    for k in SCAN_BYTES:
        md.kernel_time(k)
    md.run(steps)
    out = {}
    for k, per_atom in SCAN_BYTES.items():
        ms, n = md.kernel_time(k)
        us = 1e3 * ms / max(n, 1)
        b = natoms * per_atom + ntraj * TRAJ_RECORD_BYTES
        out["k_" + k] = {"avg_us": us, "bytes_per_launch": b, "GB/s": b / (us * 1e-6) / 1e9 if us > 0 else None,
                         "frac_of_hbm_peak": (b / (us * 1e-6) / 1e9) / HBM_PEAK_GBS if us > 0 else None}
    md.set_profiling(False)
    md.close()
    return out


def run_one_per_launch(steps, warmup, fused=False):
    import qcxms_amd
    trajs = workload(1)
    md = qcxms_amd.MdEnsemble(trajs, max_steps=5000, use_graph=True, fused=fused)
    md.run(warmup)
    n0 = md.read()["results"][0].nstep
    t0 = time.perf_counter()
    md.run(steps)
    t1 = time.perf_counter()
    n1 = md.read()["results"][0].nstep
    md.close()
    return (n1 - n0) / (t1 - t0)


def host_cores():
    """(usable, nproc, quota): CPUs this process may run on (sched affinity,
    what nproc prints), os.cpu_count(), and the cgroup CPU quota in CPUs
    (None when unlimited)."""
    usable = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    return usable, os.cpu_count() or usable, quota


def _cpu_md(trajs, steps, threads):
    """Oracle C MD + GFN2 of trajs on `threads` host threads; returns
    (trajectory-steps done, wall seconds)."""
    import ctypes as C

    import oracle_bind as ob
    ntraj = len(trajs)
    L = ob.oracle()
    L.orc_md_run_batch.argtypes = [C.c_int, ob.ip, ob.ip, ob.ip, ob.dp, ob.dp, ob.dp, ob.dp,
                                   ob.dp, ob.dp, ob.ip, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    off = np.cumsum([0] + [len(t["num"]) for t in trajs]).astype(np.int32)
    num = np.concatenate([t["num"] for t in trajs]).astype(np.int32)
    imass = np.full(len(num), -1, dtype=np.int32)
    mass = np.concatenate([t["mass"] for t in trajs])
    velof = np.ones(len(num))
    xyz = np.ascontiguousarray(np.concatenate([t["xyz"] for t in trajs]))
    velo = np.ascontiguousarray(np.concatenate([t["velo"] for t in trajs]))
    grad = np.zeros_like(xyz)
    qat = np.zeros(len(num))
    lst = np.zeros(len(num), dtype=np.int32)
    states = (ob.MdState * ntraj)()
    for s, t in enumerate(trajs):
        st = states[s]
        st.nat, st.chrg, st.nfragexit, st.max_steps, st.isec = len(t["num"]), 1, 3, 5000, 1
        st.tstep, st.tadd, st.eimp = 0.5 * 41.3413733365614, t["tadd"], t["eimp"]
        st.ax, st.ieetemp, st.etempin = 0.0, 0.0, -1.0
    opts = ob.default_opts()
    t0 = time.perf_counter()
    L.orc_md_run_batch(ntraj, ob.ptr(off, ob.ip), ob.ptr(num, ob.ip), ob.ptr(imass, ob.ip),
                       ob.ptr(mass), ob.ptr(velof), ob.ptr(xyz), ob.ptr(velo), ob.ptr(grad),
                       ob.ptr(qat), ob.ptr(lst, ob.ip), C.cast(states, C.c_void_p),
                       C.cast(C.pointer(opts), C.c_void_p), steps, threads)
    wall = time.perf_counter() - t0
    return sum(states[s].nstep for s in range(ntraj)), wall


def eigensolver_vs_lapack(n=50, reps=200):
    """Per-core speed of the oracle's generalized eigensolver (Cholesky +
    Householder tridiagonalisation + implicit QL, as tred2/tql2) against
    LAPACK's (scipy.linalg.eigh(F, S), dsygvd-class) on one thread, for an
    n x n problem shaped like a GFN2 Fock/overlap pair."""
    import ctypes as C

    import oracle_bind as ob
    from scipy import linalg
    rng = np.random.default_rng(0)
    a = rng.normal(size=(n, n))
    s = a @ a.T / n + np.eye(n)
    f = rng.normal(size=(n, n))
    f = 0.5 * (f + f.T)
    L = ob.oracle()
    L.orc_sygv.argtypes = [C.c_int, ob.dp, ob.dp, ob.dp, ob.dp]
    ev, work = np.zeros(n), np.zeros(4 * n * n + 8 * n)
    t0 = time.perf_counter()
    for _ in range(reps):
        fc, sc = f.copy(), s.copy()
        L.orc_sygv(n, ob.ptr(fc), ob.ptr(sc), ob.ptr(ev), ob.ptr(work))
    t_orc = (time.perf_counter() - t0) / reps
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for _ in range(reps):
            linalg.eigh(f, s, driver="gvd")
        t_lap = (time.perf_counter() - t0) / reps
    return dict(n=n, oracle_ms=1e3 * t_orc, lapack_ms=1e3 * t_lap, ratio=t_orc / t_lap)


def run_cpu_baseline(steps=None, target_s=20.0):
    """One trajectory per usable host CPU, each on its own thread (bin/pqcxms
    runs one single-threaded qcxms per core), for a bounded number of steps so
    that the CPU work lasts about target_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    usable, nproc, quota = host_cores()
    threads = usable
    eff = min(usable, quota) if quota else usable
    if steps is None:
        # about 42 ms per cold-started trajectory-step and thread
        steps = max(10, int(target_s / 0.042 * eff / threads))
    done, wall = _cpu_md(workload(threads), steps, threads)
    eig = eigensolver_vs_lapack()
    return dict(value=done / wall, unit="trajectory-steps/s", cores=threads, kind="port",
                nproc=nproc, cgroup_cpu_quota=quota,
                sample=f"{threads} trajectories x {steps} MD steps of the same C8H18+ workload "
                       f"(oracle C restatement, SCC cold-started every step as the reference, "
                       f"one trajectory per thread on every usable host CPU; not "
                       f"QCxMS+tblite), wall {wall:.1f} s",
                eigensolver_vs_lapack=eig)


def run_cpu_single(steps=200):
    """Config-0 latency on the CPU: one C8H18+ trajectory on one thread, a
    bounded number of steps of the 2000-step run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    done, wall = _cpu_md(workload(1), steps, 1)
    return dict(ms_per_step=1e3 * wall / max(done, 1), steps=done,
                est_2000_steps_s=2000 * wall / max(done, 1), cores=1, kind="port")


def copy_calibration(mib=256, reps=5):
    """A plain device copy of known size (for scaling the HBM traffic counters
    FETCH_SIZE / WRITE_SIZE): reps copies of mib MiB."""
    import torch
    n = mib * 1024 * 1024 // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": reps}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ntraj", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="MD steps per trajectory of the CPU baseline (default: about 20 s "
                         "of CPU work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--steady-steps", type=int, default=400,
                    help="also time this many steps after --steady-warmup steps (0: skip)")
    ap.add_argument("--steady-warmup", type=int, default=220)
    ap.add_argument("--no-one-per-launch", action="store_true")
    ap.add_argument("--no-graph-check", action="store_true",
                    help="skip the hipGraph replay of the same steps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration)")
    ap.add_argument("--fused", type=int, default=1, choices=[0, 1],
                    help="1 (default): persistent kernel running whole MD iterations per "
                         "trajectory; 0: one launch per kernel and step")
    ap.add_argument("--eigensolver", type=int, default=0, choices=[0, 1, 2],
                    help="SCC eigenvectors: 1 MFMA simultaneous rotations then Jacobi, 2 Jacobi "
                         "only, 0 library default (1)")
    ap.add_argument("--lanes", type=int, default=1,
                    help="split the ensemble into this many groups on separate streams")
    ap.add_argument("--config", default="md256",
                    choices=["md256", "uneven1024", "stretched64", "c20md", "cid256"],
                    help="md256: the headline batched MD (default); the others print the "
                         "throughput of the other workload configurations")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.config != "md256":
        if args.gpus != 1:
            sys.exit("bench.py measures one GPU only")
        if args.config == "cid256":
            out = run_cid(args.ntraj, args.steps, args.warmup)
        elif args.config == "c20md":
            out = run_c20_md(args.ntraj if args.ntraj != 256 else 4096, args.steps, args.warmup)
        else:
            out = run_singlepoints(args.config, max(1, args.steps))
        print(json.dumps(out), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs every trajectory on a single "
              "MI355X and has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    g = run_gpu(args.ntraj, args.steps, args.warmup, use_graph_check=not args.no_graph_check,
                lanes=args.lanes, fused=bool(args.fused), eigensolver=args.eigensolver)
    line = {
        "metric": METRIC,
        "value": g["value"],
        "unit": "trajectory-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched EI MD: 256 x C8H18+ (26 atoms, 50 basis functions), "
                               "GFN2-xTB FP64, etemp 5000 K, 0.5 fs leap-frog, IEE Poisson "
                               "0.6 eV/atom (<= 1.5), tadd 100 fs, 500 K Maxwell-Boltzmann",
                   "ntraj": args.ntraj, "lanes": args.lanes, "fused": bool(args.fused),
                   "eigensolver": "jacobi" if args.eigensolver == 2 else "mfma+jacobi",
                   "trajectory_steps": g["traj_steps"],
                   "running_after": g["running"],
                   "scc_iterations_per_trajectory_step": g["scc_iters_per_traj_step"]},
        "roofline": {"bound": "mfma", "kernel": "k_md_fused" if args.fused else "k_scc",
                     "achieved": g["achieved_tflops"], "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": g["achieved_tflops"] / FP64_PEAK_TFLOPS,
                     "traffic": traffic_from_profiles(
                         bool(args.fused), args.steps / max(g["dominant_launches"], 1)),
                     "flops_per_launch": g["scc_flops_per_launch"],
                     "avg_launch_ms": g["scc_avg_ms"],
                     "flop_model": {"mflop_per_trajectory_step": g["mflop_per_traj_step"],
                                    "counted_in_launch": (["setup", "scc", "int", "grad", "pair"]
                                                          if args.fused else ["scc"]),
                                    "source": "profiles/flop_model.json"},
                     "achieved_scc_only": g["achieved_scc_only"]},
        "kernels_ms_avg": {k: v["ms_avg"] for k, v in g["kernels"].items()},
        "per_step_hipgraph": {"value": g.get("per_step_hipgraph_value"),
                              "unit": "trajectory-steps/s",
                              "form": "per-step kernels, each MD step replayed as one hipGraph"},
    }
    # the integrator and connectivity scans are latency-bound: their HBM rate
    # beside the roofline of the dominant kernel
    line["bookkeeping_scans"] = scan_kernels(args.ntraj)
    if args.steady_steps > 0:
        # the same workload past heating (tadd = 100 fs = 200 steps), so that a
        # short timed window inside the heating phase is shown beside the rate
        # the trajectories keep for the rest of their run
        w = run_gpu(args.ntraj, args.steady_steps, args.steady_warmup, use_graph_check=False,
                    lanes=args.lanes, fused=bool(args.fused), eigensolver=args.eigensolver)
        line["steady_state"] = {"value": w["value"], "unit": "trajectory-steps/s",
                                "steps": args.steady_steps, "warmup": args.steady_warmup,
                                "ms_per_step": w["ms_per_step"],
                                "scc_iterations_per_trajectory_step": w["scc_iters_per_traj_step"],
                                "roofline_frac": w["achieved_tflops"] / FP64_PEAK_TFLOPS}
    if not args.no_one_per_launch:
        line["one_per_launch"] = {"value": run_one_per_launch(100, 10, bool(args.fused)),
                                  "unit": "trajectory-steps/s"}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = run_cpu_baseline(args.cpu_steps)
        line["cpu_single_trajectory"] = run_cpu_single()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
