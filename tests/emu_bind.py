"""Runs qcxms_egrad_batch through the CPU-emulated build of the HIP library
(tests/cpu_emu: kernels on host threads under AddressSanitizer).  Test
infrastructure for bounds checking the kernels without a GPU."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_DIR = os.path.join(ROOT, "tests", "cpu_emu")
EMU_BIN = os.path.join(EMU_DIR, "_build", "egrad_emu")


def build():
    subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)


def egrad_batch(systems, xyz_list, etemp, method=2, timeout=1800, stored=False, tuning=None):
    """systems: [(num, charge[, multiplicity])]; returns (rc, energy, [grad],
    [qat], stat, stderr).  stored: the gradient from stored derivative
    blocks (a persistent batch, SCC thresholds 1e-11 / 1e-9).  tuning:
    [(max_nao, eig, lds), ...], the SCC launch plan; by default one launch
    with the Jacobi eigensolver and LDS matrices (emulating the MFMA
    refinement takes minutes; the MD emulation covers it)."""
    nsys = len(systems)
    off = np.cumsum([0] + [len(s[0]) for s in systems]).astype(np.int32)
    num = np.concatenate([np.asarray(s[0], dtype=np.int32) for s in systems])
    xyz = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3) for x in xyz_list])
    chg = np.array([s[1] for s in systems], dtype=np.int32)
    mult = np.array([s[2] if len(s) > 2 else 1 for s in systems], dtype=np.int32)
    et = np.broadcast_to(np.asarray(etemp, dtype=np.float64), (nsys,)).copy()
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nsys, method], dtype=np.int32).tofile(f)
            off.tofile(f)
            num.tofile(f)
            xyz.tofile(f)
            chg.tofile(f)
            mult.tofile(f)
            et.tofile(f)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
                   QX_EMU_STORED="1" if stored else "0",
                   QX_EMU_TUNING=";".join(f"{a},{e},{l}" for a, e, l in (tuning or [(256, 0, 1)])))
        p = subprocess.run([EMU_BIN, fin, fout], capture_output=True, text=True, timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError(f"emulated run failed ({p.returncode}):\n{p.stderr[-6000:]}")
        raw = open(fout, "rb").read()
    nat = int(off[-1])
    rc = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    o = 4
    e = np.frombuffer(raw, np.float64, nsys, o).copy(); o += 8 * nsys
    g = np.frombuffer(raw, np.float64, 3 * nat, o).reshape(nat, 3).copy(); o += 24 * nat
    q = np.frombuffer(raw, np.float64, nat, o).copy(); o += 8 * nat
    st = np.frombuffer(raw, np.int32, nsys, o).copy()
    grads = [g[off[i]:off[i + 1]] for i in range(nsys)]
    qs = [q[off[i]:off[i + 1]] for i in range(nsys)]
    return rc, e, grads, qs, st, p.stderr


EMU_MD = os.path.join(EMU_DIR, "_build", "md_emu")


def md_run(trajs, nsteps, fused, use_graph=1, timeout=1800, env=None):
    """MD ensemble through the emulated library: (rc, running, xyz, velo, qat,
    [(nstep, exit_reason, epot, ekin, nfrag, check_fragmented, new_velo)]).
    env: extra environment of the driver (QX_EMU_* settings)."""
    nt = len(trajs)
    off = np.cumsum([0] + [len(t["num"]) for t in trajs]).astype(np.int32)
    cat = lambda k, dt: np.concatenate([np.asarray(t[k], dtype=dt).reshape(-1) for t in trajs])
    nat = int(off[-1])
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nt, nsteps, int(fused), int(use_graph)], dtype=np.int32).tofile(f)
            off.tofile(f)
            cat("num", np.int32).tofile(f)
            np.array([t.get("charge", 1) for t in trajs], dtype=np.int32).tofile(f)
            np.full(nat, -1, dtype=np.int32).tofile(f)
            cat("mass", np.float64).tofile(f)
            cat("xyz", np.float64).tofile(f)
            cat("velo", np.float64).tofile(f)
            np.ones(nat).tofile(f)
            np.array([t["eimp"] for t in trajs], dtype=np.float64).tofile(f)
            np.array([t["tadd"] for t in trajs], dtype=np.float64).tofile(f)
        penv = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", **(env or {}))
        p = subprocess.run([EMU_MD, fin, fout], capture_output=True, text=True, timeout=timeout, env=penv)
        if p.returncode != 0:
            raise RuntimeError(f"emulated MD failed ({p.returncode}):\n{p.stderr[-6000:]}")
        raw = open(fout, "rb").read()
    rc, running = np.frombuffer(raw, np.int32, 2, 0)
    o = 8
    xyz = np.frombuffer(raw, np.float64, 3 * nat, o).reshape(nat, 3).copy(); o += 24 * nat
    velo = np.frombuffer(raw, np.float64, 3 * nat, o).reshape(nat, 3).copy(); o += 24 * nat
    q = np.frombuffer(raw, np.float64, nat, o).copy(); o += 8 * nat
    res = []
    for _ in range(nt):
        ns, ex = np.frombuffer(raw, np.int32, 2, o); o += 8
        ep, ek = np.frombuffer(raw, np.float64, 2, o); o += 16
        nf, chk = np.frombuffer(raw, np.int32, 2, o); o += 8
        vcm = float(np.frombuffer(raw, np.float64, 1, o)[0]); o += 8
        res.append((int(ns), int(ex), float(ep), float(ek), int(nf), int(chk), vcm))
    return int(rc), int(running), xyz, velo, q, res
