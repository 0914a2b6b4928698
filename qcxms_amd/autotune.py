"""Autotuner of the SCC launch plan of single-point batches on MI355X.

For each basis-size bin it times cold-started GFN2 single points of a batch
of 256 systems of that bin (one workgroup per system fills the 256 compute
units once) under every launch choice the SCC kernel offers:
  eig  MFMA simultaneous-rotation eigen refinement before Jacobi (1) or
       Jacobi alone (0);
  lds  SCC matrices in LDS where they fit (1, up to 56 basis functions, LDS
       sized for the launch's largest system) or in global memory (0);
and keeps the fastest per bin (SCC kernel time from HIP events, median of the
repetitions): the choice for batches inside one bin.  For batches spanning
several bins it times a mixed batch (the 1024 uneven random chains, largest
first) launched once per bin against one launch with each single choice, and
keeps the fastest.  Output:
DIR/tuning.json (every timing) and DIR/tuning_table.h, the table the library
compiles in (qcxms_amd/csrc/tuning_table.h).

Run on the GPU:  python -m qcxms_amd.autotune --out DIR
"""
import argparse
import json
import os
import time

import numpy as np

from . import gpu
from . import workloads as W

BINS = [(0, 24), (24, 40), (40, 56), (56, 72), (72, 96), (96, 256)]
NSYS = 256


def nao_of(num):
    return int(sum(1 if int(z) == 1 else 4 for z in num))


def systems_in_bin(lo, hi, n, seed0=900000):
    """n random saturated C/H/N/O chains (seeded) with lo < nao <= hi; the
    largest bin is filled with C20H36O2N2+ (132 basis functions) and stretched
    C12H24O2N2+ (88) chains when the random chains run short."""
    out = []
    seed = seed0
    while len(out) < n and seed < seed0 + 200000:
        num, xyz, chg = W.random_chain(seed)
        seed += 1
        if lo < nao_of(num) <= hi:
            out.append((num, chg, xyz))
    k = 0
    while len(out) < n:
        num, xyz = (W.chain_c20h36o2n2(seed=k) if k % 2 == 0 else W.chain_c12h24o2n2(seed=k))
        if lo < nao_of(num) <= hi:
            out.append((num, 1, xyz))
        k += 1
        if k > 4 * n:
            break
    out.sort(key=lambda t: -nao_of(t[0]))
    return out


def time_plan(sysl, table, reps, mixed=(-1, 0)):
    """Median SCC kernel time (ms) per evaluation of the batch under a
    tuning table, and the batch's energies."""
    gpu.set_tuning(table, mixed)
    b = gpu.Batch([(t[0], t[1]) for t in sysl])
    try:
        b.set_warm_start(0)
        xyz = np.concatenate([t[2] for t in sysl])
        e0, _, _, st = b.egrad(xyz, 5000.0, gradient=False)
        if not np.all(st == 0):
            raise RuntimeError("SCC not converged while tuning")
        b.set_profiling(True)
        times = []
        for _ in range(reps):
            b.kernel_time("scc")
            b.egrad(xyz, 5000.0, gradient=False)
            ms, _ = b.kernel_time("scc")
            times.append(ms)
        b.set_profiling(False)
        return float(np.median(times)), e0
    finally:
        b.close()


def choices(hi):
    return [(e, l) for e in (0, 1) for l in ((0, 1) if hi <= 56 else (0,))]


def header(table, mixed):
    rows = "\n".join(f"    {{{a}, {e}, {l}}}," for a, e, l in table)
    return ("// SCC launch choices of single-point batches by basis size on MI355X, written\n"
            "// by qcxms_amd/autotune.py: {max_nao, MFMA eigen refinement, LDS matrices}\n"
            "// per bin, and {eig, lds} for batches spanning several bins (eig -1: one\n"
            "// launch per run of systems of one bin).\n"
            "#pragma once\n\nnamespace qx {\n\nconstexpr int kTuneTable[][3] = {\n"
            f"{rows}\n}};\nconstexpr int kTuneMixed[2] = {{{mixed[0]}, {mixed[1]}}};\n\n"
            "}  // namespace qx\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="tune_out")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    saved, saved_mixed = gpu.get_tuning()
    record = {"bins": [], "nsys_per_bin": NSYS, "reps": args.reps}
    table = []
    try:
        for lo, hi in BINS:
            sysl = systems_in_bin(lo, hi, NSYS)
            naos = [nao_of(t[0]) for t in sysl]
            res = {}
            ref_e = None
            for e, l in choices(hi):
                ms, en = time_plan(sysl, [(256, e, l)], args.reps)
                if ref_e is None:
                    ref_e = en
                res[f"eig{e}_lds{l}"] = {"scc_ms": ms, "max_dE_vs_first": float(np.abs(en - ref_e).max())}
                print(f"bin ({lo},{hi}] nao {min(naos)}-{max(naos)}: eig {e} lds {l}: {ms:.3f} ms",
                      flush=True)
            best = min(res, key=lambda k: res[k]["scc_ms"])
            e, l = int(best[3]), int(best[-1])
            table.append((hi, e, l))
            record["bins"].append({"lo": lo, "hi": hi, "nao_min": min(naos), "nao_max": max(naos),
                                   "nsys": len(sysl), "timings": res, "chosen": best})
        # merge neighbouring bins with the same choice
        merged = []
        for a, e, l in table:
            if merged and merged[-1][1:] == (e, l):
                merged[-1] = (a, e, l)
            else:
                merged.append((a, e, l))
        merged[-1] = (256, merged[-1][1], merged[-1][2])
        record["table"] = merged
        # the mixed batch decides: the binned table against one launch with
        # each single choice (one launch keeps the largest-first order over
        # all systems, which separate launches per bin lose)
        mixed = []
        for s in range(1024):
            num, xyz, chg = W.random_chain(5000 + s)
            mixed.append((num, chg, xyz))
        mixed.sort(key=lambda t: -nao_of(t[0]))
        cands = {"binned": (-1, 0)}
        for e, l in choices(56):
            cands[f"single_eig{e}_lds{l}"] = (e, l)
        timing, energies = {}, {}
        for name, mix in cands.items():
            timing[name], energies[name] = time_plan(mixed, merged, args.reps, mixed=mix)
            print(f"mixed 1024, {name}: {timing[name]:.2f} ms", flush=True)
        timing["untuned"], energies["untuned"] = time_plan(mixed, [(256, 0, 1)], args.reps)
        print(f"mixed 1024, untuned (one launch, Jacobi, LDS): {timing['untuned']:.2f} ms", flush=True)
        best = min(cands, key=timing.get)
        ref = energies["untuned"]
        record["mixed_1024"] = {"scc_ms": timing, "chosen": best,
                                "max_dE_vs_untuned": {k: float(np.abs(v - ref).max())
                                                      for k, v in energies.items()}}
        record["table"] = merged
        record["mixed"] = cands[best]
    finally:
        gpu.set_tuning(saved, saved_mixed)
    record["when"] = time.strftime("%Y-%m-%d %H:%M:%S")
    json.dump(record, open(os.path.join(args.out, "tuning.json"), "w"), indent=1)
    open(os.path.join(args.out, "tuning_table.h"), "w").write(header(record["table"], record["mixed"]))
    print("table", record["table"], "mixed", record["mixed"], flush=True)


if __name__ == "__main__":
    main()
