"""Race audit of the kernels under ThreadSanitizer (test infrastructure).

Builds the emulated library with -fsanitize=thread (make -C tests/cpu_emu
tsan) and runs, with every GPU thread of a block as a host thread:
  - single points of a 3-atom and a 5-atom system through k_setup, k_scc
    and k_grad under four SCC launch choices (Jacobi or MFMA refinement,
    matrices in LDS or in global memory);
  - two heating MD steps of two trajectories through the per-step kernels and
    through the persistent kernel k_md_fused.
Two accesses of one address by threads that no block or wavefront barrier
orders are reported; the summary lists each reported source line once.  A
run takes a few minutes per case, so this is a script rather than a test of
the default suite.

Run: python tests/cpu_emu/race_audit.py [--out DIR]
"""
import argparse
import collections
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import emu_bind as eb  # noqa: E402
from qcxms_amd import workloads as W  # noqa: E402


def summarize(logdir):
    sites = collections.Counter()
    nwarn = 0
    for f in glob.glob(os.path.join(logdir, "tsan.*")):
        text = open(f).read()
        for rep in text.split("WARNING: ThreadSanitizer: ")[1:]:
            nwarn += 1
            m = re.search(r"#0 (\S+).*?(\w+\.(?:h|hip|cpp)):(\d+)", rep)
            if m:
                sites[f"{m.group(2)}:{m.group(3)} ({m.group(1)[:60]})"] += 1
    return nwarn, sites


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "_build", "race_audit"))
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", HERE, "tsan"], check=True)
    eb.EMU_BIN = os.path.join(HERE, "_build", "egrad_emu_tsan")
    eb.EMU_MD = os.path.join(HERE, "_build", "md_emu_tsan")
    systems, coords = [], []
    for s in (5026, 5004):
        num, xyz, charge = W.random_chain(s)
        systems.append((num, charge, W.getspin(num, charge)))
        coords.append(xyz)
    cases = [(f"egrad eig{e} lds{l}", "egrad", [(256, e, l)]) for e in (0, 1) for l in (1, 0)]
    cases += [("md per-step", "md", 0), ("md persistent", "md", 1)]
    for name, kind, arg in cases:
        d = os.path.join(args.out, name.replace(" ", "_"))
        os.makedirs(d, exist_ok=True)
        for f in glob.glob(os.path.join(d, "tsan.*")):
            os.remove(f)
        os.environ["TSAN_OPTIONS"] = f"halt_on_error=0 report_signal_unsafe=0 history_size=2 log_path={d}/tsan"
        try:
            if kind == "egrad":
                eb.egrad_batch(systems, coords, 5000.0, tuning=arg, timeout=3000)
            else:
                trajs = []
                for s in (5004, 5026):
                    num, xyz, charge = W.random_chain(s)
                    mass = np.array([W.atomic_mass(int(z)) for z in num])
                    trajs.append(dict(num=num, xyz=xyz, velo=W.maxwell_boltzmann(mass, 800.0, s),
                                      mass=mass, eimp=0.05, tadd=100.0 * W.FSTOAU, charge=charge))
                eb.md_run(trajs, 2, fused=arg, timeout=3000)
        except RuntimeError:
            pass  # ThreadSanitizer ends a run that reported with status 66
        nwarn, sites = summarize(d)
        print(f"{name}: {nwarn} reports", flush=True)
        for site, n in sites.most_common():
            print(f"    {n:3d}  {site}", flush=True)


if __name__ == "__main__":
    main()
