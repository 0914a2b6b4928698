"""Counts the FP64 operations of the oracle's GFN2 single point by phase and
freezes the per-unit costs of the flop model that bench.py's roofline uses.

Test and measurement infrastructure (never part of the product).  Recipe:
  1. gfn2_oracle.c -> LLVM IR with -DORC_FLOP_COUNT (clang -O1, no
     vectorisation, no FMA contraction);
  2. after every double-precision fadd/fsub/fmul/fdiv/fneg, llvm.fma and
     llvm.fmuladd (2), sqrt, exp, pow, erf, log (1 each) an increment of
     orc_flops[orc_flop_phase] is inserted (phases are set by the ORC_PHASE
     markers of the source);
  3. the IR is compiled to oracle/_build/liboracle_flops.so and one single
     point per structure is counted;
  4. per-unit costs are fitted: integrals and their derivatives per
     primitive pair and per (primitive pair x AO pair) of the shell-pair
     classes, pair terms per atom pair and per atom triple (the
     coefficients reproduce the counted totals of every structure to the
     reported residual), and written to profiles/flop_model.json.

Run: python oracle/count_flops.py   (needs /opt/rocm/llvm/bin/clang)
"""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
CLANG = "/opt/rocm/llvm/bin/clang"
PHASES = ["other", "pair", "int", "scc", "grad_w", "grad_int", "grad_pair"]
NPH = 8

FLOP_RE = re.compile(r"^\s*%[\w.]+ = (fadd|fsub|fmul|fdiv|fneg)\b[^,]*?\b(<(\d+) x double>|double)\b")
CALL_RE = re.compile(r"call [^@]*@(llvm\.fmuladd\.f64|llvm\.fma\.f64|llvm\.sqrt\.f64|sqrt|exp|pow|erf|"
                     r"log|llvm\.exp\.f64|llvm\.pow\.f64|llvm\.log\.f64)\(")


def instrument(ll_in, ll_out):
    out, k, nins = [], 0, 0
    for line in open(ll_in):
        out.append(line)
        n = 0
        m = FLOP_RE.match(line)
        if m:
            n = int(m.group(3)) if m.group(3) else 1
        else:
            c = CALL_RE.search(line)
            if c and "declare" not in line:
                n = 2 if c.group(1) in ("llvm.fmuladd.f64", "llvm.fma.f64") else 1
        if n:
            k += 1
            nins += 1
            out.append(f"  %__fp{k} = load i32, ptr @orc_flop_phase\n"
                       f"  %__fa{k} = getelementptr [{NPH} x i64], ptr @orc_flops, i32 0, i32 %__fp{k}\n"
                       f"  %__fv{k} = load i64, ptr %__fa{k}\n"
                       f"  %__fw{k} = add i64 %__fv{k}, {n}\n"
                       f"  store i64 %__fw{k}, ptr %__fa{k}\n")
    out.append(f"@orc_flops = global [{NPH} x i64] zeroinitializer\n")
    open(ll_out, "w").writelines(out)
    return nins


def build():
    os.makedirs(BUILD, exist_ok=True)
    ll = os.path.join(BUILD, "gfn2_flops.ll")
    lli = os.path.join(BUILD, "gfn2_flops_counted.ll")
    subprocess.run([CLANG, "-O1", "-fno-vectorize", "-fno-slp-vectorize", "-ffp-contract=off",
                    "-fno-fast-math", "-DORC_FLOP_COUNT", "-fPIC", "-std=c11", "-S", "-emit-llvm",
                    os.path.join(HERE, "gfn2_oracle.c"), "-o", ll], check=True)
    n = instrument(ll, lli)
    so = os.path.join(BUILD, "liboracle_flops.so")
    objs = [os.path.join(BUILD, f) for f in ("md_oracle.o", "postmd_oracle.o", "ei_oracle.o")]
    subprocess.run([CLANG, "-O1", "-fPIC", "-shared", lli, *objs, "-o", so, "-lm", "-lpthread"],
                   check=True)
    return so, n


def shell_counts(num):
    """Per structure: shell pairs of each class (s-s, s-p, p-p; all ordered
    pairs ia >= ib as the oracle forms them, and those on different atoms),
    with their primitive pairs and AO pairs."""
    nprim = {1: [3], 6: [4, 4], 7: [4, 4], 8: [4, 4]}
    shells = []
    for a, z in enumerate(num):
        for l, g in enumerate(nprim[int(z)]):
            shells.append((a, l, g))
    f = dict(pp_all=0.0, pp_ao_all=0.0, sp_all=0.0, pp_off=0.0, pp_ao_off=0.0, sp_off=0.0)
    for i, (ai, li, gi) in enumerate(shells):
        for j in range(i + 1):
            aj, lj, gj = shells[j]
            npp = gi * gj
            nao = (2 * li + 1) * (2 * lj + 1)
            f["pp_all"] += npp
            f["pp_ao_all"] += npp * nao
            f["sp_all"] += 1
            if ai != aj:
                f["pp_off"] += npp
                f["pp_ao_off"] += npp * nao
                f["sp_off"] += 1
    return f, len(shells)


def structures():
    sys.path.insert(0, ROOT)
    from qcxms_amd import workloads as W
    out = [("C8H18+ (octane, seed 0)", *W.octane(seed=0), 1)]
    out.append(("C12H24O2N2+ stretched", *W.chain_c12h24o2n2(seed=1), 1))
    out.append(("C20H36O2N2+", *W.chain_c20h36o2n2(seed=1), 1))
    for s in range(5):
        num, xyz, chg = W.random_chain(100 + s)
        out.append((f"random chain {s}", num, xyz, chg))
    return out


def main():
    so, nins = build()
    L = C.CDLL(so)
    flops = (C.c_int64 * NPH).in_dll(L, "orc_flops")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    L.orc_gfn2_egrad.argtypes = [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int,
                                 C.c_double, C.c_void_p, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p,
                                 C.c_void_p, C.c_void_p]
    rows = []
    for name, num, xyz, chg in structures():
        num = np.ascontiguousarray(num, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        nat = len(num)
        for i in range(NPH):
            flops[i] = 0
        e, g, q = C.c_double(), np.zeros((nat, 3)), np.zeros(nat)
        info = ob.Gfn2Info()
        o = ob.default_opts()
        st = L.orc_gfn2_egrad(nat, num.ctypes.data_as(C.POINTER(C.c_int32)),
                              xyz.ctypes.data_as(C.POINTER(C.c_double)), chg, 5000.0, C.byref(o),
                              C.byref(e), g.ctypes.data_as(C.POINTER(C.c_double)),
                              q.ctypes.data_as(C.POINTER(C.c_double)), None, None, C.byref(info))
        counts = {PHASES[i]: int(flops[i]) for i in range(len(PHASES))}
        sc, nsh = shell_counts(num)
        rows.append(dict(name=name, nat=nat, nao=info.nao, nsh=nsh, iterations=info.iterations,
                         stat=st, counts=counts, shells=sc))
        print(f"{name}: nat {nat} nao {info.nao} it {info.iterations} " +
              " ".join(f"{k} {v / 1e6:.3f}M" for k, v in counts.items()), flush=True)

    def fit(cols, target):
        A = np.array([[r_[c] for c in cols] for r_ in cols_rows])
        y = np.array(target, dtype=float)
        coef, *_ = np.linalg.lstsq(A, y, rcond=None)
        resid = float(np.max(np.abs(A @ coef - y) / y))
        return coef, resid

    cols_rows = []
    for r in rows:
        nat = r["nat"]
        cols_rows.append(dict(pp_all=r["shells"]["pp_all"], pp_ao_all=r["shells"]["pp_ao_all"],
                              sp_all=r["shells"]["sp_all"], pp_off=r["shells"]["pp_off"],
                              pp_ao_off=r["shells"]["pp_ao_off"], sp_off=r["shells"]["sp_off"],
                              nat2=nat * (nat - 1) / 2.0, nat3=nat * (nat - 1) * (nat - 2) / 6.0,
                              nat1=float(nat), nsh2=r["nsh"] * r["nsh"] * 1.0))
    c_int, r_int = fit(["pp_all", "pp_ao_all", "sp_all"], [r["counts"]["int"] for r in rows])
    c_gint, r_gint = fit(["pp_off", "pp_ao_off", "sp_off"], [r["counts"]["grad_int"] for r in rows])
    c_pair, r_pair = fit(["nat1", "nat2", "nat3", "nsh2"],
                         [r["counts"]["pair"] + r["counts"]["grad_pair"] for r in rows])
    model = {
        "source": "oracle/count_flops.py: FP64 operations of oracle/gfn2_oracle.c counted in its "
                  "LLVM IR (clang -O1, no vectorisation, no FMA contraction; fma = 2, sqrt/exp/"
                  "pow/erf/log = 1), one GFN2 energy + gradient per structure, etemp 5000 K",
        "instrumented_ops": nins,
        "F_setup": "4/3 n^3 (Cholesky and inverse of S, once per evaluation)",
        "F_iter": "15 n^3 + 30 n^2 (eigendecomposition 9 n^3, two basis transforms 4 n^3, "
                  "density 2 n^3, Fock update and S/D/Q contractions 30 n^2)",
        "F_int": {"per_prim_pair": c_int[0], "per_prim_pair_ao_pair": c_int[1],
                  "per_shell_pair": c_int[2], "pairs": "all shell pairs ia >= ib",
                  "max_rel_residual": r_int},
        "F_grad": {"per_prim_pair": c_gint[0], "per_prim_pair_ao_pair": c_gint[1],
                   "per_shell_pair": c_gint[2], "pairs": "shell pairs on different atoms",
                   "W": "2 n^3 (energy-weighted density)", "max_rel_residual": r_gint},
        "F_pair": {"per_atom": c_pair[0], "per_atom_pair": c_pair[1],
                   "per_atom_triple": c_pair[2], "per_shell_pair_sq": c_pair[3],
                   "terms": "CN, D4 CN and weights, repulsion, ES2 gamma, D4 two-body, ATM, "
                            "multipole kernels, energy and gradient",
                   "max_rel_residual": r_pair},
        "nprim": {"1": [3], "6": [4, 4], "7": [4, 4], "8": [4, 4]},
        "structures": rows,
    }
    path = os.path.join(ROOT, "profiles", "flop_model.json")
    json.dump(model, open(path, "w"), indent=1)
    print(f"F_int {c_int} (resid {r_int:.2e}); F_grad {c_gint} (resid {r_gint:.2e}); "
          f"F_pair {c_pair} (resid {r_pair:.2e}) -> {path}")


if __name__ == "__main__":
    main()
