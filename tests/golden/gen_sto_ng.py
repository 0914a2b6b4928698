"""Least-squares STO-nG expansions (Stewart, J. Chem. Phys. 52, 431 (1970)).

A Slater function r^(n-1) exp(-zeta r) Y_lm (zeta = 1) is expanded in ng
primitive Gaussians r^l exp(-alpha_i r^2) Y_lm.  Stewart's criterion is the
least-squares fit of the normalised functions, i.e. maximal overlap between
the Slater function and the normalised contraction.  For fixed exponents the
optimal coefficients solve a linear system; the exponents are optimised in
log space.  Exponents scale with zeta^2 for other zeta.

tblite builds its basis from Stewart's tabulated fits (tblite v0.2.1,
slater_to_gauss).  Wherever Stewart's printed values are held in
tests/golden/stewart_1970.json they replace the fit, digit for digit (the
shells GFN2 uses for H, C, N, O are among them: 1s with three terms, 2s and 2p
with four); the fit fills the remaining entries.  tests/test_sto_ng.py pins the
printed values: their coefficients are the least-squares coefficients of their
exponents to the printed digits, and the overlap is stationary there.

Writes the table (published values, fits elsewhere) as data into
(tests/golden/sto_ng_fit.json holds the fits; --refit rewrites it):
  qcxms_amd/csrc/sto_ng_table.h   (product)
  oracle/sto_ng_oracle.h          (CPU oracle)
Run: python tests/golden/gen_sto_ng.py [--refit]   (--refit redoes the fits, slow)
"""
import json
import math
import os

import numpy as np
from scipy import optimize, special

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

CASES = [(n, l) for n in (1, 2, 3) for l in range(n)]
NG = range(1, 7)


def sto_gauss_overlap(n, l, alpha):
    """<r^(n-1) e^-r | r^l e^(-alpha r^2)> with r^2 dr, unnormalised.

    Closed form int_0^inf r^p exp(-r - a r^2) dr
      = Gamma(p+1) (2a)^(-(p+1)/2) exp(1/(8a)) D_{-(p+1)}(1/sqrt(2a))
    with D the parabolic cylinder function (Gradshteyn-Ryzhik 3.462.1)."""
    p = n - 1 + l + 2
    a = float(alpha)
    return (math.gamma(p + 1) * (2 * a) ** (-(p + 1) / 2) * math.exp(1 / (8 * a))
            * special.pbdv(-(p + 1), 1 / math.sqrt(2 * a))[0])


def gauss_gauss(l, a):
    """<r^l e^(-a1 r^2) | r^l e^(-a2 r^2)> for a = a1 + a2."""
    return special.gamma(l + 1.5) / (2.0 * a ** (l + 1.5))


def sto_norm2(n):
    return math.factorial(2 * n) / 2.0 ** (2 * n + 1)


def overlap_of(n, l, logalpha):
    alpha = np.exp(logalpha)
    if np.any(alpha < 5e-3) or np.any(alpha > 1e6):
        return -1.0, None, None
    b = np.array([sto_gauss_overlap(n, l, a) for a in alpha])
    g = np.array([[gauss_gauss(l, ai + aj) for aj in alpha] for ai in alpha])
    if not np.all(np.isfinite(b)) or np.linalg.cond(g) > 1e14:
        return -1.0, None, g
    c = np.linalg.solve(g, b)
    ov2 = float(b @ c) / sto_norm2(n)
    return ov2, c, g


def fit(n, l, ng, start=None):
    if start is None:
        start = np.log(np.geomspace(0.1 * 8.0 ** (ng - 1), 0.1, ng)) if ng > 1 else np.log([0.3])
    best = None
    for scale in (1.0, 0.5, 2.0, 0.25, 4.0):
        x0 = np.sort(start + math.log(scale))[::-1]
        res = optimize.minimize(lambda x: -overlap_of(n, l, x)[0], x0, method="Nelder-Mead",
                                options={"xatol": 1e-12, "fatol": 1e-16, "maxiter": 40000,
                                         "maxfev": 80000})
        res = optimize.minimize(lambda x: -overlap_of(n, l, x)[0], res.x, method="BFGS",
                                options={"gtol": 1e-13})
        if best is None or res.fun < best.fun:
            best = res
    x = np.sort(best.x)[::-1]
    ov2, c, g = overlap_of(n, l, x)
    alpha = np.exp(x)
    # coefficients for normalised primitives, contraction normalised to one
    d = c * np.sqrt(np.diag(g))
    gn = g / np.sqrt(np.outer(np.diag(g), np.diag(g)))
    d = d / math.sqrt(float(d @ gn @ d))
    if d.sum() < 0:
        d = -d
    return alpha, d, math.sqrt(ov2)


def published():
    """Stewart's printed expansions, key 'nl_ng' -> (alpha, coeff)."""
    with open(os.path.join(HERE, "stewart_1970.json")) as fh:
        data = json.load(fh)
    return {k: (v[0], v[1]) for k, v in data.items() if not k.startswith("_")}


def refit():
    """All least-squares fits (slow: minutes per entry set); writes
    sto_ng_fit.json."""
    table = {}
    prev = {}
    for (n, l) in CASES:
        for ng in NG:
            start = None
            if (n, l, ng - 1) in prev:
                a_prev = np.log(prev[(n, l, ng - 1)])
                start = np.sort(np.concatenate([a_prev, [a_prev.max() + 1.5]]))[::-1]
            alpha, d, ov = fit(n, l, ng, start)
            prev[(n, l, ng)] = alpha
            table[f"{n}{'spd'[l]}_{ng}"] = {"n": n, "l": l, "ng": ng, "alpha": alpha.tolist(),
                                            "coeff": d.tolist(), "overlap": ov}
            print(f"{n}{'spd'[l]} ng={ng} overlap={ov:.12f} alpha={alpha} coeff={d}")
    with open(os.path.join(HERE, "sto_ng_fit.json"), "w") as fh:
        json.dump(table, fh, indent=1)


def main(do_refit=False):
    if do_refit:
        refit()
    with open(os.path.join(HERE, "sto_ng_fit.json")) as fh:
        table = json.load(fh)
    for key, (alpha, coeff) in published().items():
        table[key] = dict(table[key], alpha=list(alpha), coeff=list(coeff))
        print(f"{key}: Stewart 1970")

    def header(guard, prefix, comment):
        lines = [f"/* {comment}",
                 " * STO-nG expansions of Slater functions with zeta = 1: Stewart's printed",
                 " * values (J. Chem. Phys. 52, 431 (1970)) where held in",
                 " * tests/golden/stewart_1970.json, least-squares fits by his criterion",
                 " * elsewhere; generated by tests/golden/gen_sto_ng.py.",
                 " * Index [n-1][l][ng-1]; exponents scale with zeta^2.  Coefficients",
                 " * multiply normalised primitives. */",
                 f"#ifndef {guard}", f"#define {guard}",
                 f"static const double {prefix}_alpha[3][3][6][6] = {{"]
        for n in (1, 2, 3):
            lines.append(" {")
            for l in range(3):
                rows = []
                for ng in NG:
                    key = f"{n}{'spd'[l]}_{ng}"
                    vals = table[key]["alpha"] if key in table else []
                    vals = list(vals) + [0.0] * (6 - len(vals))
                    rows.append("   {" + ", ".join(f"{v:.17g}" for v in vals) + "}")
                lines.append("  {\n" + ",\n".join(rows) + "},")
            lines.append(" },")
        lines.append("};")
        lines.append(f"static const double {prefix}_coeff[3][3][6][6] = {{")
        for n in (1, 2, 3):
            lines.append(" {")
            for l in range(3):
                rows = []
                for ng in NG:
                    key = f"{n}{'spd'[l]}_{ng}"
                    vals = table[key]["coeff"] if key in table else []
                    vals = list(vals) + [0.0] * (6 - len(vals))
                    rows.append("   {" + ", ".join(f"{v:.17g}" for v in vals) + "}")
                lines.append("  {\n" + ",\n".join(rows) + "},")
            lines.append(" },")
        lines.append("};")
        lines.append("#endif")
        return "\n".join(lines) + "\n"

    with open(os.path.join(ROOT, "qcxms_amd", "csrc", "sto_ng_table.h"), "w") as fh:
        fh.write(header("QCXMS_STO_NG_TABLE_H", "sto_ng", "Basis data for the HIP library."))
    with open(os.path.join(ROOT, "oracle", "sto_ng_oracle.h"), "w") as fh:
        fh.write(header("QCXMS_STO_NG_ORACLE_H", "orc_sto", "Basis data for the CPU oracle."))


if __name__ == "__main__":
    import sys
    main(do_refit="--refit" in sys.argv)
