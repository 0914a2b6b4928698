"""Extract the D4 reference data that GFN2-xTB dispersion needs for H, C, N, O.

Source: the reference's in-tree D4 data and tables
  include/param_d4.fh  (reference polarisabilities alphaiw, refq/refh = GFN2
                        charges, hcount, ascale, refcovcn, refcn, refsys, and the
                        secondary systems secq/sscale/secaiw)
  src/dftd4.f90:20-150 (zeff, gam, en, rcov, r4r2)
The data is copied as data only (no code); each consumer restates the D4
algorithm itself.  Writes:
  oracle/d4_data_oracle.h      (CPU oracle)
  qcxms_amd/csrc/d4_data.h     (HIP library)
Run in the container that has the reference:  python tests/golden/gen_d4_data.py
"""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
ELEMENTS = (1, 6, 7, 8)
NSEC = 8


def parse_param(path):
    txt = open(path).read().replace("&\n     &", " ").replace("&\n", " ")
    scal = {}
    arr = {}
    for m in re.finditer(r"data\s+(\w+)\s*\(([^)]*)\)\s*/([^/]*)/", txt):
        name, idx, val = m.group(1).lower(), m.group(2), m.group(3)
        parts = [p.strip() for p in idx.split(",")]
        vals = [float(v.replace("d", "e")) for v in val.replace("\n", " ").split(",") if v.strip()]
        if parts[0] == ":":
            key = tuple(int(p) for p in parts[1:])
            arr[(name,) + key] = vals
        else:
            key = tuple(int(p) for p in parts)
            scal[(name,) + key] = vals[0]
    return scal, arr


def parse_table(src, name):
    m = re.search(r"::\s*" + name + r"\(\*\)\s*=\s*\(/(.*?)/\)", src, re.S | re.I)
    body = re.sub(r"!.*", "", m.group(1))
    body = body.replace("&", " ")
    return [float(v.replace("d", "e").replace("_wp", "")) for v in body.split(",") if v.strip()]


def main():
    scal, arr = parse_param(os.path.join(REF, "include", "param_d4.fh"))
    src = open(os.path.join(REF, "src", "dftd4.f90")).read()
    tables = {n: parse_table(src, n) for n in ("zeff", "gam", "en", "rcov", "r4r2")}

    def elem_block(prefix):
        out = []
        out.append(f"/* element order: Z = {', '.join(str(z) for z in ELEMENTS)} */")
        out.append(f"static const int {prefix}_z[4] = {{{', '.join(str(z) for z in ELEMENTS)}}};")
        out.append(f"static const int {prefix}_refn[4] = {{"
                   + ", ".join(str(int(scal[("refn", z)])) for z in ELEMENTS) + "};")
        for name in ("refq", "refh", "hcount", "ascale", "refcovcn", "refcn"):
            rows = []
            for z in ELEMENTS:
                vals = [scal.get((name, j, z), 0.0) for j in range(1, 8)]
                rows.append("  {" + ", ".join(f"{v:.17g}" for v in vals) + "}")
            out.append(f"static const double {prefix}_{name}[4][7] = {{\n" + ",\n".join(rows) + "};")
        rows = []
        for z in ELEMENTS:
            vals = [int(scal.get(("refsys", j, z), 0)) for j in range(1, 8)]
            rows.append("  {" + ", ".join(str(v) for v in vals) + "}")
        out.append(f"static const int {prefix}_refsys[4][7] = {{\n" + ",\n".join(rows) + "};")
        rows = []
        for z in ELEMENTS:
            sub = []
            for j in range(1, 8):
                vals = arr.get(("alphaiw", j, z), [0.0] * 23)
                sub.append("   {" + ", ".join(f"{v:.17g}" for v in vals) + "}")
            rows.append("  {\n" + ",\n".join(sub) + "}")
        out.append(f"static const double {prefix}_alphaiw[4][7][23] = {{\n" + ",\n".join(rows) + "};")
        # secondary systems 1..8 (index = element number of the secondary)
        for name in ("secq", "sscale"):
            vals = [scal.get((name, i), 0.0) for i in range(1, NSEC + 1)]
            out.append(f"static const double {prefix}_{name}[{NSEC}] = {{"
                       + ", ".join(f"{v:.17g}" for v in vals) + "};")
        sub = []
        for i in range(1, NSEC + 1):
            vals = arr.get(("secaiw", i), [0.0] * 23)
            sub.append("  {" + ", ".join(f"{v:.17g}" for v in vals) + "}")
        out.append(f"static const double {prefix}_secaiw[{NSEC}][23] = {{\n" + ",\n".join(sub) + "};")
        for name in ("zeff", "gam", "en", "rcov", "r4r2"):
            vals = tables[name][:NSEC]
            out.append(f"/* dftd4.f90 {name}, elements 1..{NSEC} */")
            out.append(f"static const double {prefix}_{name}[{NSEC}] = {{"
                       + ", ".join(f"{v:.17g}" for v in vals) + "};")
        return "\n".join(out) + "\n"

    def header(guard, prefix, comment):
        return (f"/* {comment}\n * D4 reference data for H, C, N, O copied from the reference's\n"
                " * include/param_d4.fh and src/dftd4.f90 tables by tests/golden/gen_d4_data.py.\n"
                " * GFN2 charge set (refq/refh/secq as loaded, no mode override). */\n"
                f"#ifndef {guard}\n#define {guard}\n" + elem_block(prefix) + "#endif\n")

    with open(os.path.join(ROOT, "oracle", "d4_data_oracle.h"), "w") as fh:
        fh.write(header("QCXMS_D4_DATA_ORACLE_H", "orc_d4", "Data for the CPU oracle."))
    with open(os.path.join(ROOT, "qcxms_amd", "csrc", "d4_data.h"), "w") as fh:
        fh.write(header("QCXMS_D4_DATA_H", "d4", "Data for the HIP library."))
    print("ok", {k: v[:8] for k, v in tables.items()})


if __name__ == "__main__":
    main()
