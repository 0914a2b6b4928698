"""Golden vectors for the post-MD fragment analysis.

- gam / gam3 of src/dftd4.f90:38-72 (REAL(4) literals as REAL(8)) for H, C,
  N, O, read from the reference's own module compiled into
  oracle/_ref/libqcxms_ref.so: the tables eself (src/iniqm.f90:430-451) uses.
- qcxms.res records written by the Fortran runtime with the format of
  src/write_fragments.f90:402-441, (F10.7,i3,2i5,2i2,2x,i3,2x,20(i4,i3)), for
  EI item lists (no collision number) and CID item lists, over seeded and
  edge-case values (negative and fractional charges, fields too narrow for
  their value, isotopes stored at 100 + mass number).
Stored in tests/golden/post_md_golden.json.
Run: python tests/golden/gen_post_md_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402


def main():
    R = ob.ref()
    if R is None:
        raise SystemExit("build oracle/_ref first (make -C oracle ref)")
    R.ref_gam.restype = C.c_double
    R.ref_gam.argtypes = [C.c_int]
    R.ref_gam3.restype = C.c_double
    R.ref_gam3.argtypes = [C.c_int]
    R.ref_res_record.argtypes = [C.c_double] + [C.c_int] * 6 + [ob.ip, ob.ip, C.c_char_p, C.c_int]
    out = {"gam": {}, "gam3": {}, "records": []}
    for z in (1, 6, 7, 8):
        out["gam"][str(z)] = R.ref_gam(z)
        out["gam3"][str(z)] = R.ref_gam3(z)
    rng = np.random.default_rng(2024)
    cases = [
        (1.0, 1, 1, -1, 1, 1, [(6, 8), (1, 18)]),
        (0.6180339887, 1, 17, -1, 1, 2, [(6, 3), (1, 7)]),
        (0.3819660113, 1, 17, -1, 1, 1, [(6, 5), (1, 11)]),
        (-0.25, -1, 4, -1, 2, 3, [(8, 1)]),
        (0.0, 1, 123456, -1, 1, 1, [(1, 1)]),
        (12.5, 2, 99999, -1, 100, 12, [(6, 20), (7, 2), (8, 2), (1, 36)]),
        (1.0, 1, 3, 2, 1, 1, [(6, 8), (1, 18)]),
        (0.123456789, 1, 42, 11, 3, 2, [(6, 2), (1, 5), (113, 1)]),
        (-1e-9, 1, 5, -1, 1, 1, [(6, 1), (1, 4)]),
        (0.99999995, 1, 6, -1, 1, 1, [(8, 1), (1, 2)]),
    ]
    for _ in range(30):
        l = int(rng.integers(1, 5))
        zs = sorted(rng.choice([1, 6, 7, 8, 102, 113], size=l, replace=False).tolist())
        pairs = [(int(z), int(rng.integers(1, 40))) for z in zs]
        icoll = -1 if rng.random() < 0.6 else int(rng.integers(0, 30))
        cases.append((float(rng.uniform(-1.2, 2.2)), int(rng.choice([-2, -1, 1, 2, 3])),
                      int(rng.integers(1, 3000)), icoll, int(rng.integers(1, 4)),
                      int(rng.integers(1, 6)), pairs))
    for q, mchrg, itrj, icoll, isec, j, pairs in cases:
        l = len(pairs)
        idum2 = np.array([p[0] for p in pairs], dtype=np.int32)
        idum1 = np.array([p[1] for p in pairs], dtype=np.int32)
        buf = C.create_string_buffer(200)
        R.ref_res_record(q, mchrg, itrj, icoll, isec, j, l, ob.ptr(idum2, ob.ip), ob.ptr(idum1, ob.ip),
                         buf, 200)
        line = buf.raw.decode().rstrip()
        out["records"].append(dict(q=q, mchrg=mchrg, itrj=itrj, icoll=icoll, isec=isec, j=j,
                                   pairs=pairs, line=line))
    path = os.path.join(HERE, "post_md_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {len(out['records'])} records")
    for r in out["records"][:10]:
        print(repr(r["line"]))


if __name__ == "__main__":
    main()
