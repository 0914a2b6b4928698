"""Golden vectors for the ensemble inputs (qcxms_amd/ensemble.py).

- qcxms.start and qcxms.gs files written by the Fortran runtime with the
  write statements of src/utility.f90:358-369 and src/md.f90:151-152,
  383-386 (oracle/ref_bind.f90), on seeded values including signs, zeros and
  magnitudes from 1e-9 to 1e5: the expected text, byte for byte.
- getmaxiee / getieeab / gauss0 / poiss0 (src/iee.f90) and calctrelax
  (src/impact.f90:58-81), compiled from the reference into
  oracle/_ref/libqcxms_ref.so, on a grid of parameters.
Stored in tests/golden/ensemble_golden.json.
Run: python tests/golden/gen_ensemble_golden.py
"""
import ctypes as C
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402

dbl = C.c_double


def main():
    R = ob.ref()
    if R is None:
        raise SystemExit("build oracle/_ref first (make -C oracle ref)")
    R.ref_write_start.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, ob.dp, ob.dp, dbl, dbl]
    R.ref_write_gs.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, ob.dp, ob.dp]
    R.ref_getmaxiee.argtypes = [dbl, dbl, dbl, C.c_int, dbl, C.POINTER(dbl), C.POINTER(dbl),
                                C.POINTER(dbl)]
    R.ref_getieeab.argtypes = [dbl, C.c_int, dbl, C.c_int, dbl, C.POINTER(dbl), C.POINTER(dbl)]
    for f in (R.ref_gauss0, R.ref_poiss0):
        f.argtypes = [dbl, dbl, dbl, dbl]
        f.restype = dbl
    R.ref_calctrelax.argtypes = [C.c_int, ob.dp, C.c_int, C.c_int, dbl]
    R.ref_calctrelax.restype = dbl
    rng = np.random.default_rng(77)
    out = {"start": [], "gs": [], "gauss0": [], "poiss0": [], "getmaxiee": [], "getieeab": [],
           "calctrelax": []}
    with tempfile.TemporaryDirectory() as d:
        for case in range(6):
            nat = int(rng.integers(1, 9))
            velo = rng.normal(0, 10.0 ** rng.integers(-9, 0), (nat, 3))
            if case == 1:
                velo[0] = 0.0
            velof = rng.uniform(0, 2, nat)
            traj = int(rng.choice([1, 7, 42, 999, 1234]))
            eimp = float(rng.choice([0.0, 0.25, 1.5e-3, 12.75]))
            tadd = float(rng.uniform(0, 1e5))
            p = os.path.join(d, "qcxms.start").encode()
            R.ref_write_start(p, len(p), traj, nat, ob.ptr(np.ascontiguousarray(velo)),
                              ob.ptr(velof), eimp, tadd)
            out["start"].append(dict(traj=traj, eimp=eimp, tadd=tadd, velo=velo.tolist(),
                                     velof=velof.tolist(), text=open(p, "rb").read().decode()))
        for case in range(3):
            nat, nd = int(rng.integers(1, 6)), int(rng.integers(1, 4))
            x = rng.normal(0, 3.0, (nd, nat, 3))
            v = rng.normal(0, 1e-3, (nd, nat, 3))
            if case == 0:
                x[0, 0, 0] = -0.0
            nmax = int(rng.integers(10, 100000))
            p = os.path.join(d, "qcxms.gs").encode()
            R.ref_write_gs(p, len(p), nmax, nd, nat, ob.ptr(np.ascontiguousarray(x)),
                           ob.ptr(np.ascontiguousarray(v)))
            out["gs"].append(dict(nmax=nmax, xyz=x.tolist(), velo=v.tolist(),
                                  text=open(p, "rb").read().decode()))
    for a, b, el, x in [(0.1, 1.2, 21.0, 3.0), (0.25, 0.7, 49.0, 12.5), (0.3, 2.0, 8.0, 0.5)]:
        out["gauss0"].append([a, b, el, x, R.ref_gauss0(a, b, el, x)])
        out["poiss0"].append([a, b, el, x, R.ref_poiss0(a, b, el, x)])
    for a, b, el, ityp, exc in [(0.2, 1.4, 49.0, 1, 25.0), (0.1, 0.7, 21.0, 0, 12.3),
                                (0.3, 2.1, 34.0, 1, 40.0)]:
        r = [dbl(), dbl(), dbl()]
        R.ref_getmaxiee(a, b, el, ityp, exc, *[C.byref(v) for v in r])
        out["getmaxiee"].append([a, b, el, ityp, exc] + [v.value for v in r])
    for el, ityp, exc, nbnd, atm in [(49.0, 1, 25.0, 26, 0.6), (21.0, 1, 15.0, 11, 0.5),
                                     (49.0, 0, 25.0, 26, 0.6)]:
        a, b = dbl(), dbl()
        R.ref_getieeab(el, ityp, exc, nbnd, atm, C.byref(a), C.byref(b))
        out["getieeab"].append([el, ityp, exc, nbnd, atm, a.value, b.value])
    for n, na, i, tr in [(26, 25, 3, 2000.0), (26, 25, 25, 2000.0), (8, 7, 1, 500.0)]:
        emo = np.sort(rng.uniform(-1.0, -0.2, 10 * n))
        out["calctrelax"].append([n, na, i, tr, emo.tolist(), R.ref_calctrelax(n, ob.ptr(emo), na, i, tr)])
    path = os.path.join(HERE, "ensemble_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")
    print(out["start"][0]["text"][:300])
    print(out["gs"][0]["text"][:200])
    print(out["getieeab"], [r[-1] for r in out["calctrelax"]])


if __name__ == "__main__":
    main()
