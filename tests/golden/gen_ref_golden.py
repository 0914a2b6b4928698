"""Golden vectors from the reference's own Fortran routines.

Runs QCxMS's fragment_structure/fragmass (src/fragments.f90), ekinet
(src/mdinit.f90), impactscale (src/impact.f90), Rad/atomic_mass tables and the
in-tree D4 routines (src/dftd4.f90), compiled in place into
oracle/_ref/libqcxms_ref.so by oracle/Makefile, on seeded inputs and stores
inputs + outputs in tests/golden/ref_golden.npz.  The fixture lets the GPU
box (which has no reference tree) and later rounds check against the
reference without rebuilding it.
Run: python tests/golden/gen_ref_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402

D4_GA, D4_GC, D4_WF = 3.0, 2.0, 6.0
D4_S6, D4_S8, D4_S9, D4_A1, D4_A2, D4_ALP = 1.0, 2.7, 5.0, 0.52, 5.0, 16


def ragged(arrs):
    off = np.cumsum([0] + [len(a) for a in arrs]).astype(np.int64)
    return np.concatenate(arrs) if arrs else np.zeros(0), off


def cluster(rng, nat, spread, elems=(1, 6, 7, 8)):
    oz = rng.choice(elems, size=nat).astype(np.int32)
    xyz = rng.normal(0.0, spread, size=(nat, 3))
    return oz, xyz


def main():
    R = ob.ref()
    if R is None:
        raise SystemExit("build oracle/_ref first (make -C oracle ref)")
    rng = np.random.default_rng(20261202)
    out = {}
    zs = np.arange(1, 119, dtype=np.int32)
    out["rad"] = np.array([R.ref_rad(int(z)) for z in zs])
    out["mass"] = np.array([R.ref_atomic_mass(int(z)) for z in zs])
    out["covrad_d3"] = np.array([R.ref_covrad_d3(int(z)) for z in zs])

    # ---- fragment_structure: random clusters of varying density and size
    f_oz, f_xyz, f_frag = [], [], []
    for case in range(120):
        nat = int(rng.integers(1, 64))
        spread = float(rng.choice([1.5, 3.0, 6.0, 12.0]))
        oz, xyz = cluster(rng, nat, spread, elems=(1, 6, 7, 8, 9, 17, 35))
        f_oz.append(oz); f_xyz.append(xyz)
        f_frag.append(ob.fragment_structure(R, "ref_", oz, xyz))
    # near-threshold pairs: distance at rcov*(1 + k ulp)
    for case in range(60):
        nat = int(rng.integers(2, 12))
        oz, xyz = cluster(rng, nat, 40.0)
        i, j = rng.choice(nat, size=2, replace=False)
        rcov = 3.0 * 0.5 * (out["rad"][oz[i] - 1] + out["rad"][oz[j] - 1])
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        k = int(rng.integers(-3, 4))
        dist = rcov
        for _ in range(abs(k)):
            dist = np.nextafter(dist, np.inf if k > 0 else -np.inf)
        xyz[j] = xyz[i] + d * dist
        f_oz.append(oz); f_xyz.append(xyz)
        f_frag.append(ob.fragment_structure(R, "ref_", oz, xyz))
    out["frag_oz"], out["frag_off"] = ragged(f_oz)
    out["frag_xyz"], _ = ragged(f_xyz)
    out["frag_out"], _ = ragged(f_frag)

    # ---- fragmass on the same cases (atoms with labels up to 10 only)
    m_nfrag, m_fragm, m_fragat = [], [], []
    for oz, frag in zip(f_oz, f_frag):
        if frag.max() > 10:
            m_nfrag.append(-1); m_fragm.append(np.zeros(10)); m_fragat.append(np.zeros((10, 200), np.int32))
            continue
        mass = out["mass"][oz - 1]
        nf, fm, fa = ob.fragmass(R, "ref_", oz, frag, mass)
        m_nfrag.append(nf); m_fragm.append(fm); m_fragat.append(fa)
    out["fm_nfrag"] = np.array(m_nfrag, dtype=np.int32)
    out["fm_fragm"] = np.array(m_fragm)
    out["fm_fragat"] = np.array(m_fragat)

    # ---- ekinet and impactscale
    e_v, e_m, e_out = [], [], []
    i_v, i_m, i_f, i_par, i_out = [], [], [], [], []
    for case in range(40):
        nat = int(rng.integers(1, 62))
        mass = out["mass"][rng.choice([0, 5, 6, 7], size=nat)]
        velo = rng.normal(0.0, 1e-3, size=(nat, 3)) * np.sqrt(1836.0 / mass)[:, None]
        e, t = ob.ekinet(R, "ref_", velo, mass)
        e_v.append(velo); e_m.append(mass); e_out.append([e, t])
        velof = np.where(rng.random(nat) < 0.5, 1.0, 1.0 + rng.random(nat))
        e0 = e
        eimp = float(rng.uniform(0.05, 0.6))
        ff = float(rng.uniform(0.0, 0.05))
        v, _ = ob.impactscale(R, "ref_", velo, mass, velof, eimp, ff, e0)
        i_v.append(velo); i_m.append(mass); i_f.append(velof); i_par.append([eimp, ff, e0])
        i_out.append(v)
    out["ek_velo"], out["ek_off"] = ragged(e_v)
    out["ek_mass"], _ = ragged(e_m)
    out["ek_out"] = np.array(e_out)
    out["is_velo"], out["is_off"] = ragged(i_v)
    out["is_mass"], _ = ragged(i_m)
    out["is_velof"], _ = ragged(i_f)
    out["is_par"] = np.array(i_par)
    out["is_out"], _ = ragged(i_out)

    # ---- D4 pieces
    zt = []
    for _ in range(200):
        a, c = 3.0, float(rng.uniform(0.5, 2.0))
        qref, qmod = float(rng.uniform(0, 9)), float(rng.uniform(-0.5, 9))
        zt.append([a, c, qref, qmod, R.ref_zeta(a, c, qref, qmod), R.ref_dzeta(a, c, qref, qmod)])
    out["d4_zeta"] = np.array(zt)
    pols = rng.random((20, 23)) * 10
    R.ref_trapzd.argtypes = [ob.dp]
    out["d4_trapzd_in"] = pols
    out["d4_trapzd_out"] = np.array([R.ref_trapzd(ob.ptr(np.ascontiguousarray(p))) for p in pols])
    # reference tables after d4dim (mode 0: GFN2 charges)
    at = np.array([1, 6, 7, 8], dtype=np.int32)
    nd = C.c_int()
    R.ref_d4dim(4, ob.ptr(at, ob.ip), C.c_double(D4_GA), C.c_double(D4_GC), 0, C.byref(nd))
    tabs = {}
    for z in (1, 6, 7, 8):
        nref = C.c_int(); rq = np.zeros(7); rc = np.zeros(7); rcn = np.zeros(7, np.int32)
        ral = np.zeros((7, 23))
        R.ref_d4_tables(z, C.byref(nref), ob.ptr(rq), ob.ptr(rc), ob.ptr(rcn, ob.ip), ob.ptr(ral))
        tabs[z] = (nref.value, rq, rc, rcn, ral)
    out["d4_nref"] = np.array([tabs[z][0] for z in (1, 6, 7, 8)], dtype=np.int32)
    out["d4_refq"] = np.array([tabs[z][1] for z in (1, 6, 7, 8)])
    out["d4_refcovcn"] = np.array([tabs[z][2] for z in (1, 6, 7, 8)])
    out["d4_refc"] = np.array([tabs[z][3] for z in (1, 6, 7, 8)])
    out["d4_refal"] = np.array([tabs[z][4] for z in (1, 6, 7, 8)])
    # covalent CN and full D4 (two-body at charges q + ATM at q = 0) with gradient
    d_oz, d_xyz, d_q, d_cn, d_e, d_g = [], [], [], [], [], []
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from qcxms_amd import workloads as W
    mols = [W.octane(seed=3), W.chain_c12h24o2n2(seed=2), W.build_chain([8], noise=0.05, seed=4),
            W.build_chain([6, 7, 8, 6], noise=0.05, seed=5)]
    for oz, xyz in mols:
        nat = len(oz)
        q = rng.normal(0.0, 0.2, size=nat)
        cn = np.zeros(nat)
        R.ref_covncoord(nat, ob.ptr(oz, ob.ip), ob.ptr(np.ascontiguousarray(xyz)), C.c_double(1600.0),
                        ob.ptr(cn))
        R.ref_d4dim(nat, ob.ptr(oz, ob.ip), C.c_double(D4_GA), C.c_double(D4_GC), 0, C.byref(nd))
        g = np.zeros((nat, 3)); e = C.c_double()
        R.ref_dispgrad(nat, nd.value, ob.ptr(oz, ob.ip), ob.ptr(q), ob.ptr(np.ascontiguousarray(xyz)),
                       C.c_double(D4_S6), C.c_double(D4_S8), C.c_double(D4_S9), C.c_double(D4_A1),
                       C.c_double(D4_A2), D4_ALP, C.c_double(D4_WF), C.c_double(D4_GA),
                       C.c_double(D4_GC), ob.ptr(g), C.byref(e))
        d_oz.append(oz); d_xyz.append(xyz); d_q.append(q); d_cn.append(cn); d_e.append(e.value); d_g.append(g)
    out["disp_oz"], out["disp_off"] = ragged(d_oz)
    out["disp_xyz"], _ = ragged(d_xyz)
    out["disp_q"], _ = ragged(d_q)
    out["disp_cn"], _ = ragged(d_cn)
    out["disp_e"] = np.array(d_e)
    out["disp_g"], _ = ragged(d_g)
    np.savez_compressed(os.path.join(HERE, "ref_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_golden.npz"))


if __name__ == "__main__":
    main()
