"""ctypes bindings to the CPU oracle (oracle/_build/liboracle.so) and to the
reference's own Fortran routines compiled in place (oracle/_ref/libqcxms_ref.so).
Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libqcxms_ref.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)


def ptr(a, t=dp):
    return None if a is None else a.ctypes.data_as(t)


class SccOpts(C.Structure):
    _fields_ = [("econv", C.c_double), ("pconv", C.c_double), ("maxiter", C.c_int),
                ("broyden_damp", C.c_double), ("cold_start", C.c_int)]


class Gfn2Info(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("energy", "e_rep", "e_disp", "e_es2", "e_es3",
                                           "e_aes", "e_axc", "e_band", "e_ts")]
                + [(n, C.c_int) for n in ("iterations", "converged", "nao", "nsh", "nel_int")])


class CidState(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("nuc", "nuc0", "chrg", "ntot", "collided", "count_average")]
                + [(n, C.c_double) for n in ("tstep", "etemp", "tinit", "beta", "velo_cm", "ttime")]
                + [(n, C.c_int) for n in ("nstep", "nfrag", "gradfail", "fragmented", "total_steps",
                                          "step_counter", "check_fragmented", "collision_step", "done")]
                + [(n, C.c_double) for n in ("epot", "ekin", "lowest_com", "new_velo", "avgt",
                                             "atlast", "ecom", "velo_diff")]
                + [("fragT", C.c_double * 10), ("eint", C.c_double * 10)])


class MdState(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("nat", "chrg", "nfragexit", "max_steps", "isec")]
                + [(n, C.c_double) for n in ("tstep", "tadd", "eimp", "ax", "ieetemp", "etempin")]
                + [(n, C.c_int) for n in ("nstep", "nfrag", "fragstate", "mdok", "exit_reason")]
                + [(n, C.c_double) for n in ("tav", "epav", "ekav", "ttime", "dtime", "atlast",
                                             "epot", "ekin", "eerror", "etemp")]
                + [("fragm", C.c_double * 10), ("fragT", C.c_double * 10)]
                + [(n, C.c_int) for n in ("cid", "count_average", "check_fragmented")]
                + [("new_velo", C.c_double)])


EGRAD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, ip, dp, C.c_int, C.c_double, dp, dp, dp)

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        L = C.CDLL(ORACLE_SO)
        for name in ("orc_aatoau", "orc_amutoau", "orc_autoamu", "orc_gnorm", "orc_setetemp",
                     "orc_d4_trapzd", "orc_d4_zeta_fn", "orc_d4_dzeta_fn", "orc_d4_twobody",
                     "orc_d4_atm"):
            getattr(L, name).restype = C.c_double
        for name in ("orc_rad", "orc_covrad_d3", "orc_atomic_mass"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_int]
        L.orc_setetemp.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_d4_zeta_fn.argtypes = [C.c_double] * 4
        L.orc_d4_dzeta_fn.argtypes = [C.c_double] * 4
        L.orc_gfn2_egrad.argtypes = [C.c_int, ip, dp, C.c_int, C.c_double, C.c_void_p, dp, dp,
                                     dp, dp, dp, C.c_void_p]
        L.orc_gfn2_egrad_batch.argtypes = [C.c_int, ip, ip, dp, ip, dp, C.c_void_p, dp, dp, dp,
                                           ip, C.c_int]
        L.orc_md_run.argtypes = [C.c_void_p, ip, ip, dp, dp, dp, dp, dp, dp, ip, EGRAD_FN,
                                 C.c_void_p, C.c_int]
        L.orc_md_run_avg.argtypes = [C.c_void_p, ip, ip, dp, dp, dp, dp, dp, dp, ip, EGRAD_FN,
                                     C.c_void_p, C.c_int, dp, dp]
        L.orc_md_set_trace.argtypes = [dp, C.c_int]
        L.orc_md_set_xyz_trace.argtypes = [dp, C.c_int]
        L.orc_cid_run.argtypes = [C.c_void_p, ip, dp, dp, dp, dp, dp, ip, dp, EGRAD_FN, C.c_void_p,
                                  C.c_int]
        _oracle = L
    return _oracle


def ref():
    """Reference Fortran library, or None when it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        L = C.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        for name in ("ref_rad", "ref_covrad_d3", "ref_atomic_mass"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_int]
        for name in ("ref_zeta", "ref_dzeta", "ref_cngw"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double] * 4 if name != "ref_cngw" else [C.c_double] * 3
        L.ref_trapzd.restype = C.c_double
        _ref = L
    return _ref


def default_opts(**kw):
    o = SccOpts()
    oracle().orc_scc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def gfn2_egrad(num, xyz, charge=0, etemp=300.0, opts=None, grad=True, moments=False):
    """Oracle GFN2 single point.  Returns dict(stat, energy, grad, qat, info[, dp, qp])."""
    L = oracle()
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nat = len(num)
    e = C.c_double()
    g = np.zeros((nat, 3)) if grad else None
    q = np.zeros(nat)
    dpm = np.zeros((nat, 3)) if moments else None
    qpm = np.zeros((nat, 6)) if moments else None
    info = Gfn2Info()
    o = opts if opts is not None else default_opts()
    st = L.orc_gfn2_egrad(nat, ptr(num, ip), ptr(xyz), int(charge), float(etemp), C.byref(o),
                          C.byref(e), ptr(g), ptr(q), ptr(dpm), ptr(qpm), C.byref(info))
    out = dict(stat=st, energy=e.value, grad=g, qat=q, info=info)
    if moments:
        out["dp"] = dpm
        out["qp"] = qpm
    return out


def gfn2_orbitals(num, xyz, charge=0, etemp=300.0, opts=None):
    """Oracle single point, returning the converged orbitals: dict(stat,
    energy, eps[n] ascending, focc[n] (electrons per orbital), coeff[n, n]
    (column k = orbital k), S[n, n], aoat[n] (0-based atom))."""
    L = oracle()
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = gfn2_matrices_nao(num, xyz)
    e = C.c_double()
    eps, focc = np.zeros(n), np.zeros(n)
    coeff, S = np.zeros((n, n)), np.zeros((n, n))
    aoat = np.zeros(n, dtype=np.int32)
    o = opts if opts is not None else default_opts()
    st = L.orc_gfn2_orbitals(len(num), ptr(num, ip), ptr(xyz), int(charge), C.c_double(etemp),
                             C.byref(o), C.byref(e), ptr(eps), ptr(focc), ptr(coeff), ptr(S),
                             ptr(aoat, ip))
    return dict(stat=st, energy=e.value, eps=eps, focc=focc, coeff=coeff, S=S, aoat=aoat)


def mo_populations(coeff, S, aoat, nat):
    """write_qmo's MO populations (oracle restatement): qmo[nao, nat]."""
    n = len(aoat)
    coeff = np.ascontiguousarray(coeff, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    aoat = np.ascontiguousarray(aoat, dtype=np.int32)
    q = np.zeros((n, nat))
    oracle().orc_mo_populations(int(nat), int(n), ptr(coeff), ptr(S), ptr(aoat, ip), ptr(q))
    return q


def gfn2_matrices_nao(num, xyz):
    nao = C.c_int()
    oracle().orc_gfn2_matrices(len(num), ptr(np.ascontiguousarray(num, dtype=np.int32), ip),
                               ptr(np.ascontiguousarray(xyz, dtype=np.float64)), None, None, None,
                               None, C.byref(nao))
    return nao.value


def gfn2_matrices(num, xyz):
    L = oracle()
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nao = C.c_int()
    L.orc_gfn2_matrices(len(num), ptr(num, ip), ptr(xyz), None, None, None, None, C.byref(nao))
    n = nao.value
    s = np.zeros((n, n)); h0 = np.zeros((n, n)); d = np.zeros((3, n, n)); qd = np.zeros((6, n, n))
    L.orc_gfn2_matrices(len(num), ptr(num, ip), ptr(xyz), ptr(s), ptr(h0), ptr(d), ptr(qd),
                        C.byref(nao))
    return s, h0, d, qd


def fragment_structure(lib, prefix, oz, xyz, rcut=3.0, at1=1, at2=0):
    oz = np.ascontiguousarray(oz, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    frag = np.zeros(len(oz), dtype=np.int32)
    getattr(lib, prefix + "fragment_structure")(len(oz), ptr(oz, ip), ptr(xyz), C.c_double(rcut),
                                                at1, at2, ptr(frag, ip))
    return frag


def fragmass(lib, prefix, iat, lst, mass, imass=None):
    iat = np.ascontiguousarray(iat, dtype=np.int32)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    imass = np.full(len(iat), -1, dtype=np.int32) if imass is None else \
        np.ascontiguousarray(imass, dtype=np.int32)
    nfrag = C.c_int32()
    fragm = np.zeros(10)
    fragat = np.zeros((10, 200), dtype=np.int32)
    getattr(lib, prefix + "fragmass")(len(iat), ptr(iat, ip), ptr(lst, ip), ptr(mass),
                                      ptr(imass, ip), C.byref(nfrag), ptr(fragm), ptr(fragat, ip))
    return nfrag.value, fragm, fragat


def ekinet(lib, prefix, velo, mass):
    velo = np.ascontiguousarray(velo, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    e = C.c_double(); t = C.c_double()
    getattr(lib, prefix + "ekinet")(len(mass), ptr(velo), ptr(mass), C.byref(e), C.byref(t))
    return e.value, t.value


def impactscale(lib, prefix, velo, mass, velof, eimp, ff, e0):
    v = np.array(velo, dtype=np.float64, order="C")
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    velof = np.ascontiguousarray(velof, dtype=np.float64)
    fn = getattr(lib, prefix + "impactscale")
    fn.restype = C.c_int
    k = fn(len(mass), ptr(v), ptr(mass), ptr(velof), C.c_double(eimp), C.c_double(ff),
           C.c_double(e0))
    return v, k


def md_run(traj, steps_limit, tstep=0.5 * 41.3413733365614, max_steps=2000, nfragexit=3,
           etempin=-1.0, opts=None, charge=1, ieetemp=0.0, ax=0.0, isec=1, egrad=None,
           trace_steps=0, cid=False, new_velo=0.0, count_average=0):
    """Oracle production MD (md.f90:285-682) with GFN2 forces for one
    trajectory dict (num, xyz, velo, mass, eimp, tadd[, velof, imass]);
    cid: the MD after a CID collision (method 3, icoll >= 1) with the ion's
    centre-of-mass speed new_velo (m/s)."""
    L = oracle()
    num = np.ascontiguousarray(traj["num"], dtype=np.int32)
    nat = len(num)
    imass = np.ascontiguousarray(traj.get("imass", np.full(nat, -1)), dtype=np.int32)
    mass = np.ascontiguousarray(traj["mass"], dtype=np.float64)
    velof = np.ascontiguousarray(traj.get("velof", np.ones(nat)), dtype=np.float64)
    xyz = np.array(traj["xyz"], dtype=np.float64, order="C")
    velo = np.array(traj["velo"], dtype=np.float64, order="C")
    grad = np.zeros((nat, 3))
    qat = np.zeros(nat)
    lst = np.zeros(nat, dtype=np.int32)
    st = MdState()
    st.nat, st.chrg, st.nfragexit, st.max_steps = nat, int(charge), nfragexit, max_steps
    st.isec = int(isec)
    st.tstep, st.tadd, st.eimp = tstep, float(traj["tadd"]), float(traj["eimp"])
    st.ax, st.ieetemp, st.etempin = float(ax), float(ieetemp), etempin
    st.cid, st.count_average, st.new_velo = int(bool(cid)), int(count_average), float(new_velo)
    o = opts if opts is not None else default_opts()
    # egrad: optional Python callback (ctx, nat, iat, xyz, chrg, etemp, e, g, q)
    # wrapped as EGRAD_FN, for tests that force an exit through the energy
    fn = egrad if egrad is not None else C.cast(L.orc_gfn2_md_egrad, EGRAD_FN)
    axyz = np.zeros((nat, 3))
    achrg = np.zeros(nat)
    # per-step records: nstep, etemp for the next egrad, epot, ekin, nfrag,
    # fragT(1), fragT(2), eerror
    trace = np.full((max(int(trace_steps), 1), 8), np.nan)
    xtrace = np.full((max(int(trace_steps), 1), nat, 3), np.nan)
    if trace_steps:
        L.orc_md_set_trace(ptr(trace), int(trace_steps))
        L.orc_md_set_xyz_trace(ptr(xtrace), int(trace_steps))
    L.orc_md_run_avg(C.byref(st), ptr(num, ip), ptr(imass, ip), ptr(mass), ptr(velof), ptr(xyz),
                     ptr(velo), ptr(grad), ptr(qat), ptr(lst, ip), fn,
                     C.cast(C.pointer(o), C.c_void_p), int(steps_limit), ptr(axyz), ptr(achrg))
    if trace_steps:
        L.orc_md_set_trace(None, 0)
        L.orc_md_set_xyz_trace(None, 0)
    return dict(state=st, xyz=xyz, velo=velo, grad=grad, qat=qat, list=lst, axyz=axyz,
                achrg=achrg, trace=trace[:st.nstep] if trace_steps else None,
                xyz_trace=xtrace[:st.nstep] if trace_steps else None)


def gfn2_egrad_batch(systems, xyz_list, etemp, opts=None, nthreads=None):
    """Oracle single points over many systems [(num, charge)] with host threads.
    Returns (energy[nsys], [grad], [qat], stat[nsys])."""
    L = oracle()
    nsys = len(systems)
    off = np.cumsum([0] + [len(n) for n, _ in systems]).astype(np.int32)
    num = np.ascontiguousarray(np.concatenate([n for n, _ in systems]), dtype=np.int32)
    chg = np.array([c for _, c in systems], dtype=np.int32)
    xyz = np.ascontiguousarray(np.concatenate(xyz_list), dtype=np.float64)
    et = np.ascontiguousarray(np.broadcast_to(np.asarray(etemp, dtype=np.float64), (nsys,)))
    e = np.zeros(nsys)
    g = np.zeros((off[-1], 3))
    q = np.zeros(off[-1])
    st = np.zeros(nsys, dtype=np.int32)
    o = opts if opts is not None else default_opts()
    nt = nthreads or max(1, min(os.cpu_count() or 1, 16))
    L.orc_gfn2_egrad_batch(nsys, ptr(off, ip), ptr(num, ip), ptr(xyz), ptr(chg, ip), ptr(et),
                           C.byref(o), ptr(e), ptr(g), ptr(q), ptr(st, ip), nt)
    return (e, [g[off[i]:off[i + 1]] for i in range(nsys)],
            [q[off[i]:off[i + 1]] for i in range(nsys)], st)


# ---- post-MD fragment analysis (oracle/postmd_oracle.c) ----
def _post_sigs(L):
    if getattr(L, "_post_sigs", False):
        return
    L.orc_eself.argtypes = [C.c_int, ip, dp]
    L.orc_nel.argtypes = [C.c_int, ip, C.c_int]
    L.orc_boltz.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
    L.orc_res_record.argtypes = [C.c_double] + [C.c_int] * 6 + [ip, ip, C.c_char_p]
    L.orc_analyse_ips.argtypes = [C.c_int, ip, dp, ip, C.c_int, C.c_int, C.c_void_p, dp,
                                  C.POINTER(C.c_int)]
    L.orc_manage_fragments.argtypes = ([C.c_int, ip, dp, dp, dp, dp, ip, C.c_double, C.c_int,
                                        C.c_int, C.c_int, C.c_double, C.c_void_p]
                                       + [C.POINTER(C.c_int), dp, C.POINTER(C.c_int), ip,
                                          C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp, dp,
                                          C.POINTER(C.c_int), C.c_char_p, C.c_char_p,
                                          C.POINTER(C.c_int)])
    L._post_sigs = True


def res_record(q, mchrg, itrj, icoll, isec, j, pairs):
    L = oracle()
    _post_sigs(L)
    z = np.array([p[0] for p in pairs] or [0], dtype=np.int32)
    n = np.array([p[1] for p in pairs] or [0], dtype=np.int32)
    buf = C.create_string_buffer(200)
    L.orc_res_record(q, mchrg, itrj, icoll, isec, j, len(pairs), ptr(z, ip), ptr(n, ip), buf)
    return buf.value.decode()


def boltz(ip_mat, temp):
    """ip_mat [nfrag, nchrg] (eV) -> populations, same shape."""
    L = oracle()
    _post_sigs(L)
    a = np.asfortranarray(ip_mat, dtype=np.float64)
    out = np.zeros_like(a, order="F")
    L.orc_boltz(a.shape[0], a.shape[1], temp, a.ctypes.data_as(dp), out.ctypes.data_as(dp))
    return out


def eself(num):
    L = oracle()
    _post_sigs(L)
    z = np.ascontiguousarray(num, dtype=np.int32)
    e = C.c_double()
    st = L.orc_eself(len(z), ptr(z, ip), C.byref(e))
    return e.value, st


def analyse_ips(num, axyz, lst, nfrag, mchrg, opts=None):
    """fragip [nfrag, |mchrg|] (eV), ipok, rc."""
    L = oracle()
    _post_sigs(L)
    num = np.ascontiguousarray(num, dtype=np.int32)
    x = np.ascontiguousarray(axyz, dtype=np.float64)
    l = np.ascontiguousarray(lst, dtype=np.int32)
    fip = np.zeros(max(nfrag, 1) * abs(mchrg))
    ok = C.c_int()
    o = opts if opts is not None else default_opts()
    rc = L.orc_analyse_ips(len(num), ptr(num, ip), ptr(x), ptr(l, ip), nfrag, mchrg,
                           C.cast(C.pointer(o), C.c_void_p), ptr(fip), C.byref(ok))
    return fip.reshape(abs(mchrg), max(nfrag, 1)).T.copy(), ok.value, rc


def manage_fragments(num, xyz, axyz, qat, mass, atlast, itrj, isec, maxsec, btf, mchrg,
                     chrgcont, nfrag_ok=1, imass=None, opts=None):
    L = oracle()
    _post_sigs(L)
    num = np.ascontiguousarray(num, dtype=np.int32)
    nat = len(num)
    im = np.ascontiguousarray(imass if imass is not None else np.full(nat, -1), dtype=np.int32)
    lst = np.zeros(nat, dtype=np.int32)
    mc, cc, ok = C.c_int(mchrg), C.c_double(chrgcont), C.c_int(nfrag_ok)
    nf, tc, ipok, nres = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    q3, fq, fip = np.zeros(10), np.zeros(10), np.zeros(30)
    asave = C.create_string_buffer(160)
    res = C.create_string_buffer(160 * 10)
    o = opts if opts is not None else default_opts()
    rc = L.orc_manage_fragments(nat, ptr(num, ip), ptr(np.ascontiguousarray(xyz, dtype=np.float64)),
                                ptr(np.ascontiguousarray(axyz, dtype=np.float64)),
                                ptr(np.ascontiguousarray(qat, dtype=np.float64)),
                                ptr(np.ascontiguousarray(mass, dtype=np.float64)), ptr(im, ip),
                                atlast, itrj, isec, maxsec, btf, C.cast(C.pointer(o), C.c_void_p),
                                C.byref(mc), C.byref(cc), C.byref(ok), ptr(lst, ip), C.byref(nf),
                                C.byref(tc), ptr(q3), ptr(fq), ptr(fip), C.byref(ipok), asave, res,
                                C.byref(nres))
    recs = [res.raw[160 * k:160 * (k + 1)].split(b"\0", 1)[0].decode() for k in range(nres.value)]
    return dict(rc=rc, mchrg=mc.value, chrgcont=cc.value, nfrag_ok=ok.value, list=lst,
                nfrag=nf.value, tcont=tc.value, fragchrg3=q3, fragchrg=fq, fragip=fip,
                ipok=ipok.value, asave=asave.value.decode(), res=recs)


# ---- fragmentation generations (oracle/ei_oracle.c) ----
class EiGen(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("isec", "nat", "nstep", "exit_reason", "mdok", "fragstate",
                                        "nfrag", "max_steps", "tcont", "mchrg", "nfrag_ok", "nres",
                                        "continued", "asave_written")]
                + [(n, C.c_double) for n in ("chrgcont", "ttime", "epot")]
                + [("atoms", C.c_int32 * 64), ("list", C.c_int32 * 64), ("asave", C.c_char * 160),
                   ("res", (C.c_char * 160) * 10)])


def ei_run(traj, nmax, tstep=0.5 * 41.3413733365614, nfragexit=3, maxsec=7, btf=1.0, itrj=1,
           charge=1, opts=None, maxgen=16):
    """Oracle EI generations (main.F90:2208-2345) of one trajectory dict;
    returns a list of generation dicts."""
    L = oracle()
    L.orc_ei_run.argtypes = [C.c_int, ip, ip, dp, dp, dp, dp, C.c_int, C.c_double, C.c_double,
                             C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                             C.c_void_p, C.c_int]
    num = np.ascontiguousarray(traj["num"], dtype=np.int32)
    nat = len(num)
    imass = np.ascontiguousarray(traj.get("imass", np.full(nat, -1)), dtype=np.int32)
    mass = np.ascontiguousarray(traj["mass"], dtype=np.float64)
    velof = np.ascontiguousarray(traj.get("velof", np.ones(nat)), dtype=np.float64)
    xyz = np.ascontiguousarray(traj["xyz"], dtype=np.float64)
    velo = np.ascontiguousarray(traj["velo"], dtype=np.float64)
    gens = (EiGen * maxgen)()
    o = opts if opts is not None else default_opts()
    ng = L.orc_ei_run(nat, ptr(num, ip), ptr(imass, ip), ptr(mass), ptr(velof), ptr(xyz), ptr(velo),
                      int(charge), float(traj["eimp"]), float(traj["tadd"]), tstep, int(nmax),
                      nfragexit, maxsec, btf, itrj, C.cast(C.pointer(o), C.c_void_p),
                      C.cast(gens, C.c_void_p), maxgen)
    return [ei_gen_dict(gens[k]) for k in range(ng)]


def ei_gen_dict(g):
    d = {f: getattr(g, f) for f, _ in g._fields_ if f not in ("atoms", "list", "asave", "res")}
    d["atoms"] = np.array(g.atoms[:g.nat], dtype=np.int32)
    d["list"] = np.array(g.list[:g.nat], dtype=np.int32)
    d["asave"] = g.asave.decode()
    d["res"] = [g.res[k].value.decode() for k in range(g.nres)]
    return d


def cid_run(system, charge=1, tstep=0.5 * 41.3413733365614, etemp=5000.0, ntot=15000, opts=None,
            steps_limit=0, collided=0, count_average=0):
    """The oracle's CID collision MD (cid_oracle.c) of one system from
    qcxms_amd.cid.cid_setup.  Returns dict(state, xyz, velo, grad, qat, list,
    axyz)."""
    L = oracle()
    num = np.ascontiguousarray(system["num"], dtype=np.int32)
    nuc0, nuc = len(num), int(system["nion"])
    xyz = np.ascontiguousarray(system["xyz"], dtype=np.float64).copy()
    velo = np.ascontiguousarray(system["velo"], dtype=np.float64).copy()
    mass = np.ascontiguousarray(system["mass"], dtype=np.float64)
    grad = np.zeros((nuc0, 3))
    q = np.zeros(nuc0)
    lst = np.zeros(nuc, dtype=np.int32)
    axyz = np.zeros((nuc, 3))
    st = CidState()
    st.nuc, st.nuc0, st.chrg, st.ntot = nuc, nuc0, int(charge), int(ntot)
    st.collided, st.count_average = int(collided), int(count_average)
    st.tstep, st.etemp, st.tinit, st.beta = tstep, etemp, float(system["tinit"]), float(system["beta"])
    st.velo_cm, st.ttime = 0.0, 0.0
    o = opts if opts is not None else default_opts()
    fn = C.cast(L.orc_gfn2_md_egrad, EGRAD_FN)
    L.orc_cid_run(C.byref(st), ptr(num, ip), ptr(mass), ptr(xyz), ptr(velo), ptr(grad), ptr(q),
                  ptr(lst, ip), ptr(axyz), fn, C.cast(C.pointer(o), C.c_void_p), int(steps_limit))
    return dict(state=st, xyz=xyz, velo=velo, grad=grad, qat=q, list=lst, axyz=axyz)
