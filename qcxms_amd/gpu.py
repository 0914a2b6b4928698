"""ctypes host layer over libqcxms_gpu.so.

Mirrors the reference's bridge get_xtb_egrad (src/tblite.f90:65-175): same
argument meaning, same status codes (0 ok, 5 unknown method, -1 failure).
There is no CPU fallback: a missing library raises LibraryMissing.
"""
import ctypes as C
import os

import numpy as np

GFN1_XTB, GFN2_XTB, IPEA1_XTB = 1, 2, 11  # tblite.f90:34-40
STAT_OK, STAT_UNKNOWN_METHOD, STAT_FATAL, STAT_INVALID = 0, 5, -1, -2

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libqcxms_gpu.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lib = None


class LibraryMissing(RuntimeError):
    pass


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def library():
    """Load libqcxms_gpu.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise LibraryMissing(f"{library_path} is not built; run __graft_entry__.build()")
        L = C.CDLL(library_path)
        L.qcxms_gpu_last_error.restype = C.c_char_p
        L.qcxms_batch_create.restype = C.c_void_p
        L.qcxms_batch_create.argtypes = [C.c_int32, _ip, _ip, _ip, C.c_int32]
        L.qcxms_batch_destroy.argtypes = [C.c_void_p]
        L.qcxms_batch_set_scc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32,
                                          C.c_double, C.c_double]
        L.qcxms_batch_set_warm_start.argtypes = [C.c_void_p, C.c_int32]
        L.qcxms_set_tuning.argtypes = [C.c_int32, _ip, _ip, _ip, C.c_int32, C.c_int32]
        L.qcxms_get_tuning.argtypes = [C.c_int32, _ip, _ip, _ip, _ip, _ip, _ip]
        L.qcxms_batch_set_stored_derivatives.argtypes = [C.c_void_p, C.c_int32]
        L.qcxms_batch_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.qcxms_batch_set_profiling.argtypes = [C.c_void_p, C.c_int32]
        L.qcxms_batch_kernel_time.argtypes = [C.c_void_p, C.c_char_p, _dp, _ip]
        L.qcxms_batch_egrad.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _ip]
        L.qcxms_batch_info.argtypes = [C.c_void_p, _ip, _dp]
        L.qcxms_batch_multipoles.argtypes = [C.c_void_p, _dp, _dp]
        L.qcxms_batch_orbitals.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, _dp, _dp, _dp,
                                           _dp, _ip]
        L.qcxms_batch_mo_spectrum.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, _ip, _dp,
                                              _dp, _dp]
        L.qcxms_get_xtb_mo_spectrum.argtypes = [C.c_int32, _ip, _ip, _dp, _dp, _dp]
        L.qcxms_batch_scc_iterations_total.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.qcxms_get_xtb_egrad.argtypes = [C.c_int32, _ip, _dp, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_double, _dp, _dp, _dp, _ip]
        L.qcxms_egrad_batch.argtypes = [C.c_int32, _ip, _ip, _dp, _ip, _ip, C.c_int32, _dp, _dp,
                                        _dp, _dp, _ip]
        _lib = L
    return _lib


def last_error():
    return library().qcxms_gpu_last_error().decode()


def get_xtb_egrad(num, xyz, charge, multiplicity, method=GFN2_XTB, etemp=300.0):
    """Energy (Eh), gradient (Eh/bohr, [nat,3]), Mulliken charges and status
    for one structure (xyz in Bohr, [nat,3]) -- the reference's
    get_xtb_egrad with a fresh wavefunction per call."""
    L = library()
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nat = len(num)
    qat = np.zeros(nat)
    grad = np.zeros((nat, 3))
    e = C.c_double()
    stat = C.c_int32()
    L.qcxms_get_xtb_egrad(nat, _p(num, _ip), _p(xyz), int(charge), int(multiplicity),
                          int(method), float(etemp), _p(qat), C.byref(e), _p(grad),
                          C.byref(stat))
    return qat, e.value, grad, stat.value


def set_tuning(table, mixed=(-1, 0)):
    """SCC launch plan for batches created afterwards: table = [(max_nao,
    eig, lds), ...] by increasing max_nao ([] for one launch per batch with
    its own options); mixed = (eig, lds) for batches spanning several bins,
    eig = -1 for one launch per run of one bin."""
    n = len(table)
    a = np.array([t[0] for t in table] or [0], dtype=np.int32)
    e = np.array([t[1] for t in table] or [0], dtype=np.int32)
    l = np.array([t[2] for t in table] or [0], dtype=np.int32)
    if library().qcxms_set_tuning(n, _p(a, _ip), _p(e, _ip), _p(l, _ip), int(mixed[0]), int(mixed[1])):
        raise RuntimeError("qcxms_set_tuning failed: " + last_error())


def get_tuning():
    """(table, mixed) as set_tuning takes them."""
    cap = 32
    n, me, ml = C.c_int32(), C.c_int32(), C.c_int32()
    a, e, l = (np.zeros(cap, dtype=np.int32) for _ in range(3))
    if library().qcxms_get_tuning(cap, C.byref(n), _p(a, _ip), _p(e, _ip), _p(l, _ip), C.byref(me),
                                  C.byref(ml)):
        raise RuntimeError("qcxms_get_tuning failed: " + last_error())
    return [(int(a[k]), int(e[k]), int(l[k])) for k in range(n.value)], (me.value, ml.value)


def _mo_spectrum(call, nat):
    cap = 256
    n, h = C.c_int32(), C.c_int32()
    emo, focc, qmo = np.zeros(cap), np.zeros(cap), np.zeros(cap * nat)
    if call(cap, C.byref(n), C.byref(h), _p(emo), _p(focc), _p(qmo)):
        raise RuntimeError("MO spectrum failed: " + last_error())
    k = n.value
    return emo[:k].copy(), focc[:k].copy(), qmo[:k * nat].reshape(k, nat).copy(), h.value


def get_xtb_mo_spectrum(nat):
    """MO spectrum of the structure of the last get_xtb_egrad call:
    (emo [nao] Eh ascending, focc [nao] electrons per orbital, qmo [nao, nat]
    Mulliken MO populations normalised per orbital, ihomo 1-based)."""
    return _mo_spectrum(library().qcxms_get_xtb_mo_spectrum, nat)


class Batch:
    """Device-resident batch of systems evaluated together (one workgroup per
    system).  systems: list of (num, charge)."""

    def __init__(self, systems, method=GFN2_XTB):
        L = library()
        nums = [np.asarray(s[0], dtype=np.int32) for s in systems]
        self.offsets = np.cumsum([0] + [len(n) for n in nums]).astype(np.int32)
        self.num = np.concatenate(nums).astype(np.int32)
        self.charge = np.array([int(s[1]) for s in systems], dtype=np.int32)
        self.nsys = len(systems)
        self.natoms = int(self.offsets[-1])
        h = L.qcxms_batch_create(self.nsys, _p(self.offsets, _ip), _p(self.num, _ip),
                                 _p(self.charge, _ip), int(method))
        if not h:
            raise RuntimeError("qcxms_batch_create failed: " + last_error())
        self._h = C.c_void_p(h)

    def close(self):
        if getattr(self, "_h", None):
            library().qcxms_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scc(self, econv=1e-9, pconv=1e-8, maxiter=100, damping=0.4, jacobi_tol=1e-11):
        st = library().qcxms_batch_set_scc(self._h, econv, pconv, maxiter, damping, jacobi_tol)
        if st:
            raise RuntimeError(last_error())

    def set_warm_start(self, warm):
        library().qcxms_batch_set_warm_start(self._h, int(bool(warm)))

    def set_stored_derivatives(self, on):
        """on: the gradient contracts stored derivative blocks (more device
        memory); off (default): it recomputes them."""
        if library().qcxms_batch_set_stored_derivatives(self._h, int(bool(on))):
            raise RuntimeError("qcxms_batch_set_stored_derivatives failed: " + last_error())

    def device_bytes(self):
        n = C.c_uint64()
        library().qcxms_batch_device_bytes(self._h, C.byref(n))
        return n.value

    def set_profiling(self, on):
        library().qcxms_batch_set_profiling(self._h, int(bool(on)))

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int32()
        if library().qcxms_batch_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)):
            raise RuntimeError(last_error())
        return ms.value, n.value

    def egrad(self, xyz, etemp=300.0, gradient=True):
        """xyz: [natoms,3] Bohr (all systems concatenated).  Returns
        (energy[nsys], grad[natoms,3] or None, qat[natoms], stat[nsys])."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.natoms, 3)
        et = np.ascontiguousarray(np.broadcast_to(np.asarray(etemp, dtype=np.float64),
                                                  (self.nsys,)))
        e = np.zeros(self.nsys)
        g = np.zeros((self.natoms, 3)) if gradient else None
        q = np.zeros(self.natoms)
        st = np.zeros(self.nsys, dtype=np.int32)
        rc = library().qcxms_batch_egrad(self._h, _p(xyz), _p(et), _p(e), _p(g), _p(q),
                                         _p(st, _ip))
        if rc not in (STAT_OK, STAT_FATAL):
            raise RuntimeError(f"qcxms_batch_egrad failed ({rc}): " + last_error())
        return e, g, q, st

    def info(self):
        it = np.zeros(self.nsys, dtype=np.int32)
        parts = np.zeros((self.nsys, 10))
        library().qcxms_batch_info(self._h, _p(it, _ip), _p(parts))
        return it, parts

    def multipoles(self):
        d = np.zeros((self.natoms, 3))
        q = np.zeros((self.natoms, 6))
        library().qcxms_batch_multipoles(self._h, _p(d), _p(q))
        return d, q

    def orbitals(self, sys=0):
        """(emo [nao] Eh ascending, focc [nao], coeff [nao(AO), nao(MO)],
        overlap [nao, nao], aoat [nao] 1-based) of system sys after the last
        evaluation."""
        cap = 256
        n = C.c_int32()
        emo, focc = np.zeros(cap), np.zeros(cap)
        coeff, ovl = np.zeros(cap * cap), np.zeros(cap * cap)
        aoat = np.zeros(cap, dtype=np.int32)
        if library().qcxms_batch_orbitals(self._h, int(sys), cap, C.byref(n), _p(emo), _p(focc),
                                          _p(coeff), _p(ovl), _p(aoat, _ip)):
            raise RuntimeError("qcxms_batch_orbitals failed: " + last_error())
        k = n.value
        return (emo[:k].copy(), focc[:k].copy(), coeff[:k * k].reshape(k, k).T.copy(),
                ovl[:k * k].reshape(k, k).T.copy(), aoat[:k].copy())

    def mo_spectrum(self, sys=0):
        """(emo [nao] Eh ascending, focc [nao], qmo [nao, nat], ihomo) of system
        sys after the last evaluation, computed on the GPU (write_qmo's MO
        populations, mo_energ.f90:36-52)."""
        nat = int(self.offsets[sys + 1] - self.offsets[sys])
        return _mo_spectrum(lambda *a: library().qcxms_batch_mo_spectrum(self._h, int(sys), *a), nat)


def egrad_batch(systems, xyz_list, etemp=300.0, method=GFN2_XTB):
    """One-shot batched single points: systems [(num, charge)], xyz_list of
    [nat,3] arrays.  Returns lists of energies, gradients, charges, stats."""
    b = Batch(systems, method)
    try:
        e, g, q, st = b.egrad(np.concatenate(xyz_list), etemp)
    finally:
        b.close()
    off = b.offsets
    return (e, [g[off[i]:off[i + 1]] for i in range(len(systems))],
            [q[off[i]:off[i + 1]] for i in range(len(systems))], st)


def egrad_many(systems, xyz_list, etemp=300.0, method=GFN2_XTB):
    """Independent single points of ragged systems through qcxms_egrad_batch
    (one launch set, systems ordered largest basis first inside the library,
    default SCC thresholds): the per-trajectory egrad calls of
    src/iniqm.f90:641-655 for many structures at once.  systems:
    [(num, charge[, multiplicity])].  Returns (energy[nsys], [grad], [qat],
    stat[nsys])."""
    L = library()
    nsys = len(systems)
    off = np.cumsum([0] + [len(s[0]) for s in systems]).astype(np.int32)
    num = np.ascontiguousarray(np.concatenate([s[0] for s in systems]), dtype=np.int32)
    chg = np.array([s[1] for s in systems], dtype=np.int32)
    mult = np.array([s[2] if len(s) > 2 else 1 for s in systems], dtype=np.int32)
    xyz = np.ascontiguousarray(np.concatenate(xyz_list), dtype=np.float64)
    et = np.ascontiguousarray(np.broadcast_to(np.asarray(etemp, dtype=np.float64), (nsys,)))
    e = np.zeros(nsys)
    g = np.zeros((off[-1], 3))
    q = np.zeros(off[-1])
    st = np.zeros(nsys, dtype=np.int32)
    rc = L.qcxms_egrad_batch(nsys, _p(off, _ip), _p(num, _ip), _p(xyz), _p(chg, _ip),
                             _p(mult, _ip), int(method), _p(et), _p(q), _p(e), _p(g),
                             _p(st, _ip))
    if rc not in (STAT_OK, STAT_FATAL, STAT_UNKNOWN_METHOD):
        raise RuntimeError(f"qcxms_egrad_batch failed ({rc}): " + last_error())
    return (e, [g[off[i]:off[i + 1]] for i in range(nsys)],
            [q[off[i]:off[i + 1]] for i in range(nsys)], st)


# ---------------------------------------------------------------- MD ensemble
class MdParams(C.Structure):
    _fields_ = [("tstep", C.c_double), ("etempin", C.c_double), ("ax", C.c_double),
                ("ieetemp", C.c_double), ("nfragexit", C.c_int32), ("max_steps", C.c_int32),
                ("use_graph", C.c_int32), ("extrapolate", C.c_int32), ("econv", C.c_double),
                ("pconv", C.c_double), ("broyden_keep", C.c_int32), ("fused", C.c_int32),
                ("eigensolver", C.c_int32)]


class MdResult(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nstep", "nfrag", "fragstate", "mdok", "exit_reason",
                                          "running", "scc_iterations", "pad")]
                + [(n, C.c_double) for n in ("tav", "epav", "ekav", "ttime", "dtime", "atlast",
                                             "epot", "ekin", "eerror", "etemp")]
                + [("fragm", C.c_double * 10), ("fragT", C.c_double * 10)]
                + [("check_fragmented", C.c_int32), ("count_average", C.c_int32), ("new_velo", C.c_double)])


EXIT_REASONS = {0: "running", 1: "scf/checkqc failure", 2: "energy drift", 3: "more than 6 fragments",
                4: "nfrag > nfragexit", 5: "nothing happens", 6: "nfrag = nfragexit, more steps done",
                7: "max steps", 8: "impactscale failed"}


def _md_sigs(L):
    if getattr(L, "_md_sigs", False):
        return
    L.qcxms_md_create.restype = C.c_void_p
    L.qcxms_md_create.argtypes = [C.c_int32, _ip, _ip, _ip, _dp, _ip, _dp, _dp, _dp, _dp, _dp,
                                  C.POINTER(MdParams)]
    L.qcxms_md_destroy.argtypes = [C.c_void_p]
    L.qcxms_md_run.argtypes = [C.c_void_p, C.c_int32, _ip]
    L.qcxms_md_run_group.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, _ip]
    L.qcxms_md_read.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip, C.POINTER(MdResult)]
    L.qcxms_md_read_averages.argtypes = [C.c_void_p, _dp, _dp]
    L.qcxms_md_batch.restype = C.c_void_p
    L.qcxms_md_batch.argtypes = [C.c_void_p]
    L.qcxms_md_set_generation.argtypes = [C.c_void_p, _ip]
    L.qcxms_md_set_cid.argtypes = [C.c_void_p, _ip, _dp, _ip]
    L.qcxms_fragments_batch.argtypes = [C.c_int32, _ip, _ip, _dp, _dp, _ip, _ip, _dp]
    L.qcxms_leapfrog_batch.argtypes = [C.c_int32, _ip, _dp, _dp, C.c_double, _dp, _dp, _dp]
    L.qcxms_impactscale_batch.argtypes = [C.c_int32, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
    L._md_sigs = True


class MdEnsemble:
    """Batched EI production MD (src/md.f90:285-682) on the GPU.

    trajectories: list of dicts with num, xyz [nat,3] (Bohr), velo [nat,3]
    (au), mass [nat] (m_e), eimp (Eh), tadd (au); optional charge, imass,
    velof, isec (fragmentation generation, default 1), and for the MD after
    a CID collision (md.f90 with method 3, icoll >= 1; qcxms_md_set_cid)
    cid=True with new_velo (the ion's centre-of-mass speed, m/s) and
    count_average (0/1).  lanes > 1 splits the trajectories into that many contiguous
    groups, each an ensemble on its own stream, advanced together
    (qcxms_md_run_group) so that the device overlaps their kernels.  fused
    runs whole MD iterations per trajectory in one persistent launch.
    eigensolver: 1 refines the SCC eigenvectors by simultaneous rotations on
    the FP64 matrix cores before Jacobi, 2 Jacobi only, 0 the library
    default (1)."""

    def __init__(self, trajectories, tstep=0.5 * 41.3413733365614, max_steps=2000,
                 nfragexit=3, etempin=-1.0, ax=0.0, ieetemp=0.0, use_graph=True,
                 econv=0.0, pconv=0.0, extrapolate=0, broyden_keep=0, lanes=1, fused=False,
                 eigensolver=0):
        L = library()
        _md_sigs(L)
        nums = [np.asarray(t["num"], dtype=np.int32) for t in trajectories]
        self.offsets = np.cumsum([0] + [len(n) for n in nums]).astype(np.int32)
        self.ntraj = len(trajectories)
        self.natoms = int(self.offsets[-1])
        self.num = np.concatenate(nums).astype(np.int32)
        lanes = max(1, min(int(lanes), self.ntraj))
        bounds = [round(k * self.ntraj / lanes) for k in range(lanes + 1)]
        self._lanes = [(bounds[k], bounds[k + 1]) for k in range(lanes)]
        p = MdParams(tstep, etempin, ax, ieetemp, nfragexit, max_steps, int(bool(use_graph)),
                     int(extrapolate), econv, pconv, int(broyden_keep), int(bool(fused)),
                     int(eigensolver))
        self._hs = []
        for t0, t1 in self._lanes:
            part = trajectories[t0:t1]
            cat = lambda key, dt, default=None: np.ascontiguousarray(np.concatenate(
                [np.asarray(t.get(key, default(t) if default else None), dtype=dt).reshape(-1)
                 for t in part]))
            off = (self.offsets[t0:t1 + 1] - self.offsets[t0]).astype(np.int32)
            num = np.ascontiguousarray(self.num[self.offsets[t0]:self.offsets[t1]])
            charge = np.array([int(t.get("charge", 1)) for t in part], dtype=np.int32)
            mass = cat("mass", np.float64)
            imass = cat("imass", np.int32, lambda t: np.full(len(t["num"]), -1))
            xyz = cat("xyz", np.float64)
            velo = cat("velo", np.float64)
            velof = cat("velof", np.float64, lambda t: np.ones(len(t["num"])))
            eimp = np.array([float(t["eimp"]) for t in part])
            tadd = np.array([float(t["tadd"]) for t in part])
            h = L.qcxms_md_create(len(part), _p(off, _ip), _p(num, _ip), _p(charge, _ip), _p(mass),
                                  _p(imass, _ip), _p(xyz), _p(velo), _p(velof), _p(eimp), _p(tadd),
                                  C.byref(p))
            if not h:
                self.close()
                raise RuntimeError("qcxms_md_create failed: " + last_error())
            self._hs.append(C.c_void_p(h))
            isec = np.array([int(t.get("isec", 1)) for t in part], dtype=np.int32)
            if np.any(isec != 1) and L.qcxms_md_set_generation(self._hs[-1], _p(isec, _ip)):
                self.close()
                raise RuntimeError("qcxms_md_set_generation failed: " + last_error())
            cid = np.array([int(bool(t.get("cid", False))) for t in part], dtype=np.int32)
            if cid.any():
                vcm = np.array([float(t.get("new_velo", 0.0)) for t in part])
                cavg = np.array([int(t.get("count_average", 0)) for t in part], dtype=np.int32)
                if L.qcxms_md_set_cid(self._hs[-1], _p(cid, _ip), _p(vcm), _p(cavg, _ip)):
                    self.close()
                    raise RuntimeError("qcxms_md_set_cid failed: " + last_error())
        self._h = self._hs[0]

    def close(self):
        for h in getattr(self, "_hs", []):
            if h:
                library().qcxms_md_destroy(h)
        self._hs = []
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        n = len(self._hs)
        arr = (C.c_void_p * n)(*[h.value for h in self._hs])
        running = np.zeros(n, dtype=np.int32)
        if library().qcxms_md_run_group(arr, n, int(nsteps), _p(running, _ip)):
            raise RuntimeError("qcxms_md_run failed: " + last_error())
        return int(running.sum())

    def _atoms(self, k):
        t0, t1 = self._lanes[k]
        return slice(int(self.offsets[t0]), int(self.offsets[t1]))

    def read(self):
        xyz = np.zeros((self.natoms, 3)); velo = np.zeros((self.natoms, 3))
        grad = np.zeros((self.natoms, 3)); qat = np.zeros(self.natoms)
        lst = np.zeros(self.natoms, dtype=np.int32)
        res = []
        for k, h in enumerate(self._hs):
            a = self._atoms(k)
            t0, t1 = self._lanes[k]
            r = (MdResult * (t1 - t0))()
            parts = [np.zeros((a.stop - a.start, 3)) for _ in range(3)]
            q = np.zeros(a.stop - a.start)
            l = np.zeros(a.stop - a.start, dtype=np.int32)
            if library().qcxms_md_read(h, _p(parts[0]), _p(parts[1]), _p(parts[2]), _p(q),
                                       _p(l, _ip), r):
                raise RuntimeError("qcxms_md_read failed: " + last_error())
            xyz[a], velo[a], grad[a], qat[a], lst[a] = parts[0], parts[1], parts[2], q, l
            res.extend(list(r))
        return dict(xyz=xyz, velo=velo, grad=grad, qat=qat, list=lst, results=res)

    def read_averages(self):
        """(axyz [natoms,3], achrg [natoms]): the average geometry and
        Mulliken charges over the last avdump steps (md.f90:691-698)."""
        axyz = np.zeros((self.natoms, 3)); achrg = np.zeros(self.natoms)
        for k, h in enumerate(self._hs):
            a = self._atoms(k)
            x = np.zeros((a.stop - a.start, 3)); q = np.zeros(a.stop - a.start)
            if library().qcxms_md_read_averages(h, _p(x), _p(q)):
                raise RuntimeError("qcxms_md_read_averages failed: " + last_error())
            axyz[a], achrg[a] = x, q
        return axyz, achrg

    def batch_handle(self, lane=0):
        return library().qcxms_md_batch(self._hs[lane])

    def set_profiling(self, on):
        for k in range(len(self._hs)):
            library().qcxms_batch_set_profiling(C.c_void_p(self.batch_handle(k)), int(bool(on)))

    def kernel_time(self, name):
        """(ms, launches) summed over the lanes."""
        tot, cnt = 0.0, 0
        for k in range(len(self._hs)):
            ms = C.c_double()
            n = C.c_int32()
            library().qcxms_batch_kernel_time(C.c_void_p(self.batch_handle(k)), name.encode(),
                                              C.byref(ms), C.byref(n))
            tot += ms.value
            cnt += n.value
        return tot, cnt

    def scc_iterations_total(self):
        out = []
        for k in range(len(self._hs)):
            t0, t1 = self._lanes[k]
            o = np.zeros(t1 - t0, dtype=np.uint64)
            library().qcxms_batch_scc_iterations_total(C.c_void_p(self.batch_handle(k)),
                                                       o.ctypes.data_as(C.POINTER(C.c_uint64)))
            out.append(o)
        return np.concatenate(out)


class EiGeneration(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("traj", "isec", "nat", "nstep", "exit_reason", "mdok",
                                          "fragstate", "nfrag", "max_steps", "tcont", "mchrg",
                                          "nfrag_ok", "nres", "continued", "asave_written", "pad")]
                + [(n, C.c_double) for n in ("chrgcont", "ttime", "epot")]
                + [("atoms", C.c_int32 * 64), ("list", C.c_int32 * 64), ("asave", C.c_char * 128),
                   ("res", (C.c_char * 128) * 10)])

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_
             if f not in ("atoms", "list", "asave", "res", "pad")}
        d["atoms"] = np.array(self.atoms[:self.nat], dtype=np.int32)
        d["list"] = np.array(self.list[:self.nat], dtype=np.int32)
        d["asave"] = self.asave.decode()
        d["res"] = [self.res[k].value.decode() for k in range(self.nres)]
        return d


class EiEnsemble:
    """Fragmentation generations of many EI trajectories (src/main.F90:
    2208-2345) on the GPU: each trajectory's MD, then manage_fragments, then
    the continued fragment as the next generation, with the ensemble laid out
    again on the device every `repack` steps in which trajectories ended.
    Trajectory dicts and MD options as MdEnsemble."""

    def __init__(self, trajectories, tstep=0.5 * 41.3413733365614, max_steps=2000, nfragexit=3,
                 etempin=-1.0, ax=0.0, ieetemp=0.0, use_graph=True, econv=0.0, pconv=0.0,
                 extrapolate=0, broyden_keep=0, fused=False, eigensolver=0, maxsec=7, btf=1.0,
                 repack=100):
        L = library()
        _md_sigs(L)
        if not getattr(L, "_ei_sigs", False):
            L.qcxms_ei_create.restype = C.c_void_p
            L.qcxms_ei_create.argtypes = [C.c_int32, _ip, _ip, _ip, _dp, _ip, _dp, _dp, _dp, _dp, _dp,
                                          C.POINTER(MdParams), C.c_int32, C.c_double, C.c_int32]
            L.qcxms_ei_destroy.argtypes = [C.c_void_p]
            L.qcxms_ei_run.argtypes = [C.c_void_p, C.c_int32, _ip]
            L.qcxms_ei_generations.argtypes = [C.c_void_p, C.c_int32, C.POINTER(EiGeneration), _ip]
            L.qcxms_ei_repacks.argtypes = [C.c_void_p, _ip]
            L.qcxms_ei_steps.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
            L.qcxms_ei_md.restype = C.c_void_p
            L.qcxms_ei_md.argtypes = [C.c_void_p]
            L._ei_sigs = True
        t = trajectories
        nums = [np.asarray(x["num"], dtype=np.int32) for x in t]
        off = np.cumsum([0] + [len(n) for n in nums]).astype(np.int32)
        cat = lambda key, dt, default=None: np.ascontiguousarray(np.concatenate(
            [np.asarray(x.get(key, default(x) if default else None), dtype=dt).reshape(-1) for x in t]))
        num = np.concatenate(nums).astype(np.int32)
        charge = np.array([int(x.get("charge", 1)) for x in t], dtype=np.int32)
        p = MdParams(tstep, etempin, ax, ieetemp, nfragexit, max_steps, int(bool(use_graph)),
                     int(extrapolate), econv, pconv, int(broyden_keep), int(bool(fused)),
                     int(eigensolver))
        h = L.qcxms_ei_create(len(t), _p(off, _ip), _p(num, _ip), _p(charge, _ip),
                              _p(cat("mass", np.float64)),
                              _p(cat("imass", np.int32, lambda x: np.full(len(x["num"]), -1)), _ip),
                              _p(cat("xyz", np.float64)), _p(cat("velo", np.float64)),
                              _p(cat("velof", np.float64, lambda x: np.ones(len(x["num"])))),
                              _p(np.array([float(x["eimp"]) for x in t])),
                              _p(np.array([float(x["tadd"]) for x in t])), C.byref(p), int(maxsec),
                              float(btf), int(repack))
        if not h:
            raise RuntimeError("qcxms_ei_create failed: " + last_error())
        self._h = C.c_void_p(h)
        self.ntraj = len(t)

    def close(self):
        if getattr(self, "_h", None):
            library().qcxms_ei_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        running = C.c_int32()
        if library().qcxms_ei_run(self._h, int(nsteps), C.byref(running)):
            raise RuntimeError("qcxms_ei_run failed: " + last_error())
        return running.value

    def generations(self):
        n = C.c_int32()
        library().qcxms_ei_generations(self._h, 0, None, C.byref(n))
        buf = (EiGeneration * max(n.value, 1))()
        library().qcxms_ei_generations(self._h, n.value, buf, C.byref(n))
        return [buf[k].as_dict() for k in range(n.value)]

    def repacks(self):
        n = C.c_int32()
        library().qcxms_ei_repacks(self._h, C.byref(n))
        return n.value

    def steps(self):
        """Trajectory-steps executed so far over all generations."""
        n = C.c_uint64()
        if library().qcxms_ei_steps(self._h, C.byref(n)):
            raise RuntimeError("qcxms_ei_steps failed: " + last_error())
        return n.value

    def batch_handle(self):
        """Batch of the current layout (timing hooks)."""
        return library().qcxms_md_batch(library().qcxms_ei_md(self._h))


def fragments_batch(nums, xyzs, masses):
    """fragment_structure + fragmass masses for each system (<= 64 atoms)."""
    L = library()
    _md_sigs(L)
    off = np.cumsum([0] + [len(n) for n in nums]).astype(np.int32)
    num = np.ascontiguousarray(np.concatenate(nums), dtype=np.int32)
    xyz = np.ascontiguousarray(np.concatenate(xyzs), dtype=np.float64)
    mass = np.ascontiguousarray(np.concatenate(masses), dtype=np.float64)
    lst = np.zeros(len(num), dtype=np.int32)
    nf = np.zeros(len(nums), dtype=np.int32)
    fm = np.zeros((len(nums), 10))
    if L.qcxms_fragments_batch(len(nums), _p(off, _ip), _p(num, _ip), _p(xyz), _p(mass),
                               _p(lst, _ip), _p(nf, _ip), _p(fm)):
        raise RuntimeError(last_error())
    return [lst[off[i]:off[i + 1]] for i in range(len(nums))], nf, fm


def leapfrog_batch(grads, masses, tstep, xyzs, velos):
    L = library()
    _md_sigs(L)
    off = np.cumsum([0] + [len(m) for m in masses]).astype(np.int32)
    g = np.ascontiguousarray(np.concatenate(grads), dtype=np.float64)
    m = np.ascontiguousarray(np.concatenate(masses), dtype=np.float64)
    x = np.ascontiguousarray(np.concatenate(xyzs), dtype=np.float64)
    v = np.ascontiguousarray(np.concatenate(velos), dtype=np.float64)
    ek = np.zeros(len(masses))
    if L.qcxms_leapfrog_batch(len(masses), _p(off, _ip), _p(g), _p(m), float(tstep), _p(x), _p(v),
                              _p(ek)):
        raise RuntimeError(last_error())
    return x, v, ek


def impactscale_batch(velos, masses, velofs, eimp, ff, e0):
    L = library()
    _md_sigs(L)
    off = np.cumsum([0] + [len(m) for m in masses]).astype(np.int32)
    v = np.ascontiguousarray(np.concatenate(velos), dtype=np.float64)
    m = np.ascontiguousarray(np.concatenate(masses), dtype=np.float64)
    f = np.ascontiguousarray(np.concatenate(velofs), dtype=np.float64)
    ei = np.ascontiguousarray(eimp, dtype=np.float64)
    ffa = np.ascontiguousarray(ff, dtype=np.float64)
    e0a = np.ascontiguousarray(e0, dtype=np.float64)
    it = np.zeros(len(masses), dtype=np.int32)
    if L.qcxms_impactscale_batch(len(masses), _p(off, _ip), _p(v), _p(m), _p(f), _p(ei), _p(ffa),
                                 _p(e0a), _p(it, _ip)):
        raise RuntimeError(last_error())
    return v, it


class FragmentReport(C.Structure):
    """qcxms_fragment_report: one trajectory's manage_fragments state."""
    _fields_ = ([("mchrg", C.c_int32), ("nfrag_ok", C.c_int32), ("chrgcont", C.c_double)]
                + [(n, C.c_int32) for n in ("nfrag", "tcont", "ipok", "nres", "stat", "icoll")]
                + [("fragchrg3", C.c_double * 10), ("fragchrg", C.c_double * 10),
                   ("fragm", C.c_double * 10), ("fragip", C.c_double * 30),
                   ("asave", C.c_char * 128), ("res", (C.c_char * 128) * 10)])

    def records(self):
        return [self.res[k].value.decode() for k in range(self.nres)]


def _post_sigs(L):
    if getattr(L, "_post_sigs", False):
        return
    L.qcxms_fragment_ips_batch.argtypes = [C.c_int32, _ip, _ip, _dp, _ip, _ip, _ip, _dp, _ip, _ip]
    L.qcxms_manage_fragments_batch.argtypes = [C.c_int32, _ip, _ip, _dp, _dp, _dp, _dp, _ip, _dp,
                                               _ip, C.c_int32, C.c_int32, C.c_double, _ip,
                                               C.POINTER(FragmentReport)]
    L._post_sigs = True


def fragment_ips_batch(nums, axyzs, lists, mchrgs):
    """analyse (src/analyse.f90:14-448) for many trajectories: returns
    (fragip [ntraj, 10, 3] eV with [t, i-1, j-1] = IP of fragment i for
    charge j, ipok [ntraj], stat [ntraj])."""
    L = library()
    _post_sigs(L)
    n = len(nums)
    off = np.cumsum([0] + [len(x) for x in nums]).astype(np.int32)
    num = np.ascontiguousarray(np.concatenate(nums), dtype=np.int32)
    x = np.ascontiguousarray(np.concatenate(axyzs), dtype=np.float64)
    lst = np.ascontiguousarray(np.concatenate(lists), dtype=np.int32)
    nf = np.array([int(np.max(l)) for l in lists], dtype=np.int32)
    mc = np.ascontiguousarray(mchrgs, dtype=np.int32)
    fip = np.zeros((n, 3, 10))
    ok = np.zeros(n, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    if L.qcxms_fragment_ips_batch(n, _p(off, _ip), _p(num, _ip), _p(x), _p(lst, _ip), _p(nf, _ip),
                                  _p(mc, _ip), _p(fip), _p(ok, _ip), _p(st, _ip)):
        raise RuntimeError("qcxms_fragment_ips_batch failed: " + last_error())
    return fip.transpose(0, 2, 1).copy(), ok, st


def manage_fragments_batch(nums, xyzs, axyzs, qats, masses, atlast, itrj, isec, maxsec, btf,
                           mchrg, chrgcont, nfrag_ok=None, imasses=None, icoll=None):
    """manage_fragments (src/write_fragments.f90:12-457) for many finished
    trajectories; icoll (per trajectory) gives CID records with the
    collision number, none EI records.  Returns (lists, [FragmentReport])."""
    L = library()
    _post_sigs(L)
    n = len(nums)
    off = np.cumsum([0] + [len(x) for x in nums]).astype(np.int32)
    cat = lambda a, dt: np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=dt).reshape(-1)
                                                             for v in a]))
    num = cat(nums, np.int32)
    im = cat(imasses, np.int32) if imasses is not None else None
    rep = (FragmentReport * n)()
    for t in range(n):
        rep[t].mchrg = int(mchrg[t])
        rep[t].chrgcont = float(chrgcont[t])
        rep[t].nfrag_ok = 1 if nfrag_ok is None else int(nfrag_ok[t])
        rep[t].icoll = 0 if icoll is None else int(icoll[t])
    lst = np.zeros(len(num), dtype=np.int32)
    if L.qcxms_manage_fragments_batch(n, _p(off, _ip), _p(num, _ip), _p(cat(xyzs, np.float64)),
                                      _p(cat(axyzs, np.float64)), _p(cat(qats, np.float64)),
                                      _p(cat(masses, np.float64)), _p(im, _ip),
                                      _p(np.ascontiguousarray(atlast, dtype=np.float64)),
                                      _p(np.ascontiguousarray(itrj, dtype=np.int32), _ip),
                                      int(isec), int(maxsec), float(btf), _p(lst, _ip), rep):
        raise RuntimeError("qcxms_manage_fragments_batch failed: " + last_error())
    return [lst[off[t]:off[t + 1]] for t in range(n)], list(rep)
