"""Collision-induced dissociation (CID) on the GPU: the set-up of one
collision (src/cid.f90:24-676, first collision, icoll = 1) on the host and the
collision MD (src/cid.f90:724-1111) for many trajectories at once
(CidEnsemble, the qcxms_cid_* entry points of libqcxms_gpu.so).

The collision gas joins the QC calculation (cid.f90:657-669, 703): He, Ne,
Ar, Kr, Xe (one atom) or N2 (two N atoms, each with the atomic N mass the
reference uses).  This build has GFN2 parameters for H, C, N and O only, so
of the gases only N2 runs; the others are rejected by the library with its
unsupported-element status.
"""
import ctypes as C
import math

import numpy as np

from . import workloads as W
from .ensemble import vary_energies
from .gpu import _dp, _ip, _p, last_error, library

EVTOAU = 1.0 / W.AUTOEV
MSTOAU = 1.0 / 2.18769126364e+06   # m/s in au (xtb_mctc_convert.f90:63)

# input.f90:511-557: (IndAtom, mass in amu, van der Waals radius in Bohr, atoms)
GASES = {
    "He": (2, 4.002, 2.64560263, 1),
    "Ne": (10, 20.18, 2.91016289, 1),
    "Ar": (18, 39.948, 3.55266638, 1),
    "Kr": (36, 83.798, 3.81722665, 1),
    "Xe": (54, 131.29, 4.081786915, 1),
    "N2": (7, 14.007, 3.64, 2),
}


def center_of_mass(mass, xyz):
    """utility.f90:541-562 (sequential sums)."""
    cm = np.zeros(3)
    tm = 0.0
    for m, r in zip(mass, xyz):
        cm = cm + m * r
        tm = tm + m
    return cm / tm


def ekinet(velo, mass):
    """mdinit.f90:57-76: kinetic energy and temperature."""
    ek = 0.0
    for m, v in zip(mass, velo):
        ek = ek + m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    ek = ek * 0.5
    return ek, ek / (0.5 * 3.0 * len(mass) * W.KB)


def cid_setup(num, xyz, velo, mass, tstep, gas="N2", elab_ev=None, ecom_ev=None, exact=True,
              rand=None, rng=None, manual_dist=0):
    """The first collision's set-up (cid.f90:179-676, icoll = 1): the ion is
    centred on its centre of mass, gets the flight velocity towards the gas
    (E_LAB, or E_COM / beta, spread by 10 % unless exact), and the gas is
    placed ahead of it, off axis by the random factors rand = (f, g, lmin,
    lpos) (drawn from rng when not given) scaled by the ion's extent.  The
    random Euler rotation and rotational velocities of the ion
    (euler_rotation, rotation_velo) are the caller's: xyz and velo are used as
    given.  Returns the system of ion + gas for CidEnsemble:
    dict(num, xyz, velo, mass, nion, tinit, beta, eimpact_ev)."""
    num = np.asarray(num, dtype=np.int32)
    xyz = np.array(xyz, dtype=np.float64)
    velo = np.array(velo, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    nuc = len(num)
    ind, m_amu, _, ngas = GASES[gas]
    mgas = m_amu * W.AMUTOAU
    xyz = xyz - center_of_mass(mass, xyz)                       # cid.f90:247-251
    summass = 0.0
    for m in mass:
        summass = summass + m
    beta = mgas / (mgas + summass)                              # cid.f90:273
    eimpact = (ecom_ev / beta) if (ecom_ev or 0.0) > 0.0 else elab_ev
    if exact:
        e_velo = eimpact * EVTOAU
    else:
        e_velo = vary_energies(eimpact, 0.1, rng) * EVTOAU
    _, tinit = ekinet(velo, mass)
    fasti = math.sqrt(2 * e_velo / summass)                     # cid.f90:368
    cm = center_of_mass(mass, xyz)
    f, g, lmin, lpos = rand if rand is not None else tuple(rng.random(4))
    lowx, highx = xyz[:, 0].min(), xyz[:, 0].max()             # cid.f90:541-556
    lowy, highy = xyz[:, 1].min(), xyz[:, 1].max()
    diff1 = lowy * f if lmin < 0.5 else highy * f               # cid.f90:558-568
    diff2 = lowx * g if lpos < 0.5 else highx * g
    step_dist = manual_dist if manual_dist else min(2 * nuc * 10, 800)
    start_dist = fasti * (2 * tstep)
    start_dist = step_dist * start_dist * W.AUTOAA              # as the reference: in Angstrom
    start_dist = max(start_dist, 10.0)
    xyz_start = np.array([cm[0], cm[1], cm[2] + start_dist])
    direc = xyz_start / math.sqrt(xyz_start[0] ** 2 + xyz_start[1] ** 2 + xyz_start[2] ** 2)
    xyz_gas = np.array([xyz_start[0] + diff2 * 0.8, xyz_start[1] + diff1 * 0.8, xyz_start[2]])
    scale_velo = direc * fasti
    nuc0 = nuc + ngas
    x0 = np.zeros((nuc0, 3))
    v0 = np.zeros((nuc0, 3))
    m0 = np.zeros(nuc0)
    n0 = np.zeros(nuc0, dtype=np.int32)
    x0[:nuc], v0[:nuc], m0[:nuc], n0[:nuc] = xyz, velo + scale_velo, mass, num
    x0[nuc0 - 1] = xyz_gas
    m0[nuc:] = mgas
    n0[nuc:] = ind
    if ngas == 2:                                               # cid.f90:660-667
        x0[nuc0 - 2] = [xyz_gas[0], xyz_gas[1], xyz_gas[2] + 1.09 * W.AATOAU]
    return dict(num=n0, xyz=x0, velo=v0, mass=m0, nion=nuc, tinit=tinit, beta=beta,
                eimpact_ev=eimpact, direc=direc)


def cid_setup_next(num, xyz, velo, mass, tstep, velo_cm, direc, gas="N2", rand=None, rng=None,
                   manual_dist=0):
    """The set-up of a later collision (cid.f90:179-676 with icoll > 1): no
    rotation; the thermal energy from all velocities (the flight included,
    as ekinet sees them); the gas placed along the flight direction direc
    (main.F90:1636-1639: the centre's move over the last MD) at
    max(velo_cm dt 2 step_dist, 17) (in Angstrom, as the reference) from the
    ion's centre of mass before the ion is moved onto the origin, off axis by
    0.7 x (extent x random factor); the ion keeps its velocities (vScale 0).
    velo_cm: the ion's speed in m/s (new_velo of the last MD).  Returns the
    system of ion + gas as cid_setup does."""
    num = np.asarray(num, dtype=np.int32)
    xyz = np.array(xyz, dtype=np.float64)
    velo = np.array(velo, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    direc = np.asarray(direc, dtype=np.float64)
    nuc = len(num)
    ind, m_amu, _, ngas = GASES[gas]
    mgas = m_amu * W.AMUTOAU
    summass = 0.0
    for m in mass:
        summass = summass + m
    beta = mgas / (mgas + summass)                              # cid.f90:273
    vv = velo_cm * MSTOAU
    e_velo = 0.5 * summass * (vv * vv)                          # cid.f90:298
    _, tinit = ekinet(velo, mass)
    cm = center_of_mass(mass, xyz)
    f, g, lmin, lpos = rand if rand is not None else tuple(rng.random(4))
    lowx, highx = xyz[:, 0].min(), xyz[:, 0].max()             # cid.f90:541-556
    lowy, highy = xyz[:, 1].min(), xyz[:, 1].max()
    diff1 = lowy * f if lmin < 0.5 else highy * f               # cid.f90:558-568
    diff2 = lowx * g if lpos < 0.5 else highx * g
    step_dist = manual_dist if manual_dist else min(2 * nuc * 10, 800)
    start_dist = vv * (2 * tstep)                               # cid.f90:602-606
    start_dist = step_dist * start_dist * W.AUTOAA
    start_dist = max(start_dist, 17.0)
    xyz = xyz - cm                                              # cid.f90:609-611
    xyz_gas = np.array([cm[0] + direc[0] * start_dist + diff2 * 0.7,
                        cm[1] + direc[1] * start_dist + diff1 * 0.7,
                        cm[2] + direc[2] * start_dist])
    nuc0 = nuc + ngas
    x0 = np.zeros((nuc0, 3))
    v0 = np.zeros((nuc0, 3))
    m0 = np.zeros(nuc0)
    n0 = np.zeros(nuc0, dtype=np.int32)
    x0[:nuc], v0[:nuc], m0[:nuc], n0[:nuc] = xyz, velo, mass, num
    x0[nuc0 - 1] = xyz_gas
    m0[nuc:] = mgas
    n0[nuc:] = ind
    if ngas == 2:                                               # cid.f90:660-667
        x0[nuc0 - 2] = [xyz_gas[0], xyz_gas[1], xyz_gas[2] + 1.09 * W.AATOAU]
    return dict(num=n0, xyz=x0, velo=v0, mass=m0, nion=nuc, tinit=tinit, beta=beta,
                eimpact_ev=e_velo * W.AUTOEV, velo_cm=float(velo_cm))


class CidResult(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nstep", "nfrag", "done", "gradfail", "collided",
                                          "fragmented", "total_steps", "collision_step",
                                          "check_fragmented", "count_average")]
                + [(n, C.c_double) for n in ("ttime", "epot", "ekin", "new_velo", "avgt", "atlast",
                                             "ecom", "velo_diff", "lowest_com")]
                + [("fragT", C.c_double * 10), ("eint", C.c_double * 10)])


class CidParams(C.Structure):
    _fields_ = [("tstep", C.c_double), ("etemp", C.c_double), ("econv", C.c_double),
                ("pconv", C.c_double), ("ntot", C.c_int32), ("reserved", C.c_int32)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = library()
    L.qcxms_cid_create.restype = C.c_void_p
    L.qcxms_cid_create.argtypes = [C.c_int32, _ip, _ip, _ip, _ip, _dp, _ip, _dp, _dp, _dp, _dp, _dp,
                                   C.c_void_p]
    L.qcxms_cid_destroy.argtypes = [C.c_void_p]
    L.qcxms_cid_run.argtypes = [C.c_void_p, C.c_int32, _ip]
    L.qcxms_cid_read.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip, _dp, C.c_void_p]
    L.qcxms_cid_set_saved.argtypes = [C.c_void_p, _ip, _ip]
    L.qcxms_cid_batch.restype = C.c_void_p
    L.qcxms_cid_batch.argtypes = [C.c_void_p]
    _bound = True


class CidEnsemble:
    """Collision MD of many CID trajectories (systems from cid_setup or
    cid_setup_next) in lock-step on the GPU.  charge: the ion's charge (per
    trajectory or one); collided / count_average: the flags cid() keeps from
    a trajectory's earlier collisions (per trajectory, default 0)."""

    def __init__(self, systems, charge=1, tstep=0.5 * W.FSTOAU, etemp=5000.0, ntot=15000,
                 econv=0.0, pconv=0.0, velo_cm=None, collided=None, count_average=None):
        _bind()
        L = library()
        self.n = len(systems)
        self.offsets = np.cumsum([0] + [len(s["num"]) for s in systems]).astype(np.int32)
        self.nion = np.array([s["nion"] for s in systems], dtype=np.int32)
        self.num = np.concatenate([s["num"] for s in systems]).astype(np.int32)
        chg = np.broadcast_to(np.asarray(charge, dtype=np.int32), (self.n,)).copy()
        mass = np.concatenate([s["mass"] for s in systems]).astype(np.float64)
        xyz = np.ascontiguousarray(np.concatenate([s["xyz"] for s in systems]), dtype=np.float64)
        velo = np.ascontiguousarray(np.concatenate([s["velo"] for s in systems]), dtype=np.float64)
        tinit = np.array([s["tinit"] for s in systems], dtype=np.float64)
        beta = np.array([s["beta"] for s in systems], dtype=np.float64)
        if velo_cm is None:
            velo_cm = [float(s.get("velo_cm", 0.0)) for s in systems]
        vcm = np.ascontiguousarray(velo_cm, dtype=np.float64)
        prm = CidParams(tstep, etemp, econv, pconv, ntot, 0)
        self._h = L.qcxms_cid_create(self.n, _p(self.offsets, _ip), _p(self.nion, _ip), _p(self.num, _ip),
                                     _p(chg, _ip), _p(mass), None, _p(xyz), _p(velo), _p(tinit),
                                     _p(beta), _p(vcm), C.byref(prm))
        if not self._h:
            raise RuntimeError("qcxms_cid_create failed: " + last_error())
        if collided is not None or count_average is not None:
            flags = [None if v is None else np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.int32),
                                                                               (self.n,)))
                     for v in (collided, count_average)]
            if L.qcxms_cid_set_saved(self._h, *[None if f is None else _p(f, _ip) for f in flags]):
                self.close()
                raise RuntimeError("qcxms_cid_set_saved failed: " + last_error())

    def run(self, nsteps):
        running = C.c_int32()
        if library().qcxms_cid_run(self._h, int(nsteps), C.byref(running)):
            raise RuntimeError("qcxms_cid_run failed: " + last_error())
        return running.value

    def read(self):
        na = int(self.offsets[-1])
        xyz, velo, grad = (np.zeros((na, 3)) for _ in range(3))
        q, axyz = np.zeros(na), np.zeros((na, 3))
        lst = np.zeros(na, dtype=np.int32)
        res = (CidResult * self.n)()
        if library().qcxms_cid_read(self._h, _p(xyz), _p(velo), _p(grad), _p(q), _p(lst, _ip), _p(axyz),
                                    res):
            raise RuntimeError("qcxms_cid_read failed: " + last_error())
        return dict(xyz=xyz, velo=velo, grad=grad, qat=q, list=lst, axyz=axyz, results=list(res))

    def close(self):
        if self._h:
            library().qcxms_cid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
