"""CID runs of many trajectories: the collision loop of src/main.F90:1586-2110
with its mean-free-path MD loop (main.F90:1774-1974), every stage batched
over the trajectories that reach it:

  collision   cid_setup (first) or cid_setup_next (later collisions) and the
              collision MD of ion + gas (CidEnsemble, cid.f90:24-1111)
  fragments   manage_fragments (write_fragments.f90:12-457) on the ion
  continue    the fragment tcont (or the whole ion) recentred on its
              nuclear-charge-weighted centre, velocities kept
              (main.F90:1683-1735); a run stops at 7 atoms or fewer, or at a
              mass up to minmass (main.F90:1720-1730)
  MFP MD      isec + 1 and the production MD in its CID mode (MdEnsemble with
              cid=True) over nmax steps (main.F90:1816-1838), then fragments
              and continue again; a new fragment goes round once more, the
              whole ion leaves the loop (main.F90:1959-1963)
  next        a speed up to 800 m/s or E(COM) up to 0.85 eV ends the run
              (main.F90:1992-2008), as do new_counter fragmentations and the
              collision count (main.F90:2013-2102); otherwise the next
              collision flies along the centre's move over the last MD
              (main.F90:1636-1639).

The collision count is the user-set one (coll%max_coll, with new_counter
fragmentations ending a run when given); the automatic modes that draw it
from the cell's mean free path (FullAuto, CollAuto: collision_setup) are not
built, nor are the Euler rotation and rotational velocities of the first
collision (the caller's) and the MinPot / ConstVelo velocity scalings.  The
flags the reference keeps in SAVE variables between calls of cid() and md()
(collided, count_average) are carried per trajectory.  Every stage is logged
with its inputs and results, so that each can be checked on its own.
"""
import math

import numpy as np

from . import workloads as W
from .cid import CidEnsemble, GASES, MSTOAU, center_of_mass, cid_setup, cid_setup_next
from .gpu import MdEnsemble, manage_fragments_batch

AUTOAMU = 1.0 / W.AMUTOAU


def nmax_mfp(nuc, isec):
    """Steps of a mean-free-path MD (main.F90:1816-1829)."""
    nmax = nuc * 100
    if isec == 3:
        nmax = int(nmax * 0.75)
    if isec == 4:
        nmax = int(nmax * 0.6)
    if isec >= 5:
        nmax = int(nmax * 0.5)
    return min(max(nmax, 1000), 10000)


def continue_fragment(num, xyz, velo, mass, imass, lst, tcont):
    """main.F90:1683-1735: the atoms of fragment tcont (all atoms when tcont
    is 0) in their order, moved by their nuclear-charge-weighted centre
    sum Z x / sum Z (sequential sums), velocities kept."""
    sel = [i for i in range(len(num)) if tcont == 0 or lst[i] == tcont]
    dum = 0.0
    cema = np.zeros(3)
    for i in sel:
        dum = dum + float(num[i])
        cema = cema + xyz[i] * float(num[i])
    cema = cema / dum
    return (np.asarray(num)[sel].astype(np.int32), np.asarray(xyz)[sel] - cema, np.asarray(velo)[sel].copy(),
            np.asarray(mass)[sel].copy(), np.asarray(imass)[sel].copy())


class _Traj:
    def __init__(self, itrj, ion, charge):
        self.itrj = itrj
        self.num = np.asarray(ion["num"], dtype=np.int32)
        self.xyz = np.array(ion["xyz"], dtype=np.float64)
        self.velo = np.array(ion["velo"], dtype=np.float64)
        self.mass = np.array(ion["mass"], dtype=np.float64)
        self.imass = np.array(ion.get("imass", np.full(len(self.num), -1)), dtype=np.int32)
        self.mchrg = int(charge)
        self.chrgcont = float(charge)
        self.icoll = 0
        self.isec = 1
        self.direc = None
        self.new_velo = 0.0
        self.collided = 0          # cid.f90:125 (SAVE)
        self.count_average = 0     # cid.f90:147 (SAVE)
        self.md_count_average = 0  # md's own averaging flag (SAVE)
        self.frag_counter = 0
        self.cm1 = self.cm2 = None
        self.stop = None
        self.records = []


def _run_until_done(ens, chunk=1000, limit=200):
    """Run an ensemble until every trajectory has ended.  The reference has
    no cap (a fragment count that keeps rising again moves the end on each
    time); limit chunks bound the stage here."""
    for _ in range(limit):
        if ens.run(chunk) == 0:
            return
    raise RuntimeError(f"a stage did not end within {chunk * limit} steps")


class CidRun:
    """CID runs of ions (dicts with num, xyz, velo, mass[, imass]) with the
    collision gas gas at elab_ev (LAB frame, exact unless rng draws the 10 %
    spread), up to collisions collisions each.  rng draws the placement
    factors of every collision (cid.f90:527-530).  The GPU options (tstep,
    etemp, econv, pconv) apply to the collision and MD ensembles alike."""

    def __init__(self, ions, charge=1, gas="N2", elab_ev=40.0, collisions=3, new_counter=0, minmass=45,
                 tstep=0.5 * W.FSTOAU, etemp=5000.0, maxsec=7, btf=1.0, exact=True, rng=None,
                 econv=0.0, pconv=0.0, ntot=15000, log=False):
        self.gas, self.elab_ev, self.collisions, self.new_counter = gas, elab_ev, collisions, new_counter
        self.minmass, self.tstep, self.etemp, self.maxsec, self.btf = minmass, tstep, etemp, maxsec, btf
        self.exact, self.rng = exact, rng if rng is not None else np.random.default_rng(0)
        self.gpu = dict(econv=econv, pconv=pconv)
        self.ntot = ntot
        chg = np.broadcast_to(np.asarray(charge, dtype=np.int32), (len(ions),))
        self.trajs = [_Traj(k + 1, ion, chg[k]) for k, ion in enumerate(ions)]
        self.log = [] if log else None

    # ---- stages ----
    def _collide(self, ts):
        systems = []
        for t in ts:
            t.icoll += 1
            t.isec = 1
            rand = tuple(self.rng.random(4))
            if t.icoll == 1:
                s = cid_setup(t.num, t.xyz, t.velo, t.mass, self.tstep, gas=self.gas, elab_ev=self.elab_ev,
                              exact=self.exact, rand=rand, rng=self.rng)
                t.direc = s["direc"]
            else:
                s = cid_setup_next(t.num, t.xyz, t.velo, t.mass, self.tstep, t.new_velo, t.direc, gas=self.gas,
                                   rand=rand)
            s["rand"] = rand
            systems.append(s)
        ens = CidEnsemble(systems, charge=[t.mchrg for t in ts], tstep=self.tstep, etemp=self.etemp,
                          ntot=self.ntot, collided=[t.collided for t in ts],
                          count_average=[t.count_average for t in ts], **self.gpu)
        _run_until_done(ens)
        out = ens.read()
        ens.close()
        off = ens.offsets
        alive = []
        for k, (t, s) in enumerate(zip(ts, systems)):
            r = out["results"][k]
            nuc = s["nion"]
            a = slice(int(off[k]), int(off[k]) + nuc)
            if self.log is not None:
                self.log.append(dict(kind="collision", itrj=t.itrj, icoll=t.icoll, charge=t.mchrg,
                                     collided=t.collided, count_average=t.count_average, system=s,
                                     xyz=out["xyz"][int(off[k]):int(off[k + 1])].copy(),
                                     velo=out["velo"][int(off[k]):int(off[k + 1])].copy(),
                                     list=out["list"][a].copy(), axyz=out["axyz"][a].copy(),
                                     nstep=r.nstep, gradfail=r.gradfail, collided_out=r.collided,
                                     count_average_out=r.count_average, new_velo=r.new_velo, atlast=r.atlast,
                                     nfrag=r.nfrag))
            if r.gradfail:                                      # main.F90:1656-1657
                t.stop = "collision gradient failure"
                continue
            t.xyz, t.velo = out["xyz"][a].copy(), out["velo"][a].copy()
            t.collided, t.count_average = r.collided, r.count_average
            t.new_velo = r.new_velo                             # velo_cm = new_velo (cid.f90:1097)
            t.stage = dict(axyz=out["axyz"][a].copy(), qat=out["qat"][a].copy(), atlast=r.atlast)
            alive.append(t)
        return alive

    def _mfp_md(self, ts):
        for t in ts:
            t.isec += 1
            t.cm1 = center_of_mass(t.mass, t.xyz)               # main.F90:1779
        groups = {}
        for t in ts:
            groups.setdefault(nmax_mfp(len(t.num), t.isec), []).append(t)
        alive = []
        for nmax, g in groups.items():
            trajs = [dict(num=t.num, xyz=t.xyz, velo=t.velo, mass=t.mass, imass=t.imass, eimp=0.0, tadd=0.0,
                          charge=t.mchrg, isec=t.isec, cid=True, new_velo=t.new_velo,
                          count_average=t.md_count_average) for t in g]
            md = MdEnsemble(trajs, tstep=self.tstep, max_steps=nmax, etempin=self.etemp, **self.gpu)
            _run_until_done(md)
            out = md.read()
            axyz, achrg = md.read_averages()
            md.close()
            for k, t in enumerate(g):
                a = slice(int(md.offsets[k]), int(md.offsets[k + 1]))
                r = out["results"][k]
                if self.log is not None:
                    self.log.append(dict(kind="mfp", itrj=t.itrj, icoll=t.icoll, isec=t.isec, nmax=nmax,
                                         traj=trajs[k], xyz=out["xyz"][a].copy(), velo=out["velo"][a].copy(),
                                         list=out["list"][a].copy(), axyz=axyz[a].copy(), achrg=achrg[a].copy(),
                                         nstep=r.nstep, exit_reason=r.exit_reason, mdok=r.mdok, nfrag=r.nfrag,
                                         new_velo=r.new_velo, count_average_out=r.count_average,
                                         check_fragmented=r.check_fragmented, atlast=r.atlast, epot=r.epot))
                t.xyz, t.velo = out["xyz"][a].copy(), out["velo"][a].copy()
                t.new_velo, t.md_count_average = r.new_velo, r.count_average
                t.cm2 = center_of_mass(t.mass, t.xyz)           # main.F90:1851
                if not r.mdok:                                  # main.F90:1863-1871
                    t.stop = "MD failed"
                    continue
                t.stage = dict(axyz=axyz[a].copy(), qat=achrg[a].copy(), atlast=r.atlast)
                alive.append(t)
        return alive

    def _fragments(self, ts):
        """manage_fragments and the continuation; returns (continued with a
        new fragment, continued whole)."""
        split, whole = [], []
        by_isec = {}
        for t in ts:
            by_isec.setdefault(t.isec, []).append(t)
        for isec, g in by_isec.items():
            lists, reps = manage_fragments_batch(
                [t.num for t in g], [t.xyz for t in g], [t.stage["axyz"] for t in g], [t.stage["qat"] for t in g],
                [t.mass for t in g], [t.stage["atlast"] for t in g], [t.itrj for t in g], isec, self.maxsec,
                self.btf, [t.mchrg for t in g], [t.chrgcont for t in g], imasses=[t.imass for t in g],
                icoll=[t.icoll for t in g])
            for t, lst, rep in zip(g, lists, reps):
                t.records.extend(rep.records())
                if self.log is not None:
                    self.log.append(dict(kind="fragments", itrj=t.itrj, icoll=t.icoll, isec=isec, tcont=rep.tcont,
                                         nfrag=rep.nfrag, nfrag_ok=rep.nfrag_ok, mchrg=rep.mchrg,
                                         list=np.array(lst)))
                if not rep.nfrag_ok:                            # main.F90:1666, 1875
                    t.stop = "too many fragments"
                    continue
                t.mchrg, t.chrgcont = rep.mchrg, rep.chrgcont
                t.num, t.xyz, t.velo, t.mass, t.imass = continue_fragment(t.num, t.xyz, t.velo, t.mass, t.imass,
                                                                          lst, rep.tcont)
                if rep.tcont > 0:
                    t.frag_counter += 1
                if len(t.num) <= 7:                             # main.F90:1720-1724
                    t.stop = "small"
                    continue
                if t.mass.sum() * AUTOAMU <= self.minmass:      # main.F90:1726-1730
                    t.stop = "little mass"
                    continue
                (split if rep.tcont > 0 else whole).append(t)
        return split, whole

    def _next(self, ts):
        """The end conditions after the mean-free-path loop (main.F90:1983-2110)
        and the direction of the next collision."""
        go = []
        for t in ts:
            summass = 0.0
            for m in t.mass:
                summass = summass + m
            vv = t.new_velo * MSTOAU
            e_kin = 0.5 * summass * (vv * vv)
            mgas = GASES[self.gas][1] * W.AMUTOAU
            beta = mgas / (mgas + summass)
            e_com = beta * e_kin * W.AUTOEV
            if t.new_velo <= 800 or e_com <= 0.85:
                t.stop = "low velocity"
            elif self.new_counter > 0 and t.frag_counter >= self.new_counter:
                t.stop = "fragmentations done"
            elif t.icoll >= self.collisions:
                t.stop = "collisions done"
            else:
                d = t.cm2 - t.cm1                               # main.F90:1636-1639
                t.direc = d / math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
                go.append(t)
        return go

    def run(self):
        todo = list(self.trajs)
        while todo:
            ts = self._collide(todo)
            split, whole = self._fragments(ts)
            mfp = split + whole
            done_mfp = []
            while mfp:                                          # MFPloop (main.F90:1776-1975)
                ts = self._mfp_md(mfp)
                split, whole = self._fragments(ts)
                done_mfp.extend(whole)
                mfp = split
            todo = self._next(done_mfp)
        return self.trajs
