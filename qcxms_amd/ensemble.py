"""Ensemble inputs and outputs of QCxMS, so that GPU runs take QCxMS-prepared
trajectories and hand back what QCxMS writes.

Files (formats as the reference writes them):
- ``TMPQCXMS/TMP.<n>/qcxms.start``: wrstart / rdstart (src/utility.f90:321-
  422): ``(i4)`` trajectory number, ``(D22.14)`` impact energy (Eh),
  ``(D22.14)`` heating time tadd (au), then per atom ``(7D22.14)`` with the
  velocity (au) and the heating factor velof.
- ``TMPQCXMS/TMP.<n>/start.xyz``: the start geometry, xyz in Angstrom
  (written by mctc-lib's xyz writer in the reference).
- ``qcxms.gs``: the ground-state trajectory (src/md.f90:151-152, 383-386):
  list-directed number of dumps, then per dump and atom ``(6d16.8)`` with
  coordinates (Bohr) and velocities (au); read back at src/main.F90:668-675,
  767-777.
- ``qcxms.res``: the records of qcxms_manage_fragments_batch.

IEE model (src/iee.f90): getmaxiee, getieeab, gauss0, poiss0; the heating
time of an ionised orbital, calctrelax (src/impact.f90:58-81); the Box-Muller
energy spread vary_energies (src/boxmuller.f90:46-76) with the caller's
numpy random generator (the reference draws from the Fortran runtime's
random_number, whose stream is compiler-specific).
"""
import math
import os

import numpy as np

from . import workloads as W

_SYMBOL = {1: "H", 6: "C", 7: "N", 8: "O"}
_Z = {v: k for k, v in _SYMBOL.items()}
AUTOEV = 27.21138505        # xtb_mctc_convert.f90:27
FSTOAU = 41.3413733365614   # xtb_mctc_convert.f90


# ---------------------------------------------------------------------------
# Fortran edit descriptors
def fortran_d(v, w, d):
    """Dw.d output: 0.<d digits>D+ee (exponent of three digits without the D
    beyond 99), right-justified in w characters, asterisks if it does not fit."""
    v = float(v)
    if not math.isfinite(v):
        return "*" * w
    if v == 0.0:
        body = "0." + "0" * d + "D+00"
    else:
        m = f"{abs(v):.{d - 1}e}"        # d significant digits: x.xxxxE+yy
        mant, ex = m.split("e")
        e10 = int(ex) + 1
        digits = mant.replace(".", "")
        if abs(e10) <= 99:
            body = f"0.{digits}D{'+' if e10 >= 0 else '-'}{abs(e10):02d}"
        else:
            body = f"0.{digits}{'+' if e10 >= 0 else '-'}{abs(e10):03d}"
    if v < 0 or (v == 0.0 and math.copysign(1.0, v) < 0):
        body = "-" + body
    return body.rjust(w) if len(body) <= w else "*" * w


def fortran_i(v, w):
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


def _read_real(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def _fixed_fields(line, w, n):
    return [_read_real(line[k * w:(k + 1) * w]) for k in range(n) if line[k * w:(k + 1) * w].strip()]


# ---------------------------------------------------------------------------
# qcxms.start / start.xyz
def traj_dir(root, traj):
    """TMPQCXMS/TMP.<traj> as wrstart names it (no zero padding)."""
    return os.path.join(root, f"TMP.{int(traj)}")


def write_start(path, traj, velo, velof, eimp, tadd):
    """qcxms.start with the records of wrstart (utility.f90:358-369)."""
    velo = np.asarray(velo, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        f.write(fortran_i(traj, 4) + "\n")
        f.write(fortran_d(eimp, 22, 14) + "\n")
        f.write(fortran_d(tadd, 22, 14) + "\n")
        for j in range(len(velo)):
            f.write("".join(fortran_d(x, 22, 14) for x in (*velo[j], velof[j])) + "\n")


def read_start(path, nat):
    """rdstart (utility.f90:372-415): (itrj, eimp, tadd, velo [nat,3], velof [nat])."""
    with open(path) as f:
        lines = f.read().splitlines()
    itrj = int(lines[0].split()[0])
    eimp = _read_real(lines[1][:22])
    tadd = _read_real(lines[2][:22])
    velo = np.zeros((nat, 3))
    velof = np.zeros(nat)
    for j in range(nat):
        vals = _fixed_fields(lines[3 + j], 22, 4)
        velo[j] = vals[:3]
        velof[j] = vals[3]
    return itrj, eimp, tadd, velo, velof


def write_xyz(path, num, xyz_bohr, comment=""):
    xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3) * W.AUTOAA
    with open(path, "w") as f:
        f.write(f"{len(num)}\n{comment}\n")
        for z, (x, y, zz) in zip(num, xyz):
            f.write(f"{_SYMBOL.get(int(z), str(int(z))):<4s}{x:20.14f}{y:20.14f}{zz:20.14f}\n")


def read_xyz(path):
    """xyz file (Angstrom) -> (num [nat], xyz [nat,3] Bohr)."""
    with open(path) as f:
        lines = f.read().splitlines()
    nat = int(lines[0].split()[0])
    num, xyz = [], []
    for line in lines[2:2 + nat]:
        t = line.split()
        sym = t[0]
        num.append(int(sym) if sym.isdigit() else _Z[sym.capitalize()])
        xyz.append([float(v) for v in t[1:4]])
    return np.array(num, dtype=np.int32), np.array(xyz) * W.AATOAU


def write_trajectory_inputs(root, traj, num, xyz_bohr, velo, velof, eimp, tadd):
    """wrstart: TMPQCXMS/TMP.<traj>/start.xyz and qcxms.start."""
    d = traj_dir(root, traj)
    os.makedirs(d, exist_ok=True)
    write_xyz(os.path.join(d, "start.xyz"), num, xyz_bohr)
    write_start(os.path.join(d, "qcxms.start"), traj, velo, velof, eimp, tadd)


def load_tmpqcxms(root, charge=1):
    """Every TMP.<n> directory under root as a trajectory dict for
    qcxms_amd.MdEnsemble (num, xyz, velo, velof, mass, eimp, tadd, charge,
    itrj), ordered by trajectory number."""
    dirs = []
    for name in os.listdir(root):
        if name.startswith("TMP.") and name[4:].isdigit():
            dirs.append((int(name[4:]), os.path.join(root, name)))
    out = []
    for n, d in sorted(dirs):
        num, xyz = read_xyz(os.path.join(d, "start.xyz"))
        itrj, eimp, tadd, velo, velof = read_start(os.path.join(d, "qcxms.start"), len(num))
        out.append(dict(num=num, xyz=xyz, velo=velo, velof=velof,
                        mass=np.array([W.atomic_mass(int(z)) for z in num]), eimp=eimp,
                        tadd=tadd, charge=charge, itrj=itrj))
    return out


# ---------------------------------------------------------------------------
# qcxms.gs
def write_gs(path, nmax, xyz_dumps, velo_dumps):
    """md.f90:151-152, 383-386: list-directed nmax, then (6d16.8) per atom
    and dump."""
    with open(path, "w") as f:
        f.write(f" {int(nmax)}\n")
        for x, v in zip(xyz_dumps, velo_dumps):
            for a in range(len(x)):
                f.write("".join(fortran_d(c, 16, 8) for c in (*x[a], *v[a])) + "\n")


def read_gs(path, nat):
    """main.F90:668-675, 767-777: (ndumpGS, xyz [ndump,nat,3], velo [...]).
    Fields are read list-directed, as main does."""
    with open(path) as f:
        ndump = int(f.readline().split()[0])
        rows = []
        for line in f:
            if line.strip():
                rows.append([_read_real(t) for t in line.split()])
    rows = np.array(rows).reshape(-1, nat, 6)
    return ndump, rows[:, :, :3].copy(), rows[:, :, 3:].copy()


def sample_ground_state(trajectory, nsteps, dumpstep=4, **md_kw):
    """Ground-state MD (no internal excess energy) on the GPU, keeping a
    snapshot every dumpstep steps as md.f90 does for qcxms.gs
    (it = 0).  Returns (xyz [ndump,nat,3], velo [ndump,nat,3])."""
    from .gpu import MdEnsemble
    tr = dict(trajectory, eimp=0.0, tadd=0.0)
    md = MdEnsemble([tr], max_steps=nsteps + 1, **md_kw)
    xs, vs = [], []
    try:
        done = 0
        while done < nsteps:
            md.run(dumpstep)
            done += dumpstep
            out = md.read()
            xs.append(out["xyz"].copy())
            vs.append(out["velo"].copy())
            if not out["results"][0].running:
                break
    finally:
        md.close()
    return np.array(xs), np.array(vs)


def write_res(path, reports, append=True):
    """qcxms.res: the records of qcxms_manage_fragments_batch reports."""
    with open(path, "a" if append else "w") as f:
        for r in reports:
            for line in r.records():
                f.write(line + "\n")


# ---------------------------------------------------------------------------
# IEE model (iee.f90) and heating times (impact.f90)
def gauss0(iee_a, iee_b, ieeel, x):
    """iee.f90:80-86."""
    return math.exp(-iee_a * (x - ieeel * iee_b) ** 2 / ieeel)


def poiss0(iee_a, iee_b, ieeel, x):
    """iee.f90:121-135."""
    z = iee_b
    k = 1.0 / iee_a
    t2 = k / ieeel
    t8 = math.log(z / k * ieeel / x)
    t14 = math.exp(t2 * x * (1.0 + t8) - 1.0 * z)
    t17 = (t2 * x + 1.0) ** (-0.5)
    return t14 * t17


def getmaxiee(iee_a, iee_b, ieeel, ityp, exc):
    """iee.f90:47-76: (ieemax, pmax, E_avg) of P(E) on 0.001, 0.011, ... < exc."""
    x = 0.001
    pmax, ieemax, e_avg, m = -1.0, 0.0, 0.0, 0.0
    while True:
        val = gauss0(iee_a, iee_b, ieeel, x) if ityp == 0 else poiss0(iee_a, iee_b, ieeel, x)
        if val > pmax:
            pmax, ieemax = val, x
        x = x + 0.01
        e_avg = e_avg + val * x
        m = m + val
        if x >= exc:
            break
    return ieemax, pmax, e_avg / m


def getieeab(ieeel, ityp, exc, nbnd, ieeatm):
    """iee.f90:12-39: the parameters (a, b) whose average energy per bond
    reaches ieeatm (eV)."""
    st = 0.005
    a = b = 0.0
    for _ in range(10000):
        a = a + st
        a = min(a, float(np.float32(0.3)))
        b = b + st * 7
        _, _, e_avg = getmaxiee(a, b, ieeel, ityp, exc)
        if e_avg / nbnd >= ieeatm:
            return a, b
    raise RuntimeError("internal error inside getieeab")


def calctrelax(emo, na, i, trelax):
    """impact.f90:58-81: heating time for ionising orbital i (1-based) with
    na alpha orbitals; emo in Eh."""
    alp = 0.5 * AUTOEV
    t = 0.0
    k = i
    for j in range(i + 1, na + 1):
        t = t + trelax * math.exp(alp * (emo[k - 1] - emo[j - 1]))
        k = k + 1
    return t


def vary_energies(e_in, e_distr, rng):
    """boxmuller.f90:46-76 with uniforms from rng."""
    u1, u2 = rng.random(), rng.random()
    sigma = e_in * e_distr
    r = math.sqrt(-2.0 * math.log(u1))
    z0 = r * math.cos(2.0 * math.pi * u2)
    z1 = r * math.sin(2.0 * math.pi * u2)
    return z0 * sigma + e_in if u1 > 0.5 else z1 * sigma + e_in


# ---------------------------------------------------------------------------
# MO spectrum of the ground state and the randomised ensemble (main.F90:756-886)
SPEC_ETEMP = 1.0e-3  # K: integer occupations for the MO spectrum (reference: etemp 0)


def getspec(num, xyz, charge=0):
    """getspec (mo_spec.f90:12-119) for GFN2 through the GPU: a single point
    of the ground state, orbital energies as positive ionisation energies
    emo (Eh, -eps), the HOMO index ihomo (1-based), ehomo = emo[ihomo-1],
    mopop [nao, nat], and nb = nel/2 (electrons_amount, iniqm.f90:45-63)."""
    from .gpu import Batch
    num = np.asarray(num, dtype=np.int32)
    b = Batch([(num, charge)])
    try:
        b.set_warm_start(0)
        e, g, q, st = b.egrad(np.asarray(xyz, dtype=np.float64), SPEC_ETEMP)
        if st[0] != 0:
            raise RuntimeError("QC initialization error")
        eps, focc, mopop, ihomo = b.mo_spectrum(0)
    finally:
        b.close()
    nel = int(sum(int(z) - (0 if z <= 2 else 2) for z in num)) - charge
    nb = nel // 2
    emo = -eps
    if emo[0] - emo[ihomo - 1] < 1.0 / AUTOEV:   # eps in eV: emo(ihomo) - emo(1) < 1
        raise RuntimeError("weird epsilons")
    return dict(emo=emo, ehomo=float(emo[ihomo - 1]), ihomo=ihomo, nb=nb,
                mopop=mopop, energy=float(e[0]))


def irand(n, rng):
    """utility.f90:502-511: integer in 1..n."""
    r = int(n * rng.random() + 1)
    return min(r, n)


def momap(ihomo, emo, edum, rng):
    """utility.f90:23-65: the ionised orbital(s) whose energies plus an
    excitation into a virtual orbital best match edum (Eh); mo2 = 0 for a
    single hole.  5001 random trials, as the reference."""
    dmin, i1, i2 = np.finfo(float).max, 0, 0
    for _ in range(5001):
        mo1, mo2 = irand(ihomo, rng), irand(ihomo, rng)
        vmo = irand(ihomo // 2, rng) + ihomo
        dum = emo[mo1 - 1]
        if mo2 > ihomo // 2:
            mo2 = 0
        else:
            dum = dum + emo[mo2 - 1]
        dum = dum + emo[vmo - 1]
        delta = abs(dum - edum)
        if delta < dmin:
            dmin, i1, i2 = delta, mo1, mo2
    return i1, i2


def generate_ensemble(num, gs_xyz, gs_velo, ntraj, spec, rng, eimp0_ev=70.0, eimpw=0.10,
                      ieeatm=0.6, edistri=1, iee_a=0.0, iee_b=0.0, hacc=3.0, fimp=1.0,
                      trelax=2000.0, unity=False):
    """The EI ensemble of main.F90:663-886 from ground-state snapshots
    (gs_xyz, gs_velo [ndump, nat, 3]) and the MO spectrum of getspec:
    ntraj distinct snapshots chosen at random, each with an impact energy
    drawn by Box-Muller, an internal excess energy drawn from the IEE
    distribution by rejection, the ionised orbitals from momap, per-atom
    heating factors from their populations and a heating time from
    calctrelax.  Returns trajectory dicts (num, xyz, velo, velof, mass,
    eimp [Eh], tadd [au]) and the IEE parameters (a, b, pmax).  Random
    numbers come from rng (numpy Generator)."""
    num = np.asarray(num, dtype=np.int32)
    nuc = len(num)
    ndump = len(gs_xyz)
    if ndump <= 2 * ntraj:
        raise ValueError("Error: compute longer GS trajectory")
    emo, ehomo, ihomo, nb, mopop = (spec["emo"], spec["ehomo"], spec["ihomo"], spec["nb"],
                                    spec["mopop"])
    eimp0 = eimp0_ev / AUTOEV
    exc = (eimp0 - ehomo) * AUTOEV
    ieeel = float(ihomo + nb)
    if not (iee_a > 0 and iee_b > 0):
        iee_a, iee_b = getieeab(ieeel, edistri, exc, nuc, ieeatm)
    _, pmax, _ = getmaxiee(iee_a, iee_b, ieeel, edistri, exc)
    icalc = np.zeros(ndump, dtype=bool)
    k = 0
    while k <= ntraj:                      # main.F90:695-701 (ntraj + 1 picks)
        j = irand(ndump, rng)
        if not icalc[j - 1]:
            k += 1
            icalc[j - 1] = True
    dist = gauss0 if edistri == 0 else poiss0
    out = []
    for i in range(ndump):
        if not icalc[i]:
            continue
        if len(out) == ntraj:
            break
        while True:
            edum = vary_energies(eimp0, eimpw, rng)
            if edum >= ehomo:
                break
        edum = edum - ehomo
        while True:
            x = rng.random() * edum * AUTOEV
            if dist(iee_a, iee_b, ieeel, x) / pmax >= rng.random():
                break
        iee = fimp * x / AUTOEV
        mo1, mo2 = momap(ihomo, emo, iee + ehomo, rng)
        modum = mopop[mo1 - 1].copy()
        if mo2 > 0:
            modum = modum + mopop[mo2 - 1]
        modum[num == 1] *= hacc
        top = modum.max()
        velof = np.ones(nuc) if (nuc > 35 or unity) else modum / top
        t = calctrelax(emo, ihomo, mo1, trelax)
        if mo2 > 0:
            t = t + calctrelax(emo, ihomo, mo2, trelax)
        t = max(t / (ihomo + nb), trelax / 10.0)
        out.append(dict(num=num, xyz=np.array(gs_xyz[i]), velo=np.array(gs_velo[i]), velof=velof,
                        mass=np.array([W.atomic_mass(int(z)) for z in num]), eimp=iee,
                        tadd=t * FSTOAU))
    return out, (iee_a, iee_b, pmax)
