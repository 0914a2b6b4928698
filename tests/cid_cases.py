"""CID cases: small ions with N2 as the collision gas, set up as the first
collision of cid.f90 (qcxms_amd.cid.cid_setup) with fixed random factors."""
import numpy as np

from qcxms_amd import workloads as W
from qcxms_amd.cid import cid_setup

TSTEP = 0.5 * W.FSTOAU


def ion(seed, nheavy):
    num, xyz = W.build_chain([6] * nheavy, noise=0.02, seed=seed)
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    velo = W.maxwell_boltzmann(mass, 500.0, seed)
    return num, xyz, velo, mass


def cases():
    """name -> system (ion + N2): a propane ion and an ethane ion hit head-on
    at two energies; the energetic one breaks apart."""
    out = {}
    num, xyz, velo, mass = ion(3, 3)
    out["c3_40eV"] = cid_setup(num, xyz, velo, mass, TSTEP, gas="N2", elab_ev=40.0,
                               rand=(0.2, 0.3, 0.7, 0.2))
    num, xyz, velo, mass = ion(5, 2)
    out["c2_60eV"] = cid_setup(num, xyz, velo, mass, TSTEP, gas="N2", elab_ev=60.0,
                               rand=(0.1, 0.05, 0.3, 0.6))
    num, xyz, velo, mass = ion(7, 2)
    out["c2_400eV"] = cid_setup(num, xyz, velo, mass, TSTEP, gas="N2", elab_ev=400.0,
                                rand=(0.0, 0.0, 0.7, 0.7))
    return out


MSTOAU = 1.0 / 2.18769126364e+06   # m/s in atomic units of velocity


def md_cases():
    """name -> (trajectory, new_velo m/s, max_steps) for the production MD in
    its CID mode (md.f90 with method 3 after a collision):
      c2_rest   an ethane ion at 1000 K with no centre-of-mass motion;
      c2_drift  the same ion flying at 1500 m/s (every atom gets the same
                extra velocity), handed over with new_velo = 1500;
      c3_frag   a propane ion at 6000 K that breaks in two at step 434 and
                stays in two pieces."""
    out = {}
    num, xyz = W.build_chain([6, 6], noise=0.02, seed=5)
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    velo = W.maxwell_boltzmann(mass, 1000.0, 5)
    velo = velo - (mass[:, None] * velo).sum(0) / mass.sum()
    rest = dict(num=num, xyz=xyz, velo=velo, mass=mass, eimp=0.0, tadd=0.0)
    out["c2_rest"] = (rest, 0.0, 120)
    d = np.array([0.48, 0.6, 0.64])
    drift = dict(rest, velo=velo + 1500.0 * MSTOAU * d / np.linalg.norm(d))
    out["c2_drift"] = (drift, 1500.0, 120)
    num, xyz = W.build_chain([6] * 3, noise=0.02, seed=3)
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    out["c3_frag"] = (dict(num=num, xyz=xyz, velo=W.maxwell_boltzmann(mass, 6000.0, 3), mass=mass,
                           eimp=0.0, tadd=0.0), 0.0, 400)
    return out


def run_ions():
    """Ions for whole CID runs: a butane ion (two of them) and a pentane ion
    at 500 K, hit by N2 at 60 eV."""
    ions = []
    for seed, nheavy in ((3, 4), (5, 4), (7, 5)):
        num, xyz, velo, mass = ion(seed, nheavy)
        ions.append(dict(num=num, xyz=xyz, velo=velo, mass=mass))
    return ions
