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
