"""Trajectories that fragment early, for the fragmentation-generation
(main.F90:2208-2345) tests: backbone bonds stretched past the connectivity
cut-off of fragments.f90 with the pieces moving apart.

  split2   n-octane cation, the bond between the second and third carbon
           stretched by 0.9 A: two fragments from the first step, so the
           first MD runs to nmax, manage_fragments continues with one piece
  split4   C20H36O2N2 cation, three backbone bonds stretched by 1.2 A: four
           fragments at the first step, nfrag > nfragexit ends the MD at once
  intact   n-octane cation, no stretch: one generation
"""
import numpy as np

from qcxms_amd import workloads as W

A = W.AATOAU


def _centres(num):
    """Backbone centre of every atom (build_chain appends each heavy atom's
    hydrogens right after it)."""
    c, out = -1, []
    for z in num:
        if z != 1:
            c += 1
        out.append(c)
    return np.array(out)


def stretched(num, xyz, cuts, amount_aa, speed, seed):
    """Pull the chain apart after the backbone centres in cuts: every piece
    moves by amount_aa/2 per cut away from its neighbours and gets speed (au)
    along the local bond direction."""
    num = np.asarray(num)
    xyz = np.array(xyz, dtype=float)
    cen = _centres(num)
    heavy = np.where(num != 1)[0]
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    velo = W.maxwell_boltzmann(mass, 300.0, seed)
    for k in cuts:
        a, b = heavy[k], heavy[k + 1]
        u = (xyz[b] - xyz[a]) / np.linalg.norm(xyz[b] - xyz[a])
        right = cen > k
        xyz[right] += amount_aa * A * u
        velo[right] += speed * u
        velo[~right] -= speed * u
    return dict(num=np.asarray(num, dtype=np.int32), xyz=xyz, velo=velo, mass=mass, eimp=0.0,
                tadd=0.0)


def cases():
    out = {}
    num, xyz = W.octane(seed=5)
    out["split2"] = stretched(num, xyz, [1], 0.9, 4e-4, 11)
    num, xyz = W.chain_c20h36o2n2(seed=2)
    out["split4"] = stretched(num, xyz, [4, 9, 15], 1.2, 4e-4, 12)
    num, xyz = W.octane(seed=6)
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    out["intact"] = dict(num=num, xyz=xyz, velo=W.maxwell_boltzmann(mass, 500.0, 13), mass=mass,
                         eimp=0.3, tadd=10 * W.FSTOAU)
    return out
