"""Seeded synthetic workloads for the GFN2-xTB MD hot path.

Geometries are built from ideal bond lengths and tetrahedral angles (all-anti
zig-zag backbones), then perturbed with seeded Gaussian noise.  Randomness is
numpy's PCG64 generator with explicit seeds, so every workload is reproducible
on any machine.  Coordinates are returned in Bohr (xtb_mctc_convert.f90:24-25),
masses in electron masses (atomic_masses.f90:58-67, mass.f90:3-22).
"""
import math

import numpy as np

AUTOAA = 0.52917726
AATOAU = 1.0 / AUTOAA
AUTOEV = 27.21138505
KB = 3.166808578545117e-06
FSTOAU = 41.3413733365614
AMUTOAU = 1.660539040e-27 * (1.0 / 9.10938356e-31)

MASS_AMU = {1: 1.00794075, 6: 12.01073590, 7: 14.00670321, 8: 15.99940492}
VALENCE = {6: 4, 7: 3, 8: 2}
BOND = {  # Angstrom
    (6, 6, 1): 1.54, (6, 6, 2): 1.34, (6, 7, 1): 1.47, (6, 8, 1): 1.43, (6, 8, 2): 1.23,
    (7, 7, 1): 1.45, (7, 8, 1): 1.40, (8, 8, 1): 1.48,
    (1, 6, 1): 1.09, (1, 7, 1): 1.01, (1, 8, 1): 0.96,
}
TETRA = math.radians(109.4712206)


def bond_length(za, zb, order=1):
    a, b = sorted((za, zb))
    return BOND[(a, b, order)]


def atomic_mass(z):
    """atomic_masses.f90 value times amutoau (electron masses)."""
    return AMUTOAU * MASS_AMU[z]


def _unit(v):
    return v / np.linalg.norm(v)


def build_chain(backbone, orders=None, carbonyl=(), noise=0.0, seed=0, stretch=None):
    """Saturated (or partially unsaturated) open chain.

    backbone: heavy-atom atomic numbers in chain order.
    orders:   bond orders between consecutive backbone atoms (default 1).
    carbonyl: backbone indices (interior carbons) carrying a double-bonded O.
    stretch:  optional per-bond scale factors for the backbone bonds.
    Returns (num, xyz[Bohr]) with hydrogens appended after the heavy atoms of
    each centre in order.
    """
    nb = len(backbone)
    orders = list(orders) if orders is not None else [1] * (nb - 1)
    stretch = list(stretch) if stretch is not None else [1.0] * (nb - 1)
    # zig-zag backbone in the xy plane
    pos = [np.zeros(3)]
    half = TETRA / 2.0
    for i in range(1, nb):
        d = bond_length(backbone[i - 1], backbone[i], orders[i - 1]) * stretch[i - 1]
        sgn = 1.0 if i % 2 else -1.0
        pos.append(pos[-1] + d * np.array([math.sin(half), sgn * math.cos(half), 0.0]))
    num, xyz = [], []
    for i in range(nb):
        z = backbone[i]
        num.append(z)
        xyz.append(pos[i])
        nbrs = []
        if i > 0:
            nbrs.append((i - 1, orders[i - 1]))
        if i < nb - 1:
            nbrs.append((i + 1, orders[i]))
        used = sum(o for _, o in nbrs) + (2 if i in carbonyl else 0)
        nh = VALENCE[z] - used
        if nb == 1:
            dirs = [_unit(np.array(v)) for v in
                    ((1, 1, 1), (-1, -1, 1), (-1, 1, -1), (1, -1, -1))]
            for k in range(nh):
                num.append(1)
                xyz.append(pos[i] + bond_length(1, z) * dirs[k])
            continue
        if len(nbrs) == 2:
            a = _unit(pos[nbrs[0][0]] - pos[i])
            b = _unit(pos[nbrs[1][0]] - pos[i])
            u = _unit(-(a + b))
            w = _unit(np.cross(a, b))
            subs = []
            if i in carbonyl:
                subs.append((8, u, 2))
            elif nh == 1:
                subs.append((1, u if (orders[i - 1] == 2 or orders[i] == 2) else
                             u * math.cos(half) + w * math.sin(half), 1))
            else:
                for s in (1.0, -1.0)[:nh]:
                    subs.append((1, u * math.cos(half) + s * w * math.sin(half), 1))
        else:
            a = _unit(pos[nbrs[0][0]] - pos[i])
            e1 = _unit(np.cross(a, np.array([0.0, 0.0, 1.0])))
            if np.linalg.norm(np.cross(a, np.array([0.0, 0.0, 1.0]))) < 1e-8:
                e1 = _unit(np.cross(a, np.array([0.0, 1.0, 0.0])))
            e2 = np.cross(a, e1)
            # staggered: first substituent anti to the backbone continuation
            phi0 = 0.0
            subs = []
            for k in range(nh):
                phi = phi0 + 2.0 * math.pi * k / 3.0
                d = a * math.cos(TETRA) + math.sin(TETRA) * (math.cos(phi) * e2 + math.sin(phi) * e1)
                subs.append((1, _unit(d), 1))
        for (zs, d, order) in subs:
            num.append(zs)
            xyz.append(pos[i] + bond_length(zs, z, order) * d)
    xyz = np.array(xyz)
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        xyz = xyz + rng.normal(0.0, noise, size=xyz.shape)
    return np.array(num, dtype=np.int32), xyz * AATOAU


def octane(noise=0.02, seed=0):
    """n-octane all-anti (C8H18): 26 atoms."""
    return build_chain([6] * 8, noise=noise, seed=seed)


def chain_c12h24o2n2(seed=0, max_stretch=0.30):
    """40-atom C12H24O2N2 chain: 14-atom C/N backbone with two carbonyl O,
    backbone bonds stretched by seeded factors in [1, 1 + max_stretch]."""
    backbone = [6, 6, 6, 7, 6, 6, 6, 6, 6, 6, 7, 6, 6, 6]
    rng = np.random.default_rng(seed)
    stretch = 1.0 + max_stretch * rng.random(len(backbone) - 1)
    return build_chain(backbone, carbonyl=(2, 9), noise=0.02, seed=seed + 1000, stretch=stretch)


def chain_c20h36o2n2(seed=0, noise=0.02):
    """60-atom C20H36O2N2 chain: 22-atom C/N backbone, two carbonyl O and two
    C=C double bonds."""
    backbone = [6, 6, 6, 6, 6, 7, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 6, 6, 6, 6, 6]
    orders = [1] * (len(backbone) - 1)
    orders[1] = 2
    orders[18] = 2
    return build_chain(backbone, orders=orders, carbonyl=(7, 13), noise=noise, seed=seed)


def random_chain(seed, nheavy_max=20, hetero=0.2, noise=0.03):
    """Random saturated C/H/N/O chain with 1..nheavy_max heavy atoms, charge
    0 or +1 with p = 0.5.  Returns (num, xyz[Bohr], charge)."""
    rng = np.random.default_rng(seed)
    nh = int(rng.integers(1, nheavy_max + 1))
    backbone = []
    for _ in range(nh):
        if rng.random() < hetero:
            backbone.append(7 if rng.random() < 0.5 else 8)
        else:
            backbone.append(6)
    # an O cannot have two heavy neighbours and still be capped beyond valence 2: fine;
    num, xyz = build_chain(backbone, noise=noise, seed=int(rng.integers(0, 2**31)))
    charge = int(rng.random() < 0.5)
    return num, xyz, charge


def getspin(num, charge):
    """utility.f90:449-464"""
    j = int(np.sum(num)) - abs(charge)
    isp = 1 + j % 2
    return -1 if j < 1 else isp


def maxwell_boltzmann(masses, temp, seed):
    """Velocities (au) drawn from Maxwell-Boltzmann at temp (K)."""
    rng = np.random.default_rng(seed)
    sig = np.sqrt(KB * temp / np.asarray(masses))
    return rng.normal(0.0, 1.0, size=(len(masses), 3)) * sig[:, None]


def iee_poisson(nat, mean_ev_per_atom, max_ev_per_atom, seed):
    """Internal excess energy (Eh) drawn from a Poisson-shaped (Gamma, shape 3)
    distribution with the given mean per atom, truncated at max per atom."""
    rng = np.random.default_rng(seed)
    while True:
        e = rng.gamma(3.0, mean_ev_per_atom / 3.0)
        if e <= max_ev_per_atom:
            return e * nat / AUTOEV


def md_ensemble(ntraj, base_seed=1, mean_ev=0.6, max_ev=1.5, temp=500.0, noise=0.02,
                builder=octane, tadd_fs=100.0):
    """Per-trajectory starting states for the batched EI production MD:
    geometry with its own noise, Maxwell-Boltzmann velocities at temp, IEE and
    heating time tadd (fs)."""
    out = []
    for t in range(ntraj):
        seed = base_seed + t
        num, xyz = builder(noise=noise, seed=seed)
        masses = np.array([atomic_mass(int(z)) for z in num])
        velo = maxwell_boltzmann(masses, temp, seed + 7919)
        eimp = iee_poisson(len(num), mean_ev, max_ev, seed + 104729)
        out.append(dict(num=num, xyz=xyz, velo=velo, mass=masses, eimp=eimp,
                        tadd=tadd_fs * FSTOAU))
    return out
