"""Split-chain geometries for the post-MD fragment analysis tests: an
all-anti chain cut between backbone atoms and pulled apart along its axis."""
import numpy as np

from qcxms_amd import workloads as W


def split_chain(num, xyz, cuts, gap_bohr, noise=0.0, seed=0):
    """Translate every atom beyond each cut (between heavy atoms c and c+1,
    by x coordinate) by gap_bohr along the chain axis."""
    xyz = np.array(xyz, dtype=np.float64)
    heavy = np.nonzero(np.asarray(num) != 1)[0]
    for c in sorted(cuts, reverse=True):
        xcut = 0.5 * (xyz[heavy[c], 0] + xyz[heavy[c + 1], 0])
        xyz[xyz[:, 0] > xcut, 0] += gap_bohr
    if noise:
        xyz = xyz + np.random.default_rng(seed).normal(0.0, noise, xyz.shape)
    return np.asarray(num, dtype=np.int32), xyz


def octane_fragments(seed, ncut=1, gap=6.0):
    """C8H18 with ncut random backbone cuts, gaps of gap Angstrom."""
    rng = np.random.default_rng(seed)
    num, xyz = W.octane(noise=0.02, seed=seed)
    cuts = rng.choice(np.arange(0, 7), size=ncut, replace=False)
    return split_chain(num, xyz, cuts, gap * W.AATOAU)


def masses(num):
    return np.array([W.atomic_mass(int(z)) for z in num])
