"""MD runs that end through each exit of the EI production loop.

Every case is a small, deterministic trajectory chosen so that one exit of
md.f90:285-682 (or the initial egrad, md.f90:182-184) ends it.  The expected
(nstep, exit_reason, mdok, fragstate) follow from the reference's rules:

  init_gnorm   an H atom 0.25 A from its carbon: gnorm > 20 Eh/bohr at the
               first egrad, checkqc sets E = 0 (iniqm.f90:700-705) and md
               returns before the loop (md.f90:184): nstep 0, exit 1
  err1_step2   an H atom 0.4 A from its carbon moving onto it at 0.1 au: the
               egrad after the first leap-frog fails checkqc, so the second
               loop iteration sees Epot = 0 (md.f90:321): exit 1, mdok 0
  err2_kick    the same H at 0.05 au: the repulsive kick breaks energy
               conservation, |Eerror| > 0.1 at step 2 (md.f90:323): exit 2
  err2_frag    n-octane at 3 fs steps: the integration error exceeds 0.1 Eh
               after it split in two: exit 2 with mdok (1 < nfrag <= 4,
               md.f90:329)
  err2_isec2   the same at 2.5 fs, one fragment, second generation: mdok
               because isec > 1 (md.f90:329)
  nfrag7       seven methane molecules 8 A apart: nfrag = 7 > 6 at the first
               connectivity (md.f90:456-462): exit 3, mdok 0
  morestep     two methanes 7 A apart with nfragexit 2: nfrag = nfragexit for
               more than 250 steps (md.f90:657-668): exit 6 at step 251
  morestep_isec2  the same with nfragexit 3 in the second generation, where
               nfragexit becomes 2 (main.F90:2255): exit 6 at step 251
  fconst       the same with nfragexit 3: two fragments for more than 1000
               steps (md.f90:647-654): exit 5 at step 1001, fragstate 2
  heat_fail    n-octane nearly at rest asked to take up 5 Eh within 1 fs:
               impactscale finds no scaling in 20000 increments (impact.f90:
               42-50, the reference stops): exit 8 at step 1
  split_ieetemp  ethane cation with its C-C bond stretched to 2.1 A and the
               methyls pulled apart, heated over 40 fs with ieetemp 3000 K/Eh:
               it splits at step 1, re-joins at step 6 and splits again at
               step 27, all while heating, so setetemp must see the previous
               step's nfrag (md.f90:443-451)
"""
import numpy as np

from qcxms_amd import workloads as W

A = W.AATOAU
FS = W.FSTOAU


def _traj(num, xyz, velo, eimp=0.0, tadd=0.0, **extra):
    num = np.asarray(num, dtype=np.int32)
    mass = np.array([W.atomic_mass(int(z)) for z in num])
    d = dict(num=num, xyz=np.asarray(xyz, float), velo=np.asarray(velo, float), mass=mass,
             eimp=float(eimp), tadd=float(tadd))
    d.update(extra)
    return d


def _masses(num):
    return np.array([W.atomic_mass(int(z)) for z in num])


def _methanes(centres):
    n1, x1 = W.build_chain([6], noise=0.0)
    nums, xyzs = [], []
    for c in centres:
        nums.append(n1)
        xyzs.append(x1 + np.asarray(c, float) * A)
    return np.concatenate(nums), np.concatenate(xyzs)


def _octane_h_on_carbon(dist_aa, speed):
    num, xyz = W.octane(seed=0)
    xyz = xyz.copy()
    d = xyz[8] - xyz[0]
    u = d / np.linalg.norm(d)
    xyz[8] = xyz[0] + u * dist_aa * A
    v = W.maxwell_boltzmann(_masses(num), 300.0, 1)
    v[8] = -u * speed
    return num, xyz, v


def _split_ethane():
    n, x = W.build_chain([6, 6], noise=0.0)
    x = x.copy()
    v = W.maxwell_boltzmann(_masses(n), 300.0, 3)
    ax = (x[1] - x[0]) / np.linalg.norm(x[1] - x[0])
    shift = (2.1 - 1.54) * A / 2
    for i in range(len(n)):
        left = i == 0 or (i >= 2 and np.linalg.norm(x[i] - x[0]) < np.linalg.norm(x[i] - x[1]))
        s = -1.0 if left else 1.0
        x[i] += s * shift * ax
        v[i] += s * 0.0006 * ax
    return _traj(n, x, v, 0.3, 40 * FS)


def cases():
    """name -> (trajectory, md keyword arguments, expected
    (nstep, exit_reason, mdok, fragstate) or None where only the oracle
    fixes it)."""
    out = {}
    num, xyz, v = _octane_h_on_carbon(0.25, 0.0)
    v = W.maxwell_boltzmann(_masses(num), 300.0, 1)
    out["init_gnorm"] = (_traj(num, xyz, v, 0.1, 100 * FS), dict(max_steps=20), (0, 1, 0, 0))
    num, xyz, v = _octane_h_on_carbon(0.4, 0.1)
    out["err1_step2"] = (_traj(num, xyz, v), dict(max_steps=40), (2, 1, 0, 0))
    num, xyz, v = _octane_h_on_carbon(0.4, 0.05)
    out["err2_kick"] = (_traj(num, xyz, v), dict(max_steps=40), (2, 2, 0, 0))
    num, xyz = W.octane(seed=0)
    out["err2_frag"] = (_traj(num, xyz, W.maxwell_boltzmann(_masses(num), 1500.0, 1), 0.2, 5 * FS),
                        dict(max_steps=200, tstep=3.0 * FS), (6, 2, 1, 0))
    out["err2_isec2"] = (_traj(num, xyz, W.maxwell_boltzmann(_masses(num), 500.0, 1), isec=2),
                         dict(max_steps=300, tstep=2.5 * FS), (50, 2, 1, 0))
    ring = [[8 * np.cos(2 * np.pi * k / 7), 8 * np.sin(2 * np.pi * k / 7), 0.0] for k in range(7)]
    num, xyz = _methanes(ring)
    out["nfrag7"] = (_traj(num, xyz, W.maxwell_boltzmann(_masses(num), 300.0, 1)),
                     dict(max_steps=20), (1, 3, 0, 0))
    num, xyz = _methanes([[0, 0, 0], [7, 0, 0]])
    v = W.maxwell_boltzmann(_masses(num), 200.0, 1)
    out["morestep"] = (_traj(num, xyz, v), dict(max_steps=2000, nfragexit=2), (251, 6, 1, 1))
    out["morestep_isec2"] = (_traj(num, xyz, v, isec=2), dict(max_steps=2000, nfragexit=3),
                             (251, 6, 1, 1))
    out["fconst"] = (_traj(num, xyz, v), dict(max_steps=2000, nfragexit=3), (1001, 5, 1, 2))
    num, xyz = W.octane(seed=0)
    v = 1e-6 * W.maxwell_boltzmann(_masses(num), 300.0, 1)
    out["heat_fail"] = (_traj(num, xyz, v, 5.0, 1.0 * FS), dict(max_steps=20), (1, 8, 0, 0))
    out["split_ieetemp"] = (_split_ethane(), dict(max_steps=60, ieetemp=3000.0), (60, 7, 1, 1))
    return out
