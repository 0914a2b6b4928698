"""Kernels of libqcxms_gpu.so executed on the CPU (tests/cpu_emu: one host
thread per GPU thread, built with the host AddressSanitizer and
UndefinedBehaviorSanitizer, static and dynamic LDS filled with NaN bytes at
every block start) through qcxms_egrad_batch, against the oracle.

Covers, without a GPU: every global and LDS access of k_setup, k_scc,
k_dblocks and k_grad stays inside its buffer (also for SCC launches over a
range of the batch's systems), no value is read before it is written, and
the kernel arithmetic (including the FP64 MFMA tile layout and the wavefront
reductions) gives the oracle's energy, gradient and charges.  Systems: a
3-atom neutral (LDS-resident matrices), a 5-atom doublet cation and a
29-atom C/H/N/O cation of 59 basis functions (global-memory matrix path),
one batch.  Bars as the GPU parity tests: |dE| <= 1e-8 Eh, max |dg| <= 1e-6
Eh/bohr, max |dq| <= 1e-6 e.
"""
import numpy as np
import pytest

import emu_bind as eb
import oracle_bind as ob
from qcxms_amd import workloads as W


@pytest.mark.timeout(900)
def test_emulated_kernels_bounds_and_parity():
    eb.build()
    systems, coords = [], []
    for s in (5026, 5004, 5065):
        num, xyz, charge = W.random_chain(s)
        systems.append((num, charge, W.getspin(num, charge)))
        coords.append(xyz)
    rc, e, g, q, st, err = eb.egrad_batch(systems, coords, 5000.0)
    assert rc == 0 and np.all(st == 0), err
    re, rg, rq, rst = ob.gfn2_egrad_batch([s[:2] for s in systems], coords, 5000.0,
                                          opts=ob.default_opts(econv=1e-12, pconv=1e-10))
    assert np.all(rst == 0)
    de = np.abs(e - re).max()
    dg = max(np.abs(g[i] - rg[i]).max() for i in range(len(systems)))
    dq = max(np.abs(q[i] - rq[i]).max() for i in range(len(systems)))
    print(f"emulated kernels: dE={de:.2e} dg={dg:.2e} dq={dq:.2e}")
    assert de <= 1e-8 and dg <= 1e-6 and dq <= 1e-6


@pytest.mark.timeout(900)
def test_emulated_binned_plan_with_stored_derivative_blocks():
    """An SCC launch plan that puts each of three systems (12, 8 and 6 basis
    functions, largest first) into its own launch, with LDS sized for that
    launch and the middle one's matrices in global memory (workgroup 0 of a
    launch is not system 0), and the gradient from stored derivative blocks
    (k_dblocks): every access in bounds, and the oracle's results for every
    system.  (The eigensolver choice per bin is covered on the GPU; the MD
    emulation below covers the MFMA refinement.)"""
    eb.build()
    systems, coords = [], []
    for s in (5045, 5004, 5026):
        num, xyz, charge = W.random_chain(s)
        systems.append((num, charge, W.getspin(num, charge)))
        coords.append(xyz)
    naos = [int(sum(1 if z == 1 else 4 for z in s[0])) for s in systems]
    assert naos == sorted(naos, reverse=True) and len(set(naos)) == 3
    table = [(naos[2], 0, 1), (naos[1], 0, 0), (256, 0, 1)]
    rc, e, g, q, st, err = eb.egrad_batch(systems, coords, 5000.0, stored=True, tuning=table)
    assert rc == 0 and np.all(st == 0), err
    re, rg, rq, rst = ob.gfn2_egrad_batch([s[:2] for s in systems], coords, 5000.0,
                                          opts=ob.default_opts(econv=1e-12, pconv=1e-10))
    de = np.abs(e - re).max()
    dg = max(np.abs(g[i] - rg[i]).max() for i in range(len(systems)))
    dq = max(np.abs(q[i] - rq[i]).max() for i in range(len(systems)))
    print(f"emulated binned plan + stored blocks (naos {naos}): dE={de:.2e} dg={dg:.2e} dq={dq:.2e}")
    assert de <= 1e-8 and dg <= 1e-6 and dq <= 1e-6


@pytest.mark.timeout(900)
def test_emulated_md_heating_fused_and_stepwise():
    """Three MD steps inside the heating window (impactscale every step) of a
    5-atom and a 3-atom trajectory through the emulated library: one
    persistent launch (k_md_fused) and the per-step kernels give identical
    bits, and both follow the oracle's trajectory."""
    eb.build()
    trajs = []
    for s in (5004, 5026):
        num, xyz, charge = W.random_chain(s)
        mass = np.array([W.atomic_mass(int(z)) for z in num])
        trajs.append(dict(num=num, xyz=xyz, velo=W.maxwell_boltzmann(mass, 800.0, s), mass=mass,
                          eimp=0.05, tadd=100.0 * W.FSTOAU, charge=charge))
    steps = 3
    a = eb.md_run(trajs, steps, fused=0)
    b = eb.md_run(trajs, steps, fused=1)
    assert a[0] == 0 and b[0] == 0
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    assert [r[:2] for r in a[5]] == [r[:2] for r in b[5]]
    off = np.cumsum([0] + [len(t["num"]) for t in trajs])
    for t, tr in enumerate(trajs):
        r = ob.md_run(tr, steps_limit=0, max_steps=steps, charge=tr["charge"],
                      opts=ob.default_opts(econv=1e-12, pconv=1e-10))
        assert r["state"].nstep == a[5][t][0] == steps
        sl = slice(off[t], off[t + 1])
        dx = np.abs(a[2][sl] - r["xyz"]).max()
        dv = np.abs(a[3][sl] - r["velo"]).max()
        print(f"emulated md traj {t}: dx={dx:.2e} dv={dv:.2e}")
        assert dx < 1e-8 and dv < 1e-8
