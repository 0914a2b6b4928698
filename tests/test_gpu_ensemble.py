"""Ground-state sampling for qcxms.gs on the GPU and QCxMS-prepared inputs
driving the batched MD.

- sample_ground_state: snapshots every 4 steps of an MD without excess
  energy equal the oracle MD stopped after the same number of steps
  (|dx| <= 1e-8 bohr, |dv| <= 1e-10 au), and the qcxms.gs written from them
  reads back.
- A TMPQCXMS tree (start.xyz + qcxms.start, written as wrstart does) loaded
  with load_tmpqcxms runs the batched MD to the same state as the oracle MD
  started from the values read back.
"""
import numpy as np
import pytest

import oracle_bind as ob
from qcxms_amd import ensemble as E
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu
TIGHT = dict(econv=1e-12, pconv=1e-10)


def test_ground_state_snapshots_match_oracle(tmp_path):
    tr = W.md_ensemble(1, base_seed=9)[0]
    xs, vs = E.sample_ground_state(tr, 12, dumpstep=4, use_graph=False, econv=1e-11, pconv=1e-9)
    assert xs.shape == (3, len(tr["num"]), 3)
    g0 = dict(tr, eimp=0.0, tadd=0.0)
    for k in range(3):
        ref = ob.md_run(g0, steps_limit=4 * (k + 1), max_steps=13, opts=ob.default_opts(**TIGHT))
        dx = np.abs(xs[k] - ref["xyz"]).max()
        dv = np.abs(vs[k] - ref["velo"]).max()
        print(f"snapshot {k}: dx {dx:.2e} dv {dv:.2e}", flush=True)
        assert dx <= 1e-8 and dv <= 1e-10
    p = tmp_path / "qcxms.gs"
    E.write_gs(p, len(xs), xs, vs)
    nd, xr, vr = E.read_gs(p, len(tr["num"]))
    assert nd == 3
    np.testing.assert_allclose(xr, xs, rtol=1e-7, atol=1e-14)


def test_tmpqcxms_inputs_drive_md(tmp_path):
    trajs = W.md_ensemble(2, base_seed=13, tadd_fs=10.0)
    for k, t in enumerate(trajs, start=1):
        E.write_trajectory_inputs(tmp_path, k, t["num"], t["xyz"], t["velo"],
                                  np.ones(len(t["num"])), t["eimp"], t["tadd"])
    loaded = E.load_tmpqcxms(tmp_path)
    import qcxms_amd
    md = qcxms_amd.MdEnsemble(loaded, max_steps=6, use_graph=True, econv=1e-11, pconv=1e-9)
    md.run(8)
    out = md.read()
    md.close()
    for t, tr in enumerate(loaded):
        ref = ob.md_run(tr, steps_limit=0, max_steps=6, opts=ob.default_opts(**TIGHT))
        sl = slice(md.offsets[t], md.offsets[t + 1])
        assert out["results"][t].nstep == ref["state"].nstep == 6
        assert np.abs(out["xyz"][sl] - ref["xyz"]).max() <= 1e-8


def test_getspec_orbitals_consistent_with_scc():
    """Orbitals read back after a GPU single point (etemp 1e-3 K): C^T S C = 1,
    occupations sum to the electron count, orbital energies ascending, and
    the Mulliken charges rebuilt from orbitals and occupations equal the
    SCC's own charges; the MO populations are normalised per orbital."""
    import qcxms_amd
    num, xyz = W.octane(noise=0.02, seed=4)
    b = qcxms_amd.Batch([(num, 0)])
    b.set_warm_start(0)
    e, g, qat, st = b.egrad(xyz, E.SPEC_ETEMP)
    eps, focc, C, S, aoat = b.orbitals(0)
    b.close()
    assert st[0] == 0
    n = len(eps)
    assert n == 50 and np.all(np.diff(eps) >= 0)
    np.testing.assert_allclose(C.T @ S @ C, np.eye(n), atol=1e-9)
    assert focc.sum() == pytest.approx(50.0, abs=1e-9)
    P = (C * focc) @ C.T
    pop = (P * S).sum(axis=1)
    zval = np.array([1.0 if z == 1 else 4.0 for z in num])
    q = zval - np.array([pop[aoat == a + 1].sum() for a in range(len(num))])
    assert np.abs(q - qat).max() <= 1e-6
    spec = E.getspec(num, xyz)
    assert spec["ihomo"] == 25 and spec["nb"] == 25
    np.testing.assert_allclose(spec["mopop"].sum(axis=1), 1.0, atol=1e-12)
    assert spec["ehomo"] == pytest.approx(-eps[24], abs=1e-8)


def test_generated_ensemble_runs_on_gpu():
    """Ground-state snapshots and the MO spectrum from the GPU feed
    generate_ensemble; the trajectories run in the batched MD."""
    import qcxms_amd
    tr = W.md_ensemble(1, base_seed=2)[0]
    xs, vs = E.sample_ground_state(tr, 40, dumpstep=2)
    spec = E.getspec(tr["num"], tr["xyz"])
    trajs, (a, b, pmax) = E.generate_ensemble(tr["num"], xs, vs, 8, spec,
                                              np.random.default_rng(0))
    for t in trajs:
        t["charge"] = 1
    md = qcxms_amd.MdEnsemble(trajs, max_steps=5, use_graph=True)
    md.run(6)
    res = md.read()["results"]
    md.close()
    print(f"IEE a={a:.3f} b={b:.3f}; eimp (eV) {[round(t['eimp'] * E.AUTOEV, 2) for t in trajs]}",
          flush=True)
    assert all(r.nstep == 5 and r.exit_reason == 7 for r in res)
