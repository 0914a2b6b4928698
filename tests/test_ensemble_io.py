"""Ensemble inputs/outputs (qcxms_amd/ensemble.py) against the reference:
qcxms.start and qcxms.gs text byte-identical to the Fortran runtime executing
the reference's write statements, readers that return the written values,
the IEE model and heating times against the compiled src/iee.f90 and
src/impact.f90, a TMPQCXMS tree round trip, and the Box-Muller spread."""
import json
import os

import numpy as np
import pytest

from qcxms_amd import ensemble as E
from qcxms_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ensemble_golden.json")))


def test_start_file_matches_fortran_and_reads_back(tmp_path):
    for c in GOLD["start"]:
        p = tmp_path / "qcxms.start"
        E.write_start(p, c["traj"], c["velo"], c["velof"], c["eimp"], c["tadd"])
        assert p.read_text() == c["text"]
        itrj, eimp, tadd, velo, velof = E.read_start(p, len(c["velof"]))
        assert itrj == c["traj"]
        assert eimp == pytest.approx(c["eimp"], rel=1e-13, abs=0)
        assert tadd == pytest.approx(c["tadd"], rel=1e-13, abs=0)
        np.testing.assert_allclose(velo, c["velo"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(velof, c["velof"], rtol=1e-13, atol=0)


def test_gs_file_matches_fortran_and_reads_back(tmp_path):
    for c in GOLD["gs"]:
        x, v = np.array(c["xyz"]), np.array(c["velo"])
        p = tmp_path / "qcxms.gs"
        E.write_gs(p, c["nmax"], x, v)
        got, want = p.read_text().splitlines(), c["text"].splitlines()
        assert int(got[0]) == int(want[0]) == c["nmax"]  # list-directed header
        assert got[1:] == want[1:]
        nd, xr, vr = E.read_gs(p, x.shape[1])
        assert nd == c["nmax"]
        np.testing.assert_allclose(xr, x, rtol=1e-7, atol=1e-15)
        np.testing.assert_allclose(vr, v, rtol=1e-7, atol=1e-15)


def test_iee_model_matches_compiled_reference():
    for a, b, el, x, p in GOLD["gauss0"]:
        assert E.gauss0(a, b, el, x) == pytest.approx(p, rel=1e-14)
    for a, b, el, x, p in GOLD["poiss0"]:
        assert E.poiss0(a, b, el, x) == pytest.approx(p, rel=1e-13)
    for a, b, el, ityp, exc, emax, pmax, eavg in GOLD["getmaxiee"]:
        got = E.getmaxiee(a, b, el, ityp, exc)
        assert got[0] == pytest.approx(emax, abs=1e-12)
        assert got[1] == pytest.approx(pmax, rel=1e-12)
        assert got[2] == pytest.approx(eavg, rel=1e-12)
    for el, ityp, exc, nbnd, atm, a, b in GOLD["getieeab"]:
        ga, gb = E.getieeab(el, ityp, exc, nbnd, atm)
        assert ga == pytest.approx(a, abs=1e-12) and gb == pytest.approx(b, abs=1e-11)
    for n, na, i, tr, emo, t in GOLD["calctrelax"]:
        assert E.calctrelax(emo, na, i, tr) == pytest.approx(t, rel=1e-13, abs=0)


def test_tmpqcxms_round_trip_feeds_md_inputs(tmp_path):
    trajs = W.md_ensemble(3, base_seed=5)
    for k, t in enumerate(trajs, start=1):
        E.write_trajectory_inputs(tmp_path, k, t["num"], t["xyz"], t["velo"],
                                  np.ones(len(t["num"])), t["eimp"], t["tadd"])
    assert sorted(os.listdir(tmp_path)) == ["TMP.1", "TMP.2", "TMP.3"]
    back = E.load_tmpqcxms(tmp_path)
    for t, b in zip(trajs, back):
        assert np.array_equal(b["num"], t["num"])
        np.testing.assert_allclose(b["xyz"], t["xyz"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(b["velo"], t["velo"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(b["mass"], t["mass"], rtol=1e-15)
        assert b["eimp"] == pytest.approx(t["eimp"], rel=1e-13)


def test_box_muller_spread():
    rng = np.random.default_rng(3)
    e = np.array([E.vary_energies(2.0, 0.1, rng) for _ in range(40000)])
    assert abs(e.mean() - 2.0) < 0.003 and abs(e.std() - 0.2) < 0.003


def test_fortran_d_edge_cases():
    assert E.fortran_d(0.0, 22, 14) == "  0.00000000000000D+00"
    assert E.fortran_d(-1.0, 16, 8) == " -0.10000000D+01"
    assert E.fortran_d(1e-120, 16, 8) == "  0.10000000-119"
    assert E.fortran_i(123456, 4) == "****"


def test_momap_finds_best_orbital_triple():
    """With few orbitals the 5001 random trials cover every (mo1, mo2, vmo)
    combination, so momap returns the exhaustive optimum."""
    rng = np.random.default_rng(1)
    emo = np.array([0.9, 0.7, 0.55, 0.5, -0.05, -0.1, -0.2])  # ihomo = 4
    ihomo = 4
    for target in (0.4, 1.1, 0.62):
        best, arg = np.inf, None
        for m1 in range(1, 5):
            for m2 in range(1, 5):
                for v in range(5, 5 + ihomo // 2):
                    m2e = 0 if m2 > ihomo // 2 else m2
                    d = emo[m1 - 1] + (emo[m2 - 1] if m2e else 0.0) + emo[v - 1]
                    if abs(d - target) < best:
                        best, arg = abs(d - target), (m1, m2e)
        m1, m2 = E.momap(ihomo, emo, target, rng)
        d = emo[m1 - 1] + (emo[m2 - 1] if m2 else 0.0)
        dvs = [abs(d + emo[v - 1] - target) for v in range(5, 7)]
        assert min(dvs) == pytest.approx(best, abs=1e-15)


def _toy_spec(nat, nao):
    rng = np.random.default_rng(4)
    ihomo = nao // 2
    eps = np.sort(rng.uniform(-1.0, 0.3, nao))
    eps[:ihomo] -= 0.2
    pop = rng.random((nao, nat))
    return dict(emo=-eps, ehomo=float(-eps[ihomo - 1]), ihomo=ihomo, nb=ihomo,
                mopop=pop / pop.sum(axis=1, keepdims=True))


def test_generate_ensemble_invariants():
    num, xyz = W.octane(noise=0.0)
    nat = len(num)
    rng = np.random.default_rng(11)
    gx = xyz[None] + rng.normal(0, 0.01, (60, nat, 3))
    gv = rng.normal(0, 1e-4, (60, nat, 3))
    spec = _toy_spec(nat, 50)
    trajs, (a, b, pmax) = E.generate_ensemble(num, gx, gv, 20, spec, rng)
    assert len(trajs) == 20 and a > 0 and b > 0 and pmax > 0
    picks = {tuple(t["xyz"][0]) for t in trajs}
    assert len(picks) == 20  # distinct snapshots
    exc_max = (70.0 / E.AUTOEV) * 1.6
    for t in trajs:
        assert t["velof"].max() == pytest.approx(1.0)  # normalised (26 atoms <= 35)
        assert 0.0 < t["eimp"] < exc_max
        assert t["tadd"] >= 200.0 * E.FSTOAU - 1e-9  # at least trelax / 10
    # more than 35 atoms: uniform heating factors
    n2, x2 = W.chain_c12h24o2n2(seed=0)
    spec2 = _toy_spec(len(n2), 88)
    gx2 = x2[None] + rng.normal(0, 0.01, (12, len(n2), 3))
    t2, _ = E.generate_ensemble(n2, gx2, np.zeros_like(gx2), 5, spec2, rng)
    assert all(np.all(t["velof"] == 1.0) for t in t2)


def test_iee_rejection_sampling_follows_distribution():
    """The accepted excess energies follow P(E) on [0, E_imp - E_homo]: their
    mean equals the distribution mean of getmaxiee (to sampling error)."""
    num, xyz = W.octane(noise=0.0)
    nat = len(num)
    spec = _toy_spec(nat, 50)
    spec["ehomo"] = 0.4
    rng = np.random.default_rng(5)
    gx = np.repeat(xyz[None], 1201, axis=0)
    trajs, (a, b, pmax) = E.generate_ensemble(num, gx, np.zeros_like(gx), 600, spec, rng,
                                              eimpw=0.0)
    exc = (70.0 / E.AUTOEV - 0.4) * E.AUTOEV
    x = np.arange(0.001, exc, 0.01)
    p = np.array([E.poiss0(a, b, 25.0 + 25.0, v) for v in x])
    mean = (p * x).sum() / p.sum()
    got = np.mean([t["eimp"] * E.AUTOEV for t in trajs])
    assert got == pytest.approx(mean, rel=0.05)
