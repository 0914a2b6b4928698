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
