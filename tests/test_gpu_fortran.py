"""The Fortran drop-in module (module qcxms_tblite, replacing src/tblite.f90)
called from a Fortran program the way src/iniqm.f90:644-645 calls
get_xtb_egrad, on the GPU, against the CPU oracle (|dE| <= 1e-8 Eh,
max |dg| <= 1e-6 Eh/bohr, max |dq| <= 1e-6 e)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "egrad_driver")


def run_records(tmp_path, records):
    """records: [(num, xyz, charge, mult, method, etemp, spec)]; one
    get_xtb_egrad call per record in one process.  Returns
    [(stat, energy, qat, grad)]."""
    lines = []
    for num, xyz, charge, mult, method, etemp, spec in records:
        lines.append(f"{len(num)} {charge} {mult} {method} {float(etemp):.17e} {int(spec)}")
        lines += [f"{int(z)} {float(x):.17e} {float(y):.17e} {float(w):.17e}"
                  for z, (x, y, w) in zip(num, xyz)]
    print(f"fortran driver: {len(records)} records, nat={[len(r[0]) for r in records]}", flush=True)
    p = subprocess.run([DRIVER], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert p.returncode == 0, p.stderr
    out = p.stdout.split()
    res, i = [], 0
    for num, *_ in records:
        nat = len(num)
        stat, energy = int(out[i]), float(out[i + 1])
        rest = np.array([float(v) for v in out[i + 2:i + 2 + 4 * nat]]).reshape(nat, 4)
        res.append((stat, energy, rest[:, 0], rest[:, 1:]))
        i += 2 + 4 * nat
    assert i == len(out)
    return res


def run_driver(tmp_path, num, xyz, charge, mult, method, etemp, spec=False):
    return run_records(tmp_path, [(num, xyz, charge, mult, method, etemp, spec)])[0]


@pytest.mark.parametrize("case", ["water", "octane_cation"])
def test_fortran_get_xtb_egrad_matches_oracle(tmp_path, case):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    if case == "water":
        num = np.array([8, 1, 1], np.int32)
        xyz = np.array([[0, 0, 0.1173], [0, 0.7572, -0.4692], [0, -0.7572, -0.4692]]) * W.AATOAU
        charge, mult, etemp = 0, 1, 300.0
    else:
        num, xyz = W.octane(seed=3)
        charge, mult, etemp = 1, 2, 5000.0
    stat, e, q, g = run_driver(tmp_path, num, xyz, charge, mult, 2, etemp)
    assert stat == 0
    r = ob.gfn2_egrad(num, xyz, charge, etemp, opts=ob.default_opts(econv=1e-12, pconv=1e-10))
    assert abs(e - r["energy"]) <= 1e-8
    assert np.abs(g - r["grad"]).max() <= 1e-6
    assert np.abs(q - r["qat"]).max() <= 1e-6
    assert "GFN2-xTB" in open(os.path.join(tmp_path, "xtb.last")).read()


def test_fortran_unknown_method_status(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    num = np.array([1, 1], np.int32)
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])
    stat, e, q, g = run_driver(tmp_path, num, xyz, 0, 1, 1, 300.0)
    assert stat == 5


def _check_vs_oracle(res, num, xyz, charge, etemp):
    stat, e, q, g = res
    assert stat == 0
    r = ob.gfn2_egrad(num, xyz, charge, etemp, opts=ob.default_opts(econv=1e-12, pconv=1e-10))
    assert abs(e - r["energy"]) <= 1e-8
    assert np.abs(g - r["grad"]).max() <= 1e-6
    assert np.abs(q - r["qat"]).max() <= 1e-6


def test_fortran_repeated_calls_match_oracle(tmp_path):
    """An MD-like sequence through one process: the library keeps the device
    arrays of a structure between calls with the same atoms and charge and
    rebuilds them when they change; every call matches the oracle."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    num, x0 = W.octane(seed=3)
    rng = np.random.default_rng(7)
    wnum = np.array([8, 1, 1], np.int32)
    wxyz = np.array([[0, 0, 0.1173], [0, 0.7572, -0.4692], [0, -0.7572, -0.4692]]) * W.AATOAU
    geoms = [x0 + rng.normal(scale=0.03, size=x0.shape) for _ in range(3)]
    recs = [(num, geoms[0], 1, 2, 2, 5000.0, 0), (num, geoms[1], 1, 2, 2, 5000.0, 0),
            (wnum, wxyz, 0, 1, 2, 300.0, 0), (num, geoms[2], 1, 2, 2, 5000.0, 0),
            (num, geoms[2], 0, 1, 2, 5000.0, 0)]
    out = run_records(tmp_path, recs)
    for (n, x, c, _, _, et, _), res in zip(recs, out):
        print(f"  nat {len(n)} charge {c}: E {res[1]:.10f}", flush=True)
        _check_vs_oracle(res, n, x, c, et)


def _read_tmp_mspec(path, nat):
    v = open(path).read().split()
    nao, ihomo = int(v[0]), int(v[1])
    vals = np.array([float(t) for t in v[2:]]).reshape(nao, 2 + nat)
    return nao, ihomo, vals[:, 0], vals[:, 1], vals[:, 2:]


def test_fortran_spec_calc_writes_the_mo_spectrum(tmp_path):
    """getspec's call (etemp 0, spec_calc true; mo_spec.f90:63,
    iniqm.f90:409-410) writes tmp.mspec and qcxms.Mspec.tbxtb with the
    orbital energies (eV), occupations and MO populations of the oracle's SCC
    (integer occupations: the oracle at 1e-3 K)."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    num, xyz = W.octane(noise=0.02, seed=4)
    nat = len(num)
    stat, e, q, g = run_driver(tmp_path, num, xyz, 0, 1, 2, 0.0, spec=True)
    assert stat == 0
    o = ob.gfn2_orbitals(num, xyz, 0, 1e-3, ob.default_opts(econv=1e-12, pconv=1e-10))
    qmo_ref = ob.mo_populations(o["coeff"], o["S"], o["aoat"], nat)
    nao, ihomo, emo_ev, focc, qmo = _read_tmp_mspec(os.path.join(tmp_path, "tmp.mspec"), nat)
    print(f"tmp.mspec: nao {nao} ihomo {ihomo} HOMO {emo_ev[ihomo - 1]:.6f} eV", flush=True)
    assert (nao, ihomo) == (50, 25)
    np.testing.assert_allclose(emo_ev, o["eps"] * 27.21138505, atol=1e-5)
    np.testing.assert_allclose(focc, o["focc"], atol=1e-9)
    # populations of orbitals that are not (nearly) degenerate with a neighbour
    gap = np.minimum(np.diff(o["eps"], prepend=-np.inf), np.diff(o["eps"], append=np.inf))
    keep = gap > 1e-3
    assert keep.sum() >= 40
    np.testing.assert_allclose(qmo[keep], qmo_ref[keep], atol=1e-6)
    np.testing.assert_allclose(qmo.sum(axis=1), 1.0, atol=1e-9)
    text = open(os.path.join(tmp_path, "qcxms.Mspec.tbxtb")).read().split("\n")
    assert text[0].split() == ["50", "25"]
    assert len([t for t in text if t.strip() == ""]) >= nao
