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
DRIVER = os.path.join(ROOT, "qcxms_amd", "fortran", "_build", "egrad_driver")


def run_driver(tmp_path, num, xyz, charge, mult, method, etemp):
    lines = [f"{len(num)} {charge} {mult} {method} {float(etemp):.17e}"]
    lines += [f"{int(z)} {float(x):.17e} {float(y):.17e} {float(w):.17e}"
              for z, (x, y, w) in zip(num, xyz)]
    print(f"fortran driver: nat={len(num)} method={method}", flush=True)
    p = subprocess.run([DRIVER], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert p.returncode == 0, p.stderr
    out = p.stdout.split()
    stat = int(out[0])
    energy = float(out[1])
    rest = np.array([float(v) for v in out[2:]]).reshape(len(num), 4)
    return stat, energy, rest[:, 0], rest[:, 1:]


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
