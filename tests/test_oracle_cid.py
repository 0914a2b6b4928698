"""The oracle's CID collision MD (cid_oracle.c, cid.f90:724-1111) and the
host set-up (qcxms_amd.cid.cid_setup, cid.f90:179-676) against the rules of
the reference:
  - the set-up centres the ion, gives it the flight kinetic energy E_LAB
    (1/2 M v^2 = E_LAB exactly), puts the gas straight ahead at the reference's
    distance and offsets, and beta = m_gas / (m_gas + M);
  - a collision is five receding distance tests, so it falls on a multiple
    of 10 steps at least 50 steps in; the run then ends 800 steps of 1 fs
    later, or add_steps after the first fragmentation (cid.f90:232-234);
  - a failed gradient (checkqc) ends the run before the first step."""
import math

import numpy as np

import cid_cases as CC
import oracle_bind as ob
from qcxms_amd import cid as CID
from qcxms_amd import workloads as W

OPTS = ob.default_opts(econv=1e-10, pconv=1e-8)


def test_setup_closed_forms():
    num, xyz, velo, mass = CC.ion(3, 3)
    s = CID.cid_setup(num, xyz, velo, mass, CC.TSTEP, gas="N2", elab_ev=40.0, rand=(0.2, 0.3, 0.7, 0.2))
    nuc = len(num)
    assert s["nion"] == nuc and len(s["num"]) == nuc + 2 and list(s["num"][nuc:]) == [7, 7]
    m = s["mass"][:nuc]
    cm = CID.center_of_mass(m, s["xyz"][:nuc])
    assert np.abs(cm).max() < 1e-12
    # the flight: every ion atom gets the same velocity on top of its thermal one
    dv = s["velo"][:nuc] - velo
    assert np.abs(dv - dv[0]).max() < 1e-15
    assert abs(0.5 * m.sum() * float(dv[0] @ dv[0]) - 40.0 / W.AUTOEV) < 1e-12
    mgas = 14.007 * W.AMUTOAU
    assert abs(s["beta"] - mgas / (mgas + m.sum())) < 1e-15
    # gas straight ahead along +z, the N-N distance 1.09 A, at rest
    gas = s["xyz"][nuc:]
    assert abs(gas[0, 2] - gas[1, 2] - 1.09 * W.AATOAU) < 1e-12 and np.allclose(gas[0, :2], gas[1, :2])
    assert np.abs(s["velo"][nuc:]).max() == 0.0
    fasti = math.sqrt(2 * 40.0 / W.AUTOEV / m.sum())
    start = max(min(2 * nuc * 10, 800) * fasti * 2 * CC.TSTEP * W.AUTOAA, 10.0)
    assert abs(gas[1, 2] - start) < 1e-9
    # off-axis offsets: 0.8 x (extent x random factor) on the side the coin picks
    x0 = s["xyz"][:nuc]
    assert abs(gas[1, 0] - 0.8 * (x0[:, 0].min() * 0.3)) < 1e-12
    assert abs(gas[1, 1] - 0.8 * (x0[:, 1].max() * 0.2)) < 1e-12


def test_add_steps_rule():
    L = ob.oracle()
    assert [L.orc_cid_add_steps(n) for n in (5, 10, 11, 25, 39, 40, 57)] == [0, 0, 500, 1000, 1500, 4000, 5000]


def test_collision_then_fragmentation_extension():
    s = CC.cases()["c3_40eV"]
    st = ob.cid_run(s, charge=1, opts=OPTS)["state"]
    assert st.collision_step % 10 == 0 and st.collision_step >= 50
    assert st.fragmented == 1 and st.nfrag >= 2
    # fragmentation no later than the collision's own end; the run ends add_steps after it
    assert st.total_steps < st.collision_step + 800
    assert st.nstep == st.total_steps
    assert (st.total_steps - ob.oracle().orc_cid_add_steps(s["nion"])) >= st.collision_step


def test_collision_without_fragmentation_runs_800_steps_more():
    """A slow head-on hit that leaves the ion whole: 800 more steps of 1 fs
    after the collision (nint(800 * 2 dt / fs), dt = 0.5 fs)."""
    num, xyz, velo, mass = CC.ion(11, 2)
    s = CID.cid_setup(num, xyz, velo * 0.0, mass, CC.TSTEP, gas="N2", elab_ev=5.0, rand=(0.0, 0.0, 0.7, 0.7))
    st = ob.cid_run(s, charge=1, opts=OPTS, ntot=3000)["state"]
    assert st.collided == 1 and st.fragmented == 0 and st.nfrag == 1
    assert st.total_steps == st.collision_step + 800 and st.nstep == st.total_steps


def test_failed_gradient_stops_before_the_first_step():
    s = dict(CC.cases()["c2_60eV"])
    s["xyz"] = s["xyz"].copy()
    s["xyz"][-1] = s["xyz"][s["nion"] - 1] + np.array([0.2, 0.0, 0.0])
    st = ob.cid_run(s, charge=1, opts=OPTS)["state"]
    assert st.gradfail == 1 and st.nstep == 0
