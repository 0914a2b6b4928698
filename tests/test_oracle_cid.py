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


# ---- the production MD in its CID mode (md.f90, method 3, icoll >= 1) ----
MD_ETEMP = 5000.0


def md_cid(name, **kw):
    tr, v, max_steps = CC.md_cases()[name]
    return ob.md_run(tr, 0, max_steps=max_steps, nfragexit=1, etempin=MD_ETEMP, opts=OPTS, cid=True,
                     new_velo=v, **kw)


def test_md_cid_mode_takes_out_the_flight():
    """md.f90:247-255, 466-487: with the ion's speed handed over, Ekin and
    the temperature are those of the motion inside the ion.  The internal
    dynamics do not depend on a uniform drift (the energy is translation
    invariant), so the flying ion reports the resting one's energies, its
    positions are the resting ones shifted by v t, and the speed measured
    from the centre of mass each step is the one handed over."""
    a, b = md_cid("c2_rest"), md_cid("c2_drift")
    sa, sb = a["state"], b["state"]
    assert sa.nstep == sb.nstep == 120 and sa.exit_reason == sb.exit_reason == 7
    assert abs(sa.ekin - sb.ekin) <= 1e-9 and abs(sa.epot - sb.epot) <= 1e-9
    assert abs(sa.atlast - sb.atlast) <= 1e-6 * sa.atlast
    assert abs(sb.new_velo - 1500.0) <= 1e-6 and sa.new_velo < 1e-3
    tr = CC.md_cases()["c2_drift"][0]
    shift = (tr["velo"] - CC.md_cases()["c2_rest"][0]["velo"])[0] * 120 * CC.TSTEP
    assert np.abs(b["xyz"] - a["xyz"] - shift).max() <= 1e-8
    # without the speed handed over, the flight's kinetic energy 1/2 M v^2
    # (0.013 Eh) stays in the first step's energy; Eerror is measured against
    # the running mean of Epot + Ekin (no heating: nadd = 0), so after 120
    # steps it is off by 1/2 M v^2 / 120
    c = ob.md_run(tr, 0, max_steps=120, nfragexit=1, etempin=MD_ETEMP, opts=OPTS, cid=True, new_velo=0.0)
    e_flight = 0.5 * tr["mass"].sum() * (1500.0 * CC.MSTOAU) ** 2
    assert abs(sb.eerror - sa.eerror) <= 1e-9
    assert abs((c["state"].eerror - sb.eerror) - e_flight / 120) <= 1e-10


def test_md_cid_mode_averages_the_fragments_and_extends_the_run():
    """md.f90:496-627, 672-698: a new fragment count starts the averaged
    structure over the next cnt_steps = 50 steps and moves the end to
    add_steps after it (500 steps for 11 atoms); a count that falls back
    drops the average; the EI exits at nfragexit do not apply; the structure
    md returns is the average over the 50 steps after the last split.  The
    propane ion splits at step 103, joins at 113, splits at 348, joins at
    367 and splits for good at 434."""
    r = md_cid("c3_frag", trace_steps=3000)
    st, tr = r["state"], r["trace"]
    nf = tr[:, 4].astype(int)
    assert list(np.nonzero(np.diff(nf))[0] + 2) == [103, 113, 348, 367, 434] and nf[0] == 1
    f = 434
    assert st.nstep == f + ob.oracle().orc_cid_add_steps(11) and st.exit_reason == 7
    assert st.nfrag == 2 and st.check_fragmented == 2 and st.count_average == 0
    want = r["xyz_trace"][f - 1:f + 49].mean(axis=0)
    assert np.abs(r["axyz"] - want).max() <= 1e-12


def test_later_collision_setup_closed_forms():
    """cid.f90:179-676 with icoll > 1: the ion keeps its velocities (flight
    included) and moves onto its centre of mass; the gas sits at the old
    centre + direc * max(v dt 2 step_dist, 17) plus 0.7 x the off-axis
    offsets; beta from the ion's mass; the thermal temperature from all
    velocities."""
    num, xyz, velo, mass = CC.ion(5, 4)
    xyz = xyz + np.array([0.3, -0.2, 0.5])
    direc = np.array([0.6, 0.0, 0.8])
    rand = (0.25, 0.5, 0.7, 0.2)
    for v_ms in (9000.0, 200.0):
        s = CID.cid_setup_next(num, xyz, velo, mass, CC.TSTEP, v_ms, direc, gas="N2", rand=rand)
        nuc = len(num)
        cm_old = CID.center_of_mass(mass, xyz)
        assert np.abs(CID.center_of_mass(mass, s["xyz"][:nuc])).max() < 1e-12
        assert np.array_equal(s["velo"][:nuc], velo) and np.abs(s["velo"][nuc:]).max() == 0.0
        start = max(v_ms * CC.MSTOAU * 2 * CC.TSTEP * min(2 * nuc * 10, 800) * W.AUTOAA, 17.0)
        # lmin = 0.7 picks the highest y (times f), lpos = 0.2 the lowest x (times g)
        highy, lowx = xyz[:, 1].max(), xyz[:, 0].min()
        want = cm_old + direc * start + np.array([0.7 * lowx * 0.5, 0.7 * highy * 0.25, 0.0])
        assert np.abs(s["xyz"][-1] - want).max() < 1e-12
        assert abs(s["xyz"][-2, 2] - s["xyz"][-1, 2] - 1.09 * W.AATOAU) < 1e-12
        mgas = 14.007 * W.AMUTOAU
        assert abs(s["beta"] - mgas / (mgas + mass.sum())) < 1e-15
        ek = 0.5 * float((mass[:, None] * velo * velo).sum())
        assert abs(s["tinit"] - ek / (1.5 * nuc * W.KB)) < 1e-9 * s["tinit"]
        assert abs(s["eimpact_ev"] - 0.5 * mass.sum() * (v_ms * CC.MSTOAU) ** 2 * W.AUTOEV) < 1e-12


def test_mfp_steps_and_continuation_rules():
    """main.F90:1816-1829: nmax = 100 nuc, cut to 75 / 60 / 50 % from the
    third / fourth / fifth MD on, within 1000..10000; main.F90:1683-1735: the
    continued fragment's atoms in order, their nuclear-charge-weighted centre
    at the origin, velocities kept."""
    from qcxms_amd.cid_run import continue_fragment, nmax_mfp
    assert [nmax_mfp(n, 2) for n in (5, 14, 60, 150)] == [1000, 1400, 6000, 10000]
    assert [nmax_mfp(40, s) for s in (2, 3, 4, 5, 9)] == [4000, 3000, 2400, 2000, 2000]
    assert nmax_mfp(14, 3) == 1050 and nmax_mfp(14, 4) == 1000
    num, xyz, velo, mass = CC.ion(5, 4)
    lst = np.where(np.arange(len(num)) % 3 == 0, 2, 1).astype(np.int32)
    n2, x2, v2, m2, i2 = continue_fragment(num, xyz, velo, mass, np.full(len(num), -1), lst, 2)
    sel = lst == 2
    assert np.array_equal(n2, num[sel]) and np.array_equal(v2, velo[sel]) and np.array_equal(m2, mass[sel])
    assert np.abs((n2[:, None] * x2).sum(0)).max() < 1e-12
    assert np.abs((x2 - xyz[sel]) - (x2 - xyz[sel])[0]).max() < 1e-15
    n0, x0, _, _, _ = continue_fragment(num, xyz, velo, mass, np.full(len(num), -1), lst, 0)
    assert len(n0) == len(num) and np.abs((num[:, None] * x0).sum(0)).max() < 1e-12
