"""Post-MD fragment analysis through the library (all fragment single points
of all trajectories in one GPU batch) against the oracle restatement of
analyse / manage_fragments (oracle/postmd_oracle.c).

Cases: C8H18+ chains cut at one to three backbone bonds and pulled 4-8 A
apart (2-4 fragments, 26 atoms), a CH3 + H split whose H+ ion takes eself,
doubly charged parents (the IP-ranking branch), a single-fragment run, and
the state of a short GPU MD run (final and averaged geometry, averaged
charges) handed over as a finished trajectory.
Bars: IPs within 1e-6 eV (two energies within 1e-8 Eh each), fragment
charges within 1e-6 e; labels, fragment counts, continued fragment, new
total charge and the integer fields of every qcxms.res record bit-exact.
"""
import numpy as np
import pytest

import oracle_bind as ob
import post_md_cases as PC
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu

IP_TOL, Q_TOL = 1e-6, 1e-6
TIGHT = dict(econv=1e-12, pconv=1e-10)


def q():
    import qcxms_amd
    return qcxms_amd


def _cases():
    out = []
    for s in range(12):
        num, xyz = PC.octane_fragments(100 + s, ncut=1 + s % 3, gap=4.0 + (s % 5))
        out.append((num, xyz, 1 if s % 4 else 2))
    num = np.array([6, 1, 1, 1, 1], dtype=np.int32)
    xyz = np.array([[0, 0, 0], [12.0, 0, 0], [-0.69, 1.94, 0], [-0.69, -0.97, 1.68],
                    [-0.69, -0.97, -1.68]], dtype=np.float64)
    out.append((num, xyz, 1))
    num, xyz = PC.octane_fragments(7, 1, gap=0.0)
    out.append((num, xyz, 1))
    return out


def test_fragment_ips_match_oracle():
    cases = _cases()
    lists, nfs = [], []
    for num, xyz, _ in cases:
        lst, nf, _ = q().fragments_batch([num], [xyz], [PC.masses(num)])
        lists.append(lst[0])
        nfs.append(int(nf[0]))
    fip, ok, st = q().fragment_ips_batch([c[0] for c in cases], [c[1] for c in cases], lists,
                                         [c[2] for c in cases])
    worst = 0.0
    for t, (num, xyz, mchrg) in enumerate(cases):
        rf, rok, rrc = ob.analyse_ips(num, xyz, lists[t], nfs[t], mchrg, opts=ob.default_opts(**TIGHT))
        assert st[t] == rrc == 0
        assert ok[t] == rok
        got = fip[t, :nfs[t], :abs(mchrg)] if nfs[t] > 1 else np.zeros((1, abs(mchrg)))
        worst = max(worst, np.abs(got - rf).max())
        if nfs[t] > 1:
            assert np.all(got[:, 0] > 0)
    print(f"{len(cases)} trajectories, fragments {sorted(set(nfs))}: max |dIP| = {worst:.2e} eV",
          flush=True)
    assert worst <= IP_TOL


def _check_report(rep, lst, ref):
    assert np.array_equal(lst, ref["list"])
    assert (rep.nfrag, rep.tcont, rep.mchrg, rep.nfrag_ok, rep.ipok) == \
        (ref["nfrag"], ref["tcont"], ref["mchrg"], ref["nfrag_ok"], ref["ipok"])
    assert np.abs(np.array(rep.fragchrg3) - ref["fragchrg3"]).max() <= Q_TOL
    assert abs(rep.chrgcont - ref["chrgcont"]) <= Q_TOL
    assert np.abs(np.array(rep.fragchrg) - ref["fragchrg"]).max() <= 1e-12
    got = rep.records() + ([rep.asave.decode()] if rep.asave else [])
    want = ref["res"] + ([ref["asave"]] if ref["asave"] else [])
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert a[10:] == b[10:], (a, b)  # integer fields and element counts
        assert abs(float(a[:10]) - float(b[:10])) <= Q_TOL + 1e-7


def test_manage_fragments_match_oracle():
    cases = _cases()
    nums = [c[0] for c in cases]
    xyzs = [c[1] for c in cases]
    qats = [np.linspace(-0.2, 0.3, len(n)) for n in nums]
    masses = [PC.masses(n) for n in nums]
    atlast = np.linspace(900.0, 2400.0, len(cases))
    itrj = np.arange(1, len(cases) + 1)
    mchrg = [c[2] for c in cases]
    chrgcont = [1.0 if m == 1 else 2.0 for m in mchrg]
    lists, reps = q().manage_fragments_batch(nums, xyzs, xyzs, qats, masses, atlast, itrj, 1, 3,
                                             1.0, mchrg, chrgcont)
    for t in range(len(cases)):
        ref = ob.manage_fragments(nums[t], xyzs[t], xyzs[t], qats[t], masses[t], atlast[t],
                                  int(itrj[t]), 1, 3, 1.0, mchrg[t], chrgcont[t],
                                  opts=ob.default_opts(**TIGHT))
        _check_report(reps[t], lists[t], ref)
    print("records:", reps[0].records(), reps[0].asave.decode(), flush=True)


def test_finished_md_trajectories_to_records():
    trajs = W.md_ensemble(2, base_seed=31, mean_ev=0.6, max_ev=1.5, tadd_fs=5.0)
    md = q().MdEnsemble(trajs, max_steps=8, use_graph=True)
    md.run(10)
    out = md.read()
    axyz, achrg = md.read_averages()
    md.close()
    off = md.offsets
    nums = [np.asarray(t["num"], dtype=np.int32) for t in trajs]
    sl = [slice(off[t], off[t + 1]) for t in range(2)]
    # pull the final geometry apart so that the analysis sees fragments
    xyzs, axyzs = [], []
    for t in range(2):
        n, x = PC.split_chain(nums[t], out["xyz"][sl[t]], [2 + t], 6.0 * W.AATOAU)
        _, ax = PC.split_chain(nums[t], axyz[sl[t]], [2 + t], 6.0 * W.AATOAU)
        xyzs.append(x)
        axyzs.append(ax)
    masses = [np.asarray(t["mass"]) for t in trajs]
    atl = np.array([r.atlast for r in out["results"]])
    lists, reps = q().manage_fragments_batch(nums, xyzs, axyzs, [achrg[s] for s in sl], masses,
                                             atl, [1, 2], 1, 3, 1.0, [1, 1], [1.0, 1.0])
    for t in range(2):
        ref = ob.manage_fragments(nums[t], xyzs[t], axyzs[t], achrg[sl[t]], masses[t], atl[t],
                                  t + 1, 1, 3, 1.0, 1, 1.0, opts=ob.default_opts(**TIGHT))
        assert reps[t].nfrag >= 2
        _check_report(reps[t], lists[t], ref)
