"""SCC launch plans by basis-size bin (qcxms_set_tuning): a ragged batch run
as several launches with their own eigensolver and LDS choices, on side
streams, or as one launch with the mixed choice, gives the oracle's single
points and the same fixed point as one launch with the batch's options."""
import numpy as np
import pytest

import oracle_bind as ob
from qcxms_amd import workloads as W

pytestmark = pytest.mark.gpu


def nao_of(num):
    return int(sum(1 if int(z) == 1 else 4 for z in num))


@pytest.mark.parametrize("order", ["largest_first", "interleaved"])
def test_binned_plan_matches_oracle_and_single_launch(order):
    import qcxms_amd
    sysl = []
    for s in range(24):
        num, xyz, chg = W.random_chain(7000 + s)
        sysl.append((num, chg, xyz))
    num, xyz = W.chain_c12h24o2n2(seed=3)
    sysl.append((num, 1, xyz))
    if order == "largest_first":
        sysl.sort(key=lambda t: -nao_of(t[0]))
    else:
        # sizes alternate: every system its own run (more runs than side streams)
        sysl.sort(key=lambda t: nao_of(t[0]))
        sysl = [sysl[i // 2] if i % 2 == 0 else sysl[-1 - i // 2] for i in range(len(sysl))]
    saved = qcxms_amd.get_tuning()
    xyz = np.concatenate([t[2] for t in sysl])
    table = [(24, 1, 1), (40, 0, 1), (56, 1, 0), (256, 1, 0)]
    out = {}
    try:
        for name, tab, mixed in (("single", [], (-1, 0)), ("binned", table, (-1, 0)),
                                 ("mixed_one_launch", table, (1, 1))):
            qcxms_amd.set_tuning(tab, mixed)
            assert qcxms_amd.get_tuning() == (tab, mixed)
            b = qcxms_amd.Batch([(t[0], t[1]) for t in sysl])
            b.set_scc(econv=1e-11, pconv=1e-9)
            out[name] = b.egrad(xyz, 5000.0)
            b.close()
    finally:
        qcxms_amd.set_tuning(*saved)
    em, gm, _, stm = out["mixed_one_launch"]
    assert np.all(stm == 0) and np.abs(em - out["single"][0]).max() <= 1e-9
    e1, g1, q1, st1 = out["binned"]
    e0, g0, q0, st0 = out["single"]
    assert np.all(st0 == 0) and np.all(st1 == 0)
    print(f"{order}: max|dE| binned vs single {np.abs(e1 - e0).max():.2e}, naos "
          f"{[nao_of(t[0]) for t in sysl]}", flush=True)
    assert np.abs(e1 - e0).max() <= 1e-9
    off = 0
    for i, (num, chg, x) in enumerate(sysl):
        k = slice(off, off + len(num))
        off += len(num)
        if i % 4:
            continue
        r = ob.gfn2_egrad(num, x, chg, 5000.0, opts=ob.default_opts(econv=1e-12, pconv=1e-10))
        assert abs(e1[i] - r["energy"]) <= 1e-8
        assert np.abs(g1[k] - r["grad"]).max() <= 1e-6
        assert np.abs(q1[k] - r["qat"]).max() <= 1e-6
