"""Every exit of the EI production loop on the GPU against the oracle MD:
the same step, exit reason, mdok, fragstate, fragment count, fragment masses
(bit-exact) and fragment temperatures, through the per-step kernels and the
persistent MD kernel (md.f90:184, 321-344, 443-462, 630-679, 715-741;
iniqm.f90:684-732; impact.f90:42-50; main.F90:2255)."""
import numpy as np
import pytest

import md_exit_cases as MC
import oracle_bind as ob

pytestmark = pytest.mark.gpu
CASES = MC.cases()
GPU_SCC = dict(econv=1e-11, pconv=1e-9)
ORC_SCC = dict(econv=1e-12, pconv=1e-10)


def q():
    import qcxms_amd
    return qcxms_amd


def _oracle(name):
    tr, kw, _ = CASES[name]
    return ob.md_run(tr, steps_limit=0, isec=int(tr.get("isec", 1)), opts=ob.default_opts(**ORC_SCC),
                     trace_steps=kw["max_steps"], **kw)


def _compare(name, res, ref, lst, ref_list):
    tr, kw, want = CASES[name]
    st = ref["state"]
    print(f"{name}: gpu nstep {res.nstep} exit {res.exit_reason} mdok {res.mdok} nfrag {res.nfrag} "
          f"| oracle {st.nstep} {st.exit_reason} {st.mdok} {st.nfrag} "
          f"dEpot {abs(res.epot - st.epot):.2e}", flush=True)
    assert (res.nstep, res.exit_reason, res.mdok, res.fragstate, res.nfrag) == \
        (st.nstep, st.exit_reason, st.mdok, st.fragstate, st.nfrag)
    assert (st.nstep, st.exit_reason, st.mdok, st.fragstate) == want
    assert np.array_equal(np.array(res.fragm[:]), np.array(st.fragm[:]))
    np.testing.assert_allclose(np.array(res.fragT[:]), np.array(st.fragT[:]), rtol=1e-6, atol=1e-6)
    # labels of the last connectivity pass (a heating failure ends the first
    # step before any)
    if st.nstep > 0 and st.exit_reason != 8:
        assert np.array_equal(lst, ref_list)


@pytest.mark.parametrize("fused", [False, True], ids=["per_step", "persistent"])
def test_md_exits_match_oracle(fused):
    # one ensemble per set of MD options; every case of a set runs in it
    groups = {}
    for name, (tr, kw, _) in CASES.items():
        key = tuple(sorted(kw.items()))
        groups.setdefault(key, []).append(name)
    for key, names in sorted(groups.items()):
        kw = dict(key)
        trajs = [CASES[n][0] for n in names]
        print(f"ensemble {names} {kw} fused={fused}", flush=True)
        md = q().MdEnsemble(trajs, use_graph=False, fused=fused, **GPU_SCC, **kw)
        md.run(kw["max_steps"] + 2)
        out = md.read()
        md.close()
        for t, name in enumerate(names):
            ref = _oracle(name)
            sl = slice(md.offsets[t], md.offsets[t + 1])
            _compare(name, out["results"][t], ref, out["list"][sl], ref["list"])


def test_split_during_heating_per_step():
    """The ethane cation that splits, re-joins and splits again while it is
    heated with ieetemp > 0: electronic temperature, energies, fragment count
    and temperatures step by step (setetemp sees the previous step's nfrag)."""
    tr, kw, _ = CASES["split_ieetemp"]
    ref = _oracle("split_ieetemp")["trace"]
    md = q().MdEnsemble([tr], use_graph=False, **GPU_SCC, **kw)
    for k in range(kw["max_steps"]):
        md.run(1)
        r = md.read()["results"][0]
        row = ref[k]
        assert r.nstep == int(row[0]) and r.nfrag == int(row[4]), (k, r.nstep, r.nfrag, row)
        assert r.etemp == row[1], (k, r.etemp, row[1])
        assert abs(r.epot - row[2]) < 1e-8, (k, r.epot, row[2])
        assert abs(r.ekin - row[3]) < 1e-9
        # 1e-5 K on a 4-atom fragment is 2e-10 Eh of kinetic energy
        np.testing.assert_allclose([r.fragT[0], r.fragT[1]], row[5:7], rtol=1e-8, atol=1e-5)
    md.close()
