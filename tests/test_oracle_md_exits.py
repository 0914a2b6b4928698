"""The oracle MD ends each case of md_exit_cases through the exit the
reference's rules give (md.f90:184, 321-344, 456-462, 630-679; impact.f90:
42-50; main.F90:2255), and sets the electronic temperature from the previous
step's fragment count (md.f90:443-451)."""
import numpy as np
import pytest

import md_exit_cases as MC
import oracle_bind as ob

CASES = MC.cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_md_exit(name):
    tr, kw, want = CASES[name]
    r = ob.md_run(tr, steps_limit=0, isec=int(tr.get("isec", 1)), **kw)
    st = r["state"]
    assert (st.nstep, st.exit_reason, st.mdok, st.fragstate) == want, name


def test_setetemp_sees_previous_step_fragment_count():
    tr, kw, _ = CASES["split_ieetemp"]
    r = ob.md_run(tr, steps_limit=0, trace_steps=kw["max_steps"], **kw)
    tr_ = r["trace"]
    nfrag = tr_[:, 4].astype(int)
    # the case splits, re-joins and splits again while heating
    assert nfrag[0] == 2 and 1 in nfrag[1:30] and nfrag[-1] == 2
    nadd = int((tr["tadd"] + kw.get("tstep", 0.5 * MC.FS)) / kw.get("tstep", 0.5 * MC.FS) - 1)
    prev = np.concatenate([[1], nfrag[:-1]])
    for k, row in enumerate(tr_):
        nstep = int(row[0])
        dum = tr["eimp"] - tr["eimp"] * float(np.float32(nstep)) / float(np.float32(nadd))
        heated = dum > 0 and prev[k] <= 1
        assert (row[1] > 5000.0) == heated, (nstep, row[1], prev[k])
        if heated:
            assert abs(row[1] - (5000.0 + dum * kw["ieetemp"])) < 1e-9
