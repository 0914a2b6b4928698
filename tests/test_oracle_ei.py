"""The oracle's fragmentation generations (orc_ei_run) follow the rules of
main.F90:2208-2345: the continued generation holds exactly the atoms of
fragment tcont (labels of the finished generation), nmax shrinks as
nmax - nmax0/5 but not below nmax0/2 (nmax0/3 after fragstate 2, nmax0/10
for ten atoms or fewer), the second generation exits at two fragments, and
the charge the next generation runs with is the one manage_fragments
returns."""
import numpy as np

import ei_cases as EC
import oracle_bind as ob

NMAX = 60
OPTS = ob.default_opts(econv=1e-12, pconv=1e-10)


def _nmax_next(nmax, fragstate, nat_next):
    n = max(nmax - NMAX // 5, NMAX // 2)
    if fragstate == 2:
        n = NMAX // 3
    if nat_next <= 10:
        n = NMAX // 10
    return n


def _check_chain(gens):
    for a, b in zip(gens, gens[1:]):
        assert a["continued"] == 1 and a["mdok"] == 1 and a["tcont"] > 0
        assert b["isec"] == a["isec"] + 1
        # the next generation's atoms are the atoms of fragment tcont, in order
        want = a["atoms"][a["list"] == a["tcont"]]
        assert np.array_equal(b["atoms"], want)
        assert b["nat"] == len(want)
        assert b["max_steps"] == _nmax_next(a["max_steps"], a["fragstate"], b["nat"])
        assert b["mchrg"] >= 1
    last = gens[-1]
    assert last["continued"] == 0


def test_split_into_two_continues_with_one_piece():
    g = ob.ei_run(EC.cases()["split2"], nmax=NMAX, opts=OPTS)
    assert len(g) == 2
    assert (g[0]["nstep"], g[0]["exit_reason"], g[0]["nfrag"]) == (NMAX, 7, 2)
    _check_chain(g)
    # the second generation runs its whole (shorter) MD as one fragment
    assert (g[1]["nstep"], g[1]["exit_reason"], g[1]["nfrag"], g[1]["tcont"]) == (48, 7, 1, 0)


def test_split_into_four_ends_the_first_md_at_once():
    g = ob.ei_run(EC.cases()["split4"], nmax=NMAX, opts=OPTS)
    assert (g[0]["nstep"], g[0]["exit_reason"], g[0]["nfrag"], g[0]["fragstate"]) == (1, 4, 4, 1)
    _check_chain(g)
    # a last generation that stays in one piece keeps the statistical charge
    # it was given (write_fragments.f90:288-290: no continuation, no rescaling)
    assert g[-1]["nfrag"] == 1 and abs(g[0]["chrgcont"] - g[1]["chrgcont"]) < 1e-15


def test_generation_limit_stops_the_chain():
    """maxsec = 1: the second generation may not continue (write_fragments.f90:
    288-290 set tcont = 0 for isec = maxsec + 1)."""
    g = ob.ei_run(EC.cases()["split4"], nmax=NMAX, maxsec=1, opts=OPTS)
    assert len(g) == 2 and g[1]["isec"] == 2 and g[1]["tcont"] == 0
