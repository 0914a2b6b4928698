"""STO-nG basis: Stewart's printed expansions, pinned by the mathematics of
his least-squares criterion, and carried digit for digit by the tables the
library and the oracle build their basis from (tblite v0.2.1 builds its
basis from the same printed tables, slater_to_gauss, for tblite.f90:123).

A Slater function r^(n-1) e^(-r) Y_lm (zeta = 1) is expanded in normalised
Gaussians r^l e^(-a_i r^2) Y_lm.  For given exponents the best coefficients
solve the linear least-squares (projection) problem, and at Stewart's
exponents the overlap with the Slater function is stationary.  Both are
checked here in 30-digit arithmetic with numerical radial quadrature,
independent of the closed form used by the fit script."""
import json
import os
import re

import mpmath as mp
import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PUB = {k: v for k, v in json.load(open(os.path.join(HERE, "golden", "stewart_1970.json"))).items()
       if not k.startswith("_")}
GFN2_SHELLS = ("1s_3", "2s_4", "2p_4")  # H 1s; C, N, O 2s and 2p
mp.mp.dps = 30


def _key(key):
    nl, ng = key.split("_")
    return int(nl[0]), "spd".index(nl[1]), int(ng)


def _proj(n, l, alpha):
    """Projection of the Slater function on the Gaussians: (b, G, c)."""
    ng = len(alpha)
    p = n - 1 + l + 2
    b = mp.matrix([mp.quad(lambda r: r ** p * mp.e ** (-r - a * r * r), [0, 1, mp.inf])
                   for a in alpha])
    g = mp.matrix(ng, ng)
    for i in range(ng):
        for j in range(ng):
            g[i, j] = mp.gamma(l + mp.mpf(3) / 2) / (2 * (alpha[i] + alpha[j]) ** (l + mp.mpf(3) / 2))
    return b, g, mp.lu_solve(g, b)


def _coefficients(n, l, alpha):
    """Least-squares coefficients of normalised primitives, contraction
    normalised, overall sign positive."""
    b, g, c = _proj(n, l, alpha)
    ng = len(alpha)
    d = [c[i] * mp.sqrt(g[i, i]) for i in range(ng)]
    nrm = mp.sqrt(sum(d[i] * g[i, j] / mp.sqrt(g[i, i] * g[j, j]) * d[j]
                      for i in range(ng) for j in range(ng)))
    d = [x / nrm for x in d]
    return [-x for x in d] if sum(d) < 0 else d


def _overlap(n, l, logalpha):
    b, g, c = _proj(n, l, [mp.e ** x for x in logalpha])
    return mp.sqrt((b.T * c)[0] / (mp.factorial(2 * n) / mp.mpf(2) ** (2 * n + 1)))


def _coeff_dev(key, alpha, coeff):
    n, l, _ = _key(key)
    d = _coefficients(n, l, [mp.mpf(a) for a in alpha])
    return max(abs(float((d[i] - coeff[i]) / coeff[i])) for i in range(len(coeff)))


@pytest.mark.parametrize("key", sorted(PUB))
def test_printed_coefficients_are_least_squares_coefficients(key):
    alpha, coeff = PUB[key]
    assert _coeff_dev(key, alpha, coeff) < 2e-9


@pytest.mark.parametrize("key", GFN2_SHELLS)
def test_printed_exponents_make_the_overlap_stationary(key):
    n, l, ng = _key(key)
    la = [mp.log(mp.mpf(a)) for a in PUB[key][0]]
    h = mp.mpf("1e-10")
    for i in range(ng):
        up, dn = list(la), list(la)
        up[i] += h
        dn[i] -= h
        grad = (_overlap(n, l, up) - _overlap(n, l, dn)) / (2 * h)
        assert abs(float(grad)) < 1e-9, (key, i, float(grad))


def test_a_changed_digit_is_detected():
    """A mistake in the seventh digit of one exponent breaks the coefficient
    check by two orders of magnitude more than its tolerance."""
    alpha, coeff = PUB["2s_4"]
    bad = list(alpha)
    bad[1] = 2.000244111
    assert _coeff_dev("2s_4", bad, coeff) > 2e-7


def _header_table(path, prefix):
    text = open(path).read()
    out = {}
    for which in ("alpha", "coeff"):
        body = text.split(f"{prefix}_{which}[3][3][6][6] = {{", 1)[1].split("};", 1)[0]
        vals = [float(v) for v in re.findall(r"[-+0-9.eE]+", body)]
        out[which] = np.array(vals).reshape(3, 3, 6, 6)
    return out


@pytest.mark.parametrize("path,prefix", [("qcxms_amd/csrc/sto_ng_table.h", "sto_ng"),
                                         ("oracle/sto_ng_oracle.h", "orc_sto")])
def test_tables_carry_the_printed_digits(path, prefix):
    t = _header_table(os.path.join(ROOT, path), prefix)
    for key, (alpha, coeff) in PUB.items():
        n, l, ng = _key(key)
        assert list(t["alpha"][n - 1, l, ng - 1, :ng]) == list(alpha), key
        assert list(t["coeff"][n - 1, l, ng - 1, :ng]) == list(coeff), key


def test_basis_sizes_the_library_and_oracle_use():
    """H 1s from three primitives, C/N/O 2s and 2p from four, in both."""
    prod = open(os.path.join(ROOT, "qcxms_amd/csrc/gfn2_params.h")).read()
    orc = open(os.path.join(ROOT, "oracle/gfn2_params_oracle.h")).read()
    assert "kNprim[kElem][2] = {{3, 0}, {4, 4}, {4, 4}, {4, 4}}" in prod
    # element records: Z, shells, angular momenta, principal numbers, primitives
    assert "{1, 1, {0, 0}, {1, 0}, {3, 0}," in orc
    for z in (6, 7, 8):
        assert f"{{{z}, 2, {{0, 1}}, {{2, 2}}, {{4, 4}}," in orc


def test_least_squares_fit_reproduces_the_printed_values():
    """The fit script's criterion is Stewart's: its independent fits land on
    the printed values within the flatness of the optimum."""
    fit = json.load(open(os.path.join(HERE, "golden", "sto_ng_fit.json")))
    for key in GFN2_SHELLS + ("1s_4", "2s_3", "2p_3"):
        np.testing.assert_allclose(fit[key]["alpha"], PUB[key][0], rtol=3e-5)
        np.testing.assert_allclose(fit[key]["coeff"], PUB[key][1], rtol=3e-5, atol=2e-7)


def test_overlaps_increase_with_ng():
    fit = json.load(open(os.path.join(HERE, "golden", "sto_ng_fit.json")))
    for nl in ("1s", "2s", "2p", "3s", "3p", "3d"):
        ov = [fit[f"{nl}_{ng}"]["overlap"] for ng in range(1, 7)]
        assert all(b > a for a, b in zip(ov, ov[1:])), (nl, ov)
        assert ov[-1] > 0.9999
