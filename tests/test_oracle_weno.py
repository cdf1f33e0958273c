"""First-principles pins of the oracle's WENO7-SYMBO reconstruction.

The reference ships no tests or golden vectors for this path, so the oracle is
pinned here against mathematics that does not depend on its own formulas:
exact rational polynomial reconstruction, the Jiang-Shu smoothness integrals,
the JS (p = 1) form of the weights and limits of the nonlinear weights.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import np_oracle as N

EPS_W = 2.592e-6


def _cell_avg_poly(cells, avgs):
    """Cubic p(x) = sum c_i x^i whose averages over cells [j-1/2, j+1/2] are avgs."""
    rows = []
    for j, v in zip(cells, avgs):
        a, b = Fr(2 * j - 1, 2), Fr(2 * j + 1, 2)
        rows.append([(b ** (i + 1) - a ** (i + 1)) / (i + 1) for i in range(4)] + [Fr(v)])
    n = 4
    for c in range(n):
        piv = next(r for r in range(c, n) if rows[r][c] != 0)
        rows[c], rows[piv] = rows[piv], rows[c]
        for r in range(n):
            if r != c and rows[r][c] != 0:
                f = rows[r][c] / rows[c][c]
                rows[r] = [x - f * y for x, y in zip(rows[r], rows[c])]
    return [rows[i][n] / rows[i][i] for i in range(n)]


def _deriv(c):
    return [c[i] * i for i in range(1, len(c))]


def _int_sq(c, a, b):
    sq = [Fr(0)] * (2 * len(c) - 1)
    for i, x in enumerate(c):
        for j, y in enumerate(c):
            sq[i + j] += x * y
    return sum(co * (b ** (k + 1) - a ** (k + 1)) / (k + 1) for k, co in enumerate(sq))


def _ev(c, x):
    return sum(co * x ** i for i, co in enumerate(c))


def _js_indicator(c, cell):
    """Jiang-Shu smoothness: sum_l int_cell (p^(l))^2, l = 1..3 (unit spacing)."""
    tot, d = Fr(0), c
    for _ in range(3):
        d = _deriv(d)
        tot += _int_sq(d, cell - Fr(1, 2), cell + Fr(1, 2))
    return tot


def _rational_windows(n, seed):
    rng = np.random.default_rng(seed)
    return [[Fr(int(v), int(q)) for v, q in zip(rng.integers(-60, 60, 8), rng.integers(1, 9, 8))]
            for _ in range(n)]


def test_m_is_p_of_reversed_window_bitwise(orc):
    rng = np.random.default_rng(11)
    for scale in (1e-8, 1.0, 1e3):
        for _ in range(200):
            s = rng.standard_normal(8) * scale + rng.standard_normal()
            assert orc.weno_m(s) == orc.weno_p(s[::-1].copy())


@pytest.mark.parametrize("seed", range(6))
def test_cubic_cell_averages_reconstruct_exactly(orc, seed):
    """For cell averages of any cubic, all five candidates equal the exact
    interface value, so P and M return it whatever the weights."""
    rng = np.random.default_rng(seed)
    c = [Fr(int(v), 5) for v in rng.integers(-20, 20, 4)]
    # cell j covers [j-1/2, j+1/2]; window cells 0..7, face between cells 3 and 4
    avgs = []
    for j in range(8):
        a, b = Fr(2 * j - 1, 2), Fr(2 * j + 1, 2)
        avgs.append(sum(ci * (b ** (i + 1) - a ** (i + 1)) / (i + 1) for i, ci in enumerate(c)))
    exact = float(_ev(c, Fr(7, 2)))
    s = np.array([float(a) for a in avgs])
    scale = max(1.0, np.abs(s).max())
    assert abs(orc.weno_p(s) - exact) <= 1e-13 * scale
    assert abs(orc.weno_m(s) - exact) <= 1e-13 * scale
    assert abs(N.weno_p(list(s)) - exact) <= 1e-13 * scale


@pytest.mark.parametrize("window", _rational_windows(8, 5))
def test_smoothness_and_candidates_from_first_principles(orc, window):
    """S_k = 25920 * (Jiang-Shu indicator of candidate k over the upwind cell 3)
    and q_k = 12 * p_k(face), with p_k the cubic matching the candidate's four
    cell averages.  The reference's candidate-2 second-derivative term is the
    second difference centred on cell 4 (s3 - 2 s4 + s5) instead of cell 3
    (OCFD_Schemes.cpp:325); that term is carried over explicitly."""
    s = np.array([float(x) for x in window])
    S, wraw, q = orc.weno_p_parts(s)
    polys = [_cell_avg_poly(range(k, k + 4), window[k:k + 4]) for k in range(5)]
    IS = [25920 * _js_indicator(p, Fr(3)) for p in polys]

    def d2(j):
        return window[j] - 2 * window[j + 1] + window[j + 2]
    IS[2] += 780 * 36 * (d2(3) ** 2 - d2(2) ** 2)
    raw = [float(x) for x in IS]
    for k in range(4):
        assert S[k] == pytest.approx(raw[k], rel=1e-13, abs=1e-9)
    # the downwind indicator is replaced by the largest of all five
    assert S[4] == pytest.approx(max(raw), rel=1e-13, abs=1e-9)
    for k in range(5):
        assert q[k] == pytest.approx(float(12 * _ev(polys[k], Fr(7, 2))), rel=1e-13, abs=1e-11)
    # product form == JS p = 1 form: w_k proportional to d_k / (eps + S_k)
    d = wraw / np.prod(EPS_W + S) * (EPS_W + S)
    w_js = d / (EPS_W + S)
    np.testing.assert_allclose(wraw / wraw.sum(), w_js / w_js.sum(), rtol=1e-13)


def test_linear_weights_checksum(orc):
    """On a constant window every S_k = 0, so the unnormalised weights are
    d_k * eps^4.  The five transcribed linear weights, as printed in
    OCFD_Schemes.cpp:344-348, sum to 1 + 1e-12; a mistyped digit moves the sum."""
    S, wraw, q = orc.weno_p_parts(np.full(8, 0.5))
    assert np.all(S == 0.0)
    d = wraw / EPS_W ** 4
    assert abs(d.sum() - 1.000000000001) < 1e-14
    assert np.all(d > 0)
    # linear weights as returned by the numpy restatement's constants
    np.testing.assert_allclose(d, np.array(N.D_LIN), rtol=1e-14)


def test_small_amplitude_limit_is_linear_scheme(orc):
    """When every S_k << eps the weights tend to the linear weights, so the
    result tends to sum_k d_k q_k / (12 sum d_k) (to O(S/eps))."""
    rng = np.random.default_rng(3)
    _, wraw0, _ = orc.weno_p_parts(np.zeros(8))
    d = wraw0 / wraw0.sum()
    for _ in range(20):
        s = 1.0 + 1e-9 * rng.standard_normal(8)
        _, _, q = orc.weno_p_parts(s)
        lin = float(np.dot(d, q) / 12.0)
        assert abs(orc.weno_p(s) - lin) < 1e-14


def test_eno_property_at_a_step(orc):
    """A step between window cells 4 and 5 leaves only the candidates on
    cells 0..3 and 1..4 smooth; the reconstruction must follow the smooth
    (left) data to within the tiny weight of the non-smooth candidates."""
    s = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    assert abs(orc.weno_p(s) - 1.0) < 1e-9
    # the mirrored step with the M rule follows the right data
    assert abs(orc.weno_m(s[::-1].copy()) - 1.0) < 1e-9


@pytest.mark.parametrize("seed", range(4))
def test_c_and_numpy_restatements_agree_on_windows(orc, seed):
    rng = np.random.default_rng(100 + seed)
    for scale in (1e-6, 1.0, 50.0):
        W = rng.standard_normal((400, 8)) * scale + 1.0
        W[::7, 4:] += 3.0 * scale                      # steps
        c = np.array([orc.weno_p(w) for w in W])
        n = N.weno_p([W[:, k] for k in range(8)])
        np.testing.assert_allclose(n, c, rtol=1e-12, atol=1e-13 * scale)
        cm = np.array([orc.weno_m(w) for w in W])
        nm = N.weno_m([W[:, k] for k in range(8)])
        np.testing.assert_allclose(nm, cm, rtol=1e-12, atol=1e-13 * scale)
