"""Pins of the oracle's Steger-Warming splitting against the physical flux."""
import numpy as np
import pytest

from oracle import np_oracle as N

GAMMA = 1.4


def _random_point(rng):
    d = 0.2 + 2.0 * rng.random()
    u, v, w = rng.standard_normal(3) * 2.0
    T = 0.3 + 2.0 * rng.random()
    A = rng.standard_normal(3)
    return d, u, v, w, T, A


@pytest.mark.parametrize("eps", [0.0, 1e-3, 0.3])
def test_split_sums_to_physical_flux(orc, eps):
    """f+ + f- = A.F(U) for any smoothing epsilon, with p = rho c^2 / gamma
    and E = p/(gamma-1) + rho |u|^2 / 2 (Ma = 1, c = sqrt(T))."""
    rng = np.random.default_rng(1)
    for _ in range(300):
        d, u, v, w, T, A = _random_point(rng)
        cc = np.sqrt(T)
        fp, fm = orc.split_sw(d, u, v, w, cc, *A, gamma=GAMMA, epsl=eps)
        p = d * cc * cc / GAMMA
        F = N.euler_flux(d, u, v, w, p, *A, GAMMA)
        scale = max(1.0, max(abs(x) for x in F), np.abs(fp).max(), np.abs(fm).max())
        for m in range(5):
            assert abs(fp[m] + fm[m] - F[m]) <= 1e-13 * scale


def test_supersonic_without_smoothing_has_no_backward_flux(orc):
    """eps = 0 and V > c|A|: all eigenvalues positive, f- is exactly zero."""
    rng = np.random.default_rng(2)
    for _ in range(200):
        d, _, v, w, T, _ = _random_point(rng)
        cc = np.sqrt(T)
        u = cc * (1.05 + 3.0 * rng.random())
        fp, fm = orc.split_sw(d, u, 0.3 * v, 0.3 * w, cc, 1.0, 0.0, 0.0, gamma=GAMMA, epsl=0.0)
        if u > cc + abs(0.0):
            assert np.all(fm == 0.0)


def test_mirror_swaps_positive_and_negative_flux(orc):
    """Reversing the direction (A -> -A) turns f+ into -f- (to rounding)."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        d, u, v, w, T, A = _random_point(rng)
        cc = np.sqrt(T)
        fp, fm = orc.split_sw(d, u, v, w, cc, *A, gamma=GAMMA, epsl=1e-3)
        gp, gm = orc.split_sw(d, u, v, w, cc, *(-A), gamma=GAMMA, epsl=1e-3)
        scale = max(1.0, np.abs(fp).max(), np.abs(fm).max())
        np.testing.assert_allclose(gp, -fm, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(gm, -fp, rtol=0, atol=1e-13 * scale)


def test_c_and_numpy_split_agree(orc):
    rng = np.random.default_rng(4)
    for _ in range(200):
        d, u, v, w, T, A = _random_point(rng)
        cc = np.sqrt(T)
        fp, fm = orc.split_sw(d, u, v, w, cc, *A, gamma=GAMMA, epsl=1e-3)
        gp, gm = N.split_sw(d, u, v, w, cc, *A, GAMMA, 1e-3)
        scale = max(1.0, np.abs(fp).max(), np.abs(fm).max())
        np.testing.assert_allclose(np.array(gp), fp, rtol=0, atol=1e-14 * scale)
        np.testing.assert_allclose(np.array(gm), fm, rtol=0, atol=1e-14 * scale)
