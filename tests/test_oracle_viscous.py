"""Pins of the oracle's viscous terms (the reference's du_vis_Jacobian3d path) from
closed forms: central-difference exactness on polynomials, the discrete symbol of the
CD6/CD8 operators on a periodic shear wave (momentum and energy), Sutherland's law at
the reference temperature, uniform states, and an independent tensor-form numpy
restatement with general per-cell metrics."""
import math

import numpy as np
import pytest

from oracle import np_oracle as N
from oracle import oracle as O
from tests.helpers import ghosted

RE, PR, TREF = 1000.0, 0.72, 288.15


@pytest.mark.parametrize("order,deg", [(6, 6), (8, 8)])
def test_central_difference_exact_on_polynomials(order, deg):
    h, x0 = 0.1, 0.3
    r = order // 2
    x = x0 + h * np.arange(-r, r + 1)
    for k in range(deg + 1):
        got = O.cd(x ** k) / h
        exact = k * x0 ** (k - 1) if k else 0.0
        assert abs(got - exact) <= 1e-11 * max(1.0, abs(exact)), (order, k)
    # and not beyond: degree deg+1 is no longer exact
    got = O.cd(x ** (deg + 1)) / h
    assert abs(got - (deg + 1) * x0 ** deg) > 1e-9


def test_sutherland_at_reference_temperature():
    p = O.params(4, 4, 4)
    T = ghosted(np.ones((4, 4, 4)))
    assert np.allclose(O.amu(p, T, RE, TREF), 1.0 / RE, rtol=1e-15, atol=0)


def _symbol(order, theta):
    """kappa(theta): CD applied to exp(i theta j) = i kappa exp(i theta j)."""
    c = N._CD[order]
    return 2.0 * sum(cq * math.sin(q * theta) for q, cq in enumerate(c, start=1))


@pytest.mark.parametrize("order", [6, 8])
def test_periodic_shear_wave_closed_form(order):
    """u = A sin(theta j) along y, v = w = 0, T = 1: du1 = -mu A kappa^2/h^2 sin(theta j),
    du4 = mu A^2 kappa(theta) kappa(2 theta)/(2 h^2) cos(2 theta j), everything else 0."""
    nx, ny, nz = 6, 32, 6
    p = O.params(nx, ny, nz)
    A, theta, h = 0.3, 2 * math.pi * 3 / ny, p.hy
    j = np.arange(ny)[None, :, None] * np.ones((nz, ny, nx))
    u = A * np.sin(theta * j)
    zero = np.zeros_like(u)
    g = [ghosted(a) for a in (u, zero, zero, np.ones_like(u))]
    for a in g:
        O.ghost_fill_periodic(p, a)
    du = O.viscous_rhs(p, *g, RE, PR, TREF, order)
    mu = 1.0 / RE
    k1, k2 = _symbol(order, theta), _symbol(order, 2 * theta)
    du1 = -mu * A * k1 * k1 / (h * h) * np.sin(theta * j)
    du4 = mu * A * A * k1 * k2 / (2 * h * h) * np.cos(2 * theta * j)
    scale = np.abs(du1).max()
    assert np.abs(du[1] - du1).max() <= 1e-12 * scale
    assert np.abs(du[4] - du4).max() <= 1e-12 * np.abs(du4).max()
    assert np.abs(du[2]).max() <= 1e-12 * scale and np.abs(du[3]).max() <= 1e-12 * scale
    assert np.all(du[0] == 0.0)


def test_uniform_state_gives_zero():
    p = O.params(10, 9, 8)
    g = [ghosted(np.full((8, 9, 10), c)) for c in (0.3, -0.2, 0.1, 1.2)]
    for a in g:
        O.ghost_fill_periodic(p, a)
    for order in (6, 8):
        assert np.all(O.viscous_rhs(p, *g, RE, PR, TREF, order) == 0.0)


@pytest.mark.parametrize("order", [6, 8])
def test_c_oracle_matches_tensor_form(order):
    nx, ny, nz = 14, 11, 9
    rng = np.random.default_rng(5)
    u, v, w = [0.3 * rng.standard_normal((nz, ny, nx)) for _ in range(3)]
    T = 1 + 0.2 * rng.random((nz, ny, nx))
    met = [0.7 + 0.3 * rng.random((nz, ny, nx)) if i in (0, 4, 8) else 0.2 * rng.standard_normal((nz, ny, nx))
           for i in range(9)]
    aj = 1 + 0.3 * rng.random((nz, ny, nx))
    p = O.params(nx, ny, nz)
    g = [ghosted(a) for a in (u, v, w, T)]
    mg = [ghosted(a) for a in met]
    ajg = ghosted(aj)
    for a in g + mg + [ajg]:
        O.ghost_fill_periodic(p, a)
    dc = O.viscous_rhs(p, *g, RE, PR, TREF, order, metrics=mg, ajac=ajg)
    dn = N.viscous_rhs((u, v, w, T), met, aj, (p.hx, p.hy, p.hz), 1.4, 1.0, RE, PR, TREF, order)
    assert np.linalg.norm(dc - dn) <= 1e-13 * np.linalg.norm(dn)
    assert np.all(dc[0] == 0.0)


@pytest.mark.parametrize("r", [3, 4])
def test_closure_forms_exact_on_polynomials(r):
    """The forms next to a closed face (OCFD_bound_Scheme.cpp:82-263): at distance t from the
    face a one-sided first-order difference (t = 0), a second-order (t = 1), a fourth-order
    (t = 2) and a sixth-order (t = 3) central difference; each exact up to its degree and not
    beyond, on both faces; interior cells and open faces fall through to the central form."""
    n, h = 12, 0.07
    degree = {0: 1, 1: 2, 2: 4, 3: 6}
    for t in range(4):
        for c, lo, hi in ((t, 1, 0), (n - 1 - t, 0, 1)):
            x0 = 0.2 + c * h
            x = x0 + h * np.arange(-r, r + 1)
            for k in range(degree[t] + 2):
                got = O.cd_closure(x ** k, c, n, lo, hi)
                assert got is not None
                exact = k * x0 ** (k - 1) if k else 0.0
                err = abs(got / h - exact)
                if k <= degree[t]:
                    assert err <= 1e-10 * max(1.0, abs(exact)), (r, t, c, k)
                else:
                    assert err > 1e-8, (r, t, c, k)
            assert O.cd_closure(x, c, n, 1 - lo, 1 - hi) is None
    for c in range(4, n - 4):
        assert O.cd_closure(np.arange(2 * r + 1.0), c, n) is None


@pytest.mark.parametrize("order", [6, 8])
def test_closed_box_linear_velocity_closed_form(order):
    """A box closed on all six faces (D0_bound everywhere, no ghosts filled: the ghost
    layers hold zeros), a linear velocity field u_i = G_ij x_j and T = 1 (mu = 1/Re): every
    derivative form is exact on linear data, the stress is uniform, so du1..du3 = 0 and du4 =
    sigma : grad u (the viscous dissipation) in every cell, boundary cells included."""
    nx, ny, nz = 13, 11, 10
    p = O.params(nx, ny, nz, periodic=(0, 0, 0), d0_bound=(1,) * 6)
    G = np.array([[0.3, -0.2, 0.1], [0.15, 0.25, -0.05], [-0.1, 0.2, 0.35]])
    k, j, i = np.meshgrid(np.arange(nz) * p.hz, np.arange(ny) * p.hy, np.arange(nx) * p.hx, indexing="ij")
    X = (i, j, k)
    vel = [sum(G[a, b] * X[b] for b in range(3)) for a in range(3)]
    g = [ghosted(a) for a in vel + [np.ones((nz, ny, nx))]]
    du = O.viscous_rhs(p, *g, RE, PR, TREF, order)
    mu = 1.0 / RE
    div = np.trace(G)
    sig = mu * (G + G.T) - 2.0 / 3.0 * mu * div * np.eye(3)
    diss = float(np.sum(sig * G))
    assert np.all(du[0] == 0.0)
    for m in (1, 2, 3):
        assert np.abs(du[m]).max() <= 1e-10 * abs(diss)
    assert np.abs(du[4] - diss).max() <= 1e-10 * abs(diss)
