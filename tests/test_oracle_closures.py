"""Pins of the oracle's one-sided boundary closures (the reference's D0_bound path,
src/boundary/OCFD_bound_Scheme.cpp:540-893, and the high-face form of the fused path's rim
kernel, src/solver/OCFD_NS_Jacobian3d.cpp:1063-1085) and non-reflecting flags: which faces
they cover, exactness on linear data, the use-before-set weights of OCFD_weno5_kernel_P,
telescoping of the RHS to the boundary faces, and which ghost values the RHS reads."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import ghosted, random_state

EP = 1e-6


def test_faces_covered():
    n = 20
    line = np.random.default_rng(0).standard_normal(n + 8)
    cov_p = [f for f in range(n + 1) if O.closure("p", line, f) is not None]
    cov_m = [f for f in range(n + 1) if O.closure("m", line, f) is not None]
    assert cov_p == [0, 1, 2, 3, n - 2, n - 1, n]
    assert cov_m == [0, 1, 2, 3, n - 3, n - 2, n - 1, n]
    assert O.closure("p", line, 2, lo=0) is None and O.closure("m", line, n - 1, hi=0) is None


def test_closures_exact_on_linear_data():
    n = 16
    x = np.arange(-4, n + 4, dtype=float)
    line = 0.7 + 0.3 * x                     # cell values; face f sits at x = f - 1/2
    for f in list(range(1, 4)) + [n - 2, n - 1, n]:
        assert abs(O.closure("p", line, f) - (0.7 + 0.3 * (f - 0.5))) <= 1e-13, f
    for f in list(range(0, 4)) + [n - 3, n - 2, n - 1]:
        assert abs(O.closure("m", line, f) - (0.7 + 0.3 * (f - 0.5))) <= 1e-13, f
    # first-order ends: P takes cell 0 at face 0, M takes cell n-1 at face n
    assert O.closure("p", line, 0) == line[4] and O.closure("m", line, n) == line[4 + n - 1]


def _js_S(s):
    """Jiang-Shu WENO5 smoothness (13/3 and 1/4 scaled by 12) of the three candidates."""
    S0 = 13 * (s[2] - 2 * s[3] + s[4]) ** 2 + 3 * (3 * s[2] - 4 * s[3] + s[4]) ** 2
    S1 = 13 * (s[1] - 2 * s[2] + s[3]) ** 2 + 3 * (s[1] - s[3]) ** 2
    S2 = 13 * (s[0] - 2 * s[1] + s[2]) ** 2 + 3 * (s[0] - 4 * s[1] + 3 * s[2]) ** 2
    return S0, S1, S2


def test_weno5_p_weights_are_built_before_S1_S2():
    """OCFD_weno5_kernel_P multiplies (12 eps + S1)^2 (12 eps + S2)^2 into a0 while S1 and S2
    are still 0, and (12 eps + S2)^2 into a1 likewise: a0 = (12 eps)^4,
    a1 = (12 eps + S0)^2 (12 eps)^2, a2 = (12 eps + S0)^2 (12 eps + S1)^2 (face 3 of a left
    closure uses it)."""
    rng = np.random.default_rng(3)
    n = 16
    for _ in range(20):
        line = rng.standard_normal(n + 8)
        s = line[4:9]                          # cells 0..4 = window of face 3
        S0, S1, S2 = _js_S(s)
        e = 12 * EP
        a = [e ** 4, (e + S0) ** 2 * e ** 2, (e + S0) ** 2 * (e + S1) ** 2]
        q = [2 * s[2] + 5 * s[3] - s[4], -s[1] + 5 * s[2] + 2 * s[3], 2 * s[0] - 7 * s[1] + 11 * s[2]]
        want = sum(ai * qi for ai, qi in zip(a, q)) / (6 * sum(a))
        assert abs(O.closure("p", line, 3) - want) <= 1e-12 * max(1.0, abs(want))
        # the correct product weights would differ whenever S1 or S2 is large
        a_ok = [(e + S1) ** 2 * (e + S2) ** 2, (e + S0) ** 2 * (e + S2) ** 2, (e + S0) ** 2 * (e + S1) ** 2]
        ok = sum(ai * qi for ai, qi in zip(a_ok, q)) / (6 * sum(a_ok))
        assert abs(ok - want) > 1e-9


def _rhs(nx, ny, nz, prims, **kw):
    p = O.params(nx, ny, nz, **kw)
    g = [ghosted(a) for a in prims]
    return O.du_comput(p, *g)[0], p, g


def test_left_ghosts_never_read_and_right_ghosts_only_at_cell_n():
    nx, ny, nz = 20, 8, 8
    prims = random_state(nx, ny, nz, 5, "rough")
    kw = dict(periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0))
    p = O.params(nx, ny, nz, **kw)
    base = [ghosted(a) for a in prims]
    rng = np.random.default_rng(7)
    for a in base:                              # the caller's x ghosts: arbitrary
        a[:, :, :4] = 1.0 + 0.1 * rng.random(a[:, :, :4].shape)
        a[:, :, -4:] = 1.0 + 0.1 * rng.random(a[:, :, -4:].shape)
    ref = O.du_comput(p, *[a.copy() for a in base])[0]
    other = [a.copy() for a in base]
    for a in other:                             # change left ghosts and right ghosts n+1..n+3
        a[:, :, :4] *= 1.5
        a[:, :, -3:] *= 0.5
    assert np.array_equal(O.du_comput(p, *other)[0], ref)
    third = [a.copy() for a in base]
    for a in third:                             # change ghost cell n: only cells n-4, n-3 move
        a[:, :, 4 + nx] *= 1.3
    d3 = O.du_comput(p, *third)[0]
    changed = np.nonzero(np.any(d3 != ref, axis=(0, 1, 2)))[0]
    assert set(changed.tolist()) <= {nx - 4, nx - 3}


def test_line_sums_telescope_to_boundary_faces():
    """With closures and no non-reflecting flags, the x contributions of a line sum to
    -(F(n) - F(0))/h * Ajac, where F = P + M face fluxes at the two domain faces."""
    nx, ny, nz = 24, 6, 6
    prims = random_state(nx, ny, nz, 9, "blobs")
    p = O.params(nx, ny, nz, periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0))
    g = [ghosted(a) for a in prims]
    for a in g:                                  # periodic images everywhere (both runs read them)
        O.ghost_fill_periodic(O.params(nx, ny, nz), a)
    cc = O.sound_speed(p, g[4])
    met = O.metrics_cartesian(p)
    du = np.zeros((5, nz, ny, nx))
    # x direction only: zero metrics for y and z directions is not allowed (norm 0), so take
    # the full RHS and subtract the y, z contributions of the periodic run
    du_all = O.inviscid_rhs(p, g[0], g[1], g[2], g[3], cc, met, np.ones_like(cc), du)
    pp = O.params(nx, ny, nz)
    du_per = O.inviscid_rhs(pp, g[0], g[1], g[2], g[3], cc, met, np.ones_like(cc), np.zeros((5, nz, ny, nx)))
    # y and z parts are identical in both runs; x parts differ only through the closures
    for k in range(nz):
        for j in range(ny):
            fp = np.zeros((5, nx + 8)); fm = np.zeros((5, nx + 8))
            for i in range(-4, nx + 4):
                a, b = O.split_sw(g[0][k + 4, j + 4, i + 4], g[1][k + 4, j + 4, i + 4], g[2][k + 4, j + 4, i + 4],
                                  g[3][k + 4, j + 4, i + 4], cc[k + 4, j + 4, i + 4], 1.0, 0.0, 0.0)
                fp[:, i + 4], fm[:, i + 4] = a, b
            for m in range(5):
                F0 = O.closure("p", fp[m], 0) + O.closure("m", fm[m], 0)
                Fn = O.closure("p", fp[m], nx) + O.closure("m", fm[m], nx)
                # periodic run telescopes to zero along x, so the difference of line sums
                # is the closure run's boundary flux balance
                got = du_all[m, k, j].sum() - du_per[m, k, j].sum()
                want = -(Fn - F0) / p.hx
                assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (m, k, j)


def test_non_reflecting_flags_drop_one_pass_at_the_ends():
    nx, ny, nz = 20, 6, 6
    prims = random_state(nx, ny, nz, 2, "shock")
    base_kw = dict(periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0))
    g = [ghosted(a) for a in prims]
    for a in g:                                  # ghosts fixed once (the runs must share them)
        O.ghost_fill_periodic(O.params(nx, ny, nz), a)
    d0 = O.du_comput(O.params(nx, ny, nz, **base_kw), *[a.copy() for a in g])[0]
    d1 = O.du_comput(O.params(nx, ny, nz, non_ref=(1, 1, 0, 0, 0, 0), **base_kw), *[a.copy() for a in g])[0]
    diff = np.nonzero(np.any(d0 != d1, axis=(0, 1, 2)))[0]
    assert set(diff.tolist()) == {0, nx - 1}


def test_rim_form_differs_only_at_the_last_faces():
    """d0_bound 2 on a high face (the rim kernel's closures of x): the same faces are covered
    and every face value equals the D0_bound one except P at face n and M at faces n-1 and
    n.  Those are fixed by the mathematics of their stencils: on linear data P(n) is the
    exact value at the face (linear extrapolation, as the D0_bound form's slope-limited
    extrapolation is for monotone data: the two agree to rounding), M(n-1) the exact value
    at face n-1, and M(n) repeats M(n-1) bit for bit -- so the last cell's M difference is
    exactly zero, where D0_bound takes the first-order value of cell n-1."""
    n = 16
    rng = np.random.default_rng(11)
    for trial in range(30):
        line = rng.standard_normal(n + 8) if trial else 0.7 + 0.3 * np.arange(-4, n + 4, dtype=float)
        for side in ("p", "m"):
            for f in range(n + 1):
                a, b = O.closure(side, line, f, hi=1), O.closure(side, line, f, hi=2)
                assert (a is None) == (b is None), (side, f)
                if a is not None and not ((side == "p" and f == n) or (side == "m" and f >= n - 1)):
                    assert a == b, (side, f)
        a1, a2 = line[4 + n - 1], line[4 + n - 2]
        mono = (a1 - a2) != 0
        if mono:
            assert abs(O.closure("p", line, n, hi=2) - O.closure("p", line, n, hi=1)) <= 1e-14 * (abs(a1) + abs(a2))
        assert O.closure("m", line, n, hi=2) == O.closure("m", line, n - 1, hi=2)
        if trial == 0:                        # linear data: face f sits at x = f - 1/2
            assert abs(O.closure("p", line, n, hi=2) - (0.7 + 0.3 * (n - 0.5))) <= 1e-13
            assert abs(O.closure("m", line, n - 1, hi=2) - (0.7 + 0.3 * (n - 1.5))) <= 1e-13
            assert O.closure("m", line, n, hi=1) == a1


def test_rim_form_changes_only_the_last_two_cells():
    """With the rim kernel's x closures (d0_bound x = {1, 2}) the RHS differs from the
    D0_bound one only in the last two x cells (always in the last), and the M pass of the
    last cell contributes nothing."""
    nx, ny, nz = 20, 6, 6
    prims = random_state(nx, ny, nz, 4, "shock")
    g = [ghosted(a) for a in prims]
    for a in g:
        O.ghost_fill_periodic(O.params(nx, ny, nz), a)
    d1 = O.du_comput(O.params(nx, ny, nz, periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0)), *[a.copy() for a in g])[0]
    d2 = O.du_comput(O.params(nx, ny, nz, periodic=(0, 1, 1), d0_bound=(1, 2, 0, 0, 0, 0)), *[a.copy() for a in g])[0]
    diff = np.nonzero(np.any(d1 != d2, axis=(0, 1, 2)))[0]
    # cell n-2 changes only through M(n-1), whose two forms agree in exact arithmetic and
    # may round alike
    assert nx - 1 in diff.tolist() and set(diff.tolist()) <= {nx - 2, nx - 1}
    # the last cell: rim = D0_bound with its M pass dropped and P(n) replaced (the M faces
    # n-1 and n are equal); compare against the D0_bound run with the M flag of x set,
    # which drops the same M pass: they differ only through P(n), by rounding
    d3 = O.du_comput(O.params(nx, ny, nz, periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0), non_ref=(0, 1, 0, 0, 0, 0)),
                     *[a.copy() for a in g])[0]
    last = (slice(None), slice(None), slice(None), nx - 1)
    assert np.allclose(d2[last], d3[last], rtol=1e-12, atol=1e-12 * np.abs(d3).max())
