"""Pins of the oracle's metric computation from mesh coordinates (the reference's
comput_Jac3d): closed forms that fix the derivative stencils, the cofactor assembly,
the Jacobian and the ghost layers independently of how the oracle spells them."""
import numpy as np

from oracle import oracle as O

LAP = 4


def _index_grids(p):
    """Ghosted (nz+8, ny+8, nx+8) arrays of i*hx, j*hy, k*hz over the whole ghosted box."""
    nx, ny, nz = p.nx, p.ny, p.nz
    k, j, i = np.meshgrid(np.arange(-LAP, nz + LAP), np.arange(-LAP, ny + LAP), np.arange(-LAP, nx + LAP),
                          indexing="ij")
    return i.astype(float), j.astype(float), k.astype(float)


def _interior(a, p):
    return a[LAP:LAP + p.nz, LAP:LAP + p.ny, LAP:LAP + p.nx]


def test_affine_mesh_gives_the_adjugate_and_inverse_determinant():
    """x = A (i h) + b: every derivative form is exact for linear data, so the metric terms
    are the adjugate of A (row d: the gradient of index d times det A) and Ajac = 1/det A,
    with or without the one-sided boundary forms."""
    nx, ny, nz = 13, 11, 9
    p = O.params(nx, ny, nz, periodic=(0, 0, 0))
    A = np.array([[1.3, 0.2, -0.1], [0.1, 0.9, 0.3], [-0.2, 0.15, 1.1]])
    i, j, k = _index_grids(p)
    I, J, K = i * p.hx, j * p.hy, k * p.hz
    xyz = [A[c, 0] * I + A[c, 1] * J + A[c, 2] * K + (0.5, 0.0, -1.0)[c] for c in range(3)]
    adj = np.linalg.inv(A) * np.linalg.det(A)
    for jb in ((0, 0, 0), (1, 1, 1), (1, 0, 1)):
        out = O.comput_jacobian(p, *xyz, jacbound=jb)
        for m in range(9):
            np.testing.assert_allclose(_interior(out[m], p), adj[m // 3, m % 3], rtol=0, atol=1e-13)
        np.testing.assert_allclose(_interior(out[9], p), 1.0 / np.linalg.det(A), rtol=1e-13)


def test_periodic_direction_uses_the_discrete_cd6_symbol():
    """Cylindrical mesh: x = i hx, y = r cos(theta_j), z = r sin(theta_j), theta_j = 2 pi j / ny
    (periodic in j), r = 1 + k hz / 2 (one-sided forms in i and k).  The sixth-order
    central difference of cos(theta) over the periodic halo is -sin(theta) * s with
    s = (45 sin D - 9 sin 2D + sin 3D) / (30 hy), D = 2 pi / ny, exactly; i and k
    derivatives of linear data are exact."""
    nx, ny, nz = 10, 24, 9
    p = O.params(nx, ny, nz, periodic=(0, 1, 0))
    i, j, k = _index_grids(p)
    D = 2 * np.pi / ny
    th = j * D
    r = 1.0 + 0.5 * k * p.hz
    x, y, z = i * p.hx, r * np.cos(th), r * np.sin(th)
    for a in (x, y, z):
        O.ghost_fill_periodic(p, a)
    out = O.comput_jacobian(p, x, y, z, jacbound=(1, 0, 1))
    s = (45 * np.sin(D) - 9 * np.sin(2 * D) + np.sin(3 * D)) / (30 * p.hy)
    thi, ri = _interior(th, p), _interior(r, p)
    xi, yj, zj, yk, zk = 1.0, -ri * np.sin(thi) * s, ri * np.cos(thi) * s, 0.5 * np.cos(thi), 0.5 * np.sin(thi)
    # cofactors with xj = xk = yi = zi = 0
    expect = [yj * zk - zj * yk, 0.0, 0.0, 0.0, zk * xi, -yk * xi, 0.0, -zj * xi, yj * xi]
    for m in range(9):
        np.testing.assert_allclose(_interior(out[m], p), expect[m], rtol=0, atol=2e-13)
    np.testing.assert_allclose(_interior(out[9], p), 1.0 / (xi * (yj * zk - zj * yk)), rtol=1e-12)


def test_boundary_forms_have_their_taylor_errors():
    """x = g(i) = (i h)^4 + i h along i only (y = j hy, z = k hz): Asz = dx/di and
    Ajac = 1/(dx/di).  With the one-sided forms the derivative at cell 0 is the forward
    difference, g'(0) + h^3 (Taylor: h f''''/24 h^3); at cell 1 the central difference,
    g'(h) + h^2 g'''(h)/6 + h^4 g^(5)/120 = g'(h) + 4 h^3; cells 2 and 3 (fourth- and
    sixth-order forms) and the interior (sixth order) are exact for a quartic; the last
    cells mirror them (forward -> backward: g'(L) - h^3 ... from the same expansion)."""
    nx, ny, nz = 16, 6, 5
    p = O.params(nx, ny, nz, periodic=(0, 0, 0))
    h = p.hx
    i, j, k = _index_grids(p)
    g = lambda t: t ** 4 + t                      # noqa: E731
    dg = lambda t: 4 * t ** 3 + 1.0               # noqa: E731
    # y and z linear over the whole ghosted box: their central forms are exact there
    x, y, z = g(i * h), j * p.hy, k * p.hz
    out = O.comput_jacobian(p, x, y, z, jacbound=(1, 0, 0))
    asz = _interior(out[8], p)[0, 0, :]
    exact = dg(np.arange(nx) * h)
    n = nx - 1
    # cell 0: (g(h) - g(0)) / h = 1 + h^3 ; cell n: (g(L) - g(L-h)) / h = g'(L) - 6 L^2 h + 4 L h^2 - h^3
    L = n * h
    assert np.isclose(asz[0], 1.0 + h ** 3, rtol=1e-13)
    assert np.isclose(asz[n], dg(L) - 6 * L ** 2 * h + 4 * L * h ** 2 - h ** 3, rtol=1e-12)
    # cells 1 and n-1: central differences, error h^2 g'''/6 + h^4 g^(5)/120 = 4 t h^2
    assert np.isclose(asz[1], dg(h) + 4 * h * h ** 2, rtol=1e-12)
    assert np.isclose(asz[n - 1], dg(L - h) + 4 * (L - h) * h ** 2, rtol=1e-12)
    np.testing.assert_allclose(asz[2:n - 1], exact[2:n - 1], rtol=1e-12)
    np.testing.assert_allclose(_interior(out[9], p)[0, 0, :], 1.0 / asz, rtol=1e-15)


def test_ghost_layers_copy_the_face_cell_then_the_periodic_halo_overrides():
    """Non-periodic faces: the LAP ghost cells of each face repeat the adjacent interior
    cell; periodic directions: the periodic halo; edges and corners untouched."""
    nx, ny, nz = 9, 8, 7
    p = O.params(nx, ny, nz, periodic=(0, 1, 0))
    rng = np.random.default_rng(3)
    i, j, k = _index_grids(p)
    x = i * p.hx + 0.01 * rng.standard_normal(i.shape)
    y = np.cos(2 * np.pi * j / ny) * (1 + 0.1 * k * p.hz)
    z = np.sin(2 * np.pi * j / ny) * (1 + 0.1 * k * p.hz)
    for a in (x, y, z):
        O.ghost_fill_periodic(p, a)
    out = O.comput_jacobian(p, x, y, z, jacbound=(1, 0, 1))
    for a in out:
        c = a[LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx]
        for g in range(LAP):
            np.testing.assert_array_equal(a[LAP:LAP + nz, LAP:LAP + ny, g], c[:, :, 0])
            np.testing.assert_array_equal(a[LAP:LAP + nz, LAP:LAP + ny, LAP + nx + g], c[:, :, -1])
            np.testing.assert_array_equal(a[g, LAP:LAP + ny, LAP:LAP + nx], c[0])
            np.testing.assert_array_equal(a[LAP + nz + g, LAP:LAP + ny, LAP:LAP + nx], c[-1])
            # periodic in y
            np.testing.assert_array_equal(a[LAP:LAP + nz, g, LAP:LAP + nx], c[:, ny - LAP + g, :])
            np.testing.assert_array_equal(a[LAP:LAP + nz, LAP + ny + g, LAP:LAP + nx], c[:, g, :])
        assert a[0, 0, 0] == 1.0      # corner: the initial value
