"""Shared test helpers: ghosted arrays, seeded states, error norms."""
import numpy as np

LAP = 4


def ghosted(interior):
    """Embed an interior (nz, ny, nx) array into a zeroed ghosted (nz+8, ny+8, nx+8) array."""
    nz, ny, nx = interior.shape
    g = np.zeros((nz + 2 * LAP, ny + 2 * LAP, nx + 2 * LAP))
    g[LAP:-LAP, LAP:-LAP, LAP:-LAP] = interior
    return g


def ghosted_prims(orc, p, prims):
    """(rho, u, v, w, T) interior -> ghosted d, u, v, w, T with periodic ghosts, plus cc."""
    out = [ghosted(a) for a in prims]
    for a in out:
        orc.ghost_fill_periodic(p, a)
    cc = orc.sound_speed(p, out[4])
    return out, cc


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 over all elements."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def random_state(nx, ny, nz, seed, kind="blobs"):
    from windstencil_amd import initial as I
    if kind == "blobs":
        return I.freestream_blobs(nx, ny, nz, seed)
    if kind == "vortex":
        return I.vortex_lattice(nx, ny, nz, seed, radius_cells=3.0)
    if kind == "sod":
        return I.sod_noise(nx, ny, nz, seed)
    if kind == "shock":
        return I.blobs_with_shock(nx, ny, nz, seed)
    if kind == "rough":
        rng = np.random.default_rng(seed)
        rho = 1.0 + 0.3 * rng.random((nz, ny, nx))
        u = rng.standard_normal((nz, ny, nx))
        v = rng.standard_normal((nz, ny, nx))
        w = rng.standard_normal((nz, ny, nx))
        T = 1.0 + 0.5 * rng.random((nz, ny, nx))
        return rho, u, v, w, T
    raise ValueError(kind)
