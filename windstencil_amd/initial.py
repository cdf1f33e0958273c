"""Seeded initial states for the benchmark and test workloads.

All randomness comes from an in-repo splitmix64 stream, so a (workload, seed)
pair gives the same bits on every machine.  States are returned as primitive
variables on the interior, numpy float64 arrays of shape (nz, ny, nx):
density rho, velocity u, v, w and the code temperature T.

Non-dimensionalisation follows the reference (src/runtime/OCFD_init.cpp:379-381):
Cv = 1/(gamma (gamma-1) Ma^2), internal energy rho*Cv*T = p/(gamma-1), hence
T = gamma Ma^2 p / rho and the sound speed sqrt(T)/Ma = sqrt(gamma p / rho).
Grid spacing is h = 1/(N-1) per direction; the periodic box has length N*h.
"""
import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): a 64-bit counter-based generator."""

    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        """Double in [lo, hi) from the top 53 bits."""
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo, hi):
        """Integer in [lo, hi] (inclusive)."""
        return lo + self.next_u64() % (hi - lo + 1)

    def uniform_array(self, n, lo=-1.0, hi=1.0):
        """n doubles in [lo, hi), vectorised splitmix64 (same stream as n scalar draws)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & _MASK
        return lo + (hi - lo) * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0))


def spacing(nx, ny, nz):
    return (1.0 / (nx - 1.0), 1.0 / (ny - 1.0), 1.0 / (nz - 1.0))


def _coords(nx, ny, nz):
    hx, hy, hz = spacing(nx, ny, nz)
    x = (np.arange(nx) * hx)[None, None, :]
    y = (np.arange(ny) * hy)[None, :, None]
    z = (np.arange(nz) * hz)[:, None, None]
    return x, y, z, (nx * hx, ny * hy, nz * hz)


def _periodic_delta(a, c, L):
    d = a - c
    return d - L * np.round(d / L)


def temperature(rho, p, gamma=1.4, ama=1.0):
    return gamma * ama * ama * p / rho


def freestream_blobs(nx, ny, nz, seed, mach=2.0, nblobs=8, sigma_cells=4.0, amp=0.1, gamma=1.4,
                     direction=(1.0, 0.0, 0.0)):
    """Mach-`mach` freestream (rho = 1, p = 1/gamma, so c = 1) along `direction`, with
    `nblobs` Gaussian blobs (sigma in cells, periodic distance) scaling rho and p by
    1 + sum a_k exp(-r^2 / 2 sigma^2), a_k uniform in [-amp, amp) drawn separately for
    rho and p."""
    rng = SplitMix64(seed)
    x, y, z, L = _coords(nx, ny, nz)
    hx, hy, hz = spacing(nx, ny, nz)
    frho = np.ones((nz, ny, nx))
    fp = np.ones((nz, ny, nx))
    for _ in range(nblobs):
        cx, cy, cz = rng.uniform(0, L[0]), rng.uniform(0, L[1]), rng.uniform(0, L[2])
        ar, ap = rng.uniform(-amp, amp), rng.uniform(-amp, amp)
        r2 = ((_periodic_delta(x, cx, L[0]) / hx) ** 2 + (_periodic_delta(y, cy, L[1]) / hy) ** 2
              + (_periodic_delta(z, cz, L[2]) / hz) ** 2)
        g = np.exp(-r2 / (2.0 * sigma_cells * sigma_cells))
        frho += ar * g
        fp += ap * g
    rho = frho
    p = fp / gamma
    n = math.sqrt(sum(c * c for c in direction))
    u = np.full((nz, ny, nx), mach * direction[0] / n)
    v = np.full((nz, ny, nx), mach * direction[1] / n)
    w = np.full((nz, ny, nx), mach * direction[2] / n)
    return rho, u, v, w, temperature(rho, p, gamma)


def vortex_lattice(nx, ny, nz, seed, nvort=6, beta=5.0, mach=0.5, radius_cells=None, gamma=1.4):
    """Isentropic (Yee) vortices with axes along z at seeded (x, y) centres, superposed on a
    Mach-`mach` x-freestream (rho = 1, p = 1/gamma).  Velocity and temperature
    perturbations add; rho = (T/T_inf)^(1/(gamma-1)), p = rho T / gamma."""
    rng = SplitMix64(seed)
    x, y, z, L = _coords(nx, ny, nz)
    hx, hy, _ = spacing(nx, ny, nz)
    if radius_cells is None:
        radius_cells = max(4.0, min(nx, ny) / 16.0)
    R = radius_cells * hx
    ry = radius_cells * hy
    du = np.zeros((1, ny, nx))
    dv = np.zeros((1, ny, nx))
    dT = np.zeros((1, ny, nx))
    for _ in range(nvort):
        cx, cy = rng.uniform(0, L[0]), rng.uniform(0, L[1])
        ex = _periodic_delta(x, cx, L[0]) / R
        ey = _periodic_delta(y, cy, L[1]) / ry
        r2 = ex * ex + ey * ey
        e = np.exp(0.5 * (1.0 - r2))
        # Yee's units have c_inf = sqrt(gamma); velocities are rescaled to c_inf = 1
        du += -beta / (2.0 * math.pi * math.sqrt(gamma)) * ey * e
        dv += beta / (2.0 * math.pi * math.sqrt(gamma)) * ex * e
        dT += -(gamma - 1.0) * beta * beta / (8.0 * gamma * math.pi * math.pi) * e * e
    # temperature ratio T/T_inf (code T_inf = gamma p_inf / rho_inf = 1)
    Tr = np.maximum(1.0 + dT, 0.05)
    rho = np.broadcast_to(Tr ** (1.0 / (gamma - 1.0)), (nz, ny, nx)).copy()
    T = np.broadcast_to(Tr, (nz, ny, nx)).copy()
    u = np.broadcast_to(mach + du, (nz, ny, nx)).copy()
    v = np.broadcast_to(dv, (nz, ny, nx)).copy()
    w = np.zeros((nz, ny, nx))
    return rho, u, v, w, T


def sod_noise(nx, ny, nz, seed, x0=0.4, left=(1.0, 1.0), right=(0.125, 0.1), noise=0.01, gamma=1.4):
    """Planar Sod discontinuity at x = x0*L (rho, p) = left | right, at rest, with
    +-`noise` multiplicative seeded noise on rho and p at every cell.  Periodic, so the
    box also holds the reverse jump at x = 0."""
    rng = SplitMix64(seed)
    x, _, _, L = _coords(nx, ny, nz)
    mask = np.broadcast_to(x < x0 * L[0], (nz, ny, nx))
    rho = np.where(mask, left[0], right[0]).astype(np.float64)
    p = np.where(mask, left[1], right[1]).astype(np.float64)
    n = nx * ny * nz
    rho = rho * (1.0 + noise * rng.uniform_array(n).reshape(nz, ny, nx))
    p = p * (1.0 + noise * rng.uniform_array(n).reshape(nz, ny, nx))
    z0 = np.zeros((nz, ny, nx))
    return rho, z0.copy(), z0.copy(), z0.copy(), temperature(rho, p, gamma)


def blobs_with_shock(nx, ny, nz, seed, mach=2.0, x_shock=0.5, gamma=1.4, **kw):
    """Mach-2 blob field plus a planar normal shock at x = x_shock*L obeying the
    Rankine-Hugoniot relations for a shock of Mach `mach` travelling into the
    upstream state (rho = 1, p = 1/gamma, at rest in the shock frame minus the stream)."""
    rho, u, v, w, T = freestream_blobs(nx, ny, nz, seed, mach=mach, gamma=gamma, **kw)
    x, _, _, L = _coords(nx, ny, nz)
    M2 = mach * mach
    r_rho = (gamma + 1.0) * M2 / ((gamma - 1.0) * M2 + 2.0)
    r_p = 1.0 + 2.0 * gamma / (gamma + 1.0) * (M2 - 1.0)
    down = np.broadcast_to(x >= x_shock * L[0], (nz, ny, nx))
    p = rho * T / gamma
    rho = np.where(down, rho * r_rho, rho)
    p = np.where(down, p * r_p, p)
    u = np.where(down, u / r_rho, u)
    return rho, u, v, w, temperature(rho, p, gamma)


def blast(nx, ny, nz, seed, ratio, radius=0.25, mach=0.5, noise=0.01, gamma=1.4):
    """Spherical blast: pressure `ratio` times the ambient p = 1/gamma inside a sphere of
    `radius` times the shortest box side around a seeded centre (periodic distance), the
    same density (rho = 1) everywhere, a Mach-`mach` x-stream, and +-`noise`
    multiplicative seeded noise on rho and p.  The jump of the split fluxes across the
    sphere grows like ratio (momentum) to ratio^1.5 (energy)."""
    rng = SplitMix64(seed)
    x, y, z, L = _coords(nx, ny, nz)
    c = (rng.uniform(0, L[0]), rng.uniform(0, L[1]), rng.uniform(0, L[2]))
    r2 = (_periodic_delta(x, c[0], L[0]) ** 2 + _periodic_delta(y, c[1], L[1]) ** 2
          + _periodic_delta(z, c[2], L[2]) ** 2)
    inside = np.broadcast_to(r2 < (radius * min(L)) ** 2, (nz, ny, nx))
    n = nx * ny * nz
    rho = 1.0 + noise * rng.uniform_array(n).reshape(nz, ny, nx)
    p = np.where(inside, ratio, 1.0) / gamma * (1.0 + noise * rng.uniform_array(n).reshape(nz, ny, nx))
    u = np.full((nz, ny, nx), float(mach))
    z0 = np.zeros((nz, ny, nx))
    return rho, u, z0.copy(), z0.copy(), temperature(rho, p, gamma)


def multiblock_draw(seed, nblocks=48, lo=32, hi=256):
    """Seeded block extents for the batched workload: each dimension U{lo..hi}."""
    rng = SplitMix64(seed)
    return [(rng.randint(lo, hi), rng.randint(lo, hi), rng.randint(lo, hi)) for _ in range(nblocks)]


def block_freestream(nx, ny, nz, seed, gamma=1.4):
    """One block of the batched workload: freestream Mach U[0.3, 3] in a seeded
    direction in the x-y plane, plus Gaussian blobs."""
    rng = SplitMix64(seed ^ 0xB10C)
    mach = rng.uniform(0.3, 3.0)
    ang = rng.uniform(0.0, 2.0 * math.pi)
    return freestream_blobs(nx, ny, nz, seed, mach=mach, gamma=gamma,
                            direction=(math.cos(ang), math.sin(ang), 0.0))


WORKLOADS = {
    # name: (builder, default extents)
    "blobs_64x32x16": (lambda s: freestream_blobs(64, 32, 16, s), (64, 32, 16)),
    "vortex_600x128x128": (lambda s: vortex_lattice(600, 128, 128, s), (600, 128, 128)),
    "sod_600x256x128": (lambda s: sod_noise(600, 256, 128, s), (600, 256, 128)),
    "shock_1200x256x256": (lambda s: blobs_with_shock(1200, 256, 256, s), (1200, 256, 256)),
}
