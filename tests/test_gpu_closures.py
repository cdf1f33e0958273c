"""GPU parity of the one-sided boundary closures (the reference's D0_bound path and the high-
face form of the fused path's rim kernel) and the non-reflecting flags: non-periodic directions keep the caller's ghost layers (filled here
with arbitrary valid values, given identically to the oracle).  Anchor kernels bitwise,
tiled FP64 within 1e-12 for every tile shape and direction order, mixed within 1e-6."""
import numpy as np
import pytest

import windstencil_amd as W
from tests.gpu_util import gpu_du_comput, log, make_ctx
from tests.helpers import ghosted, random_state, rel_l2

pytestmark = pytest.mark.gpu


def _ghosted_state(orc, prims, periodic, seed):
    """Ghosted d, u, v, w, T: periodic images in periodic directions, arbitrary positive
    densities / temperatures and small velocities in the others' ghost slabs."""
    nz, ny, nx = prims[0].shape
    rng = np.random.default_rng(seed)
    g = [ghosted(a) for a in prims]
    sl = {0: (slice(None), slice(None), np.r_[0:4, nx + 4:nx + 8]),
          1: (slice(None), np.r_[0:4, ny + 4:ny + 8], slice(None)),
          2: (np.r_[0:4, nz + 4:nz + 8], slice(None), slice(None))}
    for d in range(3):
        if not periodic[d]:
            for q, a in enumerate(g):
                shape = a[sl[d]].shape
                a[sl[d]] = (1.0 + 0.2 * rng.random(shape)) if q in (0, 4) else 0.1 * rng.standard_normal(shape)
    p = orc.params(nx, ny, nz, periodic=periodic)
    for a in g:
        orc.ghost_fill_periodic(p, a)
    return g


def _upload(g):
    import torch
    nz, ny, nx = (s - 8 for s in g[0].shape)
    st = W.TorchState(nx, ny, nz)
    for k, a in zip(("d", "u", "v", "w", "T"), g):
        st.t[k].copy_(torch.from_numpy(a))
    return st


def _oracle(orc, g, **kw):
    nz, ny, nx = (s - 8 for s in g[0].shape)
    p = orc.params(nx, ny, nz, **kw)
    return orc.du_comput(p, *[a.copy() for a in g])[0]


CASES = [
    ("x both", (0, 1, 1), (1, 1, 0, 0, 0, 0), (0,) * 6),
    ("y both", (1, 0, 1), (0, 0, 1, 1, 0, 0), (0,) * 6),
    ("z both", (1, 1, 0), (0, 0, 0, 0, 1, 1), (0,) * 6),
    ("x low only", (0, 1, 1), (1, 0, 0, 0, 0, 0), (0,) * 6),
    ("x high only + y both", (0, 0, 1), (0, 1, 1, 1, 0, 0), (0,) * 6),
    ("x non-reflecting", (0, 1, 1), (1, 1, 0, 0, 0, 0), (1, 1, 0, 0, 0, 0)),
    ("all closed, non-reflecting", (0, 0, 0), (1,) * 6, (1, 0, 0, 1, 1, 1)),
    # the rim kernel's high-face form, and the fused path's closures as a whole: x {1, 2},
    # y {1, 1} with the M pass of y dropped in the last cell, z from the ghosts
    ("x high rim form", (0, 1, 1), (1, 2, 0, 0, 0, 0), (0,) * 6),
    ("fused-path rim closures", (0, 0, 1), (1, 2, 1, 1, 0, 0), (0, 0, 0, 1, 0, 0)),
]


@pytest.mark.parametrize("name,periodic,d0,nr", CASES, ids=[c[0] for c in CASES])
def test_closures_match_oracle(orc, ws, name, periodic, d0, nr):
    nx, ny, nz = 37, 26, 21
    prims = random_state(nx, ny, nz, 14, "shock")
    g = _ghosted_state(orc, prims, periodic, 3)
    kw = dict(periodic=periodic, d0_bound=d0, non_ref=nr)
    ref = _oracle(orc, g, **kw)
    actx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR, **kw)
    da = gpu_du_comput(actx, _upload(g))
    log(f"closures {name}: anchor max |diff| {np.abs(da - ref).max():.3e}")
    assert np.array_equal(da, ref)
    for tile in W.TILE_SHAPES:
        for order in W.ORDERS:
            tctx = make_ctx(nx, ny, nz, **kw)
            tctx.set_tile_shape(tile)
            tctx.set_direction_order(order)
            err = rel_l2(gpu_du_comput(tctx, _upload(g)), ref)
            log(f"  tile {tile} order {order}: rel L2 {err:.3e}")
            assert err <= 1e-12
    for order in W.ORDERS:
        mctx = make_ctx(nx, ny, nz, precision=W.WS_PREC_MIXED, **kw)
        mctx.set_direction_order(order)
        assert rel_l2(gpu_du_comput(mctx, _upload(g)), ref) <= 1e-6


def test_closures_on_a_sub_box_match_the_full_box(orc, ws):
    nx, ny, nz = 40, 20, 16
    kw = dict(periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0), non_ref=(1, 1, 0, 0, 0, 0))
    prims = random_state(nx, ny, nz, 6, "blobs")
    g = _ghosted_state(orc, prims, kw["periodic"], 4)
    for kern in (W.WS_KERNEL_ANCHOR, W.WS_KERNEL_TILED):
        ctx = make_ctx(nx, ny, nz, kernel=kern, **kw)
        full = gpu_du_comput(ctx, _upload(g))
        st = _upload(g)
        st.t["du"].fill_(5.0)
        part = gpu_du_comput(ctx, st, job=W.job((4, 4 + 2, 4 + 1), (4 + 13, 4 + 18, 4 + 15)))
        inside = (slice(None), slice(1, 15), slice(2, 18), slice(0, 13))
        assert np.array_equal(part[inside], full[inside])


def test_closure_configuration_errors(ws):
    L = W.lib()
    cfg = W.default_config(16, 16, 16, d0_bound=(1, 0, 0, 0, 0, 0))
    import ctypes
    h = ctypes.c_void_p()
    assert L.ws_ctx_create(ctypes.byref(cfg), ctypes.byref(h)) == W.WS_ERR_INVALID   # periodic x
    cfg = W.default_config(6, 16, 16, periodic=(0, 1, 1), d0_bound=(1, 1, 0, 0, 0, 0))
    assert L.ws_ctx_create(ctypes.byref(cfg), ctypes.byref(h)) == W.WS_ERR_INVALID   # too few cells
    cfg = W.default_config(16, 16, 16, periodic=(0, 1, 1), viscous=6)
    assert L.ws_ctx_create(ctypes.byref(cfg), ctypes.byref(h)) == W.WS_ERR_UNSUPPORTED
