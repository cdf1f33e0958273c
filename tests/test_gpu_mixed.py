"""GPU tests of the mixed-precision RHS: FP64 splitting, merge and flux difference,
FP32 WENO7 smoothness, weights and candidates on first differences of the split
fluxes taken relative to an FP64 base point per face.

Tolerance: RHS relative L2 <= 1e-6 against the FP64 oracle (the mixed-precision bar),
over all interior cells and all five components; the same bar for the conserved
state after SSP-RK3 steps.  Properties that do not depend on rounding hold exactly:
a uniform state gives zero and a periodic shift of the input shifts the RHS bit for
bit.
"""
import ctypes

import numpy as np
import pytest

import windstencil_amd as W
from tests.gpu_util import gpu_du_comput, log, make_ctx, oracle_du, upload_state
from tests.helpers import ghosted, random_state, rel_l2

pytestmark = pytest.mark.gpu

SMALL = [(24, 16, 12), (31, 17, 19), (64, 32, 16), (67, 45, 37)]
KINDS = ["blobs", "vortex", "sod", "shock", "rough"]
TOL_MIXED = 1e-6


def _mixed_ctx(nx, ny, nz, **kw):
    return make_ctx(nx, ny, nz, precision=W.WS_PREC_MIXED, **kw)


@pytest.mark.parametrize("tile", W.TILE_SHAPES)
@pytest.mark.parametrize("shape", SMALL)
@pytest.mark.parametrize("kind", KINDS)
def test_mixed_rhs_matches_oracle(orc, ws, shape, kind, tile):
    nx, ny, nz = shape
    log(f"mixed {kind} {shape} tile {tile}")
    prims = random_state(nx, ny, nz, 23, kind)
    ctx = _mixed_ctx(nx, ny, nz)
    ctx.set_tile_shape(tile)
    du = gpu_du_comput(ctx, upload_state(prims))
    ref = oracle_du(orc, prims)
    err = rel_l2(du, ref)
    log(f"  rel L2 {err:.3e}")
    assert err <= TOL_MIXED


def test_mixed_is_a_different_path_from_fp64(ws):
    """The precision switch reaches the mixed kernels: their result differs from the
    FP64 kernels' in the last bits but not beyond the bar."""
    nx, ny, nz = 40, 24, 20
    prims = random_state(nx, ny, nz, 12, "rough")
    a = gpu_du_comput(make_ctx(nx, ny, nz), upload_state(prims))
    b = gpu_du_comput(_mixed_ctx(nx, ny, nz), upload_state(prims))
    err = rel_l2(b, a)
    log(f"  mixed vs fp64 rel L2 {err:.3e}")
    assert 1e-12 < err <= TOL_MIXED


def test_mixed_exact_properties(ws):
    """Uniform state -> exactly zero; periodic shift -> RHS shifted bit for bit;
    totals conserved to summation rounding."""
    nx, ny, nz = 45, 32, 20
    ctx = _mixed_ctx(nx, ny, nz)
    uni = [np.full((nz, ny, nx), v) for v in (1.1, 0.4, -0.3, 0.2, 1.3)]
    assert np.all(gpu_du_comput(ctx, upload_state(uni)) == 0.0)
    prims = random_state(nx, ny, nz, 5, "shock")
    du = gpu_du_comput(ctx, upload_state(prims))
    rolled = [np.roll(a, (3, 5, 7), axis=(0, 1, 2)) for a in prims]
    dr = gpu_du_comput(ctx, upload_state(rolled))
    assert np.array_equal(dr, np.roll(du, (3, 5, 7), axis=(1, 2, 3)))
    for m in range(5):
        assert abs(du[m].sum()) <= 1e-12 * np.abs(du[m]).sum()


def test_mixed_sub_box_and_accumulate(ws):
    import torch
    nx, ny, nz = 40, 24, 20
    prims = random_state(nx, ny, nz, 3, "vortex")
    ctx = _mixed_ctx(nx, ny, nz)
    full = gpu_du_comput(ctx, upload_state(prims))
    st = upload_state(prims)
    st.t["du"].fill_(7.0)
    box = W.job((4 + 5, 4 + 3, 4 + 2), (4 + 33, 4 + 21, 4 + 17))
    part = gpu_du_comput(ctx, st, job=box)
    inside = (slice(None), slice(2, 17), slice(3, 21), slice(5, 33))
    assert np.array_equal(part[inside], full[inside])
    mask = np.ones(part.shape, bool)
    mask[inside] = False
    assert np.all(part[mask] == 7.0)
    st.t["du"].fill_(0.25)
    acc = gpu_du_comput(ctx, st, accumulate=True)
    assert rel_l2(acc - 0.25, full) <= 1e-13


def test_anchor_kernels_stay_fp64_under_mixed(orc, ws):
    nx, ny, nz = 24, 16, 12
    prims = random_state(nx, ny, nz, 2, "blobs")
    ctx = _mixed_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR)
    assert np.array_equal(gpu_du_comput(ctx, upload_state(prims)), oracle_du(orc, prims))


def _rk3(prims, nsteps, precision, graph=True, cfl=0.5):
    import torch
    nz, ny, nx = prims[0].shape
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz, precision=precision)
    st = upload_state(prims)
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), s))
    W.check(L.ws_rk3_run(ctx.handle, ctypes.byref(st.ws), nsteps, cfl, 1 if graph else 0, s), "rk3 run")
    st.sync_from_ws()
    torch.cuda.synchronize()
    return st


@pytest.mark.parametrize("kind", ["blobs", "vortex", "shock"])
def test_mixed_rk3_matches_oracle(orc, ws, kind):
    prims = random_state(48, 32, 16, 11, kind)
    log(f"mixed rk3 {kind} x10")
    st = _rk3(prims, 10, W.WS_PREC_MIXED, graph=False)
    f_gpu = st.interior("f").cpu().numpy()
    p = orc.params(48, 32, 16)
    g = [ghosted(a) for a in prims]
    f_ref = orc.pri_to_cons(p, *g)
    orc.rk3_run(p, 10, -0.5, f_ref, *g)
    err = rel_l2(f_gpu, f_ref)
    log(f"  rel L2 {err:.3e}")
    assert err <= TOL_MIXED


def test_mixed_sod_600x256x128_rhs_and_200_steps(ws):
    """The mixed-precision workload at its full size: Sod plane at x = 0.4 L with 1 %
    seeded noise, periodic.  One RHS against the FP64 kernels, then the conserved state
    after 200 SSP-RK3 steps (CFL 0.5, graph-replayed) against the FP64 run: relative
    L2, and Linf relative to max |f| over the state, within the bar."""
    import torch
    from windstencil_amd import initial as I
    nx, ny, nz = 600, 256, 128
    log("mixed 600x256x128: sod + noise")
    prims = I.sod_noise(nx, ny, nz, 20240601)
    st = upload_state(prims)
    d64 = torch.from_numpy(gpu_du_comput(make_ctx(nx, ny, nz), st)).cuda()
    dmx = torch.from_numpy(gpu_du_comput(_mixed_ctx(nx, ny, nz), st)).cuda()
    rhs_l2 = float(torch.linalg.vector_norm(dmx - d64) / torch.linalg.vector_norm(d64))
    rhs_linf = float(((dmx - d64).abs().amax(dim=(1, 2, 3)) / d64.abs().amax(dim=(1, 2, 3))).max())
    log(f"  RHS rel L2 {rhs_l2:.3e}  rel Linf {rhs_linf:.3e}")
    assert rhs_l2 <= TOL_MIXED
    del st, d64, dmx
    a = _rk3(prims, 200, W.WS_PREC_FP64)
    fa = a.interior("f").clone()
    del a
    b = _rk3(prims, 200, W.WS_PREC_MIXED)
    fb = b.interior("f")
    l2 = float(torch.linalg.vector_norm(fb - fa) / torch.linalg.vector_norm(fa))
    linf = float((fb - fa).abs().max() / fa.abs().max())
    log(f"  200 steps: state rel L2 {l2:.3e}  Linf / max|f| {linf:.3e}")
    for m in range(5):
        dm = (fb[m] - fa[m]).abs()
        log(f"    component {m}: rel L2 {float(torch.linalg.vector_norm(fb[m] - fa[m]) / torch.linalg.vector_norm(fa[m])):.3e}"
            f"  Linf {float(dm.max()):.3e}  max|f| {float(fa[m].abs().max()):.3e}")
    assert l2 <= TOL_MIXED and linf <= TOL_MIXED


@pytest.mark.parametrize("ratio", [1e1, 1e2, 1e3, 1e4, 1e6])
def test_mixed_rhs_at_strong_discontinuities(orc, ws, ratio):
    """Spherical blasts with pressure ratios 1e1..1e6 (split-flux jumps of about 1e1..1e9
    per cell): the FP32 weights of the faces next to the jump leave FP32's range from a
    ratio of about 1e3 on (tests/test_mixed_range.py), those faces fall back to FP64, and
    the mixed RHS stays finite and within the mixed bar of the FP64 oracle."""
    from windstencil_amd import initial as I
    nx, ny, nz = 48, 20, 16
    prims = I.blast(nx, ny, nz, 3, ratio)
    ref = oracle_du(orc, prims)
    for tile in (W.WS_TILE_30x16_K2, W.WS_TILE_126x2_K2, W.WS_TILE_15x16_K1):
        ctx = _mixed_ctx(nx, ny, nz)
        ctx.set_tile_shape(tile)
        du = gpu_du_comput(ctx, upload_state(prims))
        err = rel_l2(du, ref)
        log(f"blast ratio {ratio:g} tile {tile}: finite {np.isfinite(du).all()} rel L2 {err:.3e}")
        assert np.isfinite(du).all()
        assert err <= TOL_MIXED
