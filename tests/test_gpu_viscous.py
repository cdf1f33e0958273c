"""GPU parity of the viscous terms (the reference's du_vis_Jacobian3d path) against the
oracle.  With the anchor kernel family the viscous kernels follow the reference's
expression order with contraction off and its divisions, so they match the oracle bit
for bit, alone, in the whole RHS and in RK3 steps.  The tiled family's viscous kernels
scale each stencil sum by 1/(60 h) instead of dividing twice: within 1e-13 relative L2
alone, and the total RHS within 1e-12."""
import ctypes

import numpy as np
import pytest

import windstencil_amd as W
from tests.gpu_util import gpu_du_comput, log, make_ctx, upload_state
from tests.helpers import ghosted, random_state, rel_l2

pytestmark = pytest.mark.gpu

RE, PR, TREF = 500.0, 0.72, 288.15


def _vis_cfg(order, **kw):
    return dict(viscous=order, re=RE, pr=PR, ref_t=TREF, **kw)


def _oracle_full(orc, prims, order, metrics=None, ajac=None):
    nz, ny, nx = prims[0].shape
    p = orc.params(nx, ny, nz, viscous=order, re=RE, pr=PR, ref_t=TREF)
    g = [ghosted(a) for a in prims]
    du, _ = orc.du_comput(p, *g, metrics=metrics, ajac=ajac)
    return du


@pytest.mark.parametrize("kernel", [W.WS_KERNEL_ANCHOR, W.WS_KERNEL_TILED])
@pytest.mark.parametrize("order", [6, 8])
@pytest.mark.parametrize("kind", ["blobs", "vortex", "rough"])
def test_viscous_terms(orc, ws, order, kind, kernel):
    import torch
    nx, ny, nz = 33, 21, 17
    prims = random_state(nx, ny, nz, 3, kind)
    p = orc.params(nx, ny, nz)
    g = [ghosted(a) for a in prims]
    for a in g:
        orc.ghost_fill_periodic(p, a)
    ref = orc.viscous_rhs(p, g[1], g[2], g[3], g[4], RE, PR, TREF, order)
    ctx = make_ctx(nx, ny, nz, kernel=kernel, **_vis_cfg(order))
    st = upload_state(prims)
    L = W.lib()
    s = W.stream_handle()
    fl = (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T)
    W.check(L.ws_ghost_fill_periodic(ctx.handle, fl, 5, s))
    st.t["du"].zero_()
    W.check(L.ws_viscous_rhs(ctx.handle, st.ws.u, st.ws.v, st.ws.w, st.ws.T, None, st.ws.du, s))
    torch.cuda.synchronize()
    du = st.interior("du").cpu().numpy()
    err = rel_l2(du, ref)
    log(f"viscous {kind} order {order} kernel {kernel}: rel L2 {err:.3e}")
    if kernel == W.WS_KERNEL_ANCHOR:
        assert np.array_equal(du, ref)
    else:
        assert err <= 1e-13


CONST_METRICS = {
    # a sheared constant metric set (every term of the metric products in play) and a
    # non-identity diagonal one (the fused kernel's diagonal instantiation)
    "sheared": (1.0, 0.2, -0.1, 0.15, 0.9, 0.05, -0.05, 0.12, 1.1, 0.8),
    "diagonal": (2.0, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 1.5, 0.7),
}


@pytest.mark.parametrize("kernel", [W.WS_KERNEL_ANCHOR, W.WS_KERNEL_TILED])
@pytest.mark.parametrize("order", [6, 8])
@pytest.mark.parametrize("mname", list(CONST_METRICS))
def test_viscous_terms_constant_metrics(orc, ws, mname, order, kernel):
    """Constant metrics given in the configuration (no metric fields): the oracle gets the
    same values as uniform fields.  The tiled family runs the fused kernel's general form
    for the sheared set and its diagonal form for the diagonal one."""
    import torch
    met = CONST_METRICS[mname]
    nx, ny, nz = 29, 22, 19
    prims = random_state(nx, ny, nz, 17, "blobs")
    p = orc.params(nx, ny, nz)
    g = [ghosted(a) for a in prims]
    for a in g:
        orc.ghost_fill_periodic(p, a)
    fields = [np.full_like(g[0], m) for m in met]
    ref = orc.viscous_rhs(p, g[1], g[2], g[3], g[4], RE, PR, TREF, order, metrics=fields[:9], ajac=fields[9])
    ctx = make_ctx(nx, ny, nz, kernel=kernel, metric=met, **_vis_cfg(order))
    st = upload_state(prims)
    L = W.lib()
    s = W.stream_handle()
    fl = (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T)
    W.check(L.ws_ghost_fill_periodic(ctx.handle, fl, 5, s))
    st.t["du"].zero_()
    W.check(L.ws_viscous_rhs(ctx.handle, st.ws.u, st.ws.v, st.ws.w, st.ws.T, None, st.ws.du, s))
    torch.cuda.synchronize()
    du = st.interior("du").cpu().numpy()
    err = rel_l2(du, ref)
    log(f"viscous {mname} metrics order {order} kernel {kernel}: rel L2 {err:.3e}")
    if kernel == W.WS_KERNEL_ANCHOR:
        assert np.array_equal(du, ref)
    else:
        assert err <= 1e-13


@pytest.mark.parametrize("shape", [(4, 4, 4), (5, 9, 4), (37, 6, 41)])
def test_viscous_terms_small_and_ragged_extents(orc, ws, shape):
    """Extents at the halo minimum and ragged ones (partial column tiles, z chunks
    shorter than a march): the anchor family bit for bit, the tiled family within 1e-13."""
    nx, ny, nz = shape
    prims = random_state(nx, ny, nz, 21, "rough")
    ref = _oracle_full(orc, prims, 8)
    got = gpu_du_comput(make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR, **_vis_cfg(8)), upload_state(prims))
    assert np.array_equal(got, ref)
    err = rel_l2(gpu_du_comput(make_ctx(nx, ny, nz, **_vis_cfg(8)), upload_state(prims)), ref)
    log(f"  {shape}: tiled + viscous rel L2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("order", [6, 8])
def test_full_rhs_with_viscous_terms(orc, ws, order):
    nx, ny, nz = 40, 24, 20
    prims = random_state(nx, ny, nz, 8, "shock")
    ref = _oracle_full(orc, prims, order)
    actx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR, **_vis_cfg(order))
    assert np.array_equal(gpu_du_comput(actx, upload_state(prims)), ref)
    tctx = make_ctx(nx, ny, nz, **_vis_cfg(order))
    err = rel_l2(gpu_du_comput(tctx, upload_state(prims)), ref)
    log(f"  tiled + viscous order {order}: rel L2 {err:.3e}")
    assert err <= 1e-12
    mctx = make_ctx(nx, ny, nz, precision=W.WS_PREC_MIXED, **_vis_cfg(order))
    assert rel_l2(gpu_du_comput(mctx, upload_state(prims)), ref) <= 1e-6


def test_viscous_terms_with_metric_fields(orc, ws):
    """Spatially varying metric terms and Jacobian through both viscous kernels."""
    import torch
    from tests.test_gpu_rhs import _wavy_metrics
    nx, ny, nz = 30, 20, 16
    prims = random_state(nx, ny, nz, 4, "vortex")
    met = _wavy_metrics(nx, ny, nz)
    p = orc.params(nx, ny, nz)
    g = [ghosted(a) for a in met]
    for a in g:
        orc.ghost_fill_periodic(p, a)
    ref = _oracle_full(orc, prims, 6, metrics=g[:9], ajac=g[9])
    tens = [torch.from_numpy(a).cuda() for a in g]
    mf = W.WsMetricFields()
    for i in range(9):
        mf.A[i] = W.field_of(tens[i])
    mf.ajac = W.field_of(tens[9])
    actx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR, **_vis_cfg(6))
    assert np.array_equal(gpu_du_comput(actx, upload_state(prims), metrics=mf), ref)


def test_rk3_with_viscous_terms_bitwise(orc, ws):
    nx, ny, nz = 32, 20, 12
    prims = random_state(nx, ny, nz, 12, "blobs")
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR, **_vis_cfg(6))
    st = upload_state(prims)
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), s))
    W.check(L.ws_rk3_run(ctx.handle, ctypes.byref(st.ws), 5, 0.5, 1, s))
    st.sync_from_ws()
    f_gpu = st.interior("f").cpu().numpy()
    p = orc.params(nx, ny, nz, viscous=6, re=RE, pr=PR, ref_t=TREF)
    g = [ghosted(a) for a in prims]
    f_ref = orc.pri_to_cons(p, *g)
    orc.rk3_run(p, 5, -0.5, f_ref, *g)
    assert np.array_equal(f_gpu, f_ref)


def test_batch_with_viscous_terms_equals_single_blocks(ws):
    from tests.test_gpu_batch import _batch, _extents, _du
    extents = _extents(23, 4, 8, 30)
    b, prims = _batch(extents, 900, **_vis_cfg(8))
    W.check(W.lib().ws_batch_du_comput(b.handle, W.stream_handle()))
    import torch
    torch.cuda.synchronize()
    for e, pr_, st in zip(extents, prims, b.states):
        ref = gpu_du_comput(make_ctx(*e, **_vis_cfg(8)), upload_state(pr_))
        assert rel_l2(_du(st), ref) <= 1e-13, e


CLOSED = [((0, 0, 0), (1, 1, 1, 1, 1, 1)), ((0, 1, 1), (1, 1, 0, 0, 0, 0)), ((1, 1, 0), (0, 0, 0, 0, 1, 1)),
          ((1, 0, 1), (0, 0, 1, 1, 0, 0))]


@pytest.mark.parametrize("kernel", [W.WS_KERNEL_ANCHOR, W.WS_KERNEL_TILED])
@pytest.mark.parametrize("order", [6, 8])
@pytest.mark.parametrize("faces", CLOSED)
def test_viscous_terms_with_closed_faces(orc, ws, faces, order, kernel):
    """Directions closed on both faces (D0_bound) take the one-sided and low-order forms in
    their four boundary cells, for the gradients and for the flux derivatives; the ghost
    layers of those directions are never read (they hold zeros here).  Anchor kernels bit
    for bit against the oracle, the tiled family within 1e-13, alone and (with the inviscid
    closures) in the whole du_comput within 1e-12."""
    import torch
    periodic, bnd = faces
    nx, ny, nz = 29, 22, 18
    prims = random_state(nx, ny, nz, 9, "vortex")
    p = orc.params(nx, ny, nz, periodic=periodic, d0_bound=bnd)
    g = [ghosted(a) for a in prims]
    for a in g:
        orc.ghost_fill_periodic(p, a)
    ref = orc.viscous_rhs(p, g[1], g[2], g[3], g[4], RE, PR, TREF, order)
    ctx = make_ctx(nx, ny, nz, kernel=kernel, periodic=periodic, d0_bound=bnd, **_vis_cfg(order))
    st = upload_state(prims)
    L = W.lib()
    s = W.stream_handle()
    fl = (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T)
    W.check(L.ws_ghost_fill_periodic(ctx.handle, fl, 5, s))
    st.t["du"].zero_()
    W.check(L.ws_viscous_rhs(ctx.handle, st.ws.u, st.ws.v, st.ws.w, st.ws.T, None, st.ws.du, s))
    torch.cuda.synchronize()
    du = st.interior("du").cpu().numpy()
    err = rel_l2(du, ref)
    log(f"viscous closed {faces} order {order} kernel {kernel}: rel L2 {err:.3e}")
    if kernel == W.WS_KERNEL_ANCHOR:
        assert np.array_equal(du, ref)
    else:
        assert err <= 1e-13
    pf = orc.params(nx, ny, nz, periodic=periodic, d0_bound=bnd, viscous=order, re=RE, pr=PR, ref_t=TREF)
    full, _ = orc.du_comput(pf, *[ghosted(a) for a in prims])
    dg = gpu_du_comput(ctx, upload_state(prims))
    err = rel_l2(dg, full)
    log(f"  du_comput rel L2 {err:.3e}")
    if kernel == W.WS_KERNEL_ANCHOR:
        assert np.array_equal(dg, full)
    else:
        assert err <= 1e-12
