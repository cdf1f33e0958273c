"""GPU parity of the inviscid RHS against the CPU oracle (through the C ABI).

Tolerances: the anchor kernels follow the reference's operation order with no
contraction and must equal the oracle bit for bit; the tiled FP64 kernels must
agree within 1e-12 relative L2 (the drop-in bar), over all interior cells and
all five components.
"""
import ctypes

import numpy as np
import pytest

import windstencil_amd as W
from tests.gpu_util import gpu_du_comput, log, make_ctx, oracle_du, upload_state
from tests.helpers import random_state, rel_l2

pytestmark = pytest.mark.gpu

SMALL = [(24, 16, 12), (31, 17, 19), (64, 32, 16)]
KINDS = ["blobs", "vortex", "sod", "shock", "rough"]
TOL_FP64 = 1e-12


@pytest.mark.parametrize("shape", SMALL)
@pytest.mark.parametrize("kind", KINDS)
def test_anchor_is_bitwise_equal_to_oracle(orc, ws, shape, kind):
    nx, ny, nz = shape
    log(f"anchor {kind} {shape}")
    prims = random_state(nx, ny, nz, 17, kind)
    ctx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR)
    du = gpu_du_comput(ctx, upload_state(prims))
    ref = oracle_du(orc, prims)
    assert np.array_equal(du, ref), f"max |diff| {np.abs(du - ref).max()} rel L2 {rel_l2(du, ref)}"


@pytest.mark.parametrize("tile", W.TILE_SHAPES)
@pytest.mark.parametrize("shape", SMALL + [(67, 45, 37)])
@pytest.mark.parametrize("kind", KINDS)
def test_tiled_fp64_matches_oracle(orc, ws, shape, kind, tile):
    nx, ny, nz = shape
    log(f"tiled {kind} {shape} tile {tile}")
    prims = random_state(nx, ny, nz, 23, kind)
    ctx = make_ctx(nx, ny, nz)
    ctx.set_tile_shape(tile)
    du = gpu_du_comput(ctx, upload_state(prims))
    ref = oracle_du(orc, prims)
    err = rel_l2(du, ref)
    log(f"  rel L2 {err:.3e}")
    assert err <= TOL_FP64


@pytest.mark.parametrize("shape", [(4, 4, 4), (4, 9, 5), (7, 4, 6), (5, 6, 4), (129, 65, 33)])
def test_smallest_and_ragged_extents(orc, ws, shape):
    """Extents at the minimum the periodic halo allows (4 cells, where the halo is the
    whole interior) and ragged sizes far from any tile multiple: anchor bitwise, tiled
    FP64 within the drop-in bar for every tile shape, mixed within 1e-6."""
    nx, ny, nz = shape
    prims = random_state(nx, ny, nz, 17, "rough")
    ref = oracle_du(orc, prims)
    assert np.array_equal(gpu_du_comput(make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR), upload_state(prims)), ref)
    for tile in W.TILE_SHAPES:
        ctx = make_ctx(nx, ny, nz)
        ctx.set_tile_shape(tile)
        err = rel_l2(gpu_du_comput(ctx, upload_state(prims)), ref)
        log(f"  {shape} tile {tile}: rel L2 {err:.3e}")
        assert err <= TOL_FP64
    err = rel_l2(gpu_du_comput(make_ctx(nx, ny, nz, precision=W.WS_PREC_MIXED), upload_state(prims)), ref)
    assert err <= 1e-6


def test_tile_shapes_agree_and_autotune_keeps_parity(orc, ws):
    """All tile shapes give the same RHS to rounding; the autotuned context still
    matches the oracle and remembers its choice (stage order included) for the grid; the
    tuning leaves the state's f and primitives as they were."""
    import ctypes
    nx, ny, nz = 90, 64, 32
    prims = random_state(nx, ny, nz, 31, "shock")
    ref = oracle_du(orc, prims)
    outs = []
    for tile in W.TILE_SHAPES:
        ctx = make_ctx(nx, ny, nz)
        ctx.set_tile_shape(tile)
        outs.append(gpu_du_comput(ctx, upload_state(prims)))
    for o in outs[1:]:
        assert rel_l2(o, outs[0]) <= 1e-13
    ctx = make_ctx(nx, ny, nz)
    st = upload_state(prims)
    gpu_du_comput(ctx, st)
    W.check(W.lib().ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                                   W.job((0, 0, 0), (nx, ny, nz)), W.stream_handle()))
    kept = ("f", "d", "u", "v", "w", "T", "cc")
    before = {k: st.t[k].clone() for k in kept}
    ns = len(W.SCHEDULES)
    times = (ctypes.c_float * (len(W.TILE_SHAPES) * ns))()
    W.check(W.lib().ws_autotune(ctx.handle, ctypes.byref(st.ws), 3, times, W.stream_handle()))
    for k in kept:   # the stage-order timing writes scratch only
        assert st.t[k].equal(before[k]), k
    log(f"  autotune ms {list(times)}")
    chosen, sched, order, sorder = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    W.check(W.lib().ws_ctx_get_tile_shape(ctx.handle, ctypes.byref(chosen)))
    W.check(W.lib().ws_ctx_get_schedule(ctx.handle, ctypes.byref(sched)))
    W.check(W.lib().ws_ctx_get_direction_order(ctx.handle, ctypes.byref(order)))
    W.check(W.lib().ws_ctx_get_stage_order(ctx.handle, ctypes.byref(sorder)))
    sshape = ctypes.c_int32()
    W.check(W.lib().ws_ctx_get_stage_shape(ctx.handle, ctypes.byref(sshape)))
    log(f"  chose shape {chosen.value} schedule {sched.value} order {order.value} stage order {sorder.value}"
        f" stage shape {sshape.value}")
    assert times[chosen.value * ns + sched.value] == min(times)
    assert order.value in W.ORDERS and sorder.value in W.ORDERS
    assert sched.value != W.WS_SCHED_GRID or sshape.value in W.TILE_SHAPES
    ctx2 = make_ctx(nx, ny, nz)
    again, again_s, again_o, again_so = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    W.check(W.lib().ws_ctx_get_tile_shape(ctx2.handle, ctypes.byref(again)))
    W.check(W.lib().ws_ctx_get_schedule(ctx2.handle, ctypes.byref(again_s)))
    W.check(W.lib().ws_ctx_get_direction_order(ctx2.handle, ctypes.byref(again_o)))
    W.check(W.lib().ws_ctx_get_stage_order(ctx2.handle, ctypes.byref(again_so)))
    again_ss = ctypes.c_int32()
    W.check(W.lib().ws_ctx_get_stage_shape(ctx2.handle, ctypes.byref(again_ss)))
    assert (again.value, again_s.value, again_o.value, again_so.value, again_ss.value) == \
        (chosen.value, sched.value, order.value, sorder.value, sshape.value)
    assert rel_l2(gpu_du_comput(ctx2, upload_state(prims)), ref) <= TOL_FP64


@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("tile", [W.WS_TILE_30x16_K2, W.WS_TILE_62x8_K2])
def test_direction_orders_agree(orc, ws, tile, precision):
    """Launching z first (it writes du, y and x accumulate) gives the x-first RHS to
    rounding: FP64 within the drop-in bar of the oracle in both orders and 1e-13 of each
    other, mixed within 1e-6 of the oracle; with accumulate = 1 both orders add the RHS
    to what du held."""
    import torch
    nx, ny, nz = 67, 45, 37
    prims = random_state(nx, ny, nz, 41, "shock")
    ref = oracle_du(orc, prims)
    tol = TOL_FP64 if precision == W.WS_PREC_FP64 else 1e-6
    outs, accs = [], []
    for order in W.ORDERS:
        ctx = make_ctx(nx, ny, nz, precision=precision)
        ctx.set_tile_shape(tile)
        ctx.set_direction_order(order)
        st = upload_state(prims)
        du = gpu_du_comput(ctx, st)
        err = rel_l2(du, ref)
        log(f"  tile {tile} prec {precision} order {order}: rel L2 {err:.3e}")
        assert err <= tol
        outs.append(du)
        # accumulate on top of a known du
        st.interior("du").fill_(0.5)
        torch.cuda.synchronize()
        acc = gpu_du_comput(ctx, st, accumulate=True)
        accs.append(acc)
        assert rel_l2(acc - 0.5, ref) <= tol
    if precision == W.WS_PREC_FP64:
        assert rel_l2(outs[1], outs[0]) <= 1e-13
        assert rel_l2(accs[1], accs[0]) <= 1e-13


@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("tile", W.TILE_SHAPES)
def test_schedules_agree(ws, tile, precision):
    """The grid, persistent and prefetching schedules run the same per-tile code (the
    compiler may contract multiply-adds differently per kernel): their RHS agree to
    1e-13 relative L2 in FP64 (1e-11 mixed), on the full box and on a sub-box job."""
    tol = 1e-13 if precision == W.WS_PREC_FP64 else 1e-11
    for shape in [(31, 17, 19), (67, 45, 37), (130, 40, 33)]:
        nx, ny, nz = shape
        prims = random_state(nx, ny, nz, 41, "shock")
        outs, subs = [], []
        for sched in W.SCHEDULES:
            ctx = make_ctx(nx, ny, nz, precision=precision)
            ctx.set_tile_shape(tile)
            ctx.set_schedule(sched)
            outs.append(gpu_du_comput(ctx, upload_state(prims)))
            st = upload_state(prims)
            st.t["du"].fill_(3.0)
            subs.append(gpu_du_comput(ctx, st, job=W.job((4 + 3, 4 + 2, 4 + 1), (4 + nx - 2, 4 + ny - 3, 4 + nz - 1))))
        log(f"schedules {shape} tile {tile} precision {precision}")
        for o in outs[1:]:
            assert rel_l2(o, outs[0]) <= tol
        inside = (slice(None), slice(1, nz - 1), slice(2, ny - 3), slice(3, nx - 2))
        mask = np.ones(subs[0].shape, bool)
        mask[inside] = False
        for o in subs:
            assert np.all(o[mask] == 3.0)
        for o in subs[1:]:
            assert rel_l2(o[inside], subs[0][inside]) <= tol


def test_tiled_is_independent_of_pitch(ws):
    nx, ny, nz = 31, 17, 19
    prims = random_state(nx, ny, nz, 5, "rough")
    ctx = make_ctx(nx, ny, nz)
    a = gpu_du_comput(ctx, upload_state(prims, pad=0))
    b = gpu_du_comput(ctx, upload_state(prims, pad=7))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("tile", W.TILE_SHAPES)
def test_inputs_of_different_pitches_take_the_64_bit_addresses(ws, tile, precision):
    """The load rounds address the five input fields with one 32-bit offset when they share
    a pitch; a sound-speed field of another pitch makes the launch address each field on its
    own in 64 bits.  Both address the same cells, so the RHS agrees bit for bit, for every
    tile shape and both precisions."""
    nx, ny, nz = 45, 26, 19
    prims = random_state(nx, ny, nz, 17, "shock")
    ctx = make_ctx(nx, ny, nz, precision=precision)
    ctx.set_tile_shape(tile)
    a = gpu_du_comput(ctx, upload_state(prims))
    st = upload_state(prims)
    st.t["cc"] = W.ghosted_tensor(nx, ny, nz, pad=5)
    st.refresh()
    assert st.ws.cc.pitch != st.ws.d.pitch
    b = gpu_du_comput(ctx, st)
    assert np.array_equal(a, b)


def test_metric_fields(orc, ws):
    """Per-cell metric arrays (all terms 1, Ajac 0.5: the reference's Init_stat = 0
    metrics) through both kernel families."""
    import torch
    nx, ny, nz = 30, 16, 16
    prims = random_state(nx, ny, nz, 8, "rough")
    met_np = [np.ones((nz + 8, ny + 8, nx + 8))] * 9
    aj_np = np.full((nz + 8, ny + 8, nx + 8), 0.5)
    ref = oracle_du(orc, prims, metrics=met_np, ajac=aj_np)
    tens = [torch.ones((nz + 8, ny + 8, nx + 8), dtype=torch.float64, device="cuda") for _ in range(9)]
    taj = torch.full((nz + 8, ny + 8, nx + 8), 0.5, dtype=torch.float64, device="cuda")
    mf = W.WsMetricFields()
    for i in range(9):
        mf.A[i] = W.field_of(tens[i])
    mf.ajac = W.field_of(taj)
    for kern, exact in ((W.WS_KERNEL_ANCHOR, True), (W.WS_KERNEL_TILED, False)):
        log(f"metric fields kernel {kern}")
        ctx = make_ctx(nx, ny, nz, kernel=kern)
        du = gpu_du_comput(ctx, upload_state(prims), metrics=mf)
        if exact:
            assert np.array_equal(du, ref)
        else:
            assert rel_l2(du, ref) <= TOL_FP64


@pytest.mark.parametrize("order", W.ORDERS)
def test_sub_box_job_and_accumulate(ws, order):
    """A job box inside the interior updates exactly its cells, with the values of a
    full-domain run; accumulate = 1 adds to what du held (both direction orders)."""
    import torch
    nx, ny, nz = 40, 24, 20
    prims = random_state(nx, ny, nz, 3, "shock")
    ctx = make_ctx(nx, ny, nz)
    ctx.set_direction_order(order)
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


def test_bad_job_is_rejected_before_launch(ws):
    nx, ny, nz = 16, 16, 16
    ctx = make_ctx(nx, ny, nz)
    st = upload_state(random_state(nx, ny, nz, 1, "blobs"))
    L = W.lib()
    bad = W.job((2, 4, 4), (20, 20, 20))
    r = L.ws_inviscid_rhs(ctx.handle, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.cc, None, st.ws.du, bad, 0,
                          W.stream_handle())
    assert r == -1 and b"job box" in L.ws_last_error()


def test_drop_in_vortex_600x128x128_against_oracle(orc, ws):
    """The FP64 drop-in workload at its full size: isentropic-vortex lattice
    (6 vortices, strength 5, Mach 0.5), one RHS, 1e-12 relative L2."""
    from windstencil_amd import initial as I
    nx, ny, nz = 600, 128, 128
    log("vortex 600x128x128: building state")
    prims = I.vortex_lattice(nx, ny, nz, 2024)
    ctx = make_ctx(nx, ny, nz)
    du = gpu_du_comput(ctx, upload_state(prims))
    log("vortex 600x128x128: oracle")
    ref = oracle_du(orc, prims)
    err = rel_l2(du, ref)
    log(f"  rel L2 {err:.3e}")
    assert err <= TOL_FP64


def test_full_size_600x256x128_properties(ws):
    """At the benchmark size: the tiled kernel agrees with the anchor kernel, a
    uniform state gives exactly zero, totals are conserved, and a periodic shift
    of the input shifts the RHS bit for bit."""
    from windstencil_amd import initial as I
    nx, ny, nz = 600, 256, 128
    log("600x256x128: sod + noise")
    prims = I.sod_noise(nx, ny, nz, 7)
    ctx = make_ctx(nx, ny, nz)
    st = upload_state(prims)
    du = gpu_du_comput(ctx, st)
    actx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR)
    da = gpu_du_comput(actx, st)
    err = rel_l2(du, da)
    log(f"  tiled vs anchor rel L2 {err:.3e}")
    assert err <= TOL_FP64
    for m in range(5):
        assert abs(du[m].sum()) <= 1e-12 * np.abs(du[m]).sum()
    rolled = [np.roll(a, (3, 5, 7), axis=(0, 1, 2)) for a in prims]
    dr = gpu_du_comput(ctx, upload_state(rolled))
    assert np.array_equal(dr, np.roll(du, (3, 5, 7), axis=(1, 2, 3)))
    uni = [np.full((nz, ny, nx), v) for v in (1.1, 0.4, -0.3, 0.2, 1.3)]
    assert np.all(gpu_du_comput(ctx, upload_state(uni)) == 0.0)


def _wavy_metrics(nx, ny, nz):
    """Smooth, non-constant metric terms and Jacobian on the interior (a curvilinear
    mesh's worth of variation), x fastest: [Akx..Asz, Ajac]."""
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    s = lambda a, b, c: np.sin(2 * np.pi * (a * x + b * y + c * z))
    A = [1.0 + 0.2 * s(1, 1, 0), 0.1 * s(0, 1, 1), 0.05 * s(1, 0, 1),
         0.1 * s(1, 1, 1), 1.0 + 0.15 * s(0, 2, 1), -0.08 * s(2, 0, 1),
         0.06 * s(1, 2, 0), -0.1 * s(0, 1, 2), 1.0 + 0.25 * s(1, 0, 1)]
    ajac = 1.0 + 0.3 * s(1, 1, 1)
    return [np.ascontiguousarray(a) for a in A + [ajac]]


def test_curvilinear_metrics_from_jacobi_file(orc, ws, tmp_path):
    """Per-cell metric terms and Jacobian that vary in space, loaded from an
    OCFD3d-Jacobi.dat file (ws_load_jacobi fills their periodic ghosts): anchor kernels
    bitwise, tiled FP64 within 1e-12 (every tile shape), mixed within 1e-6 of the oracle
    given the same metric arrays."""
    import torch
    from tests.helpers import ghosted
    nx, ny, nz = 40, 24, 20
    prims = random_state(nx, ny, nz, 19, "blobs")
    met = _wavy_metrics(nx, ny, nz)
    coords = [np.zeros((nz, ny, nx)) for _ in range(3)]
    path = str(tmp_path / "OCFD3d-Jacobi.dat").encode()
    W.check(W.lib().ws_jacobi_write_host(path, nx, ny, nz,
                                         (ctypes.c_void_p * 13)(*[a.ctypes.data for a in coords + met])))
    p = orc.params(nx, ny, nz)
    g = [ghosted(a) for a in met]
    for a in g:
        orc.ghost_fill_periodic(p, a)
    ref = oracle_du(orc, prims, metrics=g[:9], ajac=g[9])
    tens = [torch.zeros((nz + 8, ny + 8, nx + 8), dtype=torch.float64, device="cuda") for _ in range(10)]
    mf = W.WsMetricFields()
    for i in range(9):
        mf.A[i] = W.field_of(tens[i])
    mf.ajac = W.field_of(tens[9])
    ctx = make_ctx(nx, ny, nz)
    W.check(W.lib().ws_load_jacobi(ctx.handle, path, ctypes.byref(mf), None, W.stream_handle()))
    torch.cuda.synchronize()
    for t, a in zip(tens, g):
        full = t.cpu().numpy()
        # faces agree with the oracle's periodic fill (edges and corners are never read)
        assert np.array_equal(full[4:-4, 4:-4, :], a[4:-4, 4:-4, :])
        assert np.array_equal(full[4:-4, :, 4:-4], a[4:-4, :, 4:-4])
        assert np.array_equal(full[:, 4:-4, 4:-4], a[:, 4:-4, 4:-4])
    actx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR)
    assert np.array_equal(gpu_du_comput(actx, upload_state(prims), metrics=mf), ref)
    for tile in W.TILE_SHAPES:
        tctx = make_ctx(nx, ny, nz)
        tctx.set_tile_shape(tile)
        err = rel_l2(gpu_du_comput(tctx, upload_state(prims), metrics=mf), ref)
        log(f"  curvilinear tile {tile}: rel L2 {err:.3e}")
        assert err <= TOL_FP64
    mctx = make_ctx(nx, ny, nz, precision=W.WS_PREC_MIXED)
    err = rel_l2(gpu_du_comput(mctx, upload_state(prims), metrics=mf), ref)
    log(f"  curvilinear mixed: rel L2 {err:.3e}")
    assert err <= 1e-6


def test_anchor_contexts_on_two_streams_run_concurrently(orc, ws):
    """Each context owns its anchor scratch: two contexts of different sizes running the
    anchor RHS on two streams at once (no synchronisation between their launches) both
    stay bit-identical to the oracle."""
    import torch
    L = W.lib()
    runs = []
    for (nx, ny, nz), seed, kind in (((40, 24, 20), 3, "shock"), ((33, 28, 17), 4, "rough")):
        prims = random_state(nx, ny, nz, seed, kind)
        runs.append((make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_ANCHOR), upload_state(prims), oracle_du(orc, prims),
                     torch.cuda.Stream()))
    torch.cuda.synchronize()
    log("anchor RHS of two contexts interleaved on two streams")
    for _ in range(3):
        for ctx, st, _, strm in runs:
            W.check(L.ws_du_comput(ctx.handle, ctypes.byref(st.ws), None, W.stream_handle(strm)), "du_comput")
    torch.cuda.synchronize()
    for _, st, ref, _ in runs:
        du = st.interior("du").cpu().numpy()
        assert np.array_equal(du, ref), f"rel L2 {rel_l2(du, ref)}"
