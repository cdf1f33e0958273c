"""GPU parity of the state helpers, the SSP-RK3 driver (eager and graph-replayed)
and the reference-named entry points, against the CPU oracle."""
import ctypes

import numpy as np
import pytest

import windstencil_amd as W
from windstencil_amd import LAP
from tests.gpu_util import log, make_ctx, upload_state
from tests.helpers import ghosted, random_state, rel_l2

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


def _ghosted_np(t):
    return t.cpu().numpy()[:, :, :]


def test_ghost_fill_sound_speed_and_conversions_bitwise(orc, ws):
    import torch
    nx, ny, nz = 29, 18, 13
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz)
    prims = random_state(nx, ny, nz, 4, "rough")
    st = upload_state(prims)
    p = orc.params(nx, ny, nz)
    # primitive -> conserved over the interior
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), s))
    _sync()
    g = [ghosted(a) for a in prims]
    f_ref = orc.pri_to_cons(p, *g)
    assert np.array_equal(st.interior("f").cpu().numpy(), f_ref)
    # conserved -> primitive
    st2 = W.TorchState(nx, ny, nz, pad=3)
    st2.t["f"][:, :, :, :nx].copy_(torch.from_numpy(f_ref))
    W.check(L.ws_cons_to_pri(ctx.handle, st2.ws.f, st2.ws.d, st2.ws.u, st2.ws.v, st2.ws.w, st2.ws.T,
                             st2.ws.P, s))
    _sync()
    refs = orc.cons_to_pri(p, f_ref)
    for k, r in zip(("d", "u", "v", "w", "T", "P"), refs):
        assert np.array_equal(st2.interior(k).cpu().numpy(), r[LAP:-LAP, LAP:-LAP, LAP:-LAP]), k
    # ghost fill (faces only) and sound speed over the ghosted box
    fl = (W.WsField * 1)(st2.ws.T)
    W.check(L.ws_ghost_fill_periodic(ctx.handle, fl, 1, s))
    W.check(L.ws_sound_speed(ctx.handle, st2.ws.T, st2.ws.cc,
                             W.job((0, 0, 0), (nx + 8, ny + 8, nz + 8)), s))
    _sync()
    T_ref = refs[4].copy()
    orc.ghost_fill_periodic(p, T_ref)
    T_gpu = st2.t["T"].cpu().numpy()[:, :, :nx + 8]
    assert np.array_equal(T_gpu, T_ref)
    cc_ref = orc.sound_speed(p, T_ref)
    assert np.array_equal(st2.t["cc"].cpu().numpy()[:, :, :nx + 8], cc_ref)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1), (1, 0, 0)])
def test_prepare_rhs_inputs_equals_ghost_fill_then_sound_speed(ws, periodic):
    """ws_prepare_rhs_inputs equals ws_ghost_fill_periodic of d,u,v,w,T followed by
    ws_sound_speed over the ghosted box, bit for bit on every ghosted array (edge and
    corner ghosts and non-periodic ghost layers keep the caller's values in both)."""
    import torch
    nx, ny, nz = 37, 22, 15
    L = W.lib()
    s = W.stream_handle()
    cfg = dict(periodic=list(periodic))
    if not all(periodic):
        cfg.update(d0_bound=[1 - p for p in periodic for _ in (0, 1)])
    ctx = make_ctx(nx, ny, nz, **cfg)
    prims = random_state(nx, ny, nz, 6, "blobs")
    runs = []
    for fused in (False, True):
        st = upload_state(prims)
        gen = torch.Generator(device="cpu").manual_seed(5)
        for k in ("d", "u", "v", "w", "T"):
            # ghost layers start from distinct positive values (the caller's, where kept)
            a = st.t[k]
            ghost = torch.rand(a.shape, generator=gen, dtype=torch.float64).to(a.device) + 0.5
            interior = st.interior(k).clone()
            a.copy_(ghost)
            st.interior(k).copy_(interior)
        st.t["cc"].fill_(-1.0)
        if fused:
            W.check(L.ws_prepare_rhs_inputs(ctx.handle, ctypes.byref(st.ws), s))
        else:
            fl = (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T)
            W.check(L.ws_ghost_fill_periodic(ctx.handle, fl, 5, s))
            W.check(L.ws_sound_speed(ctx.handle, st.ws.T, st.ws.cc,
                                     W.job((0, 0, 0), (nx + 2 * LAP, ny + 2 * LAP, nz + 2 * LAP)), s))
        _sync()
        # cc compared over the ghosted box (the flat pass also fills the pitch padding)
        runs.append({k: st.t[k].cpu().numpy()[..., :nx + 2 * LAP] for k in ("d", "u", "v", "w", "T", "cc")})
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


def test_time_advance_stages_bitwise(orc, ws):
    import torch
    nx, ny, nz = 20, 12, 9
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz)
    p = orc.params(nx, ny, nz)
    rng = np.random.default_rng(2)
    fn = rng.random((5, nz, ny, nx))
    du = rng.standard_normal((5, nz, ny, nx))
    f = rng.random((5, nz, ny, nx))
    st = W.TorchState(nx, ny, nz, pad=2)
    for k, a in (("fn", fn), ("du", du), ("f", f)):
        st.interior(k).copy_(torch.from_numpy(a))
    fc = f.copy()
    for krk in (1, 2, 3):
        W.check(L.ws_time_advance(ctx.handle, krk, st.ws.f, st.ws.fn, st.ws.du, W.WsSoa(), 0.013, s))
        orc.time_advance(p, krk, fc, fn, du, 0.013)
        _sync()
        assert np.array_equal(st.interior("f").cpu().numpy(), fc), krk


def _gpu_run(prims, nsteps, kernel, graph, cfl=0.5, precision=W.WS_PREC_FP64, order=None, fusion=None,
             stage_order=None, tile=None, stage_shape=None):
    nz, ny, nx = prims[0].shape
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz, kernel=kernel, precision=precision)
    if order is not None:
        ctx.set_direction_order(order)
        ctx.set_stage_order(order if stage_order is None else stage_order)
    elif stage_order is not None:
        ctx.set_stage_order(stage_order)
    if tile is not None:
        ctx.set_tile_shape(tile)
    if stage_shape is not None:
        ctx.set_stage_shape(stage_shape)
    if fusion is not None:
        ctx.set_stage_fusion(fusion)
    st = upload_state(prims)
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), s))
    W.check(L.ws_rk3_run(ctx.handle, ctypes.byref(st.ws), nsteps, cfl, 1 if graph else 0, s), "rk3 run")
    st.sync_from_ws()
    _sync()
    return st.interior("f").cpu().numpy(), st


def _oracle_run(orc, prims, nsteps, cfl=0.5):
    nz, ny, nx = prims[0].shape
    p = orc.params(nx, ny, nz)
    g = [ghosted(a) for a in prims]
    f = orc.pri_to_cons(p, *g)
    orc.rk3_run(p, nsteps, -cfl, f, *g)
    return f


def test_rk3_cpu_workload_anchor_bitwise(orc, ws):
    """The CPU-path workload: 64x32x16, Mach 2 with 8 seeded blobs, 10 SSP-RK3
    steps at CFL 0.5 (dt recomputed on the device every step)."""
    from windstencil_amd import initial as I
    prims = I.freestream_blobs(64, 32, 16, 1)
    log("rk3 anchor 64x32x16 x10")
    f_gpu, _ = _gpu_run(prims, 10, W.WS_KERNEL_ANCHOR, graph=False)
    f_ref = _oracle_run(orc, prims, 10)
    assert np.array_equal(f_gpu, f_ref), rel_l2(f_gpu, f_ref)


@pytest.mark.parametrize("kind", ["blobs", "vortex", "shock"])
def test_rk3_tiled_matches_oracle(orc, ws, kind):
    prims = random_state(48, 32, 16, 11, kind)
    log(f"rk3 tiled {kind} x10")
    f_gpu, _ = _gpu_run(prims, 10, W.WS_KERNEL_TILED, graph=False)
    f_ref = _oracle_run(orc, prims, 10)
    err = rel_l2(f_gpu, f_ref)
    log(f"  rel L2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("fusion", [1, 2, 3])
@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("order", [W.WS_ORDER_XYZ, W.WS_ORDER_ZYX])
@pytest.mark.parametrize("graph", [False, True])
def test_stage_applied_by_the_last_rhs_launch_equals_the_stage_kernel(ws, precision, order, graph, fusion):
    """RK3 steps whose last directional RHS launch applies the stage (fusion 1), also
    derives the primitives of stages 1 and 2 into the context's second primitive set
    (fusion 2), and within the run those of stage 3 as well, alternating the primitive sets
    from step to step with the CFL maxima taken per tile (fusion 3, the default), against
    the separate stage kernel reading du back: the same stage and primitive expressions on
    the same complete RHS, so the states -- f and every ghosted primitive field, ghost
    layers included -- agree bit for bit (boundary planes on, CFL step from the device)."""
    prims = random_state(45, 26, 19, 13, "shock")
    a, sa = _gpu_run(prims, 4, W.WS_KERNEL_TILED, graph=graph, precision=precision, order=order, fusion=fusion)
    b, sb = _gpu_run(prims, 4, W.WS_KERNEL_TILED, graph=graph, precision=precision, order=order, fusion=0)
    log(f"  fusion {fusion} vs separate stage: prec {precision} order {order} graph {graph}: max |diff| "
        f"{np.abs(a - b).max():.3e}, rel L2 {rel_l2(a, b):.3e}")
    assert np.array_equal(a, b)
    for k in ("d", "u", "v", "w", "T", "P", "cc"):
        assert np.array_equal(sa.t[k].cpu().numpy(), sb.t[k].cpu().numpy()), k


@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("stage_order", [W.WS_ORDER_XYZ, W.WS_ORDER_ZYX])
def test_stage_applying_launches_follow_the_stage_order(ws, precision, stage_order):
    """Fused steps launch every RHS with the stage order and the unfused ones with the
    direction order: fusion 3 with the other direction order equals the separate stage
    kernel under direction order = stage order bit for bit, and the two stage orders agree
    to rounding (FP64 within 1e-12, mixed within 1e-6, relative L2 of f)."""
    prims = random_state(46, 26, 19, 13, "shock")
    other = 1 - stage_order
    a, sa = _gpu_run(prims, 3, W.WS_KERNEL_TILED, graph=False, precision=precision, order=other, fusion=3,
                     stage_order=stage_order)
    b, sb = _gpu_run(prims, 3, W.WS_KERNEL_TILED, graph=False, precision=precision, order=stage_order, fusion=0)
    c, _ = _gpu_run(prims, 3, W.WS_KERNEL_TILED, graph=False, precision=precision, order=other, fusion=3)
    err = rel_l2(a, c)
    log(f"  prec {precision} stage order {stage_order}: vs unfused max |diff| {np.abs(a - b).max():.3e}; "
        f"vs the other stage order rel L2 {err:.3e}")
    assert np.array_equal(a, b)
    for k in ("d", "u", "v", "w", "T", "P", "cc"):
        assert np.array_equal(sa.t[k].cpu().numpy(), sb.t[k].cpu().numpy()), k
    assert 0 < err <= (1e-12 if precision == W.WS_PREC_FP64 else 1e-6)


@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
def test_stage_applying_launches_follow_the_stage_shape(ws, precision):
    """The stage-applying RHS launches run in the stage shape when one is set: fusion 3
    with tile shape 1 and stage shape 4 equals the separate stage kernels with tile shape 4
    bit for bit (every ghosted primitive field included), and stage shape -1 follows the
    tile shape."""
    prims = random_state(47, 26, 19, 13, "shock")
    a, sa = _gpu_run(prims, 3, W.WS_KERNEL_TILED, graph=False, precision=precision, order=W.WS_ORDER_ZYX,
                     fusion=3, tile=W.WS_TILE_30x16_K2, stage_shape=W.WS_TILE_126x2_K2)
    b, sb = _gpu_run(prims, 3, W.WS_KERNEL_TILED, graph=False, precision=precision, order=W.WS_ORDER_ZYX,
                     fusion=0, tile=W.WS_TILE_126x2_K2)
    c, _ = _gpu_run(prims, 3, W.WS_KERNEL_TILED, graph=False, precision=precision, order=W.WS_ORDER_ZYX,
                    fusion=3, tile=W.WS_TILE_126x2_K2, stage_shape=-1)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    for k in ("d", "u", "v", "w", "T", "P", "cc"):
        assert np.array_equal(sa.t[k].cpu().numpy(), sb.t[k].cpu().numpy()), k


@pytest.mark.parametrize("nsteps", [1, 2, 3, 5])
@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("shape", [(37, 24, 18), (64, 32, 32)])
def test_alternating_primitive_sets_end_as_the_plain_steps(ws, precision, graph, nsteps, shape):
    """Fusion 3 ends a run whose last step starts from the second primitive set (an even
    number of steps) by copying that set to the state's fields before the last launch
    writes the new primitives; an odd number of steps ends with a plain step.  Either way
    f, the ghosted primitives (ghost layers included) and P equal those of fusion 0, for
    runs of 1, 2, 3 and 5 steps (graph replays in between, CFL step from the device).  At
    64x32x32 the stage kernel of a plain step has more work-groups (256) than the
    stage-applying launch has tiles (160): a plain last step writes no CFL maxima into the
    per-tile buffer."""
    nx, ny, nz = shape
    prims = random_state(nx, ny, nz, 17, "vortex")
    log(f"fusion 3 vs 0: {nsteps} steps, precision {precision}, graph {graph}")
    a, sa = _gpu_run(prims, nsteps, W.WS_KERNEL_TILED, graph=graph, precision=precision, fusion=3)
    b, sb = _gpu_run(prims, nsteps, W.WS_KERNEL_TILED, graph=graph, precision=precision, fusion=0)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b), rel_l2(a, b)
    for k in ("d", "u", "v", "w", "T", "P", "cc"):
        assert np.array_equal(sa.t[k].cpu().numpy(), sb.t[k].cpu().numpy()), k


@pytest.mark.parametrize("fusion", [0, 2, 3])
@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
def test_run_cfl_from_stage_maxima_bitwise(ws, precision, fusion):
    """ws_rk3_run takes each step's dt (after the first) from the CFL maxima its last stage
    kernel left; steps driven one by one with a ws_cfl_dt pass over the state give the
    same state bit for bit (tiled kernels, graph replay)."""
    import torch
    prims = random_state(41, 22, 17, 21, "vortex")
    a, _ = _gpu_run(prims, 4, W.WS_KERNEL_TILED, graph=True, cfl=0.45, precision=precision, fusion=fusion)
    nz, ny, nx = prims[0].shape
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_TILED, precision=precision)
    ctx.set_stage_fusion(fusion)
    st = upload_state(prims)
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), s))
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    for _ in range(4):
        W.check(L.ws_cfl_dt(ctx.handle, ctypes.byref(st.ws), 0.45, ctypes.c_void_p(dt.data_ptr()), s))
        W.check(L.ws_rk3_step(ctx.handle, ctypes.byref(st.ws), ctypes.c_void_p(dt.data_ptr()), s))
    st.sync_from_ws()
    _sync()
    b = st.interior("f").cpu().numpy()
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nsteps", [2, 3, 4, 5])
def test_graph_replay_equals_eager(ws, nsteps):
    """Every replay of the captured step (steps 1..n-1) against the eager path."""
    prims = random_state(32, 16, 16, 9, "shock")
    log(f"graph vs eager {nsteps}")
    a, _ = _gpu_run(prims, nsteps, W.WS_KERNEL_TILED, graph=True)
    b, _ = _gpu_run(prims, nsteps, W.WS_KERNEL_TILED, graph=False)
    assert np.array_equal(a, b)


def test_reference_named_entry_points(ws):
    """The solver-facing names (du_comput, OCFD_time_advance, get_duvwT, OCFD_bc)
    reproduce one SSP-RK3 stage of the C API."""
    import torch
    nx, ny, nz = 30, 16, 12
    L = W.lib()
    s = W.stream_handle()
    prims = random_state(nx, ny, nz, 6, "blobs")
    ctx = make_ctx(nx, ny, nz, dt=2e-4)
    ref = upload_state(prims)
    W.check(L.ws_pri_to_cons(ctx.handle, ref.ws.f, ref.ws.d, ref.ws.u, ref.ws.v, ref.ws.w, ref.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), s))
    other = upload_state(prims)
    other.t["f"].copy_(ref.t["f"])
    for st in (ref, other):
        st.t["fn"].copy_(st.t["f"])
    # C API stage 1
    W.check(L.ws_du_comput(ctx.handle, ctypes.byref(ref.ws), None, s))
    W.check(L.ws_time_advance(ctx.handle, 1, ref.ws.f, ref.ws.fn, ref.ws.du, W.WsSoa(), 2e-4, s))
    W.check(L.ws_cons_to_pri(ctx.handle, ref.ws.f, ref.ws.d, ref.ws.u, ref.ws.v, ref.ws.w, ref.ws.T,
                             ref.ws.P, s))
    W.check(L.ws_bc_planes(ctx.handle, ctypes.byref(ref.ws), s))
    # reference names
    L.ws_ref_set_abort_on_error(0)
    W.check(L.ws_ref_bind(ctx.handle, ctypes.byref(other.ws), None, s))
    L.du_comput(1)
    L.OCFD_time_advance(1)
    L.get_duvwT()
    L.OCFD_bc()
    assert L.ws_ref_last_status() == 0
    _sync()
    for k in ("f", "du", "d", "u", "v", "w", "T"):
        assert torch.equal(ref.t[k], other.t[k]), k
    # an error is recorded, not aborted, after ws_ref_set_abort_on_error(0)
    L.OCFD_time_advance(7)
    assert L.ws_ref_last_status() == -1
    L.ws_ref_set_abort_on_error(1)


def test_restart_save_and_load_device_state(ws, tmp_path):
    """ws_save_restart writes the interior of d, u, v, w, T in the reference's format;
    ws_load_restart puts it back into a state of another pitch, ghosts untouched."""
    import torch
    nx, ny, nz = 23, 14, 9
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz)
    prims = random_state(nx, ny, nz, 77, "rough")
    st = upload_state(prims)
    path = str(tmp_path / "OCFD00000005.dat").encode()
    W.check(L.ws_save_restart(ctx.handle, ctypes.byref(st.ws), 5, 0.25, path, s))
    out = [np.zeros((nz, ny, nx)) for _ in range(5)]
    istep, tt = ctypes.c_int32(), ctypes.c_double()
    W.check(L.ws_restart_read_host(path, nx, ny, nz, ctypes.byref(istep), ctypes.byref(tt),
                                   (ctypes.c_void_p * 5)(*[a.ctypes.data for a in out])))
    assert istep.value == 5 and tt.value == 0.25
    for a, p in zip(out, prims):
        assert np.array_equal(a, p)
    st2 = W.TorchState(nx, ny, nz, pad=5)
    W.check(L.ws_load_restart(ctx.handle, ctypes.byref(st2.ws), path, ctypes.byref(istep), ctypes.byref(tt), s))
    _sync()
    for k, p in zip(("d", "u", "v", "w", "T"), prims):
        assert np.array_equal(st2.interior(k).cpu().numpy(), p)
        full = st2.t[k].clone()
        full[LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx] = 0
        assert float(full.abs().max()) == 0.0


def test_mem_value_init_on_pitched_boxes(ws):
    """cuda_mem_value_init_warp / ws_mem_value_init (commen_kernel.cpp:12-29) as the
    reference calls it (OCFD_init.cpp:39-48): a five-variable interior state
    (nx, ny, nz*5) and a ghosted field, both with pitches wider than the rows.  Exactly
    the box's elements take the value; the row padding keeps what it held."""
    import torch
    L = W.lib()
    nx, ny, nz, pad = 37, 11, 9, 27
    soa = torch.full((5, nz, ny, nx + pad), 7.0, dtype=torch.float64, device="cuda")
    L.ws_ref_set_abort_on_error(0)
    L.cuda_mem_value_init_warp(0.0, ctypes.c_void_p(soa.data_ptr()), nx + pad, nx, ny, nz * 5)
    assert L.ws_ref_last_status() == 0
    g = torch.full((nz + 8, ny + 8, nx + 8 + 5), 7.0, dtype=torch.float64, device="cuda")
    W.check(L.ws_mem_value_init(-2.5, ctypes.c_void_p(g.data_ptr()), nx + 13, nx + 8, ny + 8, nz + 8,
                                W.stream_handle()))
    # an empty box writes nothing; a pitch below the row is refused
    W.check(L.ws_mem_value_init(1.0, ctypes.c_void_p(g.data_ptr()), nx + 13, 0, ny, nz, W.stream_handle()))
    assert L.ws_mem_value_init(1.0, ctypes.c_void_p(g.data_ptr()), 4, 8, 2, 2, W.stream_handle()) == -1
    L.ws_ref_set_abort_on_error(1)
    _sync()
    assert bool((soa[..., :nx] == 0.0).all()) and bool((soa[..., nx:] == 7.0).all())
    assert bool((g[..., :nx + 8] == -2.5).all()) and bool((g[..., nx + 8:] == 7.0).all())


@pytest.mark.parametrize("kernel", [W.WS_KERNEL_TILED, W.WS_KERNEL_ANCHOR])
def test_du_invis_jacobian3d_x_by_name_on_a_sub_box(orc, ws, kernel):
    """The reference-named du_invis_Jacobian3d_init / du_invis_Jacobian3d_x
    (OCFD_NS_Jacobian3d.h:11-15) on a sub-box job: du over that box equals the oracle's
    RHS there (anchor kernels bit for bit, tiled within 1e-12 relative L2) and every cell
    outside the box keeps what du held."""
    import torch
    from tests.gpu_util import oracle_du
    nx, ny, nz = 41, 26, 19
    prims = random_state(nx, ny, nz, 12, "shock")
    ref = oracle_du(orc, prims)
    ctx = make_ctx(nx, ny, nz, kernel=kernel)
    st = upload_state(prims)
    st.t["du"].fill_(-9.0)
    L = W.lib()
    s = W.stream_handle()
    W.check(L.ws_ghost_fill_periodic(ctx.handle, (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T), 5, s))
    L.ws_ref_set_abort_on_error(0)
    W.check(L.ws_ref_bind(ctx.handle, ctypes.byref(st.ws), None, s))
    lo, hi = (6, 3, 2), (35, 20, 17)                     # interior cells [lo, hi)
    job = W.job(tuple(LAP + v for v in lo), tuple(LAP + v for v in hi))
    sp = ctypes.c_void_p(s.value)
    L.du_invis_Jacobian3d_init(job, ctypes.byref(sp))
    L.du_invis_Jacobian3d_x(job, None, None, ctypes.byref(sp))
    assert L.ws_ref_last_status() == 0
    L.ws_ref_set_abort_on_error(1)
    _sync()
    du = st.interior("du").cpu().numpy()
    box = (slice(None), slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
    if kernel == W.WS_KERNEL_ANCHOR:
        assert np.array_equal(du[box], ref[box])
    else:
        err = rel_l2(du[box], ref[box])
        log(f"  sub-box rel L2 {err:.3e}")
        assert err <= 1e-12
    mask = np.ones(du.shape, bool)
    mask[box] = False
    assert np.all(du[mask] == -9.0)


def test_time_advance_writes_the_bound_f_lap(ws):
    """OCFD_time_advance(3) with an f_lap bound (ws_ref_bind_f_lap) writes the new f into
    its interior as the reference's stage 3 does (OCFD_time.cpp:48-58); a pitch below the
    ghosted row is refused."""
    import torch
    nx, ny, nz = 22, 13, 10
    L = W.lib()
    s = W.stream_handle()
    ctx = make_ctx(nx, ny, nz, dt=1e-3)
    st = upload_state(random_state(nx, ny, nz, 8, "blobs"))
    for k, v in (("f", 1.0), ("fn", 2.0), ("du", 3.0)):
        st.t[k].copy_(torch.rand_like(st.t[k]) + v)
    lap = torch.full((5, nz + 8, ny + 8, nx + 8 + 3), 5.0, dtype=torch.float64, device="cuda")
    soa = W.WsSoa(lap.data_ptr(), nx + 11, ny + 8, nz + 8, 0)
    W.check(L.ws_ref_bind(ctx.handle, ctypes.byref(st.ws), None, s))
    W.check(L.ws_ref_bind_f_lap(ctypes.byref(soa)))
    L.ws_ref_set_abort_on_error(0)
    L.OCFD_time_advance(3)
    assert L.ws_ref_last_status() == 0
    L.ws_ref_set_abort_on_error(1)
    _sync()
    inner = lap[:, LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx]
    assert torch.equal(inner, st.interior("f"))
    mask = torch.ones_like(lap, dtype=torch.bool)
    mask[:, LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx] = False
    assert bool((lap[mask] == 5.0).all())
    bad = W.WsSoa(lap.data_ptr(), nx + 2, ny + 8, nz + 8, 0)
    assert L.ws_ref_bind_f_lap(ctypes.byref(bad)) == -1
    W.check(L.ws_ref_bind_f_lap(None))
