"""GPU tests of batches of uneven blocks (one persistent RHS launch per direction).

The batched kernels run the same per-tile code as the single-block kernels; the compiler
may contract multiply-adds differently in different kernels, so the batch reproduces the
per-block results to rounding (1e-13 relative L2 for an RHS, 1e-12 for the RK3 state at a
fixed dt) rather than bit for bit.  Exact: the CFL step (the minimum of the blocks') and
graph replay against eager steps.  Against the oracle the batch meets the single-block
bars (1e-12 relative L2 for FP64, bitwise for the anchor kernels).
"""
import ctypes

import numpy as np
import pytest

import windstencil_amd as W
from windstencil_amd.initial import SplitMix64
from tests.gpu_util import gpu_du_comput, log, make_ctx, oracle_du, upload_state
from tests.helpers import random_state, rel_l2

pytestmark = pytest.mark.gpu

KINDS = ["blobs", "vortex", "sod", "shock", "rough"]


def _extents(seed, n, lo, hi):
    rng = SplitMix64(seed)
    return [(rng.randint(lo, hi), rng.randint(lo, hi), rng.randint(lo, hi)) for _ in range(n)]


def _batch(extents, seed, precision=W.WS_PREC_FP64, kernel=W.WS_KERNEL_TILED, kinds=KINDS, **kw):
    prims = [random_state(*e, seed + i, kinds[i % len(kinds)]) for i, e in enumerate(extents)]
    states = [upload_state(p) for p in prims]
    base = W.default_config(8, 8, 8, precision=precision, kernel=kernel, **kw)
    return W.Batch(base, extents, states), prims


def _sync():
    import torch
    torch.cuda.synchronize()


def _du(st):
    return st.interior("du").cpu().numpy()


@pytest.mark.parametrize("schedule", W.SCHEDULES)
@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("tile", W.TILE_SHAPES)
def test_batch_rhs_equals_single_block_runs(ws, tile, precision, schedule):
    extents = _extents(7, 9, 5, 47)
    log(f"batch rhs tile {tile} precision {precision} schedule {schedule}: {extents}")
    b, prims = _batch(extents, 100, precision=precision)
    W.check(W.lib().ws_batch_set_tile_shape(b.handle, tile))
    W.check(W.lib().ws_batch_set_schedule(b.handle, schedule))
    W.check(W.lib().ws_batch_du_comput(b.handle, W.stream_handle()))
    _sync()
    for e, p, st in zip(extents, prims, b.states):
        ctx = make_ctx(*e, precision=precision)
        ctx.set_tile_shape(tile)
        ref = gpu_du_comput(ctx, upload_state(p))
        assert rel_l2(_du(st), ref) <= 1e-13, e


def test_batch_rhs_against_oracle(orc, ws):
    extents = [(37, 21, 13), (16, 40, 9), (24, 16, 30)]
    b, prims = _batch(extents, 200)
    W.check(W.lib().ws_batch_du_comput(b.handle, W.stream_handle()))
    _sync()
    for e, p, st in zip(extents, prims, b.states):
        err = rel_l2(_du(st), oracle_du(orc, p))
        log(f"  {e}: rel L2 {err:.3e}")
        assert err <= 1e-12


def test_batch_anchor_kernels_bitwise_oracle(orc, ws):
    extents = [(20, 12, 9), (11, 17, 14)]
    b, prims = _batch(extents, 300, kernel=W.WS_KERNEL_ANCHOR)
    W.check(W.lib().ws_batch_du_comput(b.handle, W.stream_handle()))
    _sync()
    for p, st in zip(prims, b.states):
        assert np.array_equal(_du(st), oracle_du(orc, p))


def test_batch_cfl_is_min_over_blocks(ws):
    import torch
    extents = _extents(11, 6, 6, 30)
    b, prims = _batch(extents, 400)
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    s = W.stream_handle()
    W.check(W.lib().ws_batch_cfl_dt(b.handle, 0.5, ctypes.c_void_p(dt.data_ptr()), s))
    each = torch.zeros(len(extents), dtype=torch.float64, device="cuda")
    for i, (e, st) in enumerate(zip(extents, b.states)):
        ctx = make_ctx(*e)
        W.check(W.lib().ws_cfl_dt(ctx.handle, ctypes.byref(st.ws), 0.5,
                                  ctypes.c_void_p(each.data_ptr() + 8 * i), s))
    _sync()
    assert float(dt[0]) == float(each.min())


def _to_cons(ctx, st):
    L = W.lib()
    nx, ny, nz = st.n
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), W.stream_handle()))


@pytest.mark.parametrize("schedule,stage_order", [(W.WS_SCHED_GRID, W.WS_ORDER_XYZ),
                                                  (W.WS_SCHED_GRID, W.WS_ORDER_ZYX),
                                                  (W.WS_SCHED_PREFETCH, W.WS_ORDER_XYZ)])
@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
def test_batch_rk3_fixed_dt_equals_single_block_runs(ws, precision, schedule, stage_order):
    """With the grid schedule the batched z launch (stage order XYZ) or x launch (ZYX)
    applies each RK3 stage itself; with the prefetching schedule the stage kernels do: all
    match single-box runs of the same stage order."""
    extents = _extents(13, 5, 6, 33)
    dt = 2e-4
    b, prims = _batch(extents, 500, precision=precision, dt=dt)
    W.check(W.lib().ws_batch_set_schedule(b.handle, schedule))
    W.check(W.lib().ws_batch_set_stage_order(b.handle, stage_order))
    for e, st in zip(extents, b.states):
        _to_cons(make_ctx(*e), st)
    log(f"batch rk3 fixed dt {extents}")
    W.check(W.lib().ws_batch_rk3_run(b.handle, 6, -1.0, 0, W.stream_handle()))
    b.sync_states()
    for e, p, st in zip(extents, prims, b.states):
        ctx = make_ctx(*e, precision=precision, dt=dt)
        ctx.set_stage_order(stage_order)
        ref = upload_state(p)
        _to_cons(ctx, ref)
        W.check(W.lib().ws_rk3_run(ctx.handle, ctypes.byref(ref.ws), 6, -1.0, 0, W.stream_handle()))
        ref.sync_from_ws()
        _sync()
        assert rel_l2(st.interior("f").cpu().numpy(), ref.interior("f").cpu().numpy()) <= 1e-12, e


@pytest.mark.parametrize("cfl", [-1.0, 0.5])
@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
def test_batch_stage_fusion_levels(ws, precision, graph, cfl):
    """Batched RK3 steps (boundary planes on; a fixed dt, or dt from the CFL step at CFL
    0.5) with the z launch applying each stage (level 1) and also deriving the primitives
    of stages 1 and 2 into every block's second primitive set (level 2, the default).
    Levels 2 and 1 run the same kernels, so f and every ghosted primitive field agree bit
    for bit.  Against the separate stage kernels (level 0) the stage-applying z launch is
    another instantiation of the tile code, which the compiler may contract differently
    (fused multiply-adds in the WENO and split arithmetic): the states agree to rounding,
    1e-14 relative L2 after 3 steps in FP64.  The mixed form rounds the window's FP64
    differences to FP32, so a last-bit difference of the two kernels' split values can flip
    one of those roundings and move that face's FP32 correction by an FP32 unit of the
    window's variation: its states agree to 1e-9.  Every state stays finite.  The CFL runs
    leave out the
    white-noise ("rough") block: at CFL 0.5 its temperature turns negative within one step
    in the oracle too (tests/test_oracle_rhs.py), so it is advanced at the small fixed dt
    only."""
    extents = _extents(19, 5, 6, 30)
    kinds = KINDS if cfl < 0 else [k for k in KINDS if k != "rough"]
    runs = []
    for level in (0, 1, 2):
        b, _ = _batch(extents, 700, precision=precision, kinds=kinds, bc_planes=1, dt=2e-4)
        W.check(W.lib().ws_batch_set_schedule(b.handle, W.WS_SCHED_GRID))
        W.check(W.lib().ws_batch_set_stage_fusion(b.handle, level))
        for e, st in zip(extents, b.states):
            _to_cons(make_ctx(*e), st)
        log(f"batch fusion {level} prec {precision} graph {graph} cfl {cfl}")
        W.check(W.lib().ws_batch_rk3_run(b.handle, 3, cfl, graph, W.stream_handle()))
        b.sync_states()
        _sync()
        runs.append([{k: st.t[k].cpu().numpy() for k in ("f", "d", "u", "v", "w", "T", "cc")}
                     for st in b.states])
        b.close()
    for i, (a, c) in enumerate(zip(runs[1], runs[2])):
        for k in a:
            assert np.array_equal(a[k], c[k]), (i, k)
    for i, (a, c) in enumerate(zip(runs[0], runs[1])):
        assert np.isfinite(a["f"]).all() and np.isfinite(c["f"]).all(), i
        err = rel_l2(c["f"], a["f"])
        log(f"  block {i}: level 1 vs 0 rel L2 {err:.2e}")
        assert err <= (1e-14 if precision == W.WS_PREC_FP64 else 1e-9), (i, err)


@pytest.mark.parametrize("stage_order", [W.WS_ORDER_XYZ, W.WS_ORDER_ZYX])
@pytest.mark.parametrize("nsteps", [2, 3, 4])
@pytest.mark.parametrize("cfl", [-1.0, 0.5])
@pytest.mark.parametrize("graph", [0, 1])
def test_batch_alternating_primitive_sets_equal_level_2(ws, graph, cfl, nsteps, stage_order):
    """Stage fusion 3 in a batch: the stage-applying launch of stage 3 (z for stage order
    XYZ, x for ZYX) also writes every block's primitives and sound speed, with the CFL
    maxima per tile (reduced per block over the block's range of that launch's tile list),
    and the primitive sets alternate from step to step; a run's last step ends in the
    blocks' fields (from the second sets: after copying them there).  The same kernels as
    level 2, so f, every ghosted primitive field and P agree bit for bit with level 2 after
    2, 3 and 4 steps (both last-step forms), eager and graph, with a fixed dt and with the
    CFL step."""
    extents = _extents(29, 5, 6, 30)
    kinds = [k for k in KINDS if k != "rough"]
    runs = []
    for level in (2, 3):
        b, _ = _batch(extents, 710, kinds=kinds, bc_planes=1, dt=2e-4)
        W.check(W.lib().ws_batch_set_schedule(b.handle, W.WS_SCHED_GRID))
        W.check(W.lib().ws_batch_set_stage_order(b.handle, stage_order))
        W.check(W.lib().ws_batch_set_stage_fusion(b.handle, level))
        for e, st in zip(extents, b.states):
            _to_cons(make_ctx(*e), st)
        log(f"batch fusion {level} vs 2: {nsteps} steps, graph {graph}, cfl {cfl}")
        W.check(W.lib().ws_batch_rk3_run(b.handle, nsteps, cfl, graph, W.stream_handle()))
        b.sync_states()
        _sync()
        runs.append([{k: st.t[k].cpu().numpy() for k in ("f", "d", "u", "v", "w", "T", "cc", "P")}
                     for st in b.states])
        b.close()
    for i, (a, c) in enumerate(zip(*runs)):
        assert np.isfinite(c["f"]).all(), i
        for k in a:
            assert np.array_equal(a[k], c[k]), (i, k)


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_stage_shape_runs_the_stage_launches(ws, graph):
    """A batch's stage-applying RHS runs in its stage shape: tile shape 1 with stage shape 4
    gives, at stage fusion 3 (every RHS of a step applies a stage), the states of tile
    shape 4 bit for bit (CFL step from the per-tile maxima of shape 4's x list), and the
    plain RHS still runs in tile shape 1."""
    extents = _extents(37, 4, 6, 30)
    kinds = [k for k in KINDS if k != "rough"]
    runs, rhs = [], []
    for tile, sshape in ((W.WS_TILE_30x16_K2, W.WS_TILE_126x2_K2), (W.WS_TILE_126x2_K2, -1),
                         (W.WS_TILE_30x16_K2, -1)):
        b, _ = _batch(extents, 720, kinds=kinds, bc_planes=1, dt=2e-4)
        L = W.lib()
        W.check(L.ws_batch_set_schedule(b.handle, W.WS_SCHED_GRID))
        W.check(L.ws_batch_set_tile_shape(b.handle, tile))
        W.check(L.ws_batch_set_stage_shape(b.handle, sshape))
        W.check(L.ws_batch_set_stage_order(b.handle, W.WS_ORDER_ZYX))
        for e, st in zip(extents, b.states):
            _to_cons(make_ctx(*e), st)
        s = W.stream_handle()
        W.check(L.ws_batch_du_comput(b.handle, s))
        _sync()
        rhs.append([_du(st) for st in b.states])
        W.check(L.ws_batch_rk3_run(b.handle, 3, 0.5, graph, s))
        b.sync_states()
        _sync()
        runs.append([{k: st.t[k].cpu().numpy() for k in ("f", "d", "u", "v", "w", "T", "cc", "P")}
                     for st in b.states])
        b.close()
    for i in range(len(extents)):
        for k in runs[0][i]:
            assert np.array_equal(runs[0][i][k], runs[1][i][k]), (i, k)
        assert np.array_equal(rhs[0][i], rhs[2][i]), i   # the plain RHS: tile shape 1


@pytest.mark.parametrize("fusion", [0, 2, 3])
@pytest.mark.parametrize("graph", [0, 1])
def test_batch_run_cfl_from_stage_maxima_bitwise(ws, graph, fusion):
    """ws_batch_rk3_run takes each step's dt (after the first) from the CFL maxima the
    blocks' last stage kernels left; steps driven one by one with ws_batch_cfl_dt passes
    over the state give the same states bit for bit."""
    import torch
    extents = _extents(23, 4, 6, 30)
    s = W.stream_handle()
    runs = []
    for manual in (False, True):
        b, _ = _batch(extents, 800, bc_planes=1)
        W.check(W.lib().ws_batch_set_schedule(b.handle, W.WS_SCHED_GRID))
        W.check(W.lib().ws_batch_set_stage_fusion(b.handle, fusion))
        for e, st in zip(extents, b.states):
            _to_cons(make_ctx(*e), st)
        if manual:
            dt = torch.zeros(1, dtype=torch.float64, device="cuda")
            for _ in range(4):
                W.check(W.lib().ws_batch_cfl_dt(b.handle, 0.4, ctypes.c_void_p(dt.data_ptr()), s))
                W.check(W.lib().ws_batch_rk3_step(b.handle, ctypes.c_void_p(dt.data_ptr()), s))
        else:
            W.check(W.lib().ws_batch_rk3_run(b.handle, 4, 0.4, graph, s))
        b.sync_states()
        _sync()
        runs.append([st.interior("f").cpu().numpy() for st in b.states])
        b.close()
    for i, (a, c) in enumerate(zip(*runs)):
        assert np.isfinite(a).all(), i
        assert np.array_equal(a, c), i


@pytest.mark.parametrize("schedule", [W.WS_SCHED_GRID, W.WS_SCHED_PREFETCH])
@pytest.mark.parametrize("nsteps", [2, 3, 5])
def test_batch_graph_replay_equals_eager(ws, nsteps, schedule):
    extents = _extents(17, 4, 6, 28)
    runs = []
    for graph in (1, 0):
        b, _ = _batch(extents, 600)
        W.check(W.lib().ws_batch_set_schedule(b.handle, schedule))
        for e, st in zip(extents, b.states):
            _to_cons(make_ctx(*e), st)
        W.check(W.lib().ws_batch_rk3_run(b.handle, nsteps, 0.5, graph, W.stream_handle()))
        b.sync_states()
        _sync()
        runs.append([st.interior("f").cpu().numpy() for st in b.states])
        b.close()
    for a, c in zip(*runs):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("precision", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
def test_batch_with_a_block_of_different_pitches_takes_the_64_bit_addresses(ws, precision):
    """A batch addresses its blocks' five input fields with one 32-bit offset only when every
    block's fields share a pitch; one block whose sound speed has another pitch makes the
    whole batch address each field in 64 bits.  The RHS and three RK3 steps (stage fusion 3,
    CFL step) agree bit for bit with the batch of common pitches."""
    extents = _extents(31, 4, 6, 30)
    kinds = [k for k in KINDS if k != "rough"]
    runs = []
    for odd in (False, True):
        prims = [random_state(*e, 900 + i, kinds[i % len(kinds)]) for i, e in enumerate(extents)]
        states = [upload_state(p) for p in prims]
        if odd:
            nx, ny, nz = extents[1]
            states[1].t["cc"] = W.ghosted_tensor(nx, ny, nz, pad=3)
            states[1].refresh()
            assert states[1].ws.cc.pitch != states[1].ws.d.pitch
        b = W.Batch(W.default_config(8, 8, 8, precision=precision, bc_planes=1, dt=2e-4), extents, states)
        W.check(W.lib().ws_batch_set_schedule(b.handle, W.WS_SCHED_GRID))
        for e, st in zip(extents, b.states):
            _to_cons(make_ctx(*e), st)
        s = W.stream_handle()
        W.check(W.lib().ws_batch_du_comput(b.handle, s))
        _sync()
        du = [_du(st) for st in b.states]
        W.check(W.lib().ws_batch_rk3_run(b.handle, 3, 0.5, 0, s))
        b.sync_states()
        _sync()
        runs.append((du, [{k: st.interior(k).cpu().numpy() for k in ("f", "d", "u", "v", "w", "T")}
                          for st in b.states]))
        b.close()
    (du_a, st_a), (du_b, st_b) = runs
    for i in range(len(extents)):
        assert np.array_equal(du_a[i], du_b[i]), i
        for k in st_a[i]:
            assert np.array_equal(st_a[i][k], st_b[i][k]), (i, k)


def test_batch_autotune_keeps_results(ws):
    """The tuned batch gives the same RHS to rounding; tuning (the stage order's timing
    included, when the grid schedule wins) leaves every block's f and primitives as they
    were."""
    extents = _extents(19, 5, 20, 60)
    b, _ = _batch(extents, 700)
    s = W.stream_handle()
    for e, st in zip(extents, b.states):
        _to_cons(make_ctx(*e), st)
    W.check(W.lib().ws_batch_du_comput(b.handle, s))
    _sync()
    before = [_du(st) for st in b.states]
    kept = ("f", "d", "u", "v", "w", "T", "cc")
    state = [{k: st.t[k].clone() for k in kept} for st in b.states]
    ns = len(W.SCHEDULES)
    times = (ctypes.c_float * (len(W.TILE_SHAPES) * ns))()
    W.check(W.lib().ws_batch_autotune(b.handle, 2, times, s))
    _sync()
    for i, st in enumerate(b.states):
        for k in kept:
            assert st.t[k].equal(state[i][k]), (i, k)
    shape, sched, sorder = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    W.check(W.lib().ws_batch_get_tile_shape(b.handle, ctypes.byref(shape)))
    W.check(W.lib().ws_batch_get_schedule(b.handle, ctypes.byref(sched)))
    W.check(W.lib().ws_batch_get_stage_order(b.handle, ctypes.byref(sorder)))
    log(f"  batch autotune ms {list(times)} -> {shape.value}/{sched.value}, stage order {sorder.value}")
    assert times[shape.value * ns + sched.value] == min(times)
    assert sorder.value in W.ORDERS
    W.check(W.lib().ws_batch_du_comput(b.handle, s))
    _sync()
    for a, st in zip(before, b.states):
        assert rel_l2(_du(st), a) <= 1e-13


def test_batch_helpers_in_one_launch_equal_per_block_launches(ws):
    """The helpers around the batched RHS (ghost fill, sound speed, the stage kernel) run as
    one launch over all blocks once the batch's work-group table exists, which the first
    call outside a stream capture builds; a batch whose first call is captured runs them
    block by block.  The RHS and an RK3 step (boundary planes on, fixed dt) come out the
    same bit for bit either way."""
    import torch
    L = W.lib()
    extents = _extents(29, 6, 5, 40)
    results = {}
    for captured in (False, True):
        outs = []
        for call in ("du", "step"):
            b, _ = _batch(extents, 900, bc_planes=1, dt=2e-4)
            W.check(L.ws_batch_set_schedule(b.handle, W.WS_SCHED_GRID))
            for e, st in zip(extents, b.states):
                _to_cons(make_ctx(*e), st)
            _sync()

            def run():
                s = W.stream_handle()
                if call == "du":
                    W.check(L.ws_batch_du_comput(b.handle, s), "batch du_comput")
                else:
                    W.check(L.ws_batch_rk3_step(b.handle, None, s), "batch rk3 step")

            if captured:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    run()
                g.replay()
            else:
                run()
            b.sync_states()
            _sync()
            outs.append([st.interior("du" if call == "du" else "f").cpu().numpy() for st in b.states])
            b.close()
        results[captured] = outs
        log(f"batch helpers captured={captured} done")
    for a, c in zip(results[False], results[True]):
        for i, (x, y) in enumerate(zip(a, c)):
            assert np.isfinite(x).all(), i
            assert np.array_equal(x, y), i
