#!/usr/bin/env python3
"""Benchmark of the inviscid RHS on one MI355X.

A step is one full RHS evaluation L(U) of the 600x256x128 periodic box
(the Sod-plus-noise workload): on-device ghost fill of rho,u,v,w,T, sound
speed over the ghosted box, and the three directional WENO7-SYMBO /
Steger-Warming kernels writing du.  Inputs are resident in HBM before the
timed region.  The headline precision (--precision, default FP64) gives `value`;
the other precision is timed on the same state in the same run and reported
under "other_precision" with its speedup and its RHS error against the first.
Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed] [--no-second] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "inviscid RHS Mcell-updates/s, FP64 vs mixed at 600x256x128; % attainable roof"
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0
ALGO_BYTES_PER_CELL = 80     # 5 FP64 state in + 5 FP64 RHS out
# the fused viscous kernel: u, v, w, T in (32 B) and du[1..4] read and written (64 B); the
# fluxes never leave the chip
VISCOUS_BYTES_PER_CELL = 96


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", default="600x256x128")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU-baseline leg")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-viscous", action="store_true", help="skip the Navier-Stokes (viscous) leg")
    ap.add_argument("--no-steps200", action="store_true",
                    help="skip the 200-step FP64-vs-mixed state comparison of the default run")
    ap.add_argument("--no-second", action="store_true",
                    help="skip timing the other precision (mixed when --precision fp64)")
    ap.add_argument("--tile", type=int, default=-1, help="force a tile shape (default: autotune)")
    ap.add_argument("--schedule", type=int, default=-1,
                    help="with --tile: force a schedule (0 grid, 1 persistent, 2 persistent + prefetch)")
    ap.add_argument("--order", type=int, default=-1,
                    help="rk3 and batched: force the stage order after tuning (0 x, y, z: z applies "
                         "the stage; 1 z, y, x: x applies it)")
    ap.add_argument("--stage-shape", type=int, default=-2,
                    help="rk3 and batched: force the tile shape of the stage-applying launches after "
                         "tuning (-1: the context's or batch's shape)")
    ap.add_argument("--workload", default="rhs", choices=["rhs", "batched", "rk3"],
                    help="rhs: the headline RHS line (default); batched: 48 uneven blocks; "
                         "rk3: SSP-RK3 steps of one box in a graph")
    ap.add_argument("--rk3-steps", type=int, default=0, help="RK3 steps of the batched/rk3 workloads "
                                                            "(default: 10 batched, 100 rk3)")
    ap.add_argument("--stage-fusion", type=int, default=3, choices=[0, 1, 2, 3],
                    help="rk3 and batched RK3: 3 (default) as 2 and, for one box, also stage 3's primitives, "
                         "sound speed and CFL maxima, alternating the primitive sets from step to step (a "
                         "batch treats 3 as 2); 2 lets the last RHS launch apply each stage and, for one box, "
                         "derive the primitives of stages 1 and 2; 1 the stage only; 0 the separate stage "
                         "kernel")
    ap.add_argument("--probe", action="store_true",
                    help="only measure the device's VALU rates (ws_probe_valu) and print them")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (for HBM counter calibration)")
    return ap.parse_args()


# Algorithmic FLOPs of one RHS per cell in the reference's formulation (add, sub,
# mul, div, sqrt, max = 1; no FMA): 3 directions x (split 137 + 5 x (WENO P 230 +
# WENO M 230) + 10 flux differences x 4) = 7431.  Counted by the CPU restatement's
# op-counting build; tests/test_bench_contract.py keeps the two in step.
ALGO_FLOPS_PER_CELL = 7431
ALGO_FLOPS_PER_DIRECTION = 2477
# The same count with each WENO reconstruction's five (eps + S_j) sums formed once instead of
# the reference's twenty (3 directions x 10 reconstructions x 15 additions fewer): the lower
# figure a fused formulation cannot go below, reported beside the headline fraction.
ALGO_FLOPS_PER_CELL_SHARED_EPS = 6981
# In the mixed RHS the WENO part (10 x 230 per direction) runs in FP32 and the split and
# the flux difference (137 + 40) in FP64; its roof is the time of those FLOPs at the
# two peaks: 2477 / (177 / 78.6 + 2300 / 157.3) = 146.8 TFLOP/s of algorithmic FLOPs.
ALGO_FLOPS_FP64_PART = 137 + 40
ALGO_FLOPS_FP32_PART = 10 * 230
# Algorithmic FLOPs of the viscous terms per cell (CD6), the same convention, counted by the
# same op-counting build: the viscosity 6, twelve gradients 12 x 9, the stress, heat flux and
# four fluxes formed for each direction 3 x 125, twelve flux derivatives added to du
# 12 x (9 + 2).
VISCOUS_FLOPS_PER_CELL_CD6 = 621


def peak_tflops(mixed):
    """The VALU roof of the RHS for its precision (algorithmic FLOPs per second)."""
    if not mixed:
        return PEAK_FP64_TFLOPS
    return ALGO_FLOPS_PER_DIRECTION / (ALGO_FLOPS_FP64_PART / PEAK_FP64_TFLOPS +
                                       ALGO_FLOPS_FP32_PART / PEAK_FP32_TFLOPS)


def measured_peaks(L, sh):
    """The device's vector ALU throughput (ws_probe_valu): FMA, multiply and add in FP64
    and scalar FP32 (tera lane-operations/s), and the attainable roofs of the two RHS forms
    from the FMA rates (2 FLOP per FMA)."""
    import windstencil_amd as W
    rate = {}
    for pname, prec in (("fp64", W.WS_PREC_FP64), ("fp32", W.WS_PREC_MIXED)):
        for oname, op in (("fma", W.WS_PROBE_FMA), ("mul", W.WS_PROBE_MUL), ("add", W.WS_PROBE_ADD)):
            v = ctypes.c_double()
            W.check(L.ws_probe_valu(prec, op, ctypes.byref(v), sh))
            rate[f"{pname}_{oname}"] = round(v.value, 2)
    f64, f32 = 2 * rate["fp64_fma"], 2 * rate["fp32_fma"]
    mixed_roof = ALGO_FLOPS_PER_DIRECTION / (ALGO_FLOPS_FP64_PART / f64 + ALGO_FLOPS_FP32_PART / f32)
    return {"tera_lane_ops_per_s": rate, "fp64_fma_tflops": round(f64, 2), "fp32_fma_tflops": round(f32, 2),
            "fp64_roof": round(f64, 2), "mixed_roof": round(mixed_roof, 2)}


def algo_flops():
    return ALGO_FLOPS_PER_CELL, ALGO_FLOPS_PER_DIRECTION


def calibrate_copy():
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n * 8, "bytes_written": n * 8}))


def cpu_baseline(prims, nx, ny, nz, budget_s):
    """The oracle (C restatement, OpenMP) on the host cores, on a bounded sample:
    whole RHS evaluations of the same state, repeated until the budget is used."""
    from oracle import oracle as O
    from tests.helpers import ghosted
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    p = O.params(nx, ny, nz, nthreads=threads)
    g = [ghosted(a) for a in prims]
    for a in g:
        O.ghost_fill_periodic(p, a)
    cc = O.sound_speed(p, g[4])
    met = O.metrics_cartesian(p)
    aj = np.ones_like(cc)
    times = []
    du = np.zeros((5, nz, ny, nx))
    O.inviscid_rhs(p, g[0], g[1], g[2], g[3], cc, met, aj, du)      # warm-up, untimed
    t_end = time.perf_counter() + budget_s
    while True:
        du = np.zeros((5, nz, ny, nx))
        t0 = time.perf_counter()
        O.inviscid_rhs(p, g[0], g[1], g[2], g[3], cc, met, aj, du)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 5:
            break
    med = statistics.median(times)
    out = {"value": round(nx * ny * nz / med / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads,
           "kind": "port",
           "gflops": round(ALGO_FLOPS_PER_CELL * nx * ny * nz / med / 1e9, 1),
           "sample": f"{len(times)} full RHS evaluations of the {nx}x{ny}x{nz} box (median, after one "
                     f"untimed), C restatement -O3 -fopenmp, {threads} threads"}
    out.update(cpu_config1(threads))
    return out


def _host_info():
    info = {"os_cpu_count": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_config1(threads):
    """The CPU-path workload in full (64x32x16 Mach 2 freestream with 8 blobs: one RHS,
    then 10 SSP-RK3 steps at CFL 0.5) on `threads` host threads, and its RHS on one
    thread; medians of 5 runs after a warm-up."""
    from oracle import oracle as O
    from tests.helpers import ghosted
    from windstencil_amd import initial as I
    nx, ny, nz = 64, 32, 16
    prims = I.freestream_blobs(nx, ny, nz, 1)

    def rhs_rate(nthreads):
        p = O.params(nx, ny, nz, nthreads=nthreads)
        g = [ghosted(a) for a in prims]
        for a in g:
            O.ghost_fill_periodic(p, a)
        cc = O.sound_speed(p, g[4])
        met, aj = O.metrics_cartesian(p), np.ones_like(cc)
        times = []
        for k in range(6):
            du = np.zeros((5, nz, ny, nx))
            t0 = time.perf_counter()
            O.inviscid_rhs(p, g[0], g[1], g[2], g[3], cc, met, aj, du)
            if k:
                times.append(time.perf_counter() - t0)
        return nx * ny * nz / statistics.median(times) / 1e6

    def run_full():
        p = O.params(nx, ny, nz, nthreads=threads)
        g = [ghosted(a) for a in prims]
        t0 = time.perf_counter()
        f = O.pri_to_cons(p, *g)
        O.du_comput(p, *g)
        O.rk3_run(p, 10, -0.5, f, *g)
        return time.perf_counter() - t0

    run_full()
    full = statistics.median(run_full() for _ in range(5))
    return {"config1": {"workload": "64x32x16 Mach 2 + 8 blobs: one RHS + 10 SSP-RK3 steps (CFL 0.5)",
                        "seconds": round(full, 4), "threads": threads,
                        "rhs_mcell_per_s": round(rhs_rate(threads), 3),
                        "rhs_mcell_per_s_1_thread": round(rhs_rate(1), 3)},
            "host": _host_info()}


def _common(args, metric, value, ms, dtype, config, **extra):
    out = {"metric": metric, "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def bench_batched(args):
    """48 periodic blocks with extents drawn from U{32..256} (seed 1), each a freestream
    of Mach U[0.3, 3] in a seeded direction with Gaussian blobs.  A step is one RHS of
    every block (per-block ghost fill and sound speed, then the batched RHS launches, one
    per direction); then RK3 steps of the whole batch replayed from one hipGraph per step."""
    import torch
    import windstencil_amd as W
    from windstencil_amd import initial as I
    L = W.lib()
    prec = W.WS_PREC_FP64 if args.precision == "fp64" else W.WS_PREC_MIXED
    extents = I.multiblock_draw(1)
    stream = torch.cuda.current_stream()
    sh = W.stream_handle(stream)
    states = []
    for i, (nx, ny, nz) in enumerate(extents):
        st = W.TorchState(nx, ny, nz)
        st.set_prims(*I.block_freestream(nx, ny, nz, 1000 + i))
        ctx = W.Context(W.default_config(nx, ny, nz))
        W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                                 W.job((0, 0, 0), (nx, ny, nz)), sh))
        states.append(st)
    b = W.Batch(W.default_config(8, 8, 8, precision=prec), extents, states)
    W.check(L.ws_batch_set_stage_fusion(b.handle, args.stage_fusion))
    N = b.cells()
    W.check(L.ws_batch_du_comput(b.handle, sh))
    torch.cuda.synchronize()
    tune = None
    if args.tile >= 0:
        W.check(L.ws_batch_set_tile_shape(b.handle, args.tile))
        if args.schedule >= 0:
            W.check(L.ws_batch_set_schedule(b.handle, args.schedule))
    else:
        ns = len(W.SCHEDULES)
        times = (ctypes.c_float * (len(W.TILE_SHAPES) * ns))()
        W.check(L.ws_batch_autotune(b.handle, 3, times, sh))
        tune = {f"{s}/{q}": round(times[s * ns + q], 4) for s in W.TILE_SHAPES for q in W.SCHEDULES}
    chosen = ctypes.c_int32()
    W.check(L.ws_batch_get_tile_shape(b.handle, ctypes.byref(chosen)))
    sched = ctypes.c_int32()
    W.check(L.ws_batch_get_schedule(b.handle, ctypes.byref(sched)))
    if args.order >= 0:
        W.check(L.ws_batch_set_stage_order(b.handle, args.order))
    if args.stage_shape >= -1:
        W.check(L.ws_batch_set_stage_shape(b.handle, args.stage_shape))
    sorder, sshape = ctypes.c_int32(), ctypes.c_int32()
    W.check(L.ws_batch_get_stage_order(b.handle, ctypes.byref(sorder)))
    W.check(L.ws_batch_get_stage_shape(b.handle, ctypes.byref(sshape)))
    for _ in range(args.warmup):
        W.check(L.ws_batch_du_comput(b.handle, sh))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        W.check(L.ws_batch_du_comput(b.handle, sh))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    # the batched RHS launches alone (inputs already ghost-filled), HIP events on the stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        W.check(L.ws_batch_inviscid_rhs(b.handle, sh))
    e1.record(stream)
    torch.cuda.synchronize()
    kern_ms = e0.elapsed_time(e1) / args.steps
    flops_cell, _ = algo_flops()
    achieved = flops_cell * N / (kern_ms * 1e-3) / 1e12
    mixed = prec == W.WS_PREC_MIXED
    nsteps = args.rk3_steps or 10
    W.check(L.ws_batch_rk3_run(b.handle, 1, 0.5, 0, sh))      # warm-up step
    t0 = time.perf_counter()
    W.check(L.ws_batch_rk3_run(b.handle, nsteps, 0.5, 1, sh))
    rk3_s = time.perf_counter() - t0
    out = _common(args, "batched inviscid RHS Mcell-updates/s, 48 uneven periodic blocks", N / (ms * 1e-3) / 1e6,
                  ms, "f64" if prec == W.WS_PREC_FP64 else "f64-state/mixed-rhs",
                  {"workload": "48 periodic blocks, extents U{32..256} (seed 1), Mach U[0.3,3] freestream + blobs",
                   "blocks": len(extents), "cells": N, "precision": args.precision},
                  roofline={"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak_tflops(mixed), 2),
                            "unit": "TFLOP/s", "frac": round(achieved / peak_tflops(mixed), 4), "traffic": None,
                            "kernel": "k_rhs_batch x3 directions (one RHS of every block)",
                            "kernel_ms": round(kern_ms, 4)},
                  tile_shape=chosen.value, schedule=sched.value, autotune_ms=tune,
                  autotune_key="tile shape/schedule (0 grid of tiles, 1 persistent, 2 persistent + prefetch)",
                  rk3={"steps": nsteps, "graph": True, "seconds": round(rk3_s, 4), "stage_order": sorder.value,
                       "stage_shape": sshape.value,
                       "mcell_updates_per_s": round(N * 3 * nsteps / rk3_s / 1e6, 2),
                       "mcell_steps_per_s": round(N * nsteps / rk3_s / 1e6, 2),
                       "note": "N x 3 stages x steps / wall time; CFL 0.5 from the per-block minimum each step"})
    print(json.dumps(out))


def _ctx_choice(ctx):
    """'shape/schedule/order/stage order/stage shape' the context runs its tiled RHS with
    (stage shape -1: the context's shape)."""
    import windstencil_amd as W
    L = W.lib()
    chosen, sched, order, sorder = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    sshape = ctypes.c_int32()
    W.check(L.ws_ctx_get_stage_shape(ctx.handle, ctypes.byref(sshape)))
    W.check(L.ws_ctx_get_tile_shape(ctx.handle, ctypes.byref(chosen)))
    W.check(L.ws_ctx_get_schedule(ctx.handle, ctypes.byref(sched)))
    W.check(L.ws_ctx_get_direction_order(ctx.handle, ctypes.byref(order)))
    W.check(L.ws_ctx_get_stage_order(ctx.handle, ctypes.byref(sorder)))
    return f"{chosen.value}/{sched.value}/{order.value}/{sorder.value}/{sshape.value}"


def bench_rk3(args):
    """SSP-RK3 steps of one periodic box, replayed from one hipGraph per step (dt from
    the CFL reduction every step): the FP64 drop-in workload (600x128x128 vortex
    lattice, 100 steps) or, with --grid 1200x256x256, Mach 2 blobs with a Mach 2 planar
    shock (1000 steps)."""
    import torch
    import windstencil_amd as W
    from windstencil_amd import initial as I
    L = W.lib()
    prec = W.WS_PREC_FP64 if args.precision == "fp64" else W.WS_PREC_MIXED
    grid = args.grid if args.grid != "600x256x128" else "600x128x128"
    nx, ny, nz = (int(v) for v in grid.split("x"))
    N = nx * ny * nz
    if (nx, ny, nz) == (600, 128, 128):
        prims, what, nsteps = I.vortex_lattice(nx, ny, nz, 2024), "six Yee vortices (beta 5) on a Mach 0.5 stream", 100
    else:
        prims, what, nsteps = I.blobs_with_shock(nx, ny, nz, 7), "Mach 2 blobs with a Mach 2 planar shock", 1000
    nsteps = args.rk3_steps or nsteps
    stream = torch.cuda.current_stream()
    sh = W.stream_handle(stream)
    ctx = W.Context(W.default_config(nx, ny, nz, precision=prec))
    ctx.set_stage_fusion(args.stage_fusion)
    st = W.TorchState(nx, ny, nz)
    st.set_prims(*prims)
    W.check(L.ws_pri_to_cons(ctx.handle, st.ws.f, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T,
                             W.job((0, 0, 0), (nx, ny, nz)), sh))
    if args.tile >= 0:
        W.check(L.ws_ctx_set_tile_shape(ctx.handle, args.tile))
    else:
        W.check(L.ws_du_comput(ctx.handle, ctypes.byref(st.ws), None, sh))
        W.check(L.ws_autotune(ctx.handle, ctypes.byref(st.ws), 3, None, sh))
    if args.order >= 0:
        W.check(L.ws_ctx_set_stage_order(ctx.handle, args.order))
    if args.stage_shape >= -1:
        W.check(L.ws_ctx_set_stage_shape(ctx.handle, args.stage_shape))
    W.check(L.ws_rk3_run(ctx.handle, ctypes.byref(st.ws), 2, 0.5, 1, sh))     # warm-up steps
    st.sync_from_ws()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    W.check(L.ws_rk3_run(ctx.handle, ctypes.byref(st.ws), nsteps, 0.5, 1, sh))
    secs = time.perf_counter() - t0
    st.sync_from_ws()
    f = st.interior("f")
    finite = bool(torch.isfinite(f).all())
    ms = secs / nsteps * 1e3
    # the step against the RHS roof: the three RHS evaluations' algorithmic FLOPs over the
    # whole step time (the memory-bound stage, ghost-fill and CFL kernels count as time)
    mixed = prec == W.WS_PREC_MIXED
    achieved = 3 * ALGO_FLOPS_PER_CELL * N / (ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak_tflops(mixed), 2),
            "unit": "TFLOP/s", "frac": round(achieved / peak_tflops(mixed), 4), "traffic": None,
            "kernel": "whole RK3 step (3 RHS of algorithmic FLOPs over the step time)"}
    out = _common(args, f"SSP-RK3 Mcell-updates/s (N x 3 stages x steps / time), {nx}x{ny}x{nz}",
                  N * 3 * nsteps / secs / 1e6, ms, "f64" if prec == W.WS_PREC_FP64 else "f64-state/mixed-rhs",
                  {"workload": f"{nsteps} SSP-RK3 steps of the {nx}x{ny}x{nz} periodic box ({what}), CFL 0.5, "
                               "one hipGraph per step", "grid": [nx, ny, nz], "precision": args.precision,
                   "cells": N},
                  roofline=roof, rk3_steps=nsteps, seconds=round(secs, 4), state_finite=finite,
                  mcell_steps_per_s=round(N * nsteps / secs / 1e6, 2),
                  stage_fusion=args.stage_fusion, **{"tile_shape/schedule/order/stage_order/stage_shape": _ctx_choice(ctx)},
                  totals=[float(v) for v in f.sum(dim=(1, 2, 3)).cpu()])
    out["steps"] = nsteps
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 "
                                   "(the multi-GPU decomposition is a design only)"}))
        sys.exit(2)
    import torch
    import windstencil_amd as W
    from windstencil_amd import initial as I

    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.probe:
        L = W.lib()
        print(json.dumps({"probe": "ws_probe_valu", **measured_peaks(L, W.stream_handle(torch.cuda.current_stream()))}))
        return
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "rk3":
        bench_rk3(args)
        return
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    N = nx * ny * nz
    L = W.lib()
    prec = W.WS_PREC_FP64 if args.precision == "fp64" else W.WS_PREC_MIXED
    ctx = W.Context(W.default_config(nx, ny, nz, precision=prec))
    prims = I.sod_noise(nx, ny, nz, 20240601)
    st = W.TorchState(nx, ny, nz)
    st.set_prims(*prims)
    stream = torch.cuda.current_stream()
    sh = W.stream_handle(stream)
    ijob = W.interior_job(nx, ny, nz)

    def rhs(ev=None):
        # one full RHS: ghost fill and sound speed (one pass), three directional kernels
        if ev is not None:
            ev[0].record(stream)
        W.check(L.ws_prepare_rhs_inputs(ctx.handle, ctypes.byref(st.ws), sh))
        if ev is not None:
            ev[1].record(stream)
        W.check(L.ws_inviscid_rhs(ctx.handle, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.cc, None, st.ws.du,
                                  ijob, 0, sh))
        if ev is not None:
            ev[2].record(stream)

    def run(precision):
        """Autotune (unless a shape is forced), warm up, then time args.steps RHS
        evaluations: wall clock over the whole loop, and HIP events per phase."""
        ctx.set_precision(precision)
        rhs()
        torch.cuda.synchronize()
        tune = None
        if args.tile >= 0:
            W.check(L.ws_ctx_set_tile_shape(ctx.handle, args.tile))
            if args.schedule >= 0:
                W.check(L.ws_ctx_set_schedule(ctx.handle, args.schedule))
        else:
            ns = len(W.SCHEDULES)
            times = (ctypes.c_float * (len(W.TILE_SHAPES) * ns))()
            W.check(L.ws_autotune(ctx.handle, ctypes.byref(st.ws), 5, times, sh))
            tune = {f"{s}/{q}": round(times[s * ns + q], 4) for s in W.TILE_SHAPES for q in W.SCHEDULES}
        chosen, sched, order = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        W.check(L.ws_ctx_get_tile_shape(ctx.handle, ctypes.byref(chosen)))
        W.check(L.ws_ctx_get_schedule(ctx.handle, ctypes.byref(sched)))
        W.check(L.ws_ctx_get_direction_order(ctx.handle, ctypes.byref(order)))
        for _ in range(args.warmup):
            rhs()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rhs()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
        for k in range(args.steps):
            rhs(evs[k])
        torch.cuda.synchronize()
        prep = statistics.mean(e[0].elapsed_time(e[1]) for e in evs)
        kern = statistics.mean(e[1].elapsed_time(e[2]) for e in evs)
        return wall / args.steps * 1e3, prep, kern, f"{chosen.value}/{sched.value}/{order.value}", tune

    other = W.WS_PREC_MIXED if prec == W.WS_PREC_FP64 else W.WS_PREC_FP64
    second = None
    if not args.no_second:
        # the other precision on the same state, for the FP64-vs-mixed comparison
        ms2, _, kern2, shape2, tune2 = run(other)
        du2 = st.interior("du").clone()
        ach2 = ALGO_FLOPS_PER_CELL * N / (kern2 * 1e-3) / 1e12
        peak2 = peak_tflops(other == W.WS_PREC_MIXED)
        second = {"precision": "mixed" if other == W.WS_PREC_MIXED else "fp64",
                  "value": round(N / (ms2 * 1e-3) / 1e6, 2), "ms_per_step": round(ms2, 4),
                  "kernel_ms": round(kern2, 4), "tile_shape/schedule/order": shape2, "autotune_ms": tune2,
                  "roofline": {"bound": "valu", "achieved": round(ach2, 3), "peak": round(peak2, 2),
                               "unit": "TFLOP/s", "frac": round(ach2 / peak2, 4)}}
    ms, prep_ms, kern_ms, shape, tune = run(prec)
    mixed = prec == W.WS_PREC_MIXED
    value = N / (ms * 1e-3) / 1e6
    flops_cell, flops_dir = algo_flops()
    achieved = flops_cell * N / (kern_ms * 1e-3) / 1e12
    # the attainable roof on this device (FMA probe), beside the published peak
    probe = measured_peaks(L, sh)
    roof_key = "mixed_roof" if mixed else "fp64_roof"
    if second is not None:
        rk = "mixed_roof" if second["precision"] == "mixed" else "fp64_roof"
        second["roofline"]["measured_roof"] = probe[rk]
        second["roofline"]["frac_of_measured"] = round(second["roofline"]["achieved"] / probe[rk], 4)
    if second is not None:
        du1 = st.interior("du")
        second["rel_l2_vs_" + args.precision] = float(torch.linalg.vector_norm(du2 - du1) /
                                                      torch.linalg.vector_norm(du1))
        mixed_ms = ms if prec == W.WS_PREC_MIXED else second["ms_per_step"]
        fp64_ms = ms if prec == W.WS_PREC_FP64 else second["ms_per_step"]
        second["speedup_mixed_over_fp64"] = round(fp64_ms / mixed_ms, 4)
        del du2

    parity = None
    if not args.no_parity:
        # fast path against the reference-order anchor kernels on the same state (before
        # the Navier-Stokes leg below overwrites du)
        du_fast = st.interior("du").clone()
        W.check(L.ws_ctx_set_kernel(ctx.handle, W.WS_KERNEL_ANCHOR))
        rhs()
        torch.cuda.synchronize()
        du_ref = st.interior("du")
        parity = float(torch.linalg.vector_norm(du_fast - du_ref) / torch.linalg.vector_norm(du_ref))
        W.check(L.ws_ctx_set_kernel(ctx.handle, W.WS_KERNEL_TILED))
        del du_fast

    steps200 = None
    if not args.no_second and not args.no_steps200:
        # the mixed form's accuracy over a run: 200 SSP-RK3 steps (CFL 0.5 from the device
        # each step, graph-replayed) of the same initial state in both precisions; the
        # conserved states compared (relative L2; Linf relative to max |f|)
        def rk3_state(precision):
            c2 = W.Context(W.default_config(nx, ny, nz, precision=precision))
            s2 = W.TorchState(nx, ny, nz)
            s2.set_prims(*prims)
            W.check(L.ws_pri_to_cons(c2.handle, s2.ws.f, s2.ws.d, s2.ws.u, s2.ws.v, s2.ws.w, s2.ws.T,
                                     W.job((0, 0, 0), (nx, ny, nz)), sh))
            W.check(L.ws_ctx_set_tile_shape(c2.handle, int(shape.split("/")[0])))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            W.check(L.ws_rk3_run(c2.handle, ctypes.byref(s2.ws), 200, 0.5, 1, sh))
            secs = time.perf_counter() - t0
            s2.sync_from_ws()
            f = s2.interior("f").clone()
            del s2, c2
            return f, secs
        fa, sa = rk3_state(W.WS_PREC_FP64)
        fb, sb = rk3_state(W.WS_PREC_MIXED)
        steps200 = {"steps": 200, "cfl": 0.5,
                    "state_rel_l2": float(torch.linalg.vector_norm(fb - fa) / torch.linalg.vector_norm(fa)),
                    "state_linf_over_max": float((fb - fa).abs().max() / fa.abs().max()),
                    "seconds_fp64": round(sa, 3), "seconds_mixed": round(sb, 3)}
        del fa, fb

    viscous = None
    if not args.no_viscous:
        # the Navier-Stokes RHS: the same inviscid RHS plus the viscous terms (CD6,
        # Re 1000), timed on the same state
        vctx = W.Context(W.default_config(nx, ny, nz, viscous=6))
        W.check(L.ws_ctx_set_tile_shape(vctx.handle, int(shape.split("/")[0])))
        W.check(L.ws_ctx_set_schedule(vctx.handle, int(shape.split("/")[1])))

        def ns_rhs(ev=None):
            W.check(L.ws_prepare_rhs_inputs(vctx.handle, ctypes.byref(st.ws), sh))
            W.check(L.ws_inviscid_rhs(vctx.handle, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.cc, None, st.ws.du,
                                      ijob, 0, sh))
            if ev is not None:
                ev[0].record(stream)
            W.check(L.ws_viscous_rhs(vctx.handle, st.ws.u, st.ws.v, st.ws.w, st.ws.T, None, st.ws.du, sh))
            if ev is not None:
                ev[1].record(stream)

        for _ in range(args.warmup):
            ns_rhs()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ns_rhs()
        torch.cuda.synchronize()
        ns_ms = (time.perf_counter() - t0) / args.steps * 1e3
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
        for k in range(args.steps):
            ns_rhs(evs[k])
        torch.cuda.synchronize()
        vis_ms = statistics.mean(e[0].elapsed_time(e[1]) for e in evs)
        vis_bytes = VISCOUS_BYTES_PER_CELL * N
        vis_tf = VISCOUS_FLOPS_PER_CELL_CD6 * N / (vis_ms * 1e-3) / 1e12
        vis_gbs = vis_bytes / (vis_ms * 1e-3) / 1e9
        # the fused kernel is bound by neither roof: it issues its own instruction stream (the
        # flux points of a ring around each tile and of the z-chunk halo are computed twice)
        # at two wavefronts per SIMD; both fractions are reported
        viscous = {"workload": "inviscid + viscous (CD6, Re 1000, Sutherland) RHS, FP64; viscous terms in one "
                               "fused kernel (k_vis_fused)",
                   "value": round(N / (ns_ms * 1e-3) / 1e6, 2), "unit": "Mcell-updates/s",
                   "ms_per_step": round(ns_ms, 4), "viscous_kernels_ms": round(vis_ms, 4),
                   "viscous_roofline": {"bound": "valu", "achieved": round(vis_tf, 3), "peak": PEAK_FP64_TFLOPS,
                                        "unit": "TFLOP/s", "frac": round(vis_tf / PEAK_FP64_TFLOPS, 4),
                                        "algo_flops_per_cell": VISCOUS_FLOPS_PER_CELL_CD6,
                                        "hbm_achieved_gbs": round(vis_gbs, 1),
                                        "hbm_frac": round(vis_gbs / PEAK_HBM_GBS, 4),
                                        "bytes_per_cell": VISCOUS_BYTES_PER_CELL}}
        del vctx

    # HBM bytes per RHS are not counted in this run: they come from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of the same kernels (scaled by a plain copy's factors)
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "rhs_traffic_600x256x128.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_rhs")
        traffic_src = f"profile-derived: profiles/rhs_traffic_600x256x128.json ({tj.get('kernel')})"

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == W.WS_PREC_FP64 else "f64-state/mixed-rhs",
        "data": "synthetic (seeded Sod plus 1% noise, periodic)",
        "config": {"workload": f"one inviscid RHS of the {nx}x{ny}x{nz} periodic box "
                               "(ghost fill + sound speed + WENO7-SYMBO/Steger-Warming in x, y, z)",
                   "grid": [nx, ny, nz], "precision": args.precision, "cells": N},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak_tflops(mixed), 2),
                     "unit": "TFLOP/s", "frac": round(achieved / peak_tflops(mixed), 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_rhs_tile x3 directions (one RHS)",
                     "measured_roof": probe[roof_key],
                     "frac_of_measured": round(achieved / probe[roof_key], 4), "fma_probe": probe,
                     "algo_flops_per_cell": flops_cell, "kernel_ms": round(kern_ms, 4),
                     "algo_flops_per_cell_shared_eps": ALGO_FLOPS_PER_CELL_SHARED_EPS,
                     "frac_shared_eps": round(achieved * ALGO_FLOPS_PER_CELL_SHARED_EPS / flops_cell
                                              / peak_tflops(mixed), 4)},
        "phases_ms": {"ghost_fill+sound_speed": round(prep_ms, 4), "rhs_kernels": round(kern_ms, 4)},
        "tile_shape/schedule/order": shape, "autotune_ms": tune,
        "autotune_key": "tile shape/schedule (0 grid of tiles, 1 persistent, 2 persistent + prefetch), "
                        "each the faster direction order; order 0 = x, y, z, 1 = z, y, x",
        "parity_rel_l2_vs_anchor": parity,
        "other_precision": second,
        "mixed_vs_fp64_200_rk3_steps": steps200,
        "navier_stokes": viscous,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prims, nx, ny, nz, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
