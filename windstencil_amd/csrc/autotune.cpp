// autotune.cpp -- pick the tile shape and schedule of the tiled kernels by timing them.
//
// Every (tile shape, schedule, direction order) runs the full inviscid RHS on the caller's
// state, in the context's precision: two untimed launches, then `reps` launches bracketed
// by HIP events on the caller's stream.  The fastest becomes the context's.  When RK3
// steps of that choice let the last RHS launch apply the stage, that stage-applying RHS is
// timed the same way in every tile shape and both direction orders (into scratch outputs)
// and the fastest becomes the stage shape and order.  The choice is remembered for the
// grid extents and precision, so later contexts of the same size and precision start from
// it.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "ws_internal.h"

using namespace ws;

namespace {

std::mutex g_mu;
struct Choice {
    int shape, schedule, order, stage_order, stage_shape;
};
// (nx, ny, nz, precision) -> choice
std::map<std::tuple<int, int, int, int>, Choice> g_best;

int launch(const Ctx &c, const ws_state *st, const ws_job &job, hipStream_t s)
{
    if (c.cfg.precision == WS_PREC_MIXED)
        return launch_rhs_mixed(c, st->d, st->u, st->v, st->w, st->cc, nullptr, st->du, job, 0, s);
    return launch_rhs_tiled(c, st->d, st->u, st->v, st->w, st->cc, nullptr, st->du, job, 0, s);
}

// Mean time of `reps` calls of fn() on s after two untimed ones; negative on error (r set).
template <class F>
float time_launches(F fn, int reps, hipEvent_t e0, hipEvent_t e1, hipStream_t s, int &r)
{
    for (int i = 0; i < 2 && !r; ++i) r = fn();
    if (r) return -1.f;
    hipEventRecord(e0, s);
    for (int i = 0; i < reps && !r; ++i) r = fn();
    hipEventRecord(e1, s);
    if (r) return -1.f;
    hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) {
        r = hip_error(e, "autotune sync");
        return -1.f;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    return ms / (float)reps;
}

// The stage shape and order: the RHS of an RK3 stage 2 (f = 3/4 fn + 1/4 (f + dt RHS))
// whose last launch writes the new f and, with every direction periodic, the primitives,
// sound speed, their periodic halo and the per-tile CFL maxima: what a fused step's stage
// launches write.  fn is the state's f (read only); f, the primitives and the maxima go to
// scratch allocated here (f a copy of the state's, so the values stay physical).  Sets
// *shape and *order to the fastest; false when the scratch cannot be had (both then stay
// as they were) or on an error (r set).
bool tune_stage(ws_ctx *c, const ws_state *st, const ws_job &job, int reps, hipEvent_t e0, hipEvent_t e1,
                hipStream_t s, int &r, int *shape_out, int *order_out)
{
    const bool periodic = c->cfg.periodic[0] && c->cfg.periodic[1] && c->cfg.periodic[2];
    const bool prims = periodic && c->fuse_stage > 1;
    const size_t fper = (size_t)st->f.pitch * c->kp.ny * c->kp.nz;
    const size_t pper = (size_t)st->d.pitch * c->kp.nyg * c->kp.nzg;
    const int saved = c->stage_order, saved_shape = c->stage_shape;
    long tiles = 0;
    for (int sh = 0; sh < WS_TILE_COUNT; ++sh)
        for (int o = 0; o < WS_ORDER_COUNT; ++o) {
            c->stage_shape = sh;
            c->stage_order = o;
            const long n = stage_launch_tiles(*c);
            if (n > tiles) tiles = n;
        }
    c->stage_order = saved;
    c->stage_shape = saved_shape;
    const size_t elems = 5 * fper + (prims ? 6 * pper + (size_t)tiles : 0);
    double *scratch = nullptr;
    if (hipMalloc(&scratch, elems * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    hipError_t e = hipMemcpyAsync(scratch, st->f.ptr, 5 * fper * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) {
        r = hip_error(e, "autotune stage scratch copy");
        hipFree(scratch);
        return false;
    }
    ws_soa f = st->f;
    f.ptr = scratch;
    StageArgs sa{f, st->f, nullptr, c->cfg.dt, 2};
    sa.planes = c->cfg.bc_planes != 0;
    if (prims) {
        ws_field *out[6] = {&sa.pd, &sa.pu, &sa.pv, &sa.pw, &sa.pT, &sa.pcc};
        for (int q = 0; q < 6; ++q) *out[q] = ws_field{scratch + 5 * fper + q * pper, st->d.pitch, 0};
        sa.halo = 1;
        sa.cfl_part = scratch + 5 * fper + 6 * pper;
    }
    auto stage = [&]() {
        return c->cfg.precision == WS_PREC_MIXED
                   ? launch_rhs_mixed_stage(*c, st->d, st->u, st->v, st->w, st->cc, st->du, job, sa, s)
                   : launch_rhs_tiled_stage(*c, st->d, st->u, st->v, st->w, st->cc, st->du, job, sa, s);
    };
    float best_ms = 1e30f;
    for (int sh = 0; sh < WS_TILE_COUNT && !r; ++sh)
        for (int o = 0; o < WS_ORDER_COUNT && !r; ++o) {
            c->stage_shape = sh;
            c->stage_order = o;
            const float ms = time_launches(stage, reps, e0, e1, s, r);
            if (!r && ms < best_ms) {
                best_ms = ms;
                *shape_out = sh;
                *order_out = o;
            }
        }
    c->stage_order = saved;
    c->stage_shape = saved_shape;
    hipStreamSynchronize(s);
    hipFree(scratch);
    return !r;
}

}  // namespace

namespace ws {

bool remembered_choice(int nx, int ny, int nz, int precision, int *shape, int *schedule, int *order,
                       int *stage_order, int *stage_shape)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_best.find(std::make_tuple(nx, ny, nz, precision));
    if (it == g_best.end()) return false;
    *shape = it->second.shape;
    *schedule = it->second.schedule;
    *order = it->second.order;
    *stage_order = it->second.stage_order;
    *stage_shape = it->second.stage_shape;
    return true;
}

}  // namespace ws

extern "C" int ws_autotune(ws_ctx *c, const ws_state *st, int32_t reps, float *times_ms, hipStream_t s)
{
    if (!c || !st || reps <= 0) return set_error(WS_ERR_INVALID, "bad autotune request");
    ws_job job;
    job.start = {(uint32_t)LAP, (uint32_t)LAP, (uint32_t)LAP};
    job.end = {(uint32_t)(c->cfg.nx + LAP), (uint32_t)(c->cfg.ny + LAP), (uint32_t)(c->cfg.nz + LAP)};
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return hip_error(e, "autotune event");
    e = hipEventCreate(&e1);
    if (e != hipSuccess) { hipEventDestroy(e0); return hip_error(e, "autotune event"); }
    const Choice saved{c->tile_shape, c->schedule, c->dir_order, c->stage_order, c->stage_shape};
    Choice best = saved;
    int r = WS_OK;
    float best_ms = 1e30f;
    auto rhs = [&]() { return launch(*c, st, job, s); };
    for (int shape = 0; shape < WS_TILE_COUNT && !r; ++shape)
        for (int sched = 0; sched < WS_SCHED_COUNT && !r; ++sched) {
            float pair_ms = 1e30f;
            for (int order = 0; order < WS_ORDER_COUNT && !r; ++order) {
                c->tile_shape = shape;
                c->schedule = sched;
                c->dir_order = order;
                const float ms = time_launches(rhs, reps, e0, e1, s, r);
                if (r) break;
                if (ms < pair_ms) pair_ms = ms;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = Choice{shape, sched, order, saved.stage_order, saved.stage_shape};
                }
            }
            if (times_ms) times_ms[shape * WS_SCHED_COUNT + sched] = pair_ms;
        }
    if (!r) {
        c->tile_shape = best.shape;
        c->schedule = best.schedule;
        c->dir_order = best.order;
        // the stage candidates differ by 1-3 %: three times the repetitions
        int sh = best.stage_shape, o = best.stage_order;
        if (fused_stage_ok(c) && tune_stage(c, st, job, 3 * reps, e0, e1, s, r, &sh, &o)) {
            best.stage_shape = sh;
            best.stage_order = o;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const Choice &use = r ? saved : best;
    c->tile_shape = use.shape;
    c->schedule = use.schedule;
    c->dir_order = use.order;
    c->stage_order = use.stage_order;
    c->stage_shape = use.stage_shape;
    if (!r) {
        std::lock_guard<std::mutex> lk(g_mu);
        g_best[std::make_tuple(c->cfg.nx, c->cfg.ny, c->cfg.nz, c->cfg.precision)] = best;
    }
    return r;
}
