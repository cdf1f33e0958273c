// autotune.cpp -- pick the tile shape and schedule of the tiled kernels by timing them.
//
// Every (tile shape, schedule, direction order) runs the full inviscid RHS on the caller's
// state, in the context's precision: two untimed launches, then `reps` launches bracketed
// by HIP events on the caller's stream.  The fastest becomes the context's and is
// remembered for the grid extents and precision, so later contexts of the same size and
// precision start from it.
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
    int shape, schedule, order;
};
// (nx, ny, nz, precision) -> choice
std::map<std::tuple<int, int, int, int>, Choice> g_best;

int launch(const Ctx &c, const ws_state *st, const ws_job &job, hipStream_t s)
{
    if (c.cfg.precision == WS_PREC_MIXED)
        return launch_rhs_mixed(c, st->d, st->u, st->v, st->w, st->cc, nullptr, st->du, job, 0, s);
    return launch_rhs_tiled(c, st->d, st->u, st->v, st->w, st->cc, nullptr, st->du, job, 0, s);
}

}  // namespace

namespace ws {

bool remembered_choice(int nx, int ny, int nz, int precision, int *shape, int *schedule, int *order)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_best.find(std::make_tuple(nx, ny, nz, precision));
    if (it == g_best.end()) return false;
    *shape = it->second.shape;
    *schedule = it->second.schedule;
    *order = it->second.order;
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
    const Choice saved{c->tile_shape, c->schedule, c->dir_order};
    Choice best = saved;
    int r = WS_OK;
    float best_ms = 1e30f;
    for (int shape = 0; shape < WS_TILE_COUNT && !r; ++shape)
        for (int sched = 0; sched < WS_SCHED_COUNT && !r; ++sched) {
            float pair_ms = 1e30f;
            for (int order = 0; order < WS_ORDER_COUNT && !r; ++order) {
                c->tile_shape = shape;
                c->schedule = sched;
                c->dir_order = order;
                for (int i = 0; i < 2 && !r; ++i) r = launch(*c, st, job, s);
                if (r) break;
                hipEventRecord(e0, s);
                for (int i = 0; i < reps && !r; ++i) r = launch(*c, st, job, s);
                hipEventRecord(e1, s);
                if (r) break;
                e = hipEventSynchronize(e1);
                if (e != hipSuccess) { r = hip_error(e, "autotune sync"); break; }
                float ms = 0.f;
                hipEventElapsedTime(&ms, e0, e1);
                ms /= (float)reps;
                if (ms < pair_ms) pair_ms = ms;
                if (ms < best_ms) { best_ms = ms; best = Choice{shape, sched, order}; }
            }
            if (times_ms) times_ms[shape * WS_SCHED_COUNT + sched] = pair_ms;
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const Choice &use = r ? saved : best;
    c->tile_shape = use.shape;
    c->schedule = use.schedule;
    c->dir_order = use.order;
    if (!r) {
        std::lock_guard<std::mutex> lk(g_mu);
        g_best[std::make_tuple(c->cfg.nx, c->cfg.ny, c->cfg.nz, c->cfg.precision)] = best;
    }
    return r;
}
