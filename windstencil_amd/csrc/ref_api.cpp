// ref_api.cpp -- the reference-named entry points over a bound context and state.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "windstencil_amd_ref.h"
#include "ws_internal.h"

namespace {

struct Bound {
    ws_ctx *ctx = nullptr;
    ws_state *st = nullptr;
    const ws_metric_fields *metrics = nullptr;
    hipStream_t stream = nullptr;
    ws_soa f_lap{};          // the ghosted copy stage 3 writes (pf_lap_d); ptr null: none
};

Bound g_bound;
int g_abort_on_error = 1;
int g_last_status = WS_OK;

bool fail(const char *fn, int code)
{
    g_last_status = code;
    if (code == WS_OK) return false;
    std::fprintf(stderr, "windstencil_amd: %s failed: %s (%s)\n", fn, ws_error_string(code), ws_last_error());
    if (g_abort_on_error) std::abort();
    return true;
}

bool bound(const char *fn)
{
    if (g_bound.ctx && g_bound.st) return true;
    ws::set_error(WS_ERR_INVALID, "no context bound: call ws_ref_bind first");
    fail(fn, WS_ERR_INVALID);
    return false;
}

ws_job lap_job()
{
    ws_config c;
    ws_ctx_get_config(g_bound.ctx, &c);
    ws_job j;
    j.start = {(uint32_t)WS_LAP, (uint32_t)WS_LAP, (uint32_t)WS_LAP};
    j.end = {(uint32_t)(c.nx + WS_LAP), (uint32_t)(c.ny + WS_LAP), (uint32_t)(c.nz + WS_LAP)};
    return j;
}

}  // namespace

extern "C" {

int ws_ref_bind(ws_ctx *ctx, ws_state *st, const ws_metric_fields *metrics, hipStream_t stream)
{
    if (!ctx || !st) return ws::set_error(WS_ERR_INVALID, "ws_ref_bind: null context or state");
    g_bound.ctx = ctx;
    g_bound.st = st;
    g_bound.metrics = metrics;
    g_bound.stream = stream;
    g_bound.f_lap = ws_soa{};
    g_last_status = WS_OK;
    return WS_OK;
}

int ws_ref_bind_f_lap(const ws_soa *f_lap)
{
    if (!g_bound.ctx) return ws::set_error(WS_ERR_INVALID, "ws_ref_bind_f_lap: no context bound");
    if (f_lap && f_lap->ptr) {
        ws_config c;
        ws_ctx_get_config(g_bound.ctx, &c);
        if (f_lap->pitch < (uint32_t)(c.nx + 2 * WS_LAP))
            return ws::set_error(WS_ERR_INVALID, "ws_ref_bind_f_lap: pitch smaller than nx+2*LAP");
        g_bound.f_lap = *f_lap;
    } else {
        g_bound.f_lap = ws_soa{};
    }
    return WS_OK;
}

void ws_ref_set_abort_on_error(int abort_on_error) { g_abort_on_error = abort_on_error; }

int ws_ref_last_status(void) { return g_last_status; }

void du_invis_Jacobian3d_init(ws_job job_in, hipStream_t *stream)
{
    if (!bound(__func__)) return;
    // the sound speed covers the job box widened by LAP on every side
    ws_job j = job_in;
    j.start.x -= WS_LAP; j.start.y -= WS_LAP; j.start.z -= WS_LAP;
    j.end.x += WS_LAP; j.end.y += WS_LAP; j.end.z += WS_LAP;
    fail(__func__, ws_sound_speed(g_bound.ctx, g_bound.st->T, g_bound.st->cc, j, stream ? *stream : g_bound.stream));
}

void du_invis_Jacobian3d_x(ws_job job_in, ws_soa *fp, ws_soa *fm, hipStream_t *stream)
{
    (void)fp; (void)fm;
    if (!bound(__func__)) return;
    const ws_state *s = g_bound.st;
    fail(__func__, ws_inviscid_rhs(g_bound.ctx, s->d, s->u, s->v, s->w, s->cc, g_bound.metrics, s->du, job_in, 0,
                                   stream ? *stream : g_bound.stream));
}

void *du_invis_Jacobian3d(void *pthread_id)
{
    (void)pthread_id;
    if (!bound(__func__)) return nullptr;
    const ws_state *s = g_bound.st;
    const ws_field g[5] = {s->d, s->u, s->v, s->w, s->T};
    if (fail(__func__, ws_ghost_fill_periodic(g_bound.ctx, g, 5, g_bound.stream))) return nullptr;
    const ws_job job = lap_job();
    hipStream_t st = g_bound.stream;
    du_invis_Jacobian3d_init(job, &st);
    if (g_last_status) return nullptr;
    du_invis_Jacobian3d_x(job, nullptr, nullptr, &st);
    return nullptr;
}

void *du_vis_Jacobian3d(void *pthread_id)
{
    (void)pthread_id;
    if (!bound(__func__)) return nullptr;
    const ws_state *s = g_bound.st;
    fail(__func__, ws_viscous_rhs(g_bound.ctx, s->u, s->v, s->w, s->T, g_bound.metrics, s->du, g_bound.stream));
    return nullptr;
}

void du_comput(int KRK)
{
    (void)KRK;
    if (!bound(__func__)) return;
    // the reference zeroes du first (OCFD_Stream.cpp:35); the RHS here overwrites du,
    // which leaves the same result without the extra pass
    du_invis_Jacobian3d(nullptr);
    if (g_last_status) return;
    ws_config c;
    ws_ctx_get_config(g_bound.ctx, &c);
    if (c.viscous) du_vis_Jacobian3d(nullptr);
}

void get_duvwT(void)
{
    if (!bound(__func__)) return;
    const ws_state *s = g_bound.st;
    fail(__func__, ws_cons_to_pri(g_bound.ctx, s->f, s->d, s->u, s->v, s->w, s->T, s->P, g_bound.stream));
}

void OCFD_time_advance(int KRK)
{
    if (!bound(__func__)) return;
    ws_config c;
    ws_ctx_get_config(g_bound.ctx, &c);
    const ws_state *s = g_bound.st;
    fail(__func__, ws_time_advance(g_bound.ctx, KRK, s->f, s->fn, s->du, g_bound.f_lap, c.dt, g_bound.stream));
}

void OCFD_bc(void)
{
    if (!bound(__func__)) return;
    fail(__func__, ws_bc_planes(g_bound.ctx, g_bound.st, g_bound.stream));
}

void cuda_mem_value_init_warp(double value, double *ptr, unsigned int pitch, unsigned int size_x,
                              unsigned int size_y, unsigned int size_z)
{
    fail(__func__, ws_mem_value_init(value, ptr, pitch, size_x, size_y, size_z, g_bound.stream));
}

void exchange_boundary_xyz_packed_dev(double *hostptr, ws_field *devptr)
{
    (void)hostptr;
    if (!bound(__func__)) return;
    if (!devptr) {
        ws::set_error(WS_ERR_INVALID, "null field");
        fail(__func__, WS_ERR_INVALID);
        return;
    }
    fail(__func__, ws_ghost_fill_periodic(g_bound.ctx, devptr, 1, g_bound.stream));
}

}  // extern "C"
