// capi.cpp -- the extern "C" entry points: context, validation, dispatch, RK3 driver.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "ws_internal.h"

namespace ws {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_error(hipError_t e, const char *what)
{
    return set_error(WS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

KParams make_kparams(const ws_config &c)
{
    KParams k{};
    k.nx = c.nx; k.ny = c.ny; k.nz = c.nz;
    k.nyg = c.ny + 2 * LAP; k.nzg = c.nz + 2 * LAP;
    k.gamma = c.gamma;
    k.c1 = 2.0 * (c.gamma - 1.0);                          // OCFD_init.cpp:385
    k.c3 = (3.0 - c.gamma) / (2.0 * (c.gamma - 1.0));      // OCFD_init.cpp:386
    k.half_rcp_gamma = 0.5 / c.gamma;
    k.gm1 = c.gamma - 1.0;
    k.eps2 = c.epsl_sw * c.epsl_sw;
    k.ama = c.ama; k.rcp_ama = 1.0 / c.ama;
    k.cv = 1.0 / (c.gamma * (c.gamma - 1.0) * c.ama * c.ama);  // OCFD_init.cpp:380
    const double h[3] = {c.hx, c.hy, c.hz};
    for (int i = 0; i < 3; ++i) { k.h[i] = h[i]; k.rh[i] = 1.0 / h[i]; }
    for (int i = 0; i < 9; ++i) k.A[i] = c.metric[i];
    for (int dir = 0; dir < 3; ++dir) {
        const double a1 = c.metric[3 * dir], a2 = c.metric[3 * dir + 1], a3 = c.metric[3 * dir + 2];
        const double sn = std::sqrt(a1 * a1 + a2 * a2 + a3 * a3);
        k.Anorm[dir] = sn;
        k.Ak[3 * dir] = a1 / sn; k.Ak[3 * dir + 1] = a2 / sn; k.Ak[3 * dir + 2] = a3 / sn;
    }
    k.ajac = c.metric[9];
    // bnd: the closure form of each face (0 none, 1 D0_bound, 2 the rim kernel's high face)
    for (int i = 0; i < 6; ++i) { k.bnd[i] = c.d0_bound[i]; k.nonref[i] = c.non_ref[i] ? 1 : 0; }
    for (int d = 0; d < 3; ++d) k.bnd_any[d] = (k.bnd[2 * d] | k.bnd[2 * d + 1]) != 0;
    return k;
}

}  // namespace ws

using namespace ws;

namespace {

bool field_ok(const ws_field &f) { return f.ptr != nullptr && f.pitch > 0; }

int check_field(const ws_ctx *c, const ws_field &f, const char *name)
{
    if (!field_ok(f)) return set_error(WS_ERR_INVALID, std::string(name) + ": null field");
    if (f.pitch < (uint32_t)(c->cfg.nx + 2 * LAP))
        return set_error(WS_ERR_INVALID, std::string(name) + ": pitch smaller than nx+2*LAP");
    return WS_OK;
}

int check_soa(const ws_ctx *c, const ws_soa &s, const char *name)
{
    if (!s.ptr || s.pitch == 0) return set_error(WS_ERR_INVALID, std::string(name) + ": null state");
    if (s.pitch < (uint32_t)c->cfg.nx) return set_error(WS_ERR_INVALID, std::string(name) + ": pitch smaller than nx");
    return WS_OK;
}

int check_interior_job(const ws_ctx *c, const ws_job &j)
{
    const uint32_t n[3] = {(uint32_t)c->cfg.nx, (uint32_t)c->cfg.ny, (uint32_t)c->cfg.nz};
    const uint32_t s[3] = {j.start.x, j.start.y, j.start.z};
    const uint32_t e[3] = {j.end.x, j.end.y, j.end.z};
    for (int i = 0; i < 3; ++i)
        if (s[i] < (uint32_t)LAP || e[i] <= s[i] || e[i] > n[i] + LAP)
            return set_error(WS_ERR_INVALID, "job box must satisfy LAP <= start < end <= n+LAP");
    return WS_OK;
}

int check_metrics(const ws_ctx *c, const ws_metric_fields *m)
{
    if (!m) return WS_OK;
    for (int i = 0; i < 9; ++i) {
        int r = check_field(c, m->A[i], "metric");
        if (r) return r;
    }
    return check_field(c, m->ajac, "ajac");
}

ws_job interior_job(const ws_ctx *c)
{
    ws_job j;
    j.start = {(uint32_t)LAP, (uint32_t)LAP, (uint32_t)LAP};
    j.end = {(uint32_t)(c->cfg.nx + LAP), (uint32_t)(c->cfg.ny + LAP), (uint32_t)(c->cfg.nz + LAP)};
    return j;
}

ws_job ghosted_job(const ws_ctx *c)
{
    ws_job j;
    j.start = {0, 0, 0};
    j.end = {(uint32_t)(c->cfg.nx + 2 * LAP), (uint32_t)(c->cfg.ny + 2 * LAP), (uint32_t)(c->cfg.nz + 2 * LAP)};
    return j;
}

int validate_config(const ws_config *cfg)
{
    if (!cfg) return set_error(WS_ERR_INVALID, "null config");
    if (cfg->nx < LAP || cfg->ny < LAP || cfg->nz < LAP)
        return set_error(WS_ERR_INVALID, "every extent must be at least LAP = 4 (periodic halo)");
    if ((long)cfg->nx * cfg->ny * cfg->nz > (1L << 33))
        return set_error(WS_ERR_INVALID, "grid too large");
    const int n[3] = {cfg->nx, cfg->ny, cfg->nz};
    for (int i = 0; i < 3; ++i) {
        if (cfg->periodic[i] != 0 && cfg->periodic[i] != 1)
            return set_error(WS_ERR_INVALID, "periodic[d] must be 0 (caller's ghosts) or 1");
        for (int side = 0; side < 2; ++side) {
            const int b = cfg->d0_bound[2 * i + side], nr = cfg->non_ref[2 * i + side];
            if ((b != 0 && b != 1 && !(side == 1 && b == 2)) || (nr != 0 && nr != 1))
                return set_error(WS_ERR_INVALID, "d0_bound entries must be 0 or 1 (2 also on a high face: the rim "
                                                 "kernel's form), non_ref entries 0 or 1");
            if (b && cfg->periodic[i])
                return set_error(WS_ERR_INVALID, "a one-sided closure (d0_bound) needs a non-periodic direction");
        }
        if ((cfg->d0_bound[2 * i] || cfg->d0_bound[2 * i + 1]) && n[i] < 8)
            return set_error(WS_ERR_INVALID, "a direction with one-sided closures needs at least 8 cells");
    }
    if (!(cfg->gamma > 1.0) || !(cfg->ama > 0.0) || !(cfg->epsl_sw >= 0.0))
        return set_error(WS_ERR_INVALID, "gamma must exceed 1, Ma must be positive, epsl_sw non-negative");
    if (!(cfg->hx > 0.0) || !(cfg->hy > 0.0) || !(cfg->hz > 0.0))
        return set_error(WS_ERR_INVALID, "grid spacings must be positive");
    for (int dir = 0; dir < 3; ++dir) {
        const double *a = cfg->metric + 3 * dir;
        if (!(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] > 0.0))
            return set_error(WS_ERR_INVALID, "constant metric of a direction has zero norm");
    }
    if (cfg->precision != WS_PREC_FP64 && cfg->precision != WS_PREC_MIXED)
        return set_error(WS_ERR_INVALID, "unknown precision");
    if (cfg->kernel != WS_KERNEL_TILED && cfg->kernel != WS_KERNEL_ANCHOR)
        return set_error(WS_ERR_INVALID, "unknown kernel family");
    if (cfg->viscous != 0 && cfg->viscous != 6 && cfg->viscous != 8)
        return set_error(WS_ERR_INVALID, "viscous must be 0 (off), 6 or 8 (central-difference order)");
    if (cfg->viscous && !(cfg->re > 0.0 && cfg->pr > 0.0 && cfg->ref_t > 0.0))
        return set_error(WS_ERR_INVALID, "viscous terms need positive Re, Pr and Ref_T");
    // the viscous derivatives need a periodic direction or closures on both of its faces (the
    // flux fields get no ghost layers from the caller)
    for (int i = 0; i < 3 && cfg->viscous; ++i)
        if (!cfg->periodic[i] && !(cfg->d0_bound[2 * i] && cfg->d0_bound[2 * i + 1]))
            return set_error(WS_ERR_UNSUPPORTED,
                             "viscous terms: every direction must be periodic or closed on both faces (d0_bound)");
    return WS_OK;
}

int alloc_field(ws_field *f, int nx, int ny, int nz)
{
    const uint32_t pitch = (uint32_t)((nx + 2 * LAP + 31) / 32 * 32);   // 256-byte rows
    const size_t n = (size_t)pitch * (ny + 2 * LAP) * (nz + 2 * LAP);
    hipError_t e = hipMalloc(&f->ptr, n * sizeof(double));
    if (e != hipSuccess) { f->ptr = nullptr; return set_error(WS_ERR_NOMEM, "field allocation failed"); }
    e = hipMemset(f->ptr, 0, n * sizeof(double));
    if (e != hipSuccess) return hip_error(e, "field clear");
    f->pitch = pitch;
    f->dummy0 = 0;
    return WS_OK;
}

int alloc_soa(ws_soa *s, int nx, int ny, int nz)
{
    const uint32_t pitch = (uint32_t)((nx + 31) / 32 * 32);
    const size_t n = (size_t)pitch * ny * nz * 5;
    hipError_t e = hipMalloc(&s->ptr, n * sizeof(double));
    if (e != hipSuccess) { s->ptr = nullptr; return set_error(WS_ERR_NOMEM, "state allocation failed"); }
    e = hipMemset(s->ptr, 0, n * sizeof(double));
    if (e != hipSuccess) return hip_error(e, "state clear");
    s->pitch = pitch;
    s->length_y = (uint32_t)ny;
    s->length_z = (uint32_t)nz;
    s->dummy0 = 0;
    return WS_OK;
}

int rhs_dispatch(ws_ctx *c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                 const ws_metric_fields *m, ws_soa du, ws_job job, int acc, hipStream_t s)
{
    if (c->cfg.kernel == WS_KERNEL_ANCHOR) return launch_rhs_anchor(*c, d, u, v, w, cc, m, du, job, acc, s);
    if (c->cfg.precision == WS_PREC_MIXED) return launch_rhs_mixed(*c, d, u, v, w, cc, m, du, job, acc, s);
    return launch_rhs_tiled(*c, d, u, v, w, cc, m, du, job, acc, s);
}

}  // namespace

namespace ws {

int check_state(const ws_ctx *c, const ws_state *st)
{
    if (!c || !st) return set_error(WS_ERR_INVALID, "null argument");
    int r;
    if ((r = check_soa(c, st->f, "f")) || (r = check_soa(c, st->fn, "fn")) || (r = check_soa(c, st->du, "du")) ||
        (r = check_field(c, st->d, "d")) || (r = check_field(c, st->u, "u")) || (r = check_field(c, st->v, "v")) ||
        (r = check_field(c, st->w, "w")) || (r = check_field(c, st->T, "T")) || (r = check_field(c, st->cc, "cc")))
        return r;
    if (st->P.ptr && (r = check_field(c, st->P, "P"))) return r;
    return WS_OK;
}

int rhs_for_ctx(ws_ctx *c, const ws_state *st, hipStream_t s)
{
    return rhs_dispatch(c, st->d, st->u, st->v, st->w, st->cc, nullptr, st->du, interior_job(c), 0, s);
}

int prepare_rhs_inputs(ws_ctx *c, const ws_state *st, hipStream_t s)
{
    return launch_prepare(*c, st->d, st->u, st->v, st->w, st->T, st->cc, s);
}

int du_from_prepared(ws_ctx *c, const ws_state *st, hipStream_t s)
{
    int r = rhs_for_ctx(c, st, s);
    if (!r && c->cfg.viscous) r = launch_viscous(*c, st->u, st->v, st->w, st->T, nullptr, st->du, s);
    return r;
}

// The stage kernel of stages 1 and 2 also writes the interior sound speed, and one ghost
// fill of the primitives and cc then prepares the next stage's RHS: the same values as the
// separate ghost fill and ghosted-box sound speed on the faces the RHS reads (ghost cells
// copy interior cells, whose sound speed is computed by the same expression).  With a
// non-periodic direction the caller's ghost layers of T feed the sound speed there, so
// every stage then prepares its inputs the separate way.
//
// With the tiled kernels, the grid schedule, constant metrics, no closures and no viscous
// terms, the last directional RHS launch applies the stage to its cells itself (the stage
// value from fn, f and the complete RHS in registers, the same expression as the stage
// kernel) and the stage kernel then only derives the primitives: du never carries the
// complete RHS, which saves its write and read-back.
//
// When moreover every direction is periodic, the last launch of stages 1 and 2 also derives
// the primitives and the sound speed itself (the stage kernel's expressions) and writes them
// to the context's second primitive set: stage 1 reads the state's fields and writes the
// second set, stage 2 reads the second set and writes the state's d, u, v, w, T and cc,
// and stage 3 ends with the stage kernel (primitives, P and the boundary planes of the
// state).  That launch also writes the periodic face halo of the six fields it fills (the
// ghost fill's copies), so the state ends the step with the bits of the unfused step.  The second set is allocated by the first eager
// step; a step captured before that runs without it.
int prim_set(ws_ctx *c, hipStream_t s, ws_field out[6])
{
    if (!c->prim_scratch) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return 0;
        ws_field f{};
        if (alloc_field(&f, c->kp.nx, c->kp.ny, c->kp.nz) != WS_OK) return 0;   // one field's size
        hipFree(f.ptr);
        const size_t per = (size_t)f.pitch * c->kp.nyg * c->kp.nzg;
        hipError_t e = hipMalloc(&c->prim_scratch, 6 * per * sizeof(double));
        if (e != hipSuccess) { c->prim_scratch = nullptr; return 0; }
        e = hipMemset(c->prim_scratch, 0, 6 * per * sizeof(double));
        if (e != hipSuccess) { hipFree(c->prim_scratch); c->prim_scratch = nullptr; return 0; }
        c->prim_pitch = f.pitch;
    }
    const size_t per = (size_t)c->prim_pitch * c->kp.nyg * c->kp.nzg;
    for (int q = 0; q < 6; ++q) out[q] = ws_field{c->prim_scratch + q * per, c->prim_pitch, 0};
    return 1;
}

// cfl_next: the stage kernel that ends the step also leaves the CFL maxima of the new
// state there (launch_stage's cfl_part).
double *cfl_stage_buffer(ws_ctx *c)
{
    if (!c->cfl_stage) {
        const long n = stage_blocks(*c);
        if (hipMalloc(&c->cfl_stage, sizeof(double) * n) != hipSuccess) c->cfl_stage = nullptr;
        c->cfl_stage_n = c->cfl_stage ? n : 0;
    }
    return c->cfl_stage;
}

bool fused_stage_ok(const ws_ctx *c)
{
    const bool closures = c->kp.bnd_any[0] || c->kp.bnd_any[1] || c->kp.bnd_any[2];
    return c->fuse_stage && c->cfg.kernel == WS_KERNEL_TILED && !c->cfg.viscous && c->schedule == WS_SCHED_GRID &&
           !closures;
}

bool alternating_steps(ws_ctx *c, hipStream_t s, bool cfl)
{
    const bool periodic = c->cfg.periodic[0] && c->cfg.periodic[1] && c->cfg.periodic[2];
    ws_field second[6];
    if (c->fuse_stage < 3 || !fused_stage_ok(c) || !periodic || !prim_set(c, s, second)) return false;
    if (!cfl) return true;
    // room for either producer's count (tiles of the stage launch; stage-kernel work-groups)
    const long tiles = stage_launch_tiles(*c), blocks = stage_blocks(*c);
    const long need = tiles > blocks ? tiles : blocks;
    if (c->cfl_tiles_n < need) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
        if (c->cfl_tiles) {
            hipDeviceSynchronize();
            hipFree(c->cfl_tiles);
        }
        if (hipMalloc(&c->cfl_tiles, sizeof(double) * need) != hipSuccess) {
            c->cfl_tiles = nullptr;
            c->cfl_tiles_n = 0;
            (void)hipGetLastError();
            return false;
        }
        c->cfl_tiles_n = need;
    }
    return true;
}

// Stage fusion 3 (modes other than STEP_PLAIN): a step from set A (the state's fields or
// the second set, ghost-filled, sound speed included) runs stage 1 into B, stage 2 into A
// and stage 3 into B, each stage's last RHS launch writing the primitives and sound speed of
// its new f and their periodic face halo (and in stage 3 the CFL maxima of each tile): the
// next step starts from B with nothing to prepare, and the
// stage kernel after stage 3 and the next step's preparation are gone.  The last step of a
// run from the second set copies that set (the stage-2 primitives, ghost layers and sound
// speed) to the state's fields and lets the stage-3 launch write the primitives and P of
// the new state over their interior: what STEP_PLAIN leaves there (whose stage kernel
// writes no sound speed after stage 3 and which fills no ghosts after it).
int rk3_step_fused(ws_ctx *c, ws_state *st, const double *dt_dev, hipStream_t s, double *cfl_next, StepMode mode,
                   bool prepare)
{
    std::swap(st->f, st->fn);
    const ws_soa nolap{};
    const bool periodic = c->cfg.periodic[0] && c->cfg.periodic[1] && c->cfg.periodic[2];
    const bool fuse = fused_stage_ok(c);
    ws_field second[6];
    const bool prims = fuse && periodic && c->fuse_stage > 1 && prim_set(c, s, second);
    const ws_field g[6] = {st->d, st->u, st->v, st->w, st->T, st->cc};
    const bool planes = c->cfg.bc_planes != 0;
    int r = WS_OK;
    if (mode != STEP_PLAIN) {
        if (!prims || c->fuse_stage < 3) return set_error(WS_ERR_INVALID, "alternating RK3 step without stage fusion 3");
        const bool from_second = mode != STEP_FROM_STATE, last = mode == STEP_LAST_FROM_SECOND;
        const ws_field *A = from_second ? second : g, *B = from_second ? g : second;
        if (prepare && !from_second) r = prepare_rhs_inputs(c, st, s);
        for (int krk = 1; krk <= 3 && !r; ++krk) {
            const ws_field *in = krk == 2 ? B : A, *out = krk == 2 ? A : B;
            StageArgs sa{st->f, st->fn, dt_dev, c->cfg.dt, krk};
            sa.planes = planes;
            sa.pd = out[0]; sa.pu = out[1]; sa.pv = out[2]; sa.pw = out[3]; sa.pT = out[4];
            if (krk == 3 && last) {
                sa.pP = st->P;
                r = launch_copy_fields(*c, second, g, 6, s);
            } else {
                sa.pcc = out[5];
                sa.halo = 1;   // the launch fills the output set's periodic halo itself
                if (krk == 3) sa.cfl_part = cfl_next;
            }
            if (!r)
                r = c->cfg.precision == WS_PREC_MIXED
                        ? launch_rhs_mixed_stage(*c, in[0], in[1], in[2], in[3], in[5], st->du, interior_job(c), sa, s)
                        : launch_rhs_tiled_stage(*c, in[0], in[1], in[2], in[3], in[5], st->du, interior_job(c), sa, s);
        }
        return r;
    }
    for (int krk = 1; krk <= 3 && !r; ++krk) {
        if (krk == 1 || !periodic) r = prepare_rhs_inputs(c, st, s);
        const bool sound = periodic && krk < 3;
        if (prims) {
            // inputs and outputs of this stage's RHS: the state's fields or the second set
            const ws_field *in = krk == 2 ? second : g;
            const ws_field *out = krk == 1 ? second : g;
            StageArgs sa{st->f, st->fn, dt_dev, c->cfg.dt, krk};
            if (krk < 3) {
                sa.planes = planes;
                sa.pd = out[0]; sa.pu = out[1]; sa.pv = out[2]; sa.pw = out[3]; sa.pT = out[4]; sa.pcc = out[5];
                sa.halo = 1;   // the launch fills the output set's periodic halo itself
            }
            if (!r)
                r = c->cfg.precision == WS_PREC_MIXED
                        ? launch_rhs_mixed_stage(*c, in[0], in[1], in[2], in[3], in[5], st->du, interior_job(c), sa, s)
                        : launch_rhs_tiled_stage(*c, in[0], in[1], in[2], in[3], in[5], st->du, interior_job(c), sa, s);
            if (!r && krk == 3) r = launch_stage(*c, 0, *st, nolap, c->cfg.dt, dt_dev, planes, false, s, cfl_next);
            continue;
        }
        if (fuse) {
            const StageArgs sa{st->f, st->fn, dt_dev, c->cfg.dt, krk};
            if (!r)
                r = c->cfg.precision == WS_PREC_MIXED
                        ? launch_rhs_mixed_stage(*c, st->d, st->u, st->v, st->w, st->cc, st->du, interior_job(c), sa, s)
                        : launch_rhs_tiled_stage(*c, st->d, st->u, st->v, st->w, st->cc, st->du, interior_job(c), sa, s);
            if (!r) r = launch_stage(*c, 0, *st, nolap, c->cfg.dt, dt_dev, planes, sound, s, krk == 3 ? cfl_next : nullptr);
        } else {
            if (!r) r = du_from_prepared(c, st, s);
            if (!r) r = launch_stage(*c, krk, *st, nolap, c->cfg.dt, dt_dev, planes, sound, s, krk == 3 ? cfl_next : nullptr);
        }
        if (!r && sound) r = launch_ghost_fill(*c, g, 6, s);
    }
    return r;
}

}  // namespace ws

extern "C" {

const char *ws_version(void) { return "windstencil_amd 0.1 (gfx950)"; }

const char *ws_error_string(int code)
{
    switch (code) {
    case WS_OK: return "ok";
    case WS_ERR_INVALID: return "invalid argument";
    case WS_ERR_UNSUPPORTED: return "unsupported";
    case WS_ERR_HIP: return "HIP runtime error";
    case WS_ERR_NOMEM: return "out of device memory";
    default: return "unknown error";
    }
}

const char *ws_last_error(void) { return g_last_error.c_str(); }

void ws_config_default(ws_config *cfg, int32_t nx, int32_t ny, int32_t nz)
{
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->nx = nx; cfg->ny = ny; cfg->nz = nz;
    cfg->periodic[0] = cfg->periodic[1] = cfg->periodic[2] = 1;
    cfg->gamma = 1.4;
    cfg->ama = 1.0;
    cfg->epsl_sw = 1e-3;
    cfg->dt = 1e-4;
    cfg->hx = 1.0 / (nx - 1.0);
    cfg->hy = 1.0 / (ny - 1.0);
    cfg->hz = 1.0 / (nz - 1.0);
    // Cartesian metrics: Akx = Aiy = Asz = Ajac = 1
    cfg->metric[0] = 1.0; cfg->metric[4] = 1.0; cfg->metric[8] = 1.0; cfg->metric[9] = 1.0;
    cfg->precision = WS_PREC_FP64;
    cfg->kernel = WS_KERNEL_TILED;
    cfg->bc_planes = 1;
    cfg->viscous = 0;
    cfg->re = 1000.0;
    cfg->pr = 0.72;
    cfg->ref_t = 288.15;
}

int ws_ctx_create(const ws_config *cfg, ws_ctx **out)
{
    if (!out) return set_error(WS_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int r = validate_config(cfg);
    if (r) return r;
    ws_ctx *c = new (std::nothrow) ws_ctx();
    if (!c) return set_error(WS_ERR_NOMEM, "host allocation failed");
    c->cfg = *cfg;
    c->kp = make_kparams(*cfg);
    int shape, sched, order, stage_order, stage_shape;
    if (remembered_choice(cfg->nx, cfg->ny, cfg->nz, cfg->precision, &shape, &sched, &order, &stage_order,
                          &stage_shape)) {
        c->tile_shape = shape;
        c->schedule = sched;
        c->dir_order = order;
        c->stage_order = stage_order;
        c->stage_shape = stage_shape;
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) { delete c; return hip_error(e, "hipGetDevice"); }
    c->cfl_blocks = 1024;
    e = hipMalloc(&c->cfl_partial, sizeof(double) * c->cfl_blocks);
    if (e != hipSuccess) { delete c; return hip_error(e, "CFL scratch allocation"); }
    e = hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { hipFree(c->cfl_partial); delete c; return hip_error(e, "capture stream"); }
    *out = c;
    return WS_OK;
}

int ws_ctx_destroy(ws_ctx *c)
{
    if (!c) return WS_OK;
    if (c->capture_stream) hipStreamDestroy(c->capture_stream);
    if (c->cfl_partial) hipFree(c->cfl_partial);
    if (c->vis_scratch || c->prim_scratch || c->cfl_stage || c->cfl_tiles || c->anchor_scratch)
        hipDeviceSynchronize();
    if (c->vis_scratch) hipFree(c->vis_scratch);
    if (c->cfl_tiles) hipFree(c->cfl_tiles);
    if (c->anchor_scratch) hipFree(c->anchor_scratch);
    if (c->prim_scratch) hipFree(c->prim_scratch);
    if (c->cfl_stage) hipFree(c->cfl_stage);
    delete c;
    return WS_OK;
}

int ws_ctx_get_config(const ws_ctx *c, ws_config *out)
{
    if (!c || !out) return set_error(WS_ERR_INVALID, "null argument");
    *out = c->cfg;
    return WS_OK;
}

int ws_ctx_set_precision(ws_ctx *c, int32_t precision)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (precision != WS_PREC_FP64 && precision != WS_PREC_MIXED) return set_error(WS_ERR_INVALID, "unknown precision");
    c->cfg.precision = precision;
    int shape, sched, order, stage_order, stage_shape;
    if (remembered_choice(c->cfg.nx, c->cfg.ny, c->cfg.nz, precision, &shape, &sched, &order, &stage_order,
                          &stage_shape)) {
        c->tile_shape = shape;
        c->schedule = sched;
        c->dir_order = order;
        c->stage_order = stage_order;
        c->stage_shape = stage_shape;
    }
    return WS_OK;
}

int ws_ctx_set_kernel(ws_ctx *c, int32_t kernel)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (kernel != WS_KERNEL_TILED && kernel != WS_KERNEL_ANCHOR) return set_error(WS_ERR_INVALID, "unknown kernel");
    c->cfg.kernel = kernel;
    return WS_OK;
}

int ws_ctx_set_tile_shape(ws_ctx *c, int32_t shape)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (shape < 0 || shape >= WS_TILE_COUNT) return set_error(WS_ERR_INVALID, "unknown tile shape");
    c->tile_shape = shape;
    return WS_OK;
}

int ws_ctx_set_schedule(ws_ctx *c, int32_t schedule)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (schedule < 0 || schedule >= WS_SCHED_COUNT) return set_error(WS_ERR_INVALID, "unknown schedule");
    c->schedule = schedule;
    return WS_OK;
}

int ws_ctx_get_schedule(const ws_ctx *c, int32_t *schedule)
{
    if (!c || !schedule) return set_error(WS_ERR_INVALID, "null argument");
    *schedule = c->schedule;
    return WS_OK;
}

int ws_ctx_set_direction_order(ws_ctx *c, int32_t order)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (order < 0 || order >= WS_ORDER_COUNT) return set_error(WS_ERR_INVALID, "unknown direction order");
    c->dir_order = order;
    return WS_OK;
}

int ws_ctx_get_direction_order(const ws_ctx *c, int32_t *order)
{
    if (!c || !order) return set_error(WS_ERR_INVALID, "null argument");
    *order = c->dir_order;
    return WS_OK;
}

int ws_ctx_set_stage_order(ws_ctx *c, int32_t order)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (order < 0 || order >= WS_ORDER_COUNT) return set_error(WS_ERR_INVALID, "unknown direction order");
    c->stage_order = order;
    return WS_OK;
}

int ws_ctx_get_stage_order(const ws_ctx *c, int32_t *order)
{
    if (!c || !order) return set_error(WS_ERR_INVALID, "null argument");
    *order = c->stage_order;
    return WS_OK;
}

int ws_ctx_set_stage_shape(ws_ctx *c, int32_t shape)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (shape < -1 || shape >= WS_TILE_COUNT) return set_error(WS_ERR_INVALID, "unknown tile shape");
    c->stage_shape = shape;
    return WS_OK;
}

int ws_ctx_get_stage_shape(const ws_ctx *c, int32_t *shape)
{
    if (!c || !shape) return set_error(WS_ERR_INVALID, "null argument");
    *shape = c->stage_shape;
    return WS_OK;
}

int ws_ctx_set_stage_fusion(ws_ctx *c, int32_t on)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (on < 0 || on > 3) return set_error(WS_ERR_INVALID, "stage fusion must be 0, 1, 2 or 3");
    c->fuse_stage = on;
    return WS_OK;
}

int ws_ctx_get_tile_shape(const ws_ctx *c, int32_t *shape)
{
    if (!c || !shape) return set_error(WS_ERR_INVALID, "null argument");
    *shape = c->tile_shape;
    return WS_OK;
}

int ws_state_alloc(ws_ctx *c, ws_state *st)
{
    if (!c || !st) return set_error(WS_ERR_INVALID, "null argument");
    std::memset(st, 0, sizeof(*st));
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    int r = WS_OK;
    ws_soa *soas[3] = {&st->f, &st->fn, &st->du};
    for (ws_soa *s : soas)
        if (!r) r = alloc_soa(s, nx, ny, nz);
    ws_field *fields[7] = {&st->d, &st->u, &st->v, &st->w, &st->T, &st->P, &st->cc};
    for (ws_field *f : fields)
        if (!r) r = alloc_field(f, nx, ny, nz);
    if (r) ws_state_free(c, st);
    return r;
}

int ws_state_free(ws_ctx *c, ws_state *st)
{
    (void)c;
    if (!st) return WS_OK;
    ws_soa *soas[3] = {&st->f, &st->fn, &st->du};
    for (ws_soa *s : soas)
        if (s->ptr) { hipFree(s->ptr); s->ptr = nullptr; }
    ws_field *fields[7] = {&st->d, &st->u, &st->v, &st->w, &st->T, &st->P, &st->cc};
    for (ws_field *f : fields)
        if (f->ptr) { hipFree(f->ptr); f->ptr = nullptr; }
    return WS_OK;
}

int ws_inviscid_rhs(ws_ctx *c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                    const ws_metric_fields *metrics, ws_soa du, ws_job job, int32_t accumulate, hipStream_t s)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    int r;
    if ((r = check_field(c, d, "d")) || (r = check_field(c, u, "u")) || (r = check_field(c, v, "v")) ||
        (r = check_field(c, w, "w")) || (r = check_field(c, cc, "cc")) || (r = check_soa(c, du, "du")) ||
        (r = check_metrics(c, metrics)) || (r = check_interior_job(c, job)))
        return r;
    return rhs_dispatch(c, d, u, v, w, cc, metrics, du, job, accumulate ? 1 : 0, s);
}

int ws_sound_speed(ws_ctx *c, ws_field T, ws_field cc, ws_job job, hipStream_t s)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    int r;
    if ((r = check_field(c, T, "T")) || (r = check_field(c, cc, "cc"))) return r;
    const uint32_t g[3] = {(uint32_t)c->cfg.nx + 2 * LAP, (uint32_t)c->cfg.ny + 2 * LAP, (uint32_t)c->cfg.nz + 2 * LAP};
    if (job.end.x > g[0] || job.end.y > g[1] || job.end.z > g[2] || job.start.x >= job.end.x ||
        job.start.y >= job.end.y || job.start.z >= job.end.z)
        return set_error(WS_ERR_INVALID, "sound speed box outside the ghosted field");
    return launch_sound_speed(*c, T, cc, job, s);
}

int ws_ghost_fill_periodic(ws_ctx *c, const ws_field *fields, int32_t n, hipStream_t s)
{
    if (!c || !fields || n <= 0) return set_error(WS_ERR_INVALID, "bad field list");
    for (int i = 0; i < n; ++i) {
        int r = check_field(c, fields[i], "ghost fill");
        if (r) return r;
    }
    return launch_ghost_fill(*c, fields, n, s);
}

static int check_prims(const ws_ctx *c, const ws_state *st)
{
    int r;
    if ((r = check_field(c, st->d, "d")) || (r = check_field(c, st->u, "u")) || (r = check_field(c, st->v, "v")) ||
        (r = check_field(c, st->w, "w")) || (r = check_field(c, st->T, "T")) || (r = check_field(c, st->cc, "cc")))
        return r;
    return WS_OK;
}

int ws_prepare_rhs_inputs(ws_ctx *c, const ws_state *st, hipStream_t s)
{
    if (!c || !st) return set_error(WS_ERR_INVALID, "null argument");
    int r = check_prims(c, st);
    return r ? r : launch_prepare(*c, st->d, st->u, st->v, st->w, st->T, st->cc, s);
}

int ws_du_comput(ws_ctx *c, const ws_state *st, const ws_metric_fields *metrics, hipStream_t s)
{
    if (!c || !st) return set_error(WS_ERR_INVALID, "null argument");
    int r;
    if ((r = ws_prepare_rhs_inputs(c, st, s))) return r;
    if ((r = ws_inviscid_rhs(c, st->d, st->u, st->v, st->w, st->cc, metrics, st->du, interior_job(c), 0, s))) return r;
    if (c->cfg.viscous) return launch_viscous(*c, st->u, st->v, st->w, st->T, metrics, st->du, s);
    return WS_OK;
}

int ws_viscous_rhs(ws_ctx *c, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields *metrics,
                   ws_soa du, hipStream_t s)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    int r;
    if ((r = check_field(c, u, "u")) || (r = check_field(c, v, "v")) || (r = check_field(c, w, "w")) ||
        (r = check_field(c, T, "T")) || (r = check_soa(c, du, "du")) || (r = check_metrics(c, metrics)))
        return r;
    return launch_viscous(*c, u, v, w, T, metrics, du, s);
}

int ws_cons_to_pri(ws_ctx *c, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w, ws_field T, ws_field P,
                   hipStream_t s)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    int r;
    if ((r = check_soa(c, f, "f")) || (r = check_field(c, d, "d")) || (r = check_field(c, u, "u")) ||
        (r = check_field(c, v, "v")) || (r = check_field(c, w, "w")) || (r = check_field(c, T, "T")))
        return r;
    if (P.ptr && (r = check_field(c, P, "P"))) return r;
    return launch_cons_to_pri(*c, f, d, u, v, w, T, P, s);
}

int ws_pri_to_cons(ws_ctx *c, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w, ws_field T, ws_job job,
                   hipStream_t s)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    int r;
    if ((r = check_soa(c, f, "f")) || (r = check_field(c, d, "d")) || (r = check_field(c, u, "u")) ||
        (r = check_field(c, v, "v")) || (r = check_field(c, w, "w")) || (r = check_field(c, T, "T")))
        return r;
    if (job.end.x > (uint32_t)c->cfg.nx || job.end.y > (uint32_t)c->cfg.ny || job.end.z > (uint32_t)c->cfg.nz ||
        job.start.x >= job.end.x || job.start.y >= job.end.y || job.start.z >= job.end.z)
        return set_error(WS_ERR_INVALID, "pri_to_cons box outside the interior");
    return launch_pri_to_cons(*c, f, d, u, v, w, T, job, s);
}

int ws_bc_planes(ws_ctx *c, const ws_state *st, hipStream_t s)
{
    if (!c || !st) return set_error(WS_ERR_INVALID, "null argument");
    const uint32_t nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    // x = 0, x = nx-1, y = 0, y = ny-1, z = 0, z = nz-1 (OCFD_boundary.cpp:38-89)
    const ws_job planes[6] = {
        {{0, 0, 0}, {1, ny, nz}},       {{nx - 1, 0, 0}, {nx, ny, nz}},
        {{0, 0, 0}, {nx, 1, nz}},       {{0, ny - 1, 0}, {nx, ny, nz}},
        {{0, 0, 0}, {nx, ny, 1}},       {{0, 0, nz - 1}, {nx, ny, nz}},
    };
    for (const ws_job &j : planes) {
        int r = launch_pri_to_cons(*c, st->f, st->d, st->u, st->v, st->w, st->T, j, s);
        if (r) return r;
    }
    return WS_OK;
}

int ws_time_advance(ws_ctx *c, int32_t krk, ws_soa f, ws_soa fn, ws_soa du, ws_soa f_lap, double dt, hipStream_t s)
{
    if (!c) return set_error(WS_ERR_INVALID, "null context");
    if (krk < 1 || krk > 3) return set_error(WS_ERR_INVALID, "krk must be 1, 2 or 3");
    int r;
    if ((r = check_soa(c, f, "f")) || (r = check_soa(c, fn, "fn")) || (r = check_soa(c, du, "du"))) return r;
    // f_lap is indexed over the ghosted box (x + LAP + pitch * (...)): a pitch below the
    // ghosted row would write into the next row
    if (f_lap.ptr && f_lap.pitch < (uint32_t)(c->cfg.nx + 2 * LAP))
        return set_error(WS_ERR_INVALID, "f_lap: pitch smaller than nx+2*LAP");
    return launch_time_advance(*c, krk, f, fn, du, f_lap, dt, nullptr, s);
}

int ws_mem_value_init(double value, double *ptr, uint32_t pitch, uint32_t sx, uint32_t sy, uint32_t sz,
                      hipStream_t s)
{
    if (!ptr || pitch < sx) return set_error(WS_ERR_INVALID, "bad fill target");
    if (sx == 0 || sy == 0 || sz == 0) return WS_OK;
    return launch_fill(value, ptr, pitch, sx, sy, sz, s);
}

int ws_cfl_dt(ws_ctx *c, const ws_state *st, double cfl, double *dt_dev, hipStream_t s)
{
    if (!c || !st || !dt_dev || !(cfl > 0.0)) return set_error(WS_ERR_INVALID, "bad CFL request");
    return launch_cfl(*c, *st, cfl, dt_dev, s);
}

int ws_rk3_step(ws_ctx *c, ws_state *st, const double *dt_dev, hipStream_t s)
{
    if (!c || !st) return set_error(WS_ERR_INVALID, "null argument");
    // pf <-> pfn (OCFD_NS_Solver.cpp:66-74): fn now holds the state at t^n; then per stage
    // du_comput, OCFD_time_advance, get_duvwT and OCFD_bc, fused
    return rk3_step_fused(c, st, dt_dev, s);
}

int ws_rk3_run(ws_ctx *c, ws_state *st, int32_t nsteps, double cfl, int32_t use_graph, hipStream_t s)
{
    if (!c || !st || nsteps < 0) return set_error(WS_ERR_INVALID, "bad RK3 request");
    double *dt_dev = nullptr;
    if (cfl > 0.0) {
        hipError_t e = hipMalloc(&dt_dev, sizeof(double));
        if (e != hipSuccess) return hip_error(e, "dt allocation");
    }
    // With a CFL step, the stage kernel that ends each step leaves the maxima of the new
    // state in c->cfl_stage, and the next step's dt is their reduction (k_cfl_final):
    // the same maximum as a k_cfl_partial pass over the state, without that pass.  With
    // stage fusion 3 the steps alternate the primitive sets (rk3_step_fused), the
    // stage-applying launch of each step but the last leaves the maxima per tile in
    // c->cfl_tiles, and the last step ends in the state's fields.
    const bool alt = nsteps > 1 && alternating_steps(c, s, dt_dev != nullptr);
    double *cfl_next = !dt_dev ? nullptr : alt ? c->cfl_tiles : cfl_stage_buffer(c);
    const int cfl_n = alt ? (int)stage_launch_tiles(*c) : (int)c->cfl_stage_n;
    auto mode_of = [&](int n) {
        if (!alt) return STEP_PLAIN;
        if (n == nsteps - 1) return (n & 1) ? STEP_LAST_FROM_SECOND : STEP_PLAIN;
        return (n & 1) ? STEP_FROM_SECOND : STEP_FROM_STATE;
    };
    // step n; the first takes its dt from a pass over the state
    auto step = [&](ws_state *S, int n, hipStream_t ss) {
        int q = WS_OK;
        if (dt_dev)
            q = (n == 0 || !cfl_next) ? ws_cfl_dt(c, S, cfl, dt_dev, ss)
                                      : launch_cfl_final(*c, cfl_next, cfl_n, cfl, dt_dev, ss);
        // the last step's maxima would never be read (and a plain last step's stage kernel
        // writes one per 256 cells, more than the per-tile buffer of an alternating run holds)
        if (!q) q = rk3_step_fused(c, S, dt_dev, ss, n == nsteps - 1 ? nullptr : cfl_next, mode_of(n), n == 0);
        return q;
    };
    int r = WS_OK;
    if (!use_graph) {
        for (int n = 0; n < nsteps && !r; ++n) r = step(st, n, s);
    } else {
        // Step 0 runs eagerly (it performs any lazy allocation); later steps replay one
        // captured graph per parity of the f/fn swap, since a step ends with f and fn
        // exchanged (and, alternating, with the primitives in the other set).  An
        // alternating run's last step, which ends differently, runs eagerly too.
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        const int replayed = alt ? nsteps - 1 : nsteps;
        for (int n = 0; n < replayed && !r; ++n) {
            if (n == 0) {
                r = step(st, 0, s);
                continue;
            }
            const int p = n & 1;
            if (!exec[p]) {
                ws_state probe = *st;
                hipGraph_t g = nullptr;
                hipStream_t cs = c->capture_stream;
                hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) { r = hip_error(e, "begin capture"); break; }
                r = step(&probe, n, cs);
                e = hipStreamEndCapture(cs, &g);
                if (r) { if (g) hipGraphDestroy(g); break; }
                if (e != hipSuccess) { r = hip_error(e, "end capture"); break; }
                e = hipGraphInstantiate(&exec[p], g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (e != hipSuccess) { r = hip_error(e, "graph instantiate"); break; }
            }
            hipError_t e = hipGraphLaunch(exec[p], s);
            if (e != hipSuccess) { r = hip_error(e, "graph launch"); break; }
            std::swap(st->f, st->fn);
        }
        if (!r && alt) r = step(st, nsteps - 1, s);
        for (int p = 0; p < 2; ++p)
            if (exec[p]) {
                hipStreamSynchronize(s);
                hipGraphExecDestroy(exec[p]);
            }
    }
    hipError_t e = hipStreamSynchronize(s);
    if (!r && e != hipSuccess) r = hip_error(e, "RK3 run");
    if (dt_dev) hipFree(dt_dev);
    return r;
}

}  // extern "C"
