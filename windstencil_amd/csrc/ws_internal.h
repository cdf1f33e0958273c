// ws_internal.h -- internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "windstencil_amd.h"

namespace ws {

constexpr int LAP = WS_LAP;

// Everything a kernel needs about the grid and the physics, passed by value.
struct KParams {
    int nx, ny, nz;            // interior extents
    int nyg, nzg;              // ghosted extents in y and z (n + 2*LAP)
    double gamma;
    double c1;                 // 2(gamma-1)
    double c3;                 // (3-gamma)/(2(gamma-1))
    double half_rcp_gamma;     // 0.5/gamma
    double gm1;                // gamma-1
    double eps2;               // epsl_sw^2
    double ama, rcp_ama;
    double cv;                 // 1/(gamma (gamma-1) Ma^2)
    double h[3];               // grid spacings
    double rh[3];              // 1/h
    // constant metrics: direction d uses A[3d..3d+2]; their norm and unit vector
    double A[9];
    double Anorm[3];
    double Ak[9];
    double ajac;
    // one-sided closures of the inviscid scheme at the domain faces (x-, x+, y-, y+, z-, z+:
    // 0 none, 1 the reference's D0_bound, 2 on a high face the rim kernel's form) and the
    // non-reflecting flags of the P (even index) and M (odd) passes of each direction
    // (Non_ref); bnd_any[d]: either side of d closed
    int bnd[6];
    int nonref[6];
    int bnd_any[3];
    // set per launch: the RHS inputs d, u, v, w, cc share d's pitch and every offset into
    // them is below 2^31 with pitch and nyg*nzg below 2^24, so one 32-bit offset with
    // 24-bit products addresses all five (prim_off32)
    int off32 = 0;
};

// whether the five RHS input fields of a box of these extents can be addressed with one
// 32-bit offset (KParams::off32)
inline bool prim_off32(const KParams &k, const ws_field &d, const ws_field &u, const ws_field &v,
                       const ws_field &w, const ws_field &cc)
{
    const unsigned long long p = d.pitch, plane = (unsigned long long)k.nyg * (unsigned long long)k.nzg;
    return u.pitch == d.pitch && v.pitch == d.pitch && w.pitch == d.pitch && cc.pitch == d.pitch &&
           p < (1ull << 24) && plane < (1ull << 24) && p * plane < (1ull << 31);
}

// The SSP-RK3 stage the last directional RHS launch applies to its cells itself (f from fn,
// f and the complete RHS, which then never goes through du): stage krk (1..3), dt from
// dt_dev when that is set.  With pd.ptr set the launch also writes the primitives d, u, v,
// w, T of the new f to the interior cells of pd, pu, pv, pw, pT (fields that no launch of
// the same RHS reads), the sound speed to pcc and the pressure to pP where those are set,
// and, with planes, re-derives f on the boundary planes from them.  With cfl_part set
// (and pd) each tile of the launch also leaves the maximum over its cells of the CFL
// speed max(|u|, |v|, |w|) + c of the new state at cfl_part[tile] (stage_launch_tiles
// entries; the expression of k_cfl_partial).
struct StageArgs {
    ws_soa f, fn;
    const double *dt_dev;
    double dt;
    int krk;
    int planes = 0;
    ws_field pd{}, pu{}, pv{}, pw{}, pT{}, pcc{}, pP{};
    double *cfl_part = nullptr;
    // with pd: also write the periodic face halo of pd..pT (and pcc) -- the copies the
    // ghost fill of those fields would make from the interior cells written here (every
    // direction periodic)
    int halo = 0;
};


struct Ctx {
    ws_config cfg;
    KParams kp;
    int device;
    // CFL scratch: per-block maxima
    double *cfl_partial = nullptr;
    int cfl_blocks = 0;
    // private stream for graph capture (the caller's stream may be the legacy
    // default stream, which cannot be captured); default priority
    hipStream_t capture_stream = nullptr;
    // tile shape of the tiled kernels, FP64 and mixed (WS_TILE_*)
    int tile_shape = WS_TILE_30x16_K2;
    // how the tiled kernels walk their tiles (WS_SCHED_*)
    int schedule = WS_SCHED_GRID;
    // order of the directional launches (WS_ORDER_*); stage_order: that of the launches
    // that apply an RK3 stage (the stage kernel's share moves into the last of them, so
    // its best order can differ from the plain RHS's)
    int dir_order = WS_ORDER_XYZ;
    int stage_order = WS_ORDER_XYZ;
    // tile shape of those stage-applying launches (-1: the context's tile_shape)
    int stage_shape = -1;
    // RK3 steps apply each stage's update in the last RHS launch where they can (1), there
    // also derive the primitives of stages 1 and 2 (2), and within ws_rk3_run those of
    // stage 3 too, alternating the primitive sets from step to step (3)
    int fuse_stage = 3;
    // viscous terms of the two-kernel form (anchor family, metric fields): twelve ghosted
    // flux fields, allocated at the first use; planes per work-group of the fused form
    double *vis_scratch = nullptr;
    size_t vis_scratch_elems = 0;
    int vis_zchunk = 64;
    // second set of ghosted primitive fields (d, u, v, w, T, cc) that RK3 stages 1 and 2
    // write their primitives to, allocated at the first eager RK3 step that can use it
    double *prim_scratch = nullptr;
    uint32_t prim_pitch = 0;
    // CFL maxima the last stage of an ws_rk3_run step leaves for the next step's dt
    // (one per stage-kernel work-group), allocated by the first eager step
    double *cfl_stage = nullptr;
    long cfl_stage_n = 0;
    // the same maxima per tile of the stage-applying RHS launch (stage fusion 3)
    double *cfl_tiles = nullptr;
    long cfl_tiles_n = 0;
    // split fluxes of the anchor kernels (10 doubles per point of the largest directional
    // sweep of the interior), allocated by the context's first anchor RHS
    double *anchor_scratch = nullptr;
    size_t anchor_scratch_bytes = 0;
};

KParams make_kparams(const ws_config &cfg);

// One block of a batch, as the batched RHS kernels see it.
struct BatchBlock {
    ws_field d, u, v, w, cc;
    ws_soa du;
    ws_soa fa, fb;       // the block's f and fn as created (RK3 steps exchange them)
    // where a stage-applying launch with BatchStage::prims writes the new primitives
    // (d, u, v, w, T, cc): the other primitive set of the block
    ws_field od, ou, ov, ow, oT, occ;
    ws_field P;          // the block's pressure (null: not kept)
    int n[3];            // interior extents
    int pad;
    double rh[3];        // 1/h per direction
};

// The RK3 stage the last batched RHS launch applies: every block's f is fb and fn fa when
// `swapped`, else f = fa and fn = fb.  prims: the launch also writes the new primitives to
// each block's od..oT, with cc the sound speed to occ, with pressure P to the block's P
// (and, with planes, re-derives f on the boundary planes); cfl_part: each tile leaves the
// maximum CFL speed of its cells at cfl_part[tile] (its index in the launch's tile list).
struct BatchStage {
    const double *dt_dev;
    double dt;
    int krk;
    int swapped;
    int prims = 0;
    int planes = 0;
    int cc = 1;
    int pressure = 0;
    double *cfl_part = nullptr;
    int halo = 0;        // as StageArgs::halo, for every block
};

// Device tile lists of a batch for one tile shape: entry (block, bx, by, bz) per tile,
// one list per direction.
struct BatchTiles {
    int shape;
    int schedule;        // WS_BATCH_*
    const BatchBlock *blocks;
    const int4 *tiles[3];
    int ntiles[3];
    const BatchStage *stage;   // non-null: the last launch applies this RK3 stage (grid schedule)
    int stage_order;           // with stage: WS_ORDER_XYZ (z applies it) or WS_ORDER_ZYX (x does)
};

// One block of a batch for the batched helper kernels (ghost fill, sound speed, the stage
// kernel and the CFL step over every block in one launch).  A launch gives each block a
// contiguous range of work-groups: [ghost_wg0, ghost_wg1) in the ghost-fill grid, one
// element of a face slab per thread, and [cell_wg0, cell_wg1) in the cell grid, 256
// interior cells per work-group in the order of the single-block kernels, so that the
// blocks' work-groups see the cells they see there.
struct AuxBlock {
    ws_field g[2][6];    // fields of a ghost fill: the state's d, u, v, w, T, cc; the second set
    ws_field P;          // the state's pressure (null: not kept)
    ws_soa fa, fb;       // the block's f and fn as created
    int n[3];
    int nyg;
    double hmin;         // min(hx, hy, hz)
    long ghost_wg0, ghost_wg1;
    long cell_wg0, cell_wg1;
};

// Batched helper launches over a device table of nb blocks (k_aux.hip).  ghost_wgs and
// cell_wgs: the table's last ghost_wg1 and cell_wg1.
// Periodic ghost fill of field set `set` (0 or 1) of every block.
int launch_ghost_fill_batch(const AuxBlock *tab, int nb, long ghost_wgs, int set, hipStream_t s);
// cc = sqrt(T) / Ma on every block's interior (set 0), then the ghost fill of set 0: the
// inputs of a batched RHS, as prepare_rhs_inputs leaves them on the cells the RHS reads.
int launch_prepare_batch(const KParams &k, const AuxBlock *tab, int nb, long ghost_wgs, long cell_wgs,
                         hipStream_t s);
// The stage kernel with krk 0 over every block (the last RHS launch already applied the
// stage to f = fb when swapped, else fa): primitives of set 0 and P, with planes the
// boundary planes, with sound the interior sound speed; cfl_part (cell_wgs entries): the
// CFL maximum of each work-group.
int launch_stage_batch(const KParams &k, const AuxBlock *tab, int nb, long cell_wgs, int swapped, bool planes,
                       bool sound, double *cfl_part, hipStream_t s);
// dts[b] = cfl * hmin / (maximum of block b's CFL partials), then *dt_dev = min over b.
// Block b's partials are cfl_part[ranges[b].x .. ranges[b].y) when ranges is set (the tiles
// of a stage-applying launch), else those of its stage-kernel work-groups.
int launch_cfl_batch(const AuxBlock *tab, int nb, const double *cfl_part, double cfl, double *dts, double *dt_dev,
                     hipStream_t s, const int2 *ranges = nullptr);

// (tile shape, schedule, direction order, stage order and shape) the autotuner found best for these
// extents and precision; false if none
bool remembered_choice(int nx, int ny, int nz, int precision, int *shape, int *schedule, int *order,
                       int *stage_order, int *stage_shape);

// error reporting (thread-local message)
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);

// ---- launchers (defined next to their kernels) ----
int launch_rhs_anchor(Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                      const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s);
int launch_rhs_tiled(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                     const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s);
int launch_rhs_mixed(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                     const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s);
// tiles of the context's launch that applies an RK3 stage (the last direction of its order,
// in its tile shape): the entries that launch writes to StageArgs::cfl_part
long stage_launch_tiles(const Ctx &c);
// the RHS with the RK3 stage `sa` applied by its last launch (grid schedule; the caller
// checks that the configuration allows it: tiled kernels, constant metrics, no closures,
// no viscous terms); du holds partial sums afterwards
int launch_rhs_tiled_stage(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du,
                           ws_job job, const StageArgs &sa, hipStream_t s);
int launch_rhs_mixed_stage(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du,
                           ws_job job, const StageArgs &sa, hipStream_t s);

int launch_rhs_batch_tiled(const KParams &k, const BatchTiles &bt, hipStream_t s);
int launch_rhs_batch_mixed(const KParams &k, const BatchTiles &bt, hipStream_t s);

// library-internal helpers of the C API (capi.cpp)
int check_state(const ws_ctx *c, const ws_state *st);
// ghost fill of d, u, v, w, T and sound speed over the ghosted box
int prepare_rhs_inputs(ws_ctx *c, const ws_state *st, hipStream_t s);
// the RHS over the interior with the context's kernel and precision (overwrites du)
int rhs_for_ctx(ws_ctx *c, const ws_state *st, hipStream_t s);
// du_comput on inputs already prepared (ghosted primitives and sound speed): the inviscid
// RHS, plus the viscous terms when the configuration has them
int du_from_prepared(ws_ctx *c, const ws_state *st, hipStream_t s);
// whether RK3 steps can let the last RHS launch apply the stage (stage fusion on, tiled
// kernels, grid schedule, no closures, no viscous terms)
bool fused_stage_ok(const ws_ctx *c);
// How a step of stage fusion 3 treats the primitive sets (rk3_step_fused): STEP_PLAIN starts
// and ends with the primitives in the state's fields (every level); STEP_FROM_STATE and
// STEP_FROM_SECOND start from the state's fields or the context's second set, prepared, and
// end in the other one, prepared for the next step; STEP_LAST_FROM_SECOND starts from the
// second set and ends in the state's fields as STEP_PLAIN leaves them.
enum StepMode { STEP_PLAIN = 0, STEP_FROM_STATE, STEP_FROM_SECOND, STEP_LAST_FROM_SECOND };
// one SSP-RK3 step of st with fused stage kernels (the same arithmetic as du_comput,
// time_advance, cons_to_pri and bc_planes in turn).  cfl_next: where the step leaves the
// CFL maxima of the new state (cfl_stage for STEP_PLAIN, cfl_tiles for the other modes).
// prepare: STEP_FROM_STATE prepares the state's fields first (the first step of a run).
int rk3_step_fused(ws_ctx *c, ws_state *st, const double *dt_dev, hipStream_t s, double *cfl_next = nullptr,
                   StepMode mode = STEP_PLAIN, bool prepare = true);
// whether ws_rk3_run can alternate the primitive sets (stage fusion 3 and every condition of
// the primitives-deriving fusion), allocating what that needs (outside a stream capture)
bool alternating_steps(ws_ctx *c, hipStream_t s, bool cfl);
// the context's second set of ghosted primitive fields (d, u, v, w, T, cc), allocated at
// the first call outside a stream capture; 0 if it is not available
int prim_set(ws_ctx *c, hipStream_t s, ws_field out[6]);
// the context's buffer for the CFL maxima of the stage kernel that ends a step
// (stage_blocks entries), allocated at the first call (outside a stream capture); null if
// it cannot be allocated
double *cfl_stage_buffer(ws_ctx *c);

// viscous terms added to du[1..4] (k_viscous.hip); allocates the context's scratch once
int launch_viscous(Ctx &c, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields *m, ws_soa du,
                   hipStream_t s);

int launch_sound_speed(const Ctx &c, ws_field T, ws_field cc, ws_job job, hipStream_t s);
// metric terms Akx..Asz and Ajac from ghosted mesh coordinates (k_metrics.hip)
int launch_comput_jacobian(const Ctx &c, const ws_field *coords, const ws_metric_fields &mf, const int32_t *jacbound,
                           hipStream_t s);
// ghost fill of d, u, v, w, T, then the sound speed over the ghosted box
int launch_prepare(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field T, ws_field cc,
                   hipStream_t s);
int launch_ghost_fill(const Ctx &c, const ws_field *f, int n, hipStream_t s);
// to[q] = from[q] over the whole ghosted box (ghost layers included), q < n <= 8
int launch_copy_fields(const Ctx &c, const ws_field *from, const ws_field *to, int n, hipStream_t s);
int launch_cons_to_pri(const Ctx &c, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w,
                       ws_field T, ws_field P, hipStream_t s);
int launch_pri_to_cons(const Ctx &c, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w,
                       ws_field T, ws_job job, hipStream_t s);
int launch_time_advance(const Ctx &c, int krk, ws_soa f, ws_soa fn, ws_soa du, ws_soa f_lap,
                        double dt, const double *dt_dev, hipStream_t s);
// one fused RK3 stage on st: f from (f, fn, du), the primitives (and P) from the new f, with
// planes the boundary-plane re-derivation of f, with sound the interior sound speed in cc
// cfl_part: the last stage of a step also writes one CFL maximum per work-group there
// (stage_blocks entries) for launch_cfl_final
int launch_stage(const Ctx &c, int krk, const ws_state &st, ws_soa f_lap, double dt, const double *dt_dev,
                 bool planes, bool sound, hipStream_t s, double *cfl_part = nullptr);
long stage_blocks(const Ctx &c);
int launch_cfl_final(const Ctx &c, const double *partial, int n, double cfl, double *dt_dev, hipStream_t s);
int launch_fill(double value, double *ptr, uint32_t pitch, uint32_t sx, uint32_t sy, uint32_t sz,
                hipStream_t s);
int launch_cfl(Ctx &c, const ws_state &st, double cfl, double *dt_dev, hipStream_t s);
// *out = min(v[0..n-1]) (one work-group, no atomics)
int launch_min_reduce(const double *v, int n, double *out, hipStream_t s);

}  // namespace ws

// the opaque handle of the C ABI is the internal context
struct ws_ctx : public ws::Ctx {};
