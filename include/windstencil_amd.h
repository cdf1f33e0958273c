/*
 * windstencil_amd.h -- C ABI of the MI355X inviscid-RHS engine.
 *
 * One GPU (gfx950), one process.  Every entry point is extern "C", takes plain
 * descriptors and pointers, is stream-ordered on the hipStream_t it is given,
 * performs no host synchronisation unless its comment says so, and returns
 * WS_OK (0) or a negative WS_ERR_* code; ws_last_error() describes the last
 * failure of the calling thread.
 *
 * Data layout (kept identical to the reference so that its buffers can be
 * handed over unchanged):
 *   ws_field  ghosted scalar field, element (x,y,z), x,y,z in [0, n+2*LAP):
 *             ptr[x + pitch*(y + (ny+2*LAP)*z)]            (get_Field_LAP,
 *             include/windstencil/cuda_utility.h:21)
 *   ws_soa    5-variable interior state, element (x,y,z,m), x,y,z in [0,n):
 *             ptr[x + pitch*(y + ny*(z + nz*m))]           (get_SoA,
 *             include/windstencil/cuda_utility.h:29)
 *   ws_job    half-open box [start, end) in ghosted coordinates
 *             (cudaJobPackage, include/windstencil/cuda_commen.h:95-104)
 * Pitches are in elements.  LAP (ghost depth) is 4.
 */
#ifndef WINDSTENCIL_AMD_H
#define WINDSTENCIL_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WS_LAP 4
#define WS_NVARS 5

/* error codes */
#define WS_OK 0
#define WS_ERR_INVALID (-1)      /* bad argument (null pointer, bad size, bad box) */
#define WS_ERR_UNSUPPORTED (-2)  /* valid request this build does not implement */
#define WS_ERR_HIP (-3)          /* a HIP runtime call failed */
#define WS_ERR_NOMEM (-4)        /* device allocation failed */

/* precision of the RHS arithmetic.  MIXED: splitting, merge and flux difference in FP64;
 * WENO7 smoothness, weights and candidates in FP32 on first differences of the split
 * fluxes taken relative to an FP64 base point per face (error ~1e-7 of the local flux
 * variation; the FP32 weight products hold for flux differences up to ~1e3).  Applies
 * to the tiled kernels; the anchor kernels are always FP64. */
#define WS_PREC_FP64 0
#define WS_PREC_MIXED 1

/* RHS kernel family */
#define WS_KERNEL_TILED 0        /* LDS-tiled directional kernels (default, fast path) */
#define WS_KERNEL_ANCHOR 1       /* per-direction split + per-cell WENO kernels, reference operation
                                    order, no contraction: bitwise equal to the CPU restatement */

/* Tile shapes of the tiled FP64 kernels (cells per line x lines per block, faces per thread) */
#define WS_TILE_15x16_K1 0       /* x: 15x16, y/z: 16x15, one face per thread */
#define WS_TILE_30x16_K2 1       /* x: 30x16, y/z: 32x15, two faces per thread (default) */
#define WS_TILE_30x16_A16_K2 2   /* x: 30x16, y/z: 32x16 on 16-aligned columns, two faces per thread */
#define WS_TILE_62x8_K2 3        /* x: 62x8, y/z: 32x15, two faces per thread */
#define WS_TILE_126x2_K2 4       /* x: 126x2 (two-wavefront work-groups), y/z: 32x15 */
#define WS_TILE_COUNT 5

/* ws_probe_valu operations */
#define WS_PROBE_FMA 0
#define WS_PROBE_MUL 1
#define WS_PROBE_ADD 2

/* how the tiled kernels walk their tiles (ws_ctx_set_schedule, ws_batch_set_schedule) */
#define WS_SCHED_GRID 0        /* one tile per work-group, grid = tile count */
#define WS_SCHED_PERSISTENT 1  /* resident work-groups, each a contiguous run of tiles */
#define WS_SCHED_PREFETCH 2    /* as 1, the next tile's loads issued during this tile's
                                  reconstruction */
#define WS_SCHED_COUNT 3

/* order of the three directional RHS launches (ws_ctx_set_direction_order): the first
 * writes du, the other two accumulate; the orders agree to rounding */
#define WS_ORDER_XYZ 0
#define WS_ORDER_ZYX 1
#define WS_ORDER_COUNT 2

/* Ghosted scalar field; layout of cudaField under __HIPCC__ (cuda_commen.h:48-54). */
typedef struct ws_field {
    double *ptr;
    uint32_t pitch;
    uint32_t dummy0;
} ws_field;

/* Interior five-variable state; layout of cudaSoA under __HIPCC__ (cuda_commen.h:64-71). */
typedef struct ws_soa {
    double *ptr;
    uint32_t pitch;
    uint32_t length_y;
    uint32_t length_z;
    uint32_t dummy0;
} ws_soa;

typedef struct ws_dim3 {
    uint32_t x, y, z;
} ws_dim3;

/* Half-open box; layout of cudaJobPackage (two dim3). */
typedef struct ws_job {
    ws_dim3 start;
    ws_dim3 end;
} ws_job;

/* Per-cell metric terms (all ghosted fields): Akx,Aky,Akz (x), Aix,Aiy,Aiz (y),
 * Asx,Asy,Asz (z) and the Jacobian factor Ajac.  The reference keeps them in
 * the globals pAkx_d .. pAjac_d (src/runtime/OCFD_init.cpp:412-424). */
typedef struct ws_metric_fields {
    ws_field A[9];
    ws_field ajac;
} ws_metric_fields;

/* Run configuration: what the reference keeps in host globals and __constant__
 * symbols (src/runtime/parameters_d.cpp:16-31, OCFD_init.cpp:29-31,379-396). */
typedef struct ws_config {
    int32_t nx, ny, nz;          /* interior extents of the block (one GPU: the domain) */
    int32_t periodic[3];         /* 1: periodic (ghost fill copies periodic images); 0: the caller owns the
                                    ghost layers of that direction (the reference's Iperiodic) */
    double gamma;                /* Gamma */
    double ama;                  /* Ama (reference Mach number) */
    double epsl_sw;              /* EPSL_SW, Steger-Warming smoothing */
    double dt;                   /* DT, time step used by the reference-named wrappers */
    double hx, hy, hz;           /* grid spacings (the reference sets 1/(N_global-1)) */
    double metric[10];           /* constant metrics Akx..Asz, Ajac, used when no metric fields are given */
    int32_t precision;           /* WS_PREC_FP64 or WS_PREC_MIXED */
    int32_t kernel;              /* WS_KERNEL_TILED or WS_KERNEL_ANCHOR */
    int32_t bc_planes;           /* 1: OCFD_bc re-derivation of the boundary planes after each stage */
    int32_t viscous;             /* 0: inviscid RHS only; 6 or 8: ws_du_comput adds the viscous terms with
                                    CD6 / CD8 central differences (Scheme_vis_ID 203 / 204) */
    double re, pr, ref_t;        /* Re, Pr, Ref_T of the viscous terms (Sutherland viscosity) */
    int32_t d0_bound[6];         /* 1: one-sided closure of the inviscid scheme at that domain face (x-, x+,
                                    y-, y+, z-, z+; the reference's D0_bound, src/boundary/OCFD_bound_Scheme.cpp:
                                    432-893); needs periodic[d] = 0 and at least 8 cells.  2 (high faces only):
                                    the form of the fused path's rim kernel (split_Jac3d_Stager_Warming_ker_
                                    boundary, src/solver/OCFD_NS_Jacobian3d.cpp:1063-1085), as 1 except the last
                                    P face 1.5 f(n-1) - 0.5 f(n-2) and the last two M faces 0.5 f(n-1) +
                                    0.5 f(n-2).  The rim kernel's x closures are {1, 2}, its y closures {1, 1}
                                    with non_ref of the y M pass (it drops the last cell's M difference); the
                                    viscous terms treat 2 as 1 */
    int32_t non_ref[6];          /* 1: non-reflecting flag of the P (even index) / M (odd) pass of a
                                    direction (the reference's Non_ref: the P pass skips the first cell, the M
                                    pass the last, where that side has the closure) */
} ws_config;

typedef struct ws_ctx ws_ctx;

/* Solver state: the device buffers the reference allocates in
 * opencfd_mem_init_dev (src/runtime/OCFD_init.cpp:411-446). */
typedef struct ws_state {
    ws_soa f, fn, du;                       /* pf_d, pfn_d, pdu_d */
    ws_field d, u, v, w, T, P, cc;          /* pd_d .. pP_d, pcc_d */
} ws_state;

/* State analysis (ws_ana_state).  First offending cells in interior coordinates (ghost
 * cells are negative or >= n); WS_ANA_NONE when there is none. */
#define WS_ANA_NONE (-2147483647 - 1)
typedef struct ws_ana_report {
    int64_t nan_T;              /* ghosted cells whose T is NaN */
    int64_t neg_T;              /* ghosted cells with T < 0 */
    int64_t neg_d;              /* ghosted cells with density < 0 (no counterpart) */
    int64_t neg_jac;            /* ghosted cells with Ajac < 0 */
    int32_t first_nan_T[3], first_neg_T[3], first_neg_d[3], first_neg_jac[3];
    double cfl;                 /* dt * max(|u|,|v|,|w|) + c / min(h) over the interior (0 if dt <= 0) */
    int32_t fatal;              /* NaN T or more than 10 cells with T < 0: where the reference aborts */
} ws_ana_report;

/* ---- library ------------------------------------------------------------ */
const char *ws_version(void);
const char *ws_error_string(int code);
const char *ws_last_error(void);
void ws_config_default(ws_config *cfg, int32_t nx, int32_t ny, int32_t nz);

/* ---- context (no counterpart: the reference uses globals) ---------------- */
int ws_ctx_create(const ws_config *cfg, ws_ctx **out);
int ws_ctx_destroy(ws_ctx *ctx);
int ws_ctx_get_config(const ws_ctx *ctx, ws_config *out);
int ws_ctx_set_precision(ws_ctx *ctx, int32_t precision);
int ws_ctx_set_kernel(ws_ctx *ctx, int32_t kernel);
int ws_ctx_set_tile_shape(ws_ctx *ctx, int32_t shape);
int ws_ctx_get_tile_shape(const ws_ctx *ctx, int32_t *shape);
/* WS_SCHED_*; the persistent schedules apply without metric fields (with them the
 * grid schedule runs) */
int ws_ctx_set_schedule(ws_ctx *ctx, int32_t schedule);
int ws_ctx_get_schedule(const ws_ctx *ctx, int32_t *schedule);
/* WS_ORDER_*: the launch that writes du reads nothing of it, the two after it
 * read-modify-write du, and which direction pays for that read shifts the time per
 * precision (the autotuner picks) */
int ws_ctx_set_direction_order(ws_ctx *ctx, int32_t order);
int ws_ctx_get_direction_order(const ws_ctx *ctx, int32_t *order);
/* WS_ORDER_* of the RHS launches whose last one applies an RK3 stage (stage fusion): that
 * launch also carries the stage's reads and writes, so the faster order can differ from
 * the plain RHS's (the autotuner picks; the default is WS_ORDER_XYZ).  Stage-applying
 * steps equal the unfused ones bit for bit when this order equals the direction order,
 * to rounding otherwise.  No counterpart in the reference. */
int ws_ctx_set_stage_order(ws_ctx *ctx, int32_t order);
int ws_ctx_get_stage_order(const ws_ctx *ctx, int32_t *order);
/* WS_TILE_* of those stage-applying launches, or -1 (the default): the context's tile
 * shape.  Chosen by the autotuner with the stage order; the same bit-for-bit and rounding
 * statements hold with "tile shape" for "direction order".  No counterpart in the
 * reference. */
int ws_ctx_set_stage_shape(ws_ctx *ctx, int32_t shape);
int ws_ctx_get_stage_shape(const ws_ctx *ctx, int32_t *shape);
/* 2: RK3 steps let the last directional RHS launch apply each stage's update (tiled
 * kernels, grid schedule, constant metrics, no closures, no viscous terms; same result as
 * the separate stage kernel, du then holds partial sums) and, with every direction
 * periodic, derive the primitives of stages 1 and 2 in that launch too (into a second set
 * of six ghosted fields the context allocates at the first eager step; same state bits at
 * the end of the step); 3 (default): as 2, and within ws_rk3_run (two steps or more) the
 * primitives, sound speed and CFL maxima of stage 3 as well, the primitives alternating
 * between the state's fields and the second set from step to step (no stage kernel after
 * stage 3, no preparation before the next step; the run's last step ends in the state's
 * fields with the bits of 2, ghost layers included; ws_rk3_step runs as 2); 1: the update
 * only; 0: separate update.  Batches treat 3 as 2. */
int ws_ctx_set_stage_fusion(ws_ctx *ctx, int32_t on);

/* Allocate / free the solver state with the reference's extents
 * (replaces opencfd_mem_init_dev, src/runtime/OCFD_init.cpp:411-446). */
int ws_state_alloc(ws_ctx *ctx, ws_state *out);
int ws_state_free(ws_ctx *ctx, ws_state *st);

/* ---- hot path ------------------------------------------------------------ */

/* Inviscid RHS of the interior box `job` from ghosted primitives d,u,v,w and sound
 * speed cc into du: du = RHS (accumulate = 0) or du += RHS (accumulate = 1).
 * Replaces split_Jac3d_Stager_Warming_ker + split_Jac3d_Stager_Warming_ker_boundary
 * (include/windstencil/OCFD_split.h:15-38) as launched by du_invis_Jacobian3d_x
 * (src/solver/OCFD_NS_Jacobian3d.cpp:240-284).  Ghosts of d,u,v,w,cc (LAP deep along
 * each direction, faces only) must be filled.  metrics == NULL: constant metrics from
 * the config.  job must satisfy LAP <= start < end <= n+LAP in every direction. */
int ws_inviscid_rhs(ws_ctx *ctx, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                    const ws_metric_fields *metrics, ws_soa du, ws_job job, int32_t accumulate,
                    hipStream_t stream);

/* Sound speed cc = sqrt(T)/Ma over the box (ghosted coordinates).
 * Replaces du_invis_Jacobian3d_init / sound_speed_kernel
 * (src/solver/OCFD_NS_Jacobian3d.cpp:24-33, src/kernels/commen_kernel.cpp:145-154). */
int ws_sound_speed(ws_ctx *ctx, ws_field T, ws_field cc, ws_job job, hipStream_t stream);

/* On-device periodic LAP-deep face halo of n fields: x, then y, then z, each over the
 * interior of the other two axes; edges and corners untouched.  Replaces
 * exchange_boundary_xyz_packed_dev on one rank with periodic neighbours
 * (src/mpi/OCFD_mpi_dev.cpp:123-128,187-290; src/mpi/OCFD_mpi.c:171-197). */
int ws_ghost_fill_periodic(ws_ctx *ctx, const ws_field *fields, int32_t nfields, hipStream_t stream);

/* The RHS inputs of a stage: the periodic face halo of d,u,v,w,T (as
 * ws_ghost_fill_periodic), then cc = sqrt(T)/Ma over the whole ghosted box (as
 * ws_sound_speed, as one flat pass over the field when T and cc share a layout),
 * bit-identical to those two calls in turn.  Replaces the exchanges
 * and du_invis_Jacobian3d_init at the head of du_invis_Jacobian3d
 * (src/solver/OCFD_Stream.cpp:187-194, src/solver/OCFD_NS_Jacobian3d.cpp:24-33). */
int ws_prepare_rhs_inputs(ws_ctx *ctx, const ws_state *st, hipStream_t stream);

/* Full RHS of one stage from the state's primitives: ws_prepare_rhs_inputs, then the
 * inviscid RHS into du (overwritten), then the viscous terms when the configuration has
 * them.  Replaces du_comput (src/solver/OCFD_Stream.cpp:27-52,185-227). */
int ws_du_comput(ws_ctx *ctx, const ws_state *st, const ws_metric_fields *metrics, hipStream_t stream);

/* ---- state helpers ------------------------------------------------------- */

/* Conserved -> primitive on the interior (get_duvwT, src/kernels/commen_kernel.cpp:78-119).
 * P may have a null ptr. */
int ws_cons_to_pri(ws_ctx *ctx, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w,
                   ws_field T, ws_field P, hipStream_t stream);

/* Primitive -> conserved over the interior box `job` given in interior coordinates
 * (pri_to_cons_kernel, src/kernels/commen_kernel.cpp:32-66). */
int ws_pri_to_cons(ws_ctx *ctx, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w,
                   ws_field T, ws_job job, hipStream_t stream);

/* OCFD_bc for periodic boxes: conserved variables of the six boundary planes from the
 * primitives (src/boundary/OCFD_boundary.cpp:38-89). */
int ws_bc_planes(ws_ctx *ctx, const ws_state *st, hipStream_t stream);

/* SSP-RK3 stage krk (1..3) over the interior (OCFD_time_advance_ker1..3,
 * src/solver/OCFD_time.cpp:15-65).  f_lap.ptr may be null (stage 3 then skips it). */
int ws_time_advance(ws_ctx *ctx, int32_t krk, ws_soa f, ws_soa fn, ws_soa du, ws_soa f_lap,
                    double dt, hipStream_t stream);

/* Fill a pitched box with a value (cuda_mem_value_init_warp,
 * src/kernels/commen_kernel.cpp:12-29). */
int ws_mem_value_init(double value, double *ptr, uint32_t pitch, uint32_t size_x, uint32_t size_y,
                      uint32_t size_z, hipStream_t stream);

/* CFL step dt = cfl*min(h)/max_cells max_dir(|u_dir| + c) written to *dt_dev (a device
 * double).  No counterpart: the reference uses a fixed DT. */
int ws_cfl_dt(ws_ctx *ctx, const ws_state *st, double cfl, double *dt_dev, hipStream_t stream);

/* Time every (tile shape, schedule, direction order) of the tiled kernels, in the context's
 * precision, on this state (two untimed and `reps` timed RHS launches each, HIP events on
 * `stream`), keep the fastest for the context and remember it for contexts of the same
 * extents and precision.  When RK3 steps of that choice apply their stages in the last RHS
 * launch (ws_ctx_set_stage_fusion), the stage order and tile shape are then timed the same
 * way (every shape in both orders, 3 * reps launches each): the RHS of an RK3 stage 2 whose last launch writes the
 * new f, primitives, sound speed and halo into scratch the call allocates and frees (the
 * state's f is read, not written).  times_ms
 * (WS_TILE_COUNT * WS_SCHED_COUNT entries, shape-major, each the faster of the two orders)
 * may be NULL.  Overwrites st->du; synchronises the stream.  No counterpart in the
 * reference. */
int ws_autotune(ws_ctx *ctx, const ws_state *st, int32_t reps, float *times_ms, hipStream_t stream);

/* Vector ALU throughput of the current device in tera lane-operations per second: eight
 * independent chains of one operation per lane over 16 work-groups of 256 lanes per
 * compute unit, best of five timed launches after 40 untimed ones.  precision
 * WS_PREC_FP64 or WS_PREC_MIXED (scalar FP32); op one of WS_PROBE_*.  FMA counts one
 * operation (two FLOP).  The attainable VALU roof beside the published peak.
 * Synchronises the stream.  No counterpart in the reference (its measurement asks for
 * this check). */
int ws_probe_valu(int32_t precision, int32_t op, double *tera_ops, hipStream_t stream);

/* Failure detection on the device, one pass over the ghosted box: cells whose T is NaN or
 * negative (the reference's ana_NAN_and_NT, src/runtime/OCFD_ana.cpp:115-225, which scans
 * T on the host and aborts on a NaN or on more than 10 negative cells: report->fatal), cells
 * with Ajac < 0 (ana_Jac, :35-56; metrics NULL checks the context's constant Ajac), cells
 * with negative density, and the CFL number of the time step dt (dt <= 0: skipped).  Counts
 * and first cells go to *report; the call returns WS_OK whatever it finds and synchronises
 * the stream. */
int ws_ana_state(ws_ctx *ctx, const ws_state *st, const ws_metric_fields *metrics, double dt,
                 ws_ana_report *report, hipStream_t stream);

/* ws_ana_state without any device allocation: the partial records live in a fixed device
 * array of the library (513 records, about 37 KB per GPU, held until the process ends), so
 * the call makes no hipMalloc or hipFree (ws_ana_state allocates and frees its records, and
 * the free waits for the device).  The array's size does not depend on the box (the pass is
 * grid-strided), so no box is too large for it; boxes beyond 2^40 ghosted cells are refused
 * as by ws_ana_state.  Every call on one GPU uses that GPU's array: calls that overlap on
 * one GPU (two host threads, or two streams) race and must be serialised by the caller.
 * Same arguments, report and synchronisation as ws_ana_state.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight solver loops that check
 * the state every step, code under stream capture, harnesses that forbid hipMalloc). */
int ws_ana_state_noalloc(ws_ctx *ctx, const ws_state *st, const ws_metric_fields *metrics, double dt,
                         ws_ana_report *report, hipStream_t stream);

/* Viscous terms of the RHS added to du components 1..4 (the reference's du_vis_Jacobian3d,
 * src/solver/OCFD_Stream.cpp:229-276, with du_viscous_Jacobian3d_{init,[xyz]_init,[xyz]_final},
 * src/solver/OCFD_NS_Jacobian3d.cpp:455-789): Sutherland viscosity (get_Amu,
 * src/kernels/commen_kernel.cpp:123-141), central-difference gradients of u, v, w, T, the
 * stress and heat flux of each direction, their periodic exchange and central derivative,
 * du += Ajac * that.  u, v, w, T must have their periodic ghosts filled; cfg.viscous gives
 * the order (6 or 8).  Uses a device scratch of 12 ghosted fields owned by the context,
 * allocated at the first call. */
int ws_viscous_rhs(ws_ctx *ctx, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields *metrics,
                   ws_soa du, hipStream_t stream);

/* ---- time integration ---------------------------------------------------- */

/* One SSP-RK3 step (the reference loop body, src/solver/OCFD_NS_Solver.cpp:64-90, with
 * all three stages): swap f/fn, then per stage du_comput, time advance, get_duvwT and
 * (if enabled) OCFD_bc.  dt_dev: device double holding dt, or NULL to use cfg.dt.  du is
 * scratch here: with stage fusion (ws_ctx_set_stage_fusion) it does not end up holding
 * the last stage's RHS. */
int ws_rk3_step(ws_ctx *ctx, ws_state *st, const double *dt_dev, hipStream_t stream);

/* nsteps RK3 steps.  cfl > 0: dt recomputed on the device at the start of every step
 * with ws_cfl_dt; cfl <= 0: cfg.dt.  use_graph = 1 captures one step in a hipGraph
 * and replays it.  Synchronises the stream before returning.  After the first step the
 * CFL maxima come from the stage kernel that ends the previous step (same dt bits as a
 * ws_cfl_dt pass; the context keeps one double per 256 cells for them). */
int ws_rk3_run(ws_ctx *ctx, ws_state *st, int32_t nsteps, double cfl, int32_t use_graph,
               hipStream_t stream);

/* ---- restart files --------------------------------------------------------------
 * The reference's restart format OCFD%08d.dat (written by OCFD_save, src/io/OCFD_IO.c:153-226
 * with the per-plane records of write_3d, src/io/OCFD_IO_mpi.c:605-701; read by read_file,
 * src/io/OCFD_IO.c:24-90, and read_3d, src/io/OCFD_IO_mpi.c:563-604): a Fortran-style record
 * int32 12 | int32 Istep | float64 tt | int32 12, then rho, u, v, w, T as NZ records each
 * (one per z plane, x fastest) framed by int32 NX*NY*8 markers; native byte order.  Only the
 * interior is stored; ghost layers are filled by the next ws_du_comput. */

/* "OCFD%08d.dat" for step into buf. */
int ws_restart_name(int32_t step, char *buf, int32_t buflen);
/* Host arrays: fields[0..4] = rho, u, v, w, T, nx*ny*nz each, x fastest. */
int ws_restart_write_host(const char *path, int32_t nx, int32_t ny, int32_t nz, int32_t istep, double tt,
                          const double *const *fields);
/* Fails (WS_ERR_INVALID) unless the file holds exactly this grid's records. */
int ws_restart_read_host(const char *path, int32_t nx, int32_t ny, int32_t nz, int32_t *istep, double *tt,
                         double *const *fields);
/* The interior of st's d, u, v, w, T to / from a restart file (OCFD_save / read_file for
 * one process).  Synchronise `stream` first; the copies themselves are blocking. */
int ws_save_restart(ws_ctx *ctx, const ws_state *st, int32_t istep, double tt, const char *path,
                    hipStream_t stream);
int ws_load_restart(ws_ctx *ctx, ws_state *st, const char *path, int32_t *istep, double *tt,
                    hipStream_t stream);

/* The reference's metric file OCFD3d-Jacobi.dat (read by Init_Jacobian3d,
 * src/solver/OCFD_Comput_Jacobian3d.cpp:138-205): no header; x, y, z, Akx, Aky, Akz, Aix,
 * Aiy, Aiz, Asx, Asy, Asz, Ajac as NZ plane records each (same records as above).
 * Host arrays: fields[0..12] in that order, nx*ny*nz each, x fastest. */
int ws_jacobi_write_host(const char *path, int32_t nx, int32_t ny, int32_t nz, const double *const *fields);
int ws_jacobi_read_host(const char *path, int32_t nx, int32_t ny, int32_t nz, double *const *fields);
/* Load the metric terms into metrics->A[0..8] (Akx..Asz) and metrics->ajac, and the mesh
 * coordinates into coords[0..2] unless coords is NULL, then fill their periodic face
 * ghosts (the reference's exchange_boundary_xyz after the read).  The fields are
 * ghosted device fields of the context's extents. */
int ws_load_jacobi(ws_ctx *ctx, const char *path, const ws_metric_fields *metrics, const ws_field *coords,
                   hipStream_t stream);

/* OCFD3d-Mesh.dat: the mesh coordinates x, y, z, each as NZ plane records (the layout of
 * the Jacobi file's first three fields), no header; read by Init_Jacobian3d when no Jacobi
 * file exists (src/solver/OCFD_Comput_Jacobian3d.cpp:120-135). */
int ws_mesh_write_host(const char *path, int32_t nx, int32_t ny, int32_t nz, const double *const *xyz);
int ws_mesh_read_host(const char *path, int32_t nx, int32_t ny, int32_t nz, double *const *xyz);
/* Metric terms from ghosted mesh coordinates coords[0..2] (their ghost layers as the
 * caller left them): metrics->A[0..8] = Akx..Asz, the cofactors of d(x,y,z)/d(i,j,k),
 * and metrics->ajac = 1/det, from sixth-order central differences divided by the spacing;
 * jacbound[d] = 1 (the reference's Jacbound, at least 8 cells) uses the one-sided and
 * low-order forms on the four cells next to each face of direction d.  Face ghosts repeat
 * the adjacent cell; periodic directions then get the periodic face halo.  Replaces
 * Comput_Jacobian3d / comput_Jac3d (src/solver/OCFD_Comput_Jacobian3d.cpp:215-429).
 * jacbound NULL means all 0. */
int ws_comput_jacobian(ws_ctx *ctx, const ws_field *coords, const ws_metric_fields *metrics,
                       const int32_t *jacbound, hipStream_t stream);
/* Init_Jacobian3d's mesh branch: every coordinate and metric value set to 1.0, the mesh
 * file read into the interior of coords, the coordinates' periodic face halo filled, then
 * ws_comput_jacobian (src/solver/OCFD_Comput_Jacobian3d.cpp:33-145). */
int ws_load_mesh(ws_ctx *ctx, const char *path, const ws_field *coords, const ws_metric_fields *metrics,
                 const int32_t *jacbound, hipStream_t stream);

/* ---- batches of blocks ------------------------------------------------------
 * Many independent periodic blocks of uneven size advanced together.  No counterpart
 * in the reference, which runs one block per process (its multi-block runs are MPI
 * decompositions of one grid).  Every block has the batch's physics, metrics and
 * precision, its own extents, spacing h = 1/(n-1) per direction (the reference's
 * convention, src/runtime/OCFD_init.cpp) and a caller-owned state.  The RHS of all
 * blocks is one launch per direction over a tile list; one dt (the minimum
 * of the blocks' CFL steps) advances every block. */
typedef struct ws_batch ws_batch;


/* extents: 3*nblocks (nx, ny, nz per block); states: nblocks states whose buffers stay
 * owned by the caller and must outlive the batch.  base supplies everything else
 * (its nx, ny, nz, hx, hy, hz are ignored). */
int ws_batch_create(const ws_config *base, int32_t nblocks, const int32_t *extents, const ws_state *states,
                    ws_batch **out);
int ws_batch_destroy(ws_batch *batch);
/* Block i's state as the batch holds it now (RK3 steps exchange f and fn). */
int ws_batch_block(const ws_batch *batch, int32_t i, ws_state *st);
int64_t ws_batch_cells(const ws_batch *batch);
int ws_batch_set_precision(ws_batch *batch, int32_t precision);
int ws_batch_set_kernel(ws_batch *batch, int32_t kernel);
int ws_batch_set_tile_shape(ws_batch *batch, int32_t shape);
int ws_batch_get_tile_shape(const ws_batch *batch, int32_t *shape);
int ws_batch_set_schedule(ws_batch *batch, int32_t schedule);
/* As ws_ctx_set_stage_fusion, for the batch's RK3 steps (grid schedule, no viscous terms):
 * 1 lets the batched z launch apply the stage, 2 (and 3, the default, which acts as 2 here)
 * also derive the primitives of stages 1 and 2 there (each block's context then holds a
 * second primitive set). */
int ws_batch_set_stage_fusion(ws_batch *batch, int32_t on);
int ws_batch_get_schedule(const ws_batch *batch, int32_t *schedule);
/* As ws_ctx_set_stage_order, for the batch's stage-applying RHS launches (WS_ORDER_XYZ:
 * the z launch applies the stage, the default; WS_ORDER_ZYX: the x launch does). */
int ws_batch_set_stage_order(ws_batch *batch, int32_t order);
int ws_batch_get_stage_order(const ws_batch *batch, int32_t *order);
/* As ws_ctx_set_stage_shape, for the batch (-1, the default: the batch's tile shape). */
int ws_batch_set_stage_shape(ws_batch *batch, int32_t shape);
int ws_batch_get_stage_shape(const ws_batch *batch, int32_t *shape);
/* RHS of every block from ghost-filled d,u,v,w and cc (overwrites du). */
int ws_batch_inviscid_rhs(ws_batch *batch, hipStream_t stream);
/* ws_du_comput for every block: ghost fill, sound speed, then the batched RHS. */
int ws_batch_du_comput(ws_batch *batch, hipStream_t stream);
/* *dt_dev = min over blocks of cfl*min(h)/max(|u_d| + c). */
int ws_batch_cfl_dt(ws_batch *batch, double cfl, double *dt_dev, hipStream_t stream);
/* One SSP-RK3 step of every block (as ws_rk3_step). */
int ws_batch_rk3_step(ws_batch *batch, const double *dt_dev, hipStream_t stream);
/* nsteps batch RK3 steps, dt from the CFL reduction (cfl > 0) or cfg.dt; use_graph = 1
 * captures a whole step (CFL included) in one hipGraph and replays it. */
int ws_batch_rk3_run(ws_batch *batch, int32_t nsteps, double cfl, int32_t use_graph, hipStream_t stream);
/* Time every (tile shape, schedule) pair of the batched RHS and keep the fastest (as
 * ws_autotune); with the grid schedule and stage fusion on, then the stage shape and order, timed on
 * a stage-2 RHS whose last launch writes into scratch the call allocates (a copy of every
 * block's f) and the blocks' second primitive sets (the blocks' f and primitives are left
 * as they were).  times_ms: WS_TILE_COUNT * WS_SCHED_COUNT entries, shape-major, or NULL. */
int ws_batch_autotune(ws_batch *batch, int32_t reps, float *times_ms, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
