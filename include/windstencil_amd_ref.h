/*
 * windstencil_amd_ref.h -- the reference's own inviscid-RHS call surface.
 *
 * The reference solver (src/solver) calls these functions by name and reads
 * grid extents, physical constants and device buffers from globals.  Here
 * the same names operate on a context and a solver state bound once with
 * ws_ref_bind(); the solver code above them needs no change beyond that call
 * and replacing its buffer descriptors by ws_state (identical layout).
 *
 * Error behaviour mirrors the reference under DEBUG_MODE (CUDA_CALL /
 * CUDA_LAUNCH print and MPI_Abort, include/windstencil/cuda_commen.h:14-46):
 * on failure a message goes to stderr and the process aborts, unless
 * ws_ref_set_abort_on_error(0) was called, in which case the failure is
 * recorded for ws_ref_last_status() and the call returns.
 */
#ifndef WINDSTENCIL_AMD_REF_H
#define WINDSTENCIL_AMD_REF_H

#include "windstencil_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bind the context and state the reference-named functions act on.  The state is
 * referenced, not copied, so pointer swaps the caller makes (pf_d <-> pfn_d,
 * src/solver/OCFD_NS_Solver.cpp:66-74) are seen.  metrics may be NULL (constant
 * metrics from the config).  `stream` is used by the functions that take none. */
int ws_ref_bind(ws_ctx *ctx, ws_state *st, const ws_metric_fields *metrics, hipStream_t stream);
/* The ghosted copy of f that stage 3 of OCFD_time_advance also writes (pf_lap_d,
 * src/solver/OCFD_time.cpp:48-58; the filter step reads it, src/runtime/OCFD_filtering.cpp:94-176):
 * a five-variable SoA over the ghosted box, pitch >= nx + 2*LAP (smaller pitches are refused).
 * NULL or a null ptr: stage 3 writes no copy (the default after ws_ref_bind). */
int ws_ref_bind_f_lap(const ws_soa *f_lap);
void ws_ref_set_abort_on_error(int abort_on_error);
int ws_ref_last_status(void);

/* src/solver/OCFD_NS_Jacobian3d.cpp:24-33, include/windstencil/OCFD_NS_Jacobian3d.h:11 */
void du_invis_Jacobian3d_init(ws_job job_in, hipStream_t *stream);
/* src/solver/OCFD_NS_Jacobian3d.cpp:133-414, OCFD_NS_Jacobian3d.h:15.  fp and fm are
 * accepted and ignored: split fluxes never leave LDS here. */
void du_invis_Jacobian3d_x(ws_job job_in, ws_soa *fp, ws_soa *fm, hipStream_t *stream);
/* src/solver/OCFD_Stream.cpp:185-227, include/windstencil/OCFD_Stream.h:19 */
void *du_invis_Jacobian3d(void *pthread_id);
/* src/solver/OCFD_Stream.cpp:229-276, OCFD_Stream.h:25: the viscous terms added to du
 * (requires cfg.viscous = 6 or 8; ghosts of u, v, w, T filled by the inviscid part) */
void *du_vis_Jacobian3d(void *pthread_id);
/* src/solver/OCFD_Stream.cpp:27-52, OCFD_Stream.h:10: the inviscid part, then the viscous
 * part when cfg.viscous is set */
void du_comput(int KRK);
/* src/kernels/commen_kernel.cpp:112-119, include/windstencil/commen_kernel.h:22 */
void get_duvwT(void);
/* src/solver/OCFD_time.cpp:67-91, include/windstencil/OCFD_time.h:11 */
void OCFD_time_advance(int KRK);
/* src/boundary/OCFD_boundary.cpp:18-92 (periodic case: boundary-plane re-derivation) */
void OCFD_bc(void);
/* src/kernels/commen_kernel.cpp:23-29, commen_kernel.h:11 */
void cuda_mem_value_init_warp(double value, double *ptr, unsigned int pitch, unsigned int size_x,
                              unsigned int size_y, unsigned int size_z);
/* src/mpi/OCFD_mpi_dev.cpp:123-128 on one periodic rank: an on-device face halo.
 * hostptr is ignored (no host staging). */
void exchange_boundary_xyz_packed_dev(double *hostptr, ws_field *devptr);

#ifdef __cplusplus
}
#endif
#endif
