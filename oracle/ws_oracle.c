/*
 * ws_oracle.c -- CPU FP64 oracle for the inviscid-RHS path (plain C, OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load this library, and only as the checker.
 * The shipped HIP library never links or calls it.
 *
 * Restates the reference's periodic (D0_bound = 0) semantics of the
 * decomposed WENO7-SYMBO path:
 *   split      src/solver/OCFD_NS_Jacobian3d.cpp:69-127
 *   WENO       src/kernels/OCFD_Schemes.cpp:306-368, 491-553
 *   faces/du   src/kernels/OCFD_Schemes.cpp:221-296, 461-484, 639-662
 *   prims      src/kernels/commen_kernel.cpp:32-66, 78-110, 145-154
 *   halo       src/mpi/OCFD_mpi_dev.cpp:187-290 with src/mpi/OCFD_mpi.c:171-197
 *   RK3        src/solver/OCFD_time.cpp:15-65, loop src/solver/OCFD_NS_Solver.cpp:64-90
 *   OCFD_bc    src/boundary/OCFD_boundary.cpp:38-89
 * Build with -ffp-contract=off so that every operation rounds as written.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ws_oracle.h"
#include "ws_oracle_cell.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#define LAP 4

static inline size_t gx(const orc_params *p) { return (size_t)p->nx + 2 * LAP; }
static inline size_t gy(const orc_params *p) { return (size_t)p->ny + 2 * LAP; }
static inline size_t gz(const orc_params *p) { return (size_t)p->nz + 2 * LAP; }

/* ghosted index, (i,j,k) in interior coordinates, each may lie in [-LAP, n+LAP) */
static inline size_t G(const orc_params *p, long i, long j, long k)
{
    return (size_t)(i + LAP) + gx(p) * ((size_t)(j + LAP) + gy(p) * (size_t)(k + LAP));
}

/* interior SoA index */
static inline size_t S(const orc_params *p, long i, long j, long k, int m)
{
    return (size_t)i + (size_t)p->nx * ((size_t)j + (size_t)p->ny * ((size_t)k + (size_t)p->nz * (size_t)m));
}

static void set_threads(const orc_params *p)
{
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

double orc_weno7_symbo_p(const double *s8) { return orc_cell_weno7_symbo_p(s8); }
double orc_weno7_symbo_m(const double *s8) { return orc_cell_weno7_symbo_m(s8); }

void orc_weno7_symbo_p_parts(const double *s8, double Sv[5], double wts[5], double q[5])
{
    double parts[15];
    int k;
    (void)orc_cell_weno7_symbo_p_ex(s8, parts);
    for (k = 0; k < 5; k++) {
        Sv[k] = parts[k];
        wts[k] = parts[5 + k];
        q[k] = parts[10 + k];
    }
}

void orc_split_sw(double d, double u, double v, double w, double cc,
                  double A1, double A2, double A3,
                  double gamma, double epsl_sw, double fp[5], double fm[5])
{
    const double C1 = 2.0 * (gamma - 1.0);
    const double C3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    orc_cell_split_sw(d, u, v, w, cc, A1, A2, A3, gamma, epsl_sw, C1, C3, fp, fm);
}

/* ---------------------------------------------------------------- prims */

static double cv_of(const orc_params *p)
{
    /* src/runtime/OCFD_init.cpp:380 */
    return 1.0 / (p->gamma * (p->gamma - 1.0) * p->ama * p->ama);
}

void orc_cons_to_pri(const orc_params *p, const double *f,
                     double *d, double *u, double *v, double *w, double *T, double *P)
{
    const double Cv = cv_of(p);
    long k;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (k = 0; k < p->nz; k++) {
        long i, j;
        for (j = 0; j < p->ny; j++)
            for (i = 0; i < p->nx; i++) {
                double out[6];
                size_t g = G(p, i, j, k);
                orc_cell_cons_to_pri(f[S(p, i, j, k, 0)], f[S(p, i, j, k, 1)], f[S(p, i, j, k, 2)],
                                     f[S(p, i, j, k, 3)], f[S(p, i, j, k, 4)], Cv, p->gamma, out);
                d[g] = out[0]; u[g] = out[1]; v[g] = out[2]; w[g] = out[3]; T[g] = out[4];
                if (P) P[g] = out[5];
            }
    }
}

void orc_pri_to_cons_box(const orc_params *p, double *f,
                         const double *d, const double *u, const double *v, const double *w, const double *T,
                         const int start[3], const int end[3])
{
    const double Cv = cv_of(p);
    long i, j, k;
    for (k = start[2]; k < end[2]; k++)
        for (j = start[1]; j < end[1]; j++)
            for (i = start[0]; i < end[0]; i++) {
                double c[5];
                size_t g = G(p, i, j, k);
                int m;
                orc_cell_pri_to_cons(d[g], u[g], v[g], w[g], T[g], Cv, c);
                for (m = 0; m < 5; m++) f[S(p, i, j, k, m)] = c[m];
            }
}

/* Periodic LAP-deep face halo: x, then y, then z, each over the interior of
 * the other two dimensions (edges and corners are left untouched, as the
 * packed exchange does: src/mpi/OCFD_mpi_dev.cpp:187-290). */
void orc_ghost_fill_periodic(const orc_params *p, double *fld)
{
    const long nx = p->nx, ny = p->ny, nz = p->nz;
    long i, j, k;
    if (p->periodic[0])
        for (k = 0; k < nz; k++)
            for (j = 0; j < ny; j++)
                for (i = 0; i < LAP; i++) {
                    fld[G(p, nx + i, j, k)] = fld[G(p, i, j, k)];          /* right ghost <- left interior */
                    fld[G(p, i - LAP, j, k)] = fld[G(p, nx - LAP + i, j, k)]; /* left ghost <- right interior */
                }
    if (p->periodic[1])
        for (k = 0; k < nz; k++)
            for (j = 0; j < LAP; j++)
                for (i = 0; i < nx; i++) {
                    fld[G(p, i, ny + j, k)] = fld[G(p, i, j, k)];
                    fld[G(p, i, j - LAP, k)] = fld[G(p, i, ny - LAP + j, k)];
                }
    if (p->periodic[2])
        for (k = 0; k < LAP; k++)
            for (j = 0; j < ny; j++)
                for (i = 0; i < nx; i++) {
                    fld[G(p, i, j, nz + k)] = fld[G(p, i, j, k)];
                    fld[G(p, i, j, k - LAP)] = fld[G(p, i, j, nz - LAP + k)];
                }
}

/* cc over the whole ghosted box (du_invis_Jacobian3d_init launches the
 * sound-speed kernel on [start-LAP, end+LAP), OCFD_NS_Jacobian3d.cpp:24-33) */
void orc_sound_speed(const orc_params *p, const double *T, double *cc)
{
    const size_t n = gx(p) * gy(p) * gz(p);
    size_t q;
    for (q = 0; q < n; q++) cc[q] = orc_cell_sound_speed(T[q], p->ama);
}

/* ---------------------------------------------------------------- RHS */

/* ---------------------------------------------------------------- boundary closures */

/* sign(1, x) and minmod2 (src/kernels/OCFD_Schemes.cpp:822-835) */
static double bs_sign1(double x) { return x >= 0 ? 1.0 : -1.0; }
static double bs_minmod2(double a, double b)
{
    return 0.5 * (bs_sign1(a) + bs_sign1(b)) * fmin(fabs(a), fabs(b));
}

/* Reduced WENO5 reconstructions on a 5-point window s[0..4] (no linear weights), used by
 * the one-sided closures.  OCFD_weno5_kernel_P (OCFD_Schemes.cpp:1110-1141) builds a0 and a1
 * from S1 and S2 before they are assigned (both still 0) -- reproduced as written. */
static double w5_p(const double *s)
{
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, tmp, a0, a1, a2, q03, q13, q23;
    const double ep = 1e-6;
    tmp = s[2] - 2.0 * s[3] + s[4]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[3] + s[4]; S0 += 3 * tmp * tmp;
    a0 = ((12.0 * ep + S1) * (12.0 * ep + S1)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    q03 = (2.0 * s[2] + 5.0 * s[3] - s[4]);
    tmp = s[1] - 2.0 * s[2] + s[3]; S1 += 13 * tmp * tmp;
    tmp = s[1] - s[3]; S1 += 3 * tmp * tmp;
    a1 = ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    q13 = (-s[1] + 5.0 * s[2] + 2.0 * s[3]);
    tmp = s[0] - 2.0 * s[1] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[0] - 4.0 * s[1] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    a2 = ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S1) * (12.0 * ep + S1));
    q23 = (2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2]);
    return (a0 * q03 + a1 * q13 + a2 * q23) / (6.0 * (a0 + a1 + a2));
}

/* OCFD_weno5_kernel_M, OCFD_Schemes.cpp:1170-1201 */
static double w5_m(const double *s)
{
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, tmp, a0, a1, a2, q03, q13, q23;
    const double ep = 1e-6;
    tmp = s[2] - 2.0 * s[1] + s[0]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[1] + s[0]; S0 += 3 * tmp * tmp;
    a0 = 1.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    q03 = (2.0 * s[2] + 5.0 * s[1] - s[0]);
    tmp = s[3] - 2.0 * s[2] + s[1]; S1 += 13 * tmp * tmp;
    tmp = s[3] - s[1]; S1 += 3 * tmp * tmp;
    a1 = 1.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    q13 = (-s[3] + 5.0 * s[2] + 2.0 * s[1]);
    tmp = s[4] - 2.0 * s[3] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[4] - 4.0 * s[3] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    q23 = (2.0 * s[4] - 7.0 * s[3] + 11.0 * s[2]);
    return (a0 * q03 + a1 * q13 + a2 * q23) / (6.0 * (a0 + a1 + a2));
}

/* one-candidate and two-candidate closures, src/boundary/OCFD_bound_Scheme.cpp:540-719 */
static double w5_p_lift(const double *s)
{
    double S0 = 0.0, tmp, a0, q03;
    const double ep = 1e-6;
    tmp = s[2] - 2.0 * s[3] + s[4]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[3] + s[4]; S0 += 3 * tmp * tmp;
    a0 = 1.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    q03 = (2.0 * s[2] + 5.0 * s[3] - s[4]);
    return a0 * q03 / (6.0 * a0);
}

static double w5_p_lift_plus(const double *s)
{
    double S0 = 0.0, S1 = 0.0, tmp, a0, a1, q03, q13;
    const double ep = 1e-6;
    tmp = s[2] - 2.0 * s[3] + s[4]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[3] + s[4]; S0 += 3 * tmp * tmp;
    a0 = 1.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    q03 = (2.0 * s[2] + 5.0 * s[3] - s[4]);
    tmp = s[1] - 2.0 * s[2] + s[3]; S1 += 13 * tmp * tmp;
    tmp = s[1] - s[3]; S1 += 3 * tmp * tmp;
    a1 = 1.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    q13 = (-s[1] + 5.0 * s[2] + 2.0 * s[3]);
    return (a0 * q03 + a1 * q13) / (6.0 * (a0 + a1));
}

static double w5_p_right_plus(const double *s)
{
    double S1 = 0.0, S2 = 0.0, tmp, a1, a2, q13, q23;
    const double ep = 1e-6;
    tmp = s[1] - 2.0 * s[2] + s[3]; S1 += 13 * tmp * tmp;
    tmp = s[1] - s[3]; S1 += 3 * tmp * tmp;
    a1 = 1.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    q13 = (-s[1] + 5.0 * s[2] + 2.0 * s[3]);
    tmp = s[0] - 2.0 * s[1] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[0] - 4.0 * s[1] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    q23 = (2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2]);
    return (a1 * q13 + a2 * q23) / (6.0 * (a1 + a2));
}

static double w5_m_lift(const double *s)
{
    double S2 = 0.0, tmp, a2, q23;
    const double ep = 1e-6;
    tmp = s[4] - 2.0 * s[3] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[4] - 4.0 * s[3] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    q23 = (2.0 * s[4] - 7.0 * s[3] + 11.0 * s[2]);
    return a2 * q23 / (6.0 * a2);
}

static double w5_m_lift_plus(const double *s)
{
    double S1 = 0.0, S2 = 0.0, tmp, a1, a2, q13, q23;
    const double ep = 1e-6;
    tmp = s[3] - 2.0 * s[2] + s[1]; S1 += 13 * tmp * tmp;
    tmp = s[3] - s[1]; S1 += 3 * tmp * tmp;
    a1 = 1.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    q13 = (-s[3] + 5.0 * s[2] + 2.0 * s[1]);
    tmp = s[4] - 2.0 * s[3] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[4] - 4.0 * s[3] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    q23 = (2.0 * s[4] - 7.0 * s[3] + 11.0 * s[2]);
    return (a1 * q13 + a2 * q23) / (6.0 * (a1 + a2));
}

static double w5_m_right_plus(const double *s)
{
    double S0 = 0.0, S1 = 0.0, tmp, a0, a1, q03, q13;
    const double ep = 1e-6;
    tmp = s[2] - 2.0 * s[1] + s[0]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[1] + s[0]; S0 += 3 * tmp * tmp;
    a0 = 1.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    q03 = (2.0 * s[2] + 5.0 * s[1] - s[0]);
    tmp = s[3] - 2.0 * s[2] + s[1]; S1 += 13 * tmp * tmp;
    tmp = s[3] - s[1]; S1 += 3 * tmp * tmp;
    a1 = 1.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    q13 = (-s[3] + 5.0 * s[2] + 2.0 * s[1]);
    return (a0 * q03 + a1 * q13) / (6.0 * (a0 + a1));
}

/* P face f (between cells f-1 and f) of a line with n cells and one-sided closures at the
 * low (lo) / high (hi) end, OCFD_bound_scheme_kernel_p (OCFD_bound_Scheme.cpp:722-807): the
 * kernel's thread c = f holds cell f-1 and the window s(f-4 .. f+3); the 5-point closures
 * use cells f-3 .. f+1.  s points at cell 0.  Returns 1 if a closure applies.  hi == 2: the
 * high end of the fused path's rim kernel (split_Jac3d_Stager_Warming_ker_boundary,
 * src/solver/OCFD_NS_Jacobian3d.cpp:1063-1072), whose last face is 1.5 s(n-1) - 0.5 s(n-2). */
static int closure_p(const double *s, long n, long f, int lo, int hi, double *out)
{
    if (lo && f <= 3) {
        if (f == 0) *out = s[0];
        else if (f == 1) *out = w5_p_lift(s + f - 3);
        else if (f == 2) *out = w5_p_lift_plus(s + f - 3);
        else *out = w5_p(s + f - 3);
        return 1;
    }
    if (hi && f >= n - 2) {
        if (f == n && hi == 2) *out = 1.5 * s[n - 1] - 0.5 * s[n - 2];
        else if (f == n) *out = s[n - 1] + 0.5 * bs_minmod2(s[n - 1] - s[n - 2], s[n - 1] - s[n - 2]);
        else if (f == n - 1) *out = w5_p_right_plus(s + f - 3);
        else *out = w5_p(s + f - 3);
        return 1;
    }
    return 0;
}

/* M face f, OCFD_bound_scheme_kernel_m (OCFD_bound_Scheme.cpp:808-893): thread c = f holds
 * cell f and the window s(f-4 .. f+3); the 5-point closures use cells f-2 .. f+2.  hi == 2:
 * the rim kernel's high end (OCFD_NS_Jacobian3d.cpp:1075-1085), whose last two faces both
 * take 0.5 s(n-1) + 0.5 s(n-2). */
static int closure_m(const double *s, long n, long f, int lo, int hi, double *out)
{
    if (lo && f <= 3) {
        if (f == 0) *out = w5_m_lift(s + f - 2);
        else if (f == 1) *out = w5_m_lift_plus(s + f - 2);
        else *out = w5_m(s + f - 2);
        return 1;
    }
    if (hi && f >= n - 3) {
        if (hi == 2 && f >= n - 1) *out = 0.5 * s[n - 1] + 0.5 * s[n - 2];
        else if (f == n) *out = s[n - 1];
        else if (f == n - 1) *out = s[n - 1] - 0.5 * bs_minmod2(s[n - 1] - s[n - 2], s[n - 1] - s[n - 2]);
        else if (f == n - 2) *out = w5_m_right_plus(s + f - 2);
        else *out = w5_m(s + f - 2);
        return 1;
    }
    return 0;
}

int orc_closure_p(const double *s, long n, long f, int lo, int hi, double *out)
{
    return closure_p(s, n, f, lo, hi, out);
}

int orc_closure_m(const double *s, long n, long f, int lo, int hi, double *out)
{
    return closure_m(s, n, f, lo, hi, out);
}

/* One direction of the RHS on one line: split at the n+2*LAP points, then the
 * P faces for all five variables (OCFD_weno7_SYMBO_P_kernel + put_du_p_kernel),
 * then the M faces (OCFD_weno7_SYMBO_M_kernel + put_du_m_kernel).
 * Face f (f = 0..n) sits between cells f-1 and f and reconstructs from the
 * split values at f-4 .. f+3.  prim[]/metric pointers are pre-offset per
 * point by the caller through the `at` index array. */
static void rhs_line(const orc_params *p, int dir, long n, const size_t *at, const size_t *du_at, double h,
                     const double *d, const double *u, const double *v, const double *w, const double *cc,
                     const double *A1, const double *A2, const double *A3, const double *ajac,
                     double *du, double *buf)
{
    const double C1 = 2.0 * (p->gamma - 1.0);
    const double C3 = (3.0 - p->gamma) / (2.0 * (p->gamma - 1.0));
    const long np = n + 2 * LAP;
    double *fpv = buf;              /* [5][np] */
    double *fmv = buf + 5 * np;     /* [5][np] */
    double *hP = buf + 10 * np;     /* [n+1] */
    double *hM = hP + (n + 1);      /* [n+1] */
    long q, f, c;
    int m;

    for (q = 0; q < np; q++) {
        double fp5[5], fm5[5];
        size_t g = at[q];
        orc_cell_split_sw(d[g], u[g], v[g], w[g], cc[g], A1[g], A2[g], A3[g],
                          p->gamma, p->epsl_sw, C1, C3, fp5, fm5);
        for (m = 0; m < 5; m++) { fpv[m * np + q] = fp5[m]; fmv[m * np + q] = fm5[m]; }
    }
    const int lo = p->d0_bound[2 * dir], hi = p->d0_bound[2 * dir + 1];
    /* non-reflecting flags (put_du_p/m_kernel, OCFD_Schemes.cpp:221-296): no P update of
     * the first cell, no M update of the last, when that side has the closure */
    const int skipP = lo && p->non_ref[2 * dir], skipM = hi && p->non_ref[2 * dir + 1];
    for (m = 0; m < 5; m++) {
        for (f = 0; f <= n; f++)
            if (!closure_p(&fpv[m * np + LAP], n, f, lo, hi, &hP[f])) hP[f] = orc_cell_weno7_symbo_p(&fpv[m * np + f]);
        for (c = 0; c < n; c++) {
            size_t g = at[c + LAP];
            if (skipP && c == 0) continue;
            du[du_at[c] + m] = orc_cell_flux_diff(du[du_at[c] + m], ajac[g], hP[c + 1], hP[c], h);
        }
    }
    for (m = 0; m < 5; m++) {
        for (f = 0; f <= n; f++)
            if (!closure_m(&fmv[m * np + LAP], n, f, lo, hi, &hM[f])) hM[f] = orc_cell_weno7_symbo_m(&fmv[m * np + f]);
        for (c = 0; c < n; c++) {
            size_t g = at[c + LAP];
            if (skipM && c == n - 1) continue;
            du[du_at[c] + m] = orc_cell_flux_diff(du[du_at[c] + m], ajac[g], hM[c + 1], hM[c], h);
        }
    }
}

/* The du of each line is gathered into a small [n][5] buffer, updated by
 * rhs_line and scattered back, so that rhs_line is direction-agnostic. */
static void rhs_direction(const orc_params *p, int dir,
                          const double *d, const double *u, const double *v, const double *w, const double *cc,
                          const double *const metrics[9], const double *ajac, double *du)
{
    const long nx = p->nx, ny = p->ny, nz = p->nz;
    const long n = dir == 0 ? nx : dir == 1 ? ny : nz;
    const long n1 = dir == 0 ? ny : nx;            /* the two other dims, outer loop last */
    const long n2 = dir == 2 ? ny : nz;
    const double h = dir == 0 ? p->hx : dir == 1 ? p->hy : p->hz;
    const size_t N = (size_t)nx * ny * nz;
    const double *A1 = metrics[3 * dir + 0];
    const double *A2 = metrics[3 * dir + 1];
    const double *A3 = metrics[3 * dir + 2];
    long b;
    set_threads(p);
#pragma omp parallel
    {
        const long np = n + 2 * LAP;
        size_t *at = (size_t *)malloc(sizeof(size_t) * np);
        size_t *du_at = (size_t *)malloc(sizeof(size_t) * n);
        double *buf = (double *)malloc(sizeof(double) * (10 * np + 2 * (n + 1)));
        double *dl = (double *)malloc(sizeof(double) * 5 * n);   /* du of this line, [n][5] */
#pragma omp for schedule(static)
        for (b = 0; b < n1 * n2; b++) {
            long a1 = b % n1, a2 = b / n1, q, c;
            int m;
            for (q = 0; q < np; q++) {
                long t = q - LAP;
                at[q] = dir == 0 ? G(p, t, a1, a2) : dir == 1 ? G(p, a1, t, a2) : G(p, a1, a2, t);
            }
            for (c = 0; c < n; c++) {
                size_t base = dir == 0 ? S(p, c, a1, a2, 0) : dir == 1 ? S(p, a1, c, a2, 0) : S(p, a1, a2, c, 0);
                du_at[c] = (size_t)(5 * c);
                for (m = 0; m < 5; m++) dl[5 * c + m] = du[base + (size_t)m * N];
            }
            rhs_line(p, dir, n, at, du_at, h, d, u, v, w, cc, A1, A2, A3, ajac, dl, buf);
            for (c = 0; c < n; c++) {
                size_t base = dir == 0 ? S(p, c, a1, a2, 0) : dir == 1 ? S(p, a1, c, a2, 0) : S(p, a1, a2, c, 0);
                for (m = 0; m < 5; m++) du[base + (size_t)m * N] = dl[5 * c + m];
            }
        }
        free(at); free(du_at); free(buf); free(dl);
    }
}

void orc_inviscid_rhs(const orc_params *p,
                      const double *d, const double *u, const double *v, const double *w, const double *cc,
                      const double *const metrics[9], const double *ajac, double *du)
{
    int dir;
    for (dir = 0; dir < 3; dir++)
        rhs_direction(p, dir, d, u, v, w, cc, metrics, ajac, du);
}

void orc_du_comput(const orc_params *p, double *d, double *u, double *v, double *w, double *T, double *cc,
                   const double *const metrics[9], const double *ajac, double *du)
{
    memset(du, 0, sizeof(double) * 5 * (size_t)p->nx * p->ny * p->nz);
    orc_ghost_fill_periodic(p, d);
    orc_ghost_fill_periodic(p, u);
    orc_ghost_fill_periodic(p, v);
    orc_ghost_fill_periodic(p, w);
    orc_ghost_fill_periodic(p, T);
    orc_sound_speed(p, T, cc);
    orc_inviscid_rhs(p, d, u, v, w, cc, metrics, ajac, du);
    if (p->viscous == 6 || p->viscous == 8)
        orc_viscous_rhs(p, p->re, p->pr, p->ref_t, p->viscous, u, v, w, T, metrics, ajac, du);
}

/* ---------------------------------------------------------------- time */

void orc_time_advance(const orc_params *p, int krk, double *f, const double *fn, const double *du, double dt)
{
    const size_t n = 5 * (size_t)p->nx * p->ny * p->nz;
    size_t q;
    if (krk == 1) {
        for (q = 0; q < n; q++) f[q] = fn[q] + dt * du[q];
    } else if (krk == 2) {
        const double tmp1 = 3.0 / 4.0, tmp2 = 1.0 / 4.0;
        for (q = 0; q < n; q++) f[q] = tmp1 * fn[q] + tmp2 * (f[q] + dt * du[q]);
    } else {
        const double tmp1 = 1.0 / 3.0, tmp2 = 2.0 / 3.0;
        for (q = 0; q < n; q++) f[q] = tmp1 * fn[q] + tmp2 * (f[q] + dt * du[q]);
    }
}

void orc_bc_planes(const orc_params *p, double *f,
                   const double *d, const double *u, const double *v, const double *w, const double *T)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    int s[3], e[3];
    /* x = 0, x = nx-1, y = 0, y = ny-1, z = 0, z = nz-1, in that order */
    s[0] = 0;      s[1] = 0;      s[2] = 0;      e[0] = 1;  e[1] = ny; e[2] = nz; orc_pri_to_cons_box(p, f, d, u, v, w, T, s, e);
    s[0] = nx - 1;                               e[0] = nx;                       orc_pri_to_cons_box(p, f, d, u, v, w, T, s, e);
    s[0] = 0;      s[1] = 0;                     e[0] = nx; e[1] = 1;             orc_pri_to_cons_box(p, f, d, u, v, w, T, s, e);
    s[1] = ny - 1;                               e[1] = ny;                       orc_pri_to_cons_box(p, f, d, u, v, w, T, s, e);
    s[1] = 0;      s[2] = 0;                     e[1] = ny; e[2] = 1;             orc_pri_to_cons_box(p, f, d, u, v, w, T, s, e);
    s[2] = nz - 1;                               e[2] = nz;                       orc_pri_to_cons_box(p, f, d, u, v, w, T, s, e);
}

double orc_cfl_dt(const orc_params *p, double cfl,
                  const double *u, const double *v, const double *w, const double *T)
{
    double smax = 0.0, hmin;
    long i, j, k;
    for (k = 0; k < p->nz; k++)
        for (j = 0; j < p->ny; j++)
            for (i = 0; i < p->nx; i++) {
                size_t g = G(p, i, j, k);
                double c = orc_cell_sound_speed(T[g], p->ama);
                double a = fabs(u[g]) + c, b = fabs(v[g]) + c, e = fabs(w[g]) + c;
                smax = fmax(smax, fmax(a, fmax(b, e)));
            }
    hmin = fmin(p->hx, fmin(p->hy, p->hz));
    return cfl * hmin / smax;
}

void orc_rk3_run(const orc_params *p, int nsteps, double dt,
                 double *f, double *d, double *u, double *v, double *w, double *T)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    const size_t NG = gx(p) * gy(p) * gz(p);
    double *fn = (double *)malloc(sizeof(double) * 5 * N);
    double *du = (double *)malloc(sizeof(double) * 5 * N);
    double *cc = (double *)calloc(NG, sizeof(double));
    double *one = (double *)malloc(sizeof(double) * NG);
    double *zero = (double *)calloc(NG, sizeof(double));
    const double *metrics[9];
    size_t q;
    int step, krk;
    for (q = 0; q < NG; q++) one[q] = 1.0;
    metrics[0] = one;  metrics[1] = zero; metrics[2] = zero;
    metrics[3] = zero; metrics[4] = one;  metrics[5] = zero;
    metrics[6] = zero; metrics[7] = zero; metrics[8] = one;

    for (step = 0; step < nsteps; step++) {
        double dts = dt > 0.0 ? dt : orc_cfl_dt(p, -dt, u, v, w, T);
        /* pf <-> pfn swap: fn holds the state at t^n (OCFD_NS_Solver.cpp:66-74) */
        memcpy(fn, f, sizeof(double) * 5 * N);
        for (krk = 1; krk <= 3; krk++) {
            orc_du_comput(p, d, u, v, w, T, cc, metrics, one, du);
            orc_time_advance(p, krk, f, fn, du, dts);
            orc_cons_to_pri(p, f, d, u, v, w, T, NULL);
            if (p->bc_planes) orc_bc_planes(p, f, d, u, v, w, T);
        }
    }
    free(fn); free(du); free(cc); free(one); free(zero);
}

/* ---------------------------------------------------------------- viscous terms */

double orc_cd6(const double *s)
{
    return orc_cell_cd6(s);
}

double orc_cd8(const double *s)
{
    return orc_cell_cd8(s);
}

void orc_amu(const orc_params *p, double re, double ref_t, const double *T, double *amu)
{
    const double tsb = 110.4 / ref_t;
    const double c0 = 1.0 / re * (1.0 + 110.4 / ref_t);
    long i, j, k;
    for (k = 0; k < p->nz; k++)
        for (j = 0; j < p->ny; j++)
            for (i = 0; i < p->nx; i++) {
                amu[S(p, i, j, k, 0)] = orc_cell_amu(T[G(p, i, j, k)], c0, tsb);
            }
}

/* One-sided and low-order forms of the viscous derivatives in the four cells next to a closed
 * domain face (D0_bound): OCFD_D0bound_scheme_kernel (src/boundary/OCFD_bound_Scheme.cpp:82-263),
 * which OCFD_CD6_kernel / OCFD_CD8_kernel (src/kernels/OCFD_Schemes.cpp:106-160) consult before
 * the central form; the flags come from OCFD_bound (:432-452).  s(q) = f(c + q); c is the cell's
 * index along the direction, n the extent.  The low face's cases are checked first.  Returns 1
 * and the undivided derivative in *tmp when a form covers the cell, else 0. */
int orc_cd_closure(const double *s0, int r, long c, long n, int lo, int hi, double *tmp)
{
    const double *s = s0 + r;   /* s[q] = f(c + q), q in [-r, r] */
    if (lo) {
        if (c == 0) { *tmp = (s[1] - s[0]); return 1; }
        if (c == 1) { *tmp = (s[1] - s[-1]) * 0.5; return 1; }
        if (c == 2) { *tmp = (s[-2] - 8.0 * s[-1] + 8.0 * s[1] - s[2]) / 12.0; return 1; }
        if (c == 3) { *tmp = (s[3] - s[-3] - 9.0 * (s[2] - s[-2]) + 45.0 * (s[1] - s[-1])) / 60.0; return 1; }
    }
    if (hi) {
        if (c == n - 1) { *tmp = (s[0] - s[-1]); return 1; }
        if (c == n - 2) { *tmp = (s[1] - s[-1]) * 0.5; return 1; }
        if (c == n - 3) { *tmp = (s[-2] - 8.0 * s[-1] + 8.0 * s[1] - s[2]) / 12.0; return 1; }
        if (c == n - 4) { *tmp = (s[3] - s[-3] - 9.0 * (s[2] - s[-2]) + 45.0 * (s[1] - s[-1])) / 60.0; return 1; }
    }
    return 0;
}

/* central derivative of ghosted field f along dir at interior cell (i,j,k), divided by h (the
 * order's central form, OCFD_kernel_CD6/_CD8, or a closure form next to a closed face) */
static double cd_at(const orc_params *p, int order, int dir, const double *f, long i, long j, long k)
{
    double s[9];
    const int r = order / 2;
    int q;
    for (q = -r; q <= r; q++) {
        const long a = i + (dir == 0 ? q : 0), b = j + (dir == 1 ? q : 0), c = k + (dir == 2 ? q : 0);
        s[q + r] = f[G(p, a, b, c)];   /* the closure forms read no point past a closed face */
    }
    const double h = dir == 0 ? p->hx : dir == 1 ? p->hy : p->hz;
    const long n = dir == 0 ? p->nx : dir == 1 ? p->ny : p->nz;
    const long c = dir == 0 ? i : dir == 1 ? j : k;
    double tmp;
    if (!orc_cd_closure(s, r, c, n, p->d0_bound[2 * dir], p->d0_bound[2 * dir + 1], &tmp))
        tmp = order == 8 ? orc_cd8(s) : orc_cd6(s);
    return tmp / h;
}

void orc_viscous_rhs(const orc_params *p, double re, double pr, double ref_t, int order,
                     const double *u, const double *v, const double *w, const double *T,
                     const double *const metrics[9], const double *ajac, double *du)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz, NG = gx(p) * gy(p) * gz(p);
    const double cp = 1.0 / ((p->gamma - 1.0) * p->ama * p->ama);
    const double vis_c = cp / pr;
    /* gradients: g[4*dir + var], var = u, v, w, T (interior arrays) */
    double *g = (double *)malloc(sizeof(double) * 12 * N);
    double *amu = (double *)malloc(sizeof(double) * N);
    double *ev = (double *)calloc(4 * NG, sizeof(double));
    const double *fld[4] = {u, v, w, T};
    int dir, var, m;
    long i, j, k;
    set_threads(p);
    orc_amu(p, re, ref_t, T, amu);
    for (dir = 0; dir < 3; dir++)
        for (var = 0; var < 4; var++) {
            double *out = g + (size_t)(4 * dir + var) * N;
#pragma omp parallel for private(i, j) schedule(static)
            for (k = 0; k < p->nz; k++)
                for (j = 0; j < p->ny; j++)
                    for (i = 0; i < p->nx; i++) out[S(p, i, j, k, 0)] = cd_at(p, order, dir, fld[var], i, j, k);
        }
    for (dir = 0; dir < 3; dir++) {
#pragma omp parallel for private(i, j) schedule(static)
        for (k = 0; k < p->nz; k++)
            for (j = 0; j < p->ny; j++)
                for (i = 0; i < p->nx; i++) {
                    const size_t c = S(p, i, j, k, 0), gg = G(p, i, j, k);
                    double gv[12], A[9], e4[4];
                    int q;
                    for (q = 0; q < 12; q++) gv[q] = g[(size_t)q * N + c];
                    for (q = 0; q < 9; q++) A[q] = metrics[q][gg];
                    orc_cell_vis_flux(gv, u[gg], v[gg], w[gg], amu[c], vis_c, A, ajac[gg], dir, e4);
                    for (q = 0; q < 4; q++) ev[(size_t)q * NG + gg] = e4[q];
                }
        for (m = 0; m < 4; m++) orc_ghost_fill_periodic(p, ev + (size_t)m * NG);
        for (m = 0; m < 4; m++) {
            const double *e = ev + (size_t)m * NG;
#pragma omp parallel for private(i, j) schedule(static)
            for (k = 0; k < p->nz; k++)
                for (j = 0; j < p->ny; j++)
                    for (i = 0; i < p->nx; i++)
                        du[S(p, i, j, k, m + 1)] += ajac[G(p, i, j, k)] * cd_at(p, order, dir, e, i, j, k);
        }
    }
    free(g);
    free(amu);
    free(ev);
}

/* ---- metrics from mesh coordinates (src/solver/OCFD_Comput_Jacobian3d.cpp:215-429) ---------- */

/* Derivative of ghosted field f along direction dir, interior cell (i, j, k), divided by the
 * direction's spacing: the sixth-order central form OCFD_dx0_CD6_kernel
 * (src/kernels/OCFD_Schemes.cpp:1229-1297); with bound set, the four cells next to each
 * face use the one-sided/low-order forms of OCFD_Dx0_bound_kernel_m/_p
 * (src/boundary/OCFD_bound_Scheme.cpp:268-307, the y and z copies to :390), selected per
 * direction by Jacbound (OCFD_dx0_jac, src/solver/OCFD_Schemes_Choose.cpp:655-700). */
static double metric_deriv(const orc_params *p, const double *f, int dir, int bound, long i, long j, long k)
{
    const long n = dir == 0 ? p->nx : dir == 1 ? p->ny : p->nz;
    const long c = dir == 0 ? i : dir == 1 ? j : k;
    const double h = dir == 0 ? p->hx : dir == 1 ? p->hy : p->hz;
#define FV(o) f[G(p, dir == 0 ? (o) : i, dir == 1 ? (o) : j, dir == 2 ? (o) : k)]
    if (bound) {
        const long t = n - 1;
        if (c == 0) return (FV(1) - FV(0)) / h;
        if (c == 1) return (FV(2) - FV(0)) * 0.5 / h;
        if (c == 2) return (FV(0) - 8.0 * FV(1) + 8.0 * FV(3) - FV(4)) / (12.0 * h);
        if (c == 3) return (FV(6) - FV(0) - 9.0 * (FV(5) - FV(1)) + 45.0 * (FV(4) - FV(2))) / (60.0 * h);
        if (c == t) return (FV(t) - FV(t - 1)) / h;
        if (c == t - 1) return (FV(t) - FV(t - 2)) * 0.5 / h;
        if (c == t - 2) return (FV(t - 4) - 8.0 * FV(t - 3) + 8.0 * FV(t - 1) - FV(t)) / (12.0 * h);
        if (c == t - 3) return (FV(t) - FV(t - 6) - 9.0 * (FV(t - 1) - FV(t - 5)) + 45.0 * (FV(t - 2) - FV(t - 4))) / (60.0 * h);
    }
    return (45.0 * (FV(c + 1) - FV(c - 1)) - 9.0 * (FV(c + 2) - FV(c - 2)) + (FV(c + 3) - FV(c - 3))) / (60.0 * h);
#undef FV
}

/* Metric terms from ghosted mesh coordinates xyz[0..2] (face ghosts filled as the caller's
 * exchange left them): out[0..8] = Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz and
 * out[9] = Ajac (ghosted fields).  Interior cells take the cofactors of d(x,y,z)/d(i,j,k)
 * and Ajac = 1/det (comput_Jac3d_kernal); each face's ghost layers copy the adjacent
 * interior cell; then the periodic face halo overwrites the periodic directions
 * (exchange_boundary_xyz_packed_dev in Comput_Jacobian3d). */
void orc_comput_jacobian(const orc_params *p, const double *const *xyz, const int jacbound[3], double *const *out)
{
    const long nx = p->nx, ny = p->ny, nz = p->nz;
    long i, j, k;
    int m, q;
#pragma omp parallel for private(j, i, m, q) collapse(2)
    for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                double dv[3][3];   /* dv[c][d]: derivative of coordinate c along direction d */
                for (m = 0; m < 3; m++)
                    for (q = 0; q < 3; q++) dv[m][q] = metric_deriv(p, xyz[m], q, jacbound[q], i, j, k);
                const double xi1 = dv[0][0], xj1 = dv[0][1], xk1 = dv[0][2];
                const double yi1 = dv[1][0], yj1 = dv[1][1], yk1 = dv[1][2];
                const double zi1 = dv[2][0], zj1 = dv[2][1], zk1 = dv[2][2];
                double v[10];
                v[9] = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                              xi1 * zj1 * yk1);
                v[0] = (yj1 * zk1 - zj1 * yk1);
                v[1] = (zj1 * xk1 - xj1 * zk1);
                v[2] = (xj1 * yk1 - yj1 * xk1);
                v[3] = (yk1 * zi1 - zk1 * yi1);
                v[4] = (zk1 * xi1 - xk1 * zi1);
                v[5] = (xk1 * yi1 - yk1 * xi1);
                v[6] = (yi1 * zj1 - zi1 * yj1);
                v[7] = (zi1 * xj1 - xi1 * zj1);
                v[8] = (xi1 * yj1 - yi1 * xj1);
                for (m = 0; m < 10; m++) {
                    double *o = out[m];
                    long g;
                    o[G(p, i, j, k)] = v[m];
                    if (i == 0) for (g = 1; g <= LAP; g++) o[G(p, -g, j, k)] = v[m];
                    if (i == nx - 1) for (g = 1; g <= LAP; g++) o[G(p, nx - 1 + g, j, k)] = v[m];
                    if (j == 0) for (g = 1; g <= LAP; g++) o[G(p, i, -g, k)] = v[m];
                    if (j == ny - 1) for (g = 1; g <= LAP; g++) o[G(p, i, ny - 1 + g, k)] = v[m];
                    if (k == 0) for (g = 1; g <= LAP; g++) o[G(p, i, j, -g)] = v[m];
                    if (k == nz - 1) for (g = 1; g <= LAP; g++) o[G(p, i, j, nz - 1 + g)] = v[m];
                }
            }
    for (m = 0; m < 10; m++) orc_ghost_fill_periodic(p, out[m]);
}
