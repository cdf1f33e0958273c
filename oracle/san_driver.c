/*
 * san_driver.c -- runs every entry point of the CPU oracle once on small, ragged grids
 * with exactly sized heap buffers, for a build under AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle san_driver`).  Any out-of-bounds access,
 * use of uninitialised heap memory the sanitizers see, or undefined arithmetic aborts
 * the run; a clean run prints a checksum and "ok".
 *
 * TEST INFRASTRUCTURE ONLY (the oracle's own check); nothing in the library uses it.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ws_oracle.h"

static double checksum;

static double *alloc(size_t n)
{
    double *a = (double *)calloc(n, sizeof(double));
    if (!a) {
        fprintf(stderr, "allocation of %zu doubles failed\n", n);
        exit(2);
    }
    return a;
}

static void add_sum(const double *a, size_t n)
{
    for (size_t i = 0; i < n; ++i) checksum += a[i];
}

static orc_params params(int nx, int ny, int nz)
{
    orc_params p;
    memset(&p, 0, sizeof p);
    p.nx = nx; p.ny = ny; p.nz = nz;
    p.gamma = 1.4; p.ama = 1.0; p.epsl_sw = 1e-3;
    p.hx = 1.0 / (nx - 1); p.hy = 1.0 / (ny - 1); p.hz = 1.0 / (nz - 1);
    p.nthreads = 1;
    p.re = 1000.0; p.pr = 0.72; p.ref_t = 288.15;
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1;
    return p;
}

/* smooth positive primitives on the whole ghosted box */
static void fill_prims(const orc_params *p, double *d, double *u, double *v, double *w, double *T)
{
    const int nxg = p->nx + 8, nyg = p->ny + 8, nzg = p->nz + 8;
    for (int k = 0; k < nzg; ++k)
        for (int j = 0; j < nyg; ++j)
            for (int i = 0; i < nxg; ++i) {
                const size_t q = (size_t)i + (size_t)nxg * ((size_t)j + (size_t)nyg * (size_t)k);
                const double x = 0.37 * i, y = 0.53 * j, z = 0.71 * k;
                d[q] = 1.0 + 0.1 * sin(x + y);
                u[q] = 0.5 + 0.2 * cos(y - z);
                v[q] = -0.3 + 0.1 * sin(x * z * 0.1);
                w[q] = 0.2 * cos(x + z);
                T[q] = 1.0 + 0.05 * sin(x - y + z);
            }
}

static void run_case(int nx, int ny, int nz, int periodic, int viscous)
{
    orc_params p = params(nx, ny, nz);
    p.viscous = viscous;
    p.bc_planes = !periodic;
    if (!periodic) {
        p.periodic[0] = 0; p.periodic[1] = 1; p.periodic[2] = 0;
        for (int f = 0; f < 6; ++f) p.d0_bound[f] = (f / 2 != 1);
        p.non_ref[1] = 1; p.non_ref[4] = 1;
    }
    const size_t ng = (size_t)(nx + 8) * (ny + 8) * (nz + 8);
    const size_t ni = (size_t)nx * ny * nz;
    double *d = alloc(ng), *u = alloc(ng), *v = alloc(ng), *w = alloc(ng), *T = alloc(ng);
    double *P = alloc(ng), *cc = alloc(ng), *ajac = alloc(ng);
    double *met[9];
    for (int m = 0; m < 9; ++m) met[m] = alloc(ng);
    for (size_t q = 0; q < ng; ++q) {
        ajac[q] = 1.0;
        met[0][q] = met[4][q] = met[8][q] = 1.0;
    }
    double *f = alloc(5 * ni), *fn = alloc(5 * ni), *du = alloc(5 * ni);
    fill_prims(&p, d, u, v, w, T);
    const int s0[3] = {0, 0, 0}, e0[3] = {nx, ny, nz};
    orc_pri_to_cons_box(&p, f, d, u, v, w, T, s0, e0);
    orc_cons_to_pri(&p, f, d, u, v, w, T, P);
    double *fl[5] = {d, u, v, w, T};
    for (int m = 0; m < 5; ++m) orc_ghost_fill_periodic(&p, fl[m]);
    orc_sound_speed(&p, T, cc);
    orc_inviscid_rhs(&p, d, u, v, w, cc, (const double *const *)met, ajac, du);
    add_sum(du, 5 * ni);
    orc_du_comput(&p, d, u, v, w, T, cc, (const double *const *)met, ajac, du);
    add_sum(du, 5 * ni);
    memcpy(fn, f, 5 * ni * sizeof(double));
    for (int krk = 1; krk <= 3; ++krk) orc_time_advance(&p, krk, f, fn, du, 1e-4);
    orc_bc_planes(&p, f, d, u, v, w, T);
    checksum += orc_cfl_dt(&p, 0.5, u, v, w, T);
    orc_cons_to_pri(&p, f, d, u, v, w, T, P);
    orc_rk3_run(&p, 2, -0.5, f, d, u, v, w, T);
    add_sum(f, 5 * ni);
    if (viscous) {
        double *amu = alloc(ni);
        orc_amu(&p, p.re, p.ref_t, T, amu);
        add_sum(amu, ni);
        free(amu);
    }
    free(d); free(u); free(v); free(w); free(T); free(P); free(cc); free(ajac);
    for (int m = 0; m < 9; ++m) free(met[m]);
    free(f); free(fn); free(du);
}

static void run_metrics(int nx, int ny, int nz)
{
    orc_params p = params(nx, ny, nz);
    p.periodic[0] = 0; p.periodic[1] = 1; p.periodic[2] = 0;
    const size_t ng = (size_t)(nx + 8) * (ny + 8) * (nz + 8);
    double *xyz[3], *out[10];
    for (int c = 0; c < 3; ++c) xyz[c] = alloc(ng);
    for (int m = 0; m < 10; ++m) out[m] = alloc(ng);
    const int nxg = nx + 8, nyg = ny + 8, nzg = nz + 8;
    for (int k = 0; k < nzg; ++k)
        for (int j = 0; j < nyg; ++j)
            for (int i = 0; i < nxg; ++i) {
                const size_t q = (size_t)i + (size_t)nxg * ((size_t)j + (size_t)nyg * (size_t)k);
                const double a = (i - 4) * p.hx, b = (j - 4) * p.hy, c = (k - 4) * p.hz;
                xyz[0][q] = a + 0.05 * sin(6.0 * b);
                xyz[1][q] = b + 0.05 * sin(6.0 * c);
                xyz[2][q] = c + 0.05 * sin(6.0 * a);
            }
    const int jacbound[3] = {1, 0, 1};
    orc_comput_jacobian(&p, (const double *const *)xyz, jacbound, out);
    for (int m = 0; m < 10; ++m) add_sum(out[m], ng);
    for (int c = 0; c < 3; ++c) free(xyz[c]);
    for (int m = 0; m < 10; ++m) free(out[m]);
}

static void run_scalars(void)
{
    double s8[8], s9[9], S[5], wts[5], q[5], fp[5], fm[5];
    for (int i = 0; i < 9; ++i) {
        s9[i] = 1.0 + 0.1 * i * i;
        if (i < 8) s8[i] = s9[i];
    }
    checksum += orc_weno7_symbo_p(s8) + orc_weno7_symbo_m(s8) + orc_cd6(s9) + orc_cd8(s9);
    orc_weno7_symbo_p_parts(s8, S, wts, q);
    orc_split_sw(1.0, 0.5, 0.1, -0.2, 1.1, 1.0, 0.0, 0.0, 1.4, 1e-3, fp, fm);
    for (int k = 0; k < 5; ++k) checksum += S[k] + wts[k] + q[k] + fp[k] + fm[k];
    /* one-sided closures on a line of 10 cells with LAP ghosts either side */
    const long n = 10;
    double *line = alloc((size_t)n + 8);
    for (long i = 0; i < n + 8; ++i) line[i] = 1.0 + 0.01 * i * i;
    for (long face = 0; face <= n; ++face) {
        double val;
        if (orc_closure_p(line + 4, n, face, 1, 1, &val)) checksum += val;
        if (orc_closure_m(line + 4, n, face, 1, 1, &val)) checksum += val;
    }
    free(line);
}

int main(void)
{
    run_scalars();
    run_case(12, 9, 7, 1, 0);
    run_case(12, 9, 7, 1, 6);
    run_case(10, 11, 9, 1, 8);
    run_case(12, 9, 10, 0, 0);
    run_metrics(12, 9, 10);
    if (!isfinite(checksum)) {
        fprintf(stderr, "non-finite checksum\n");
        return 1;
    }
    printf("checksum %.17g\nok\n", checksum);
    return 0;
}
