// opcount.cpp -- algorithmic FLOP count of the inviscid RHS, per cell.
//
// TEST INFRASTRUCTURE ONLY (part of the CPU oracle).  Compiles the oracle's
// cell formulas (ws_oracle_cell.h) over a counting scalar and prints, as JSON,
// the number of floating-point operations in each piece and per cell.
// Convention: add, sub, mul, div, sqrt and max count 1 each; a sign flip
// counts 0; there are no fused multiply-adds in the reference-form formulas.
#include <cmath>
#include <cstdio>

static long g_ops = 0;

struct Counted {
    double v;
    Counted() : v(0.0) {}
    Counted(double x) : v(x) {}
};
static inline Counted operator+(Counted a, Counted b) { ++g_ops; return Counted(a.v + b.v); }
static inline Counted operator-(Counted a, Counted b) { ++g_ops; return Counted(a.v - b.v); }
static inline Counted operator*(Counted a, Counted b) { ++g_ops; return Counted(a.v * b.v); }
static inline Counted operator/(Counted a, Counted b) { ++g_ops; return Counted(a.v / b.v); }
static inline Counted operator-(Counted a) { return Counted(-a.v); }
static inline Counted &operator+=(Counted &a, Counted b) { ++g_ops; a.v += b.v; return a; }
static inline Counted &operator/=(Counted &a, Counted b) { ++g_ops; a.v /= b.v; return a; }
static inline Counted csqrt(Counted a) { ++g_ops; return Counted(std::sqrt(a.v)); }
static inline Counted cfmax(Counted a, Counted b) { ++g_ops; return Counted(a.v > b.v ? a.v : b.v); }

typedef Counted orc_real;
#define ORC_REAL_DEFINED 1
#define ORC_SQRT(x) csqrt(x)
#define ORC_FMAX(a, b) cfmax((a), (b))
#include "ws_oracle_cell.h"

int main()
{
    orc_real s[8];
    for (int i = 0; i < 8; ++i) s[i] = Counted(1.0 + 0.1 * i * i);

    g_ops = 0; (void)orc_cell_weno7_symbo_p(s); const long weno_p = g_ops;
    g_ops = 0; (void)orc_cell_weno7_symbo_m(s); const long weno_m = g_ops;

    orc_real fp[5], fm[5];
    g_ops = 0;
    orc_cell_split_sw(Counted(1.0), Counted(0.5), Counted(0.1), Counted(0.2), Counted(1.2),
                      Counted(1.0), Counted(0.0), Counted(0.0), 1.4, 1e-3, 0.8, 2.0, fp, fm);
    const long split = g_ops;

    g_ops = 0; (void)orc_cell_flux_diff(Counted(0.0), Counted(1.0), Counted(2.0), Counted(1.0), 0.1);
    const long fluxdiff = g_ops;

    orc_real out[6];
    g_ops = 0; orc_cell_cons_to_pri(Counted(1.0), Counted(0.1), Counted(0.2), Counted(0.3), Counted(2.5), 1.0, 1.4, out);
    const long c2p = g_ops;
    g_ops = 0; (void)orc_cell_sound_speed(Counted(1.0), 1.0); const long snd = g_ops;

    // per cell and RHS: three directions, each one split, five P and five M
    // reconstructions (one face per cell per quantity) and ten flux differences
    const long per_dir = split + 5 * (weno_p + weno_m) + 10 * fluxdiff;
    const long rhs = 3 * per_dir;
    const long prep = c2p + snd;
    // viscous terms per cell (orc_viscous_rhs): the viscosity, twelve gradients (central
    // form / h), the stress and heat flux and the four fluxes formed for each of the three
    // directions, then twelve flux derivatives (/ h), each added to du as du += Ajac * d
    g_ops = 0; (void)orc_cell_amu(Counted(1.2), 1.4, 0.38); const long amu = g_ops;
    orc_real st[9], gr[12], am[9], ev[4];
    for (int i = 0; i < 9; ++i) { st[i] = Counted(0.3 * i * i); am[i] = Counted(0.1 + i); }
    for (int i = 0; i < 12; ++i) gr[i] = Counted(0.2 * i - 1.0);
    g_ops = 0; (void)orc_cell_cd6(st); const long cd6 = g_ops + 1;     // + the division by h
    g_ops = 0; (void)orc_cell_cd8(st); const long cd8 = g_ops + 1;
    g_ops = 0;
    orc_cell_vis_flux(gr, Counted(0.5), Counted(0.1), Counted(0.2), Counted(0.01), 3.5, am, Counted(1.0), 1, ev);
    const long visflux = g_ops;
    const long vis6 = amu + 12 * cd6 + 3 * visflux + 12 * (cd6 + 2);
    const long vis8 = amu + 12 * cd8 + 3 * visflux + 12 * (cd8 + 2);
    // the reference forms each weight from four (2.592e-6 + S_j) sums, 20 per reconstruction
    // (OCFD_Schemes.cpp:344-352); five suffice when each sum is formed once
    const long shared_eps = rhs - 3 * 10 * (20 - 5);
    std::printf("{\"weno_p\": %ld, \"weno_m\": %ld, \"split\": %ld, \"flux_diff\": %ld, "
                "\"cons_to_pri\": %ld, \"sound_speed\": %ld, \"per_direction\": %ld, "
                "\"ALGO_FLOPS_PER_CELL\": %ld, \"ALGO_FLOPS_PER_CELL_SHARED_EPS\": %ld, "
                "\"prep_flops_per_cell\": %ld, \"vis_amu\": %ld, \"vis_cd6\": %ld, \"vis_cd8\": %ld, "
                "\"vis_flux_per_direction\": %ld, \"VISCOUS_FLOPS_PER_CELL_CD6\": %ld, "
                "\"VISCOUS_FLOPS_PER_CELL_CD8\": %ld}\n",
                weno_p, weno_m, split, fluxdiff, c2p, snd, per_dir, rhs, shared_eps, prep, amu, cd6, cd8, visflux,
                vis6, vis8);
    return 0;
}
