// rhs_tile.h -- the LDS-tiled inviscid RHS kernel template, shared by the FP64 path
// (k_rhs_tiled.hip) and the mixed-precision path (k_rhs_mixed.hip).
//
// One launch per direction.  A block owns a plane tile of W lines of L cells:
//   x-direction: lines along x (rows), W rows along y, at one z;
//   y-direction: lines along y, W columns along x, at one z;
//   z-direction: lines along z, W columns along x, at one y.
// Phase 1 splits every point of the tile's lines, halo included (L+8 points per
// line), into LDS: 10 doubles per point.  Phase 2 gives each thread K consecutive
// faces of one line; for every variable it reconstructs them with the positive and
// the negative rule (sharing differences between the K faces) and merges each pair.
// Phase 3 turns face fluxes into cell updates: the first of the three launches writes
// du, the other two accumulate into it, and in an RK3 step the last one may instead
// apply the stage to f (STG).
// MIX selects the reconstruction: FP64 (weno_k, the pair merged over one reciprocal)
// or FP32 relative to FP64 base points (weno_k_mixed); splitting, the merge and the
// flux difference are FP64 either way.
// The shapes are template parameters; the library picks one per grid (autotune.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "ws_device.h"
#include "ws_tiles.h"

namespace ws {

namespace {


struct Box {
    int s[3];   // start (ghosted coordinates)
    int e[3];   // end
};

template <int DIR, int L, int W, int K>
struct Tile {
    static constexpr int NL = L + 2 * LAP;                // points per line
    static constexpr int F = L + 1;                        // faces per line
    static constexpr int T = (F + K - 1) / K;              // threads per line
    static constexpr int NT = W * T;                       // threads per block
    // every tile stores its lines line-major (point p of line l at l * LS + p), so a
    // thread's window is contiguous and is read with 128-bit LDS loads.  x tiles fill
    // their lines along x (LS = NL); y/z tiles fill them across lines (consecutive
    // threads write the same point of consecutive lines), so their line stride is the
    // smallest LS >= NL with LS = 2 mod 4 doubles: those writes spread over the banks
    // and every line still starts 16-byte aligned
    static constexpr int LS = DIR == 0 ? NL : NL + (6 - NL % 4) % 4;
    static constexpr int QS = W * LS;                      // LDS entries per quantity
    static constexpr int NPTS = W * NL;
    static constexpr int ROUNDS = (NPTS + NT - 1) / NT;   // split points per thread (upper bound)
    // y/z tiles of 16 columns start on a multiple of 16 ghosted columns, so that their
    // rows begin on 128-byte lines of the fields (rows of the reference layout are
    // line-aligned when the pitch is a multiple of 16 elements)
    static constexpr int XALIGN = tile_xalign(DIR, W);
    // the last thread of a line may read up to K+6 points past its first face, i.e.
    // past the line's end (into the next line, or for the last line past the buffer);
    // the pad keeps those reads inside the buffer (their values only reach faces that
    // are never stored)
    static constexpr int PAD = K + 8;
    static_assert(NT <= 320, "block too large");
    static_assert(5 * W * F <= 10 * QS, "face buffer must fit in the split buffer");
    // LDS index of point p (0..NL-1) of line l (0..W-1)
    __device__ static int pidx(int l, int p) { return l * LS + p; }
    // face buffer index of face f of line l.  The flux difference reads, per lane, the two
    // faces of one cell, with consecutive lanes on consecutive cells of a line in x tiles
    // and on consecutive lines in y/z tiles (so that their du rows stay coalesced): x
    // tiles store faces line-major, y/z tiles face-major, and either way the lanes of
    // one read touch consecutive doubles (a line-major y/z buffer with F = L + 1 put the
    // lanes (l, i) and (l + 1, i - 1) on one bank).
    __device__ static int fidx(int l, int f) { return DIR == 0 ? l * F + f : f * W + l; }
};

// Where a tile lies: x0 always tiles x; the line axis is x (DIR 0), y (DIR 1) or z
// (DIR 2).  Cells of the tile inside the box: lim_line along the line, lines
// lo_across .. lim_across-1 across (only the first aligned y/z tile has lo_across > 0).
struct TileGeo {
    int x0, y0, z0;
    int lim_line, lim_across, lo_across;
};

template <int DIR, int L, int W, int K>
__device__ __forceinline__ TileGeo tile_geo(const Box &box, int bx, int by, int bz)
{
    using TL = Tile<DIR, L, W, K>;
    TileGeo g;
    if (DIR == 0) {
        g.x0 = box.s[0] + bx * L;
        g.y0 = box.s[1] + by * W;
        g.z0 = box.s[2] + bz;
    } else if (DIR == 1) {
        g.x0 = box.s[0] / TL::XALIGN * TL::XALIGN + bx * W;
        g.y0 = box.s[1] + by * L;
        g.z0 = box.s[2] + bz;
    } else {
        g.x0 = box.s[0] / TL::XALIGN * TL::XALIGN + bx * W;
        g.z0 = box.s[2] + by * L;
        g.y0 = box.s[1] + bz;
    }
    g.lim_line = DIR == 0 ? box.e[0] - g.x0 : DIR == 1 ? box.e[1] - g.y0 : box.e[2] - g.z0;
    g.lim_across = DIR == 0 ? box.e[1] - g.y0 : box.e[0] - g.x0;
    g.lo_across = DIR == 0 ? 0 : box.s[0] - g.x0;
    return g;
}

// The primitive values (rho, u, v, w, c) of this thread's phase-1 points.
template <int DIR, int L, int W, int K>
struct TileLoads {
    double v[Tile<DIR, L, W, K>::ROUNDS][5];
};

// point pt of the tile: line l, point p (x fastest in memory for every direction)
template <int DIR, int L, int W, int K>
__device__ __forceinline__ void tile_point_at(int pt, int &l, int &p)
{
    using TL = Tile<DIR, L, W, K>;
    if (DIR == 0) { p = pt % TL::NL; l = pt / TL::NL; }
    else          { l = pt % W;      p = pt / W; }
}

// point r of this thread
template <int DIR, int L, int W, int K>
__device__ __forceinline__ void tile_point(int r, int &l, int &p)
{
    tile_point_at<DIR, L, W, K>((int)threadIdx.x + r * Tile<DIR, L, W, K>::NT, l, p);
}

template <int DIR, int L, int W, int K>
__device__ __forceinline__ bool tile_point_ok(const TileGeo &g, int r, int l, int p)
{
    using TL = Tile<DIR, L, W, K>;
    // a line needs points -LAP .. lim_line+LAP-1 of lines lo_across .. lim_across-1
    return (int)threadIdx.x + r * TL::NT < TL::NPTS && l >= g.lo_across && l < g.lim_across &&
           p < g.lim_line + 2 * LAP;
}

template <int DIR>
__device__ __forceinline__ void tile_point_xyz(const TileGeo &g, int l, int p, int &x, int &y, int &z)
{
    if (DIR == 0)      { x = g.x0 - LAP + p; y = g.y0 + l; z = g.z0; }
    else if (DIR == 1) { x = g.x0 + l; y = g.y0 - LAP + p; z = g.z0; }
    else               { x = g.x0 + l; y = g.y0; z = g.z0 - LAP + p; }
}

// Cell coordinates of phase-3 cell c (0 .. W*L-1 in the tile's cell order) for loads that
// every lane issues: clamped to the tile's last cell and, across and along the line, to
// its cells inside the box (the values of a clamped lane are never used).
template <int DIR, int L, int W, int K>
__device__ __forceinline__ void tile_cell_clamped(const TileGeo &g, int c, int &cx, int &cy, int &cz)
{
    c = min(c, W * L - 1);
    int l, i;
    if (DIR == 0) { i = c % L; l = c / L; }
    else          { l = c % W; i = c / W; }
    l = min(max(l, g.lo_across), min(g.lim_across, W) - 1);
    i = min(i, min(g.lim_line, L) - 1);
    if (DIR == 0)      { cx = g.x0 + i; cy = g.y0 + l; cz = g.z0; }
    else if (DIR == 1) { cx = g.x0 + l; cy = g.y0 + i; cz = g.z0; }
    else               { cx = g.x0 + l; cy = g.y0; cz = g.z0 + i; }
}

// Phase 1a: issue the loads of this thread's points (nothing waits on them here).  Every
// lane issues every round's loads; a lane without a point reads one of the tile instead
// (past the tile's points the last one, outside the box the nearest; the value is never
// used).  With no lane-dependent branch around them the number of loads in flight is the
// same on every path, so the split of round r waits for its own loads only while the
// later rounds' (and an accumulating launch's du) loads are still in flight.
template <int DIR, int L, int W, int K>
__device__ __forceinline__ void tile_load(int nyg, const ws_field &d, const ws_field &u, const ws_field &v,
                                          const ws_field &w, const ws_field &cc, const TileGeo &g,
                                          TileLoads<DIR, L, W, K> &ld, int off32)
{
    using TL = Tile<DIR, L, W, K>;
    const int lhi = min(g.lim_across, W) - 1, phi = min(g.lim_line + 2 * LAP, TL::NL) - 1;
    if (off32) {
        // one 32-bit offset for the five fields (KParams::off32: a uniform branch)
#pragma unroll
        for (int r = 0; r < TL::ROUNDS; ++r) {
            int l, p, x, y, z;
            tile_point_at<DIR, L, W, K>(min((int)threadIdx.x + r * TL::NT, TL::NPTS - 1), l, p);
            l = min(max(l, g.lo_across), lhi);
            p = min(p, phi);
            tile_point_xyz<DIR>(g, l, p, x, y, z);
            const unsigned o = (unsigned)x + __umul24(d.pitch, (unsigned)y + __umul24((unsigned)nyg, (unsigned)z));
            ld.v[r][0] = d.ptr[o];
            ld.v[r][1] = u.ptr[o];
            ld.v[r][2] = v.ptr[o];
            ld.v[r][3] = w.ptr[o];
            ld.v[r][4] = cc.ptr[o];
            __builtin_amdgcn_sched_barrier(0);
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < TL::ROUNDS; ++r) {
        int l, p, x, y, z;
        tile_point_at<DIR, L, W, K>(min((int)threadIdx.x + r * TL::NT, TL::NPTS - 1), l, p);
        l = min(max(l, g.lo_across), lhi);
        p = min(p, phi);
        tile_point_xyz<DIR>(g, l, p, x, y, z);
        ld.v[r][0] = d.ptr[fidx(d.pitch, nyg, x, y, z)];
        ld.v[r][1] = u.ptr[fidx(u.pitch, nyg, x, y, z)];
        ld.v[r][2] = v.ptr[fidx(v.pitch, nyg, x, y, z)];
        ld.v[r][3] = w.ptr[fidx(w.pitch, nyg, x, y, z)];
        ld.v[r][4] = cc.ptr[fidx(cc.pitch, nyg, x, y, z)];
        // rounds go out in order (the scheduler would otherwise issue them in any order,
        // and the wait for round 0 would then cover whatever went out before it)
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Phases 1b-3 of one tile with the block's LDS buffer sp (10 * QS + PAD doubles):
// split the loaded points into LDS, call next() (the persistent kernels issue the next
// tile's loads there, into the registers just consumed), reconstruct the faces, then
// write the flux difference of every cell.  Every thread of the block calls it.
// With STG the launch is the last of the three and applies the RK3 stage of `sa` to its
// cells: f = stage(fn, f, du + this direction's terms), du itself is not written.
// Returns, when the launch applies a stage with sa.cfl_part set, the maximum CFL speed of
// this thread's cells (0 otherwise).
template <int DIR, int L, int W, int K, bool FIELDS, bool ACC, bool MIX, bool BND, bool STG, typename Next>
__device__ __forceinline__ double tile_compute(const KParams &k, const ws_metric_fields &mf, const ws_soa &du,
                                               const TileGeo &g, TileLoads<DIR, L, W, K> &ld, double *sp,
                                               const StageArgs &sa, Next next)
{
    static_assert(!(STG && (BND || !ACC)), "the stage is applied by an accumulating launch without closures");
    using TL = Tile<DIR, L, W, K>;
    const int t = threadIdx.x;
    const double cA1 = k.A[3 * DIR + 0], cA2 = k.A[3 * DIR + 1], cA3 = k.A[3 * DIR + 2];
    const double cSn = k.Anorm[DIR];
    const double cK1 = k.Ak[3 * DIR + 0], cK2 = k.Ak[3 * DIR + 1], cK3 = k.Ak[3 * DIR + 2];

    // ---- phase 3 prologue: accumulating launches read du; those loads go out with the
    // tile's primitive loads, so one wait at the split covers both (issued later, the
    // wait the compiler places before the reconstruction exposes their whole latency)
    constexpr int CPT = (W * L + TL::NT - 1) / TL::NT;    // cells per thread (upper bound)
    const size_t vstride = (size_t)du.pitch * k.ny * k.nz;
    double duo[ACC ? CPT : 1][5];
    if (ACC) {
        // issued by every lane, like the primitive loads, so that the split does not wait
        // for them
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
            int cx, cy, cz;
            tile_cell_clamped<DIR, L, W, K>(g, t + r * TL::NT, cx, cy, cz);
            const double *dp = du.ptr + (size_t)(cx - LAP) +
                               (size_t)du.pitch * ((size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP));
#pragma unroll
            for (int m = 0; m < 5; ++m) duo[r][m] = dp[m * vstride];
        }
    }

    // ---- phase 1b: split every point of the tile into LDS --------------------------------
#pragma unroll
    for (int r = 0; r < TL::ROUNDS; ++r) {
        int l, p;
        tile_point<DIR, L, W, K>(r, l, p);
        if (!tile_point_ok<DIR, L, W, K>(g, r, l, p)) continue;
        const double dd = ld.v[r][0], uu = ld.v[r][1], vv = ld.v[r][2], ww = ld.v[r][3], c = ld.v[r][4];
        double fp[5], fm[5];
        if (FIELDS) {
            int x, y, z;
            tile_point_xyz<DIR>(g, l, p, x, y, z);
            const int nyg = k.nyg;
            const ws_field &f1 = mf.A[3 * DIR + 0], &f2 = mf.A[3 * DIR + 1], &f3 = mf.A[3 * DIR + 2];
            const double A1 = f1.ptr[fidx(f1.pitch, nyg, x, y, z)];
            const double A2 = f2.ptr[fidx(f2.pitch, nyg, x, y, z)];
            const double A3 = f3.ptr[fidx(f3.pitch, nyg, x, y, z)];
            const double sn = sqrt(fma(A3, A3, fma(A2, A2, A1 * A1)));
            const double rs = 1.0 / sn;
            split_fast(k, dd, uu, vv, ww, c, A1, A2, A3, sn, A1 * rs, A2 * rs, A3 * rs, fp, fm);
        } else {
            split_fast(k, dd, uu, vv, ww, c, cA1, cA2, cA3, cSn, cK1, cK2, cK3, fp, fm);
        }
        const int li = TL::pidx(l, p);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            sp[m * TL::QS + li] = fp[m];
            sp[(5 + m) * TL::QS + li] = fm[m];
        }
    }
    next();
    __syncthreads();

    // ---- phase 2: K faces of one line per thread, all five variables --------------------
    const int line = t / TL::T, seg = t % TL::T;
    const int f0 = seg * K;                               // first face of this thread
    const bool active = (t < TL::NT) && (line >= g.lo_across) && (line < g.lim_across) && (f0 <= g.lim_line);
    double Fv[5][K];
    // one-sided closures (BND): global index of the tile's first cell along the line
    const int n_dom = DIR == 0 ? k.nx : DIR == 1 ? k.ny : k.nz;
    const int gl0 = (DIR == 0 ? g.x0 : DIR == 1 ? g.y0 : g.z0) - LAP;
    const bool lo = BND && k.bnd[2 * DIR];
    const int hi = BND ? k.bnd[2 * DIR + 1] : 0;     // the high face's closure form (0: none)
    const bool nr_lo = lo && k.nonref[2 * DIR], nr_hi = hi && k.nonref[2 * DIR + 1];
    double nfix[BND ? 5 : 1];       // P(n) - P(n-1) of the line's last cell (non-reflecting M)
    // MIX: bit m*K + j set when face j of variable m had FP32 weights out of FP32's range
    unsigned bad_mix = 0;
    if (active) {
        const int base = TL::pidx(line, f0);
        const bool near = BND && ((lo && gl0 + f0 <= 3) || (hi && gl0 + f0 + K - 1 >= n_dom - 3));
#pragma unroll 1
        for (int m = 0; m < 5; ++m) {
            const double *qp = sp + m * TL::QS + base;
            const double *qm = sp + (5 + m) * TL::QS + base;
            auto cellP = [&](int cg) { return sp[m * TL::QS + TL::pidx(line, cg - gl0 + LAP)]; };
            auto cellM = [&](int cg) { return sp[(5 + m) * TL::QS + TL::pidx(line, cg - gl0 + LAP)]; };
            // faces covered by a closure: P + M with either side replaced (FP64 closures)
            auto closed = [&](int j, auto wenoP, auto wenoM) {
                const int fg = gl0 + f0 + j;
                double pv, mv;
                const bool cp = closure_p(cellP, n_dom, fg, lo, hi, pv);
                const bool cm = closure_m(cellM, n_dom, fg, lo, hi, mv);
                if (cp || cm) Fv[m][j] = (cp ? pv : wenoP()) + (cm ? mv : wenoM());
                if (BND && nr_hi && fg == n_dom) {
                    double a = 0.0, b = 0.0;
                    closure_p(cellP, n_dom, n_dom - 1, lo, hi, a);
                    closure_p(cellP, n_dom, n_dom, lo, hi, b);
                    nfix[m] = b - a;
                }
            };
            // a stored face f (f <= lim_line) reads only points f..f+7, all written in
            // phase 1; faces past lim_line read stale LDS and are not stored
            if (MIX) {
                // FP32 reconstruction relative to FP64 base points; the merge is FP64
                double bP[K], bM[K];
                float cP[K], cM[K];
                unsigned badP, badM;
                weno_k_mixed<false, K>([&](int j) { return qp[j]; }, bP, cP, badP);
                weno_k_mixed<true, K>([&](int j) { return qm[j]; }, bM, cM, badM);
#pragma unroll
                for (int j = 0; j < K; ++j) Fv[m][j] = (bP[j] + bM[j]) + (double)(cP[j] + cM[j]);
                bad_mix |= (badP | badM) << (m * K);
                if (near)
#pragma unroll
                    for (int j = 0; j < K; ++j)
                        closed(j, [&] { return bP[j] + (double)cP[j]; }, [&] { return bM[j] + (double)cM[j]; });
            } else {
                double nP[K], dP[K], nM[K], dM[K], bP[K], bM[K];
                weno_k<false, K>([&](int j) { return qp[j]; }, nP, dP, bP);
                weno_k<true, K>([&](int j) { return qm[j]; }, nM, dM, bM);
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const double den = dP[j] * dM[j];
                    Fv[m][j] = (bP[j] + bM[j]) + fma(nP[j], dM[j], nM[j] * dP[j]) * recip_nr(den);
                }
                if (near)
#pragma unroll
                    for (int j = 0; j < K; ++j)
                        closed(j, [&] { return bP[j] + nP[j] / dP[j]; }, [&] { return bM[j] + nM[j] / dM[j]; });
            }
        }
        // Faces whose FP32 weights left FP32's range (strong discontinuities of the split
        // flux): both halves again in FP64 from the split values still in LDS, merged as the
        // FP64 kernel merges them (a half a closure covers keeps its closure value).  Done
        // after the variable loop, when the FP32 registers are dead, and skipped by every
        // wave without such a face.
        if (MIX && bad_mix) {
#pragma unroll 1
            for (int m = 0; m < 5; ++m) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    if (!((bad_mix >> (m * K + j)) & 1u)) continue;
                    const double *qp = sp + m * TL::QS + base, *qm = sp + (5 + m) * TL::QS + base;
                    double nP, dP, nM, dM, bP, bM, val;
                    const int fg = gl0 + f0 + j;
                    if (BND && ((lo && fg <= 3) || (hi && fg >= n_dom - 3))) {
                        auto cellP = [&](int cg) { return sp[m * TL::QS + TL::pidx(line, cg - gl0 + LAP)]; };
                        auto cellM = [&](int cg) { return sp[(5 + m) * TL::QS + TL::pidx(line, cg - gl0 + LAP)]; };
                        double pv, mv;
                        if (!closure_p(cellP, n_dom, fg, lo, hi, pv)) { face64<false>(qp + j, nP, dP, bP); pv = bP + nP / dP; }
                        if (!closure_m(cellM, n_dom, fg, lo, hi, mv)) { face64<true>(qm + j, nM, dM, bM); mv = bM + nM / dM; }
                        val = pv + mv;
                    } else {
                        face64<false>(qp + j, nP, dP, bP);
                        face64<true>(qm + j, nM, dM, bM);
                        val = (bP + bM) + fma(nP, dM, nM * dP) * recip_nr(dP * dM);
                    }
#pragma unroll
                    for (int q = 0; q < 5; ++q)
                        if (q == m) Fv[q][j] = val;
                }
            }
        }
    }
    __syncthreads();
    // The stage's fn and f values of this thread's phase-3 cells go out here, together:
    // loaded inside the stage arithmetic, each sat behind its own stage-number branch and
    // was waited on alone.
    double fnv[STG ? CPT : 1][5], fvv[STG ? CPT : 1][5];
    if (STG) {
        const size_t nstride = (size_t)sa.fn.pitch * k.ny * k.nz, fstride = (size_t)sa.f.pitch * k.ny * k.nz;
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
            int cx, cy, cz;
            tile_cell_clamped<DIR, L, W, K>(g, t + r * TL::NT, cx, cy, cz);   // every lane loads
            const size_t cell = (size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP);
            const double *np = sa.fn.ptr + (size_t)(cx - LAP) + (size_t)sa.fn.pitch * cell;
            const double *fp = sa.f.ptr + (size_t)(cx - LAP) + (size_t)sa.f.pitch * cell;
#pragma unroll
            for (int m = 0; m < 5; ++m) fnv[r][m] = np[m * nstride];
            if (sa.krk != 1)
#pragma unroll
                for (int m = 0; m < 5; ++m) fvv[r][m] = fp[m * fstride];
        }
    }
    // face fluxes reuse the split buffer (its first 5*W*F entries; the M split values, from
    // 5*QS on, stay intact for the non-reflecting fix of phase 3)
    double *nfix_lds = sp + 5 * W * TL::F;        // [W][5], free space below 5*QS
    if (active) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (f0 + j > g.lim_line || f0 + j >= TL::F) break;
#pragma unroll
            for (int m = 0; m < 5; ++m) sp[m * W * TL::F + TL::fidx(line, f0 + j)] = Fv[m][j];
            if (BND && nr_hi && gl0 + f0 + j == n_dom)
#pragma unroll
                for (int m = 0; m < 5; ++m) nfix_lds[line * 5 + m] = nfix[m];
        }
    }
    __syncthreads();

    // ---- phase 3: flux difference of each cell ---------------------------------------------
    const double scale0 = -k.rh[DIR];
    const double sdt = STG ? (sa.dt_dev ? *sa.dt_dev : sa.dt) : 0.0;
    double cfl_max = 0.0;
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
        const int c = t + r * TL::NT;
        if (c >= W * L) break;
        int l, i;
        if (DIR == 0) { i = c % L; l = c / L; }
        else          { l = c % W; i = c / W; }
        if (l < g.lo_across || l >= g.lim_across || i >= g.lim_line) continue;
        int cx, cy, cz;
        if (DIR == 0)      { cx = g.x0 + i; cy = g.y0 + l; cz = g.z0; }
        else if (DIR == 1) { cx = g.x0 + l; cy = g.y0 + i; cz = g.z0; }
        else               { cx = g.x0 + l; cy = g.y0; cz = g.z0 + i; }
        const double aj = FIELDS ? mf.ajac.ptr[fidx(mf.ajac.pitch, k.nyg, cx, cy, cz)] : k.ajac;
        const double scale = aj * scale0;
        double *dp = du.ptr + (size_t)(cx - LAP) +
                     (size_t)du.pitch * ((size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP));
        const int fl = TL::fidx(l, i), fr = TL::fidx(l, i + 1);
        // non-reflecting flags: the first cell takes only the M pass, the last only the P pass
        const bool first_only_m = BND && nr_lo && gl0 + i == 0;
        const bool last_only_p = BND && nr_hi && gl0 + i == n_dom - 1;
        double rv[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double val = scale * (sp[m * W * TL::F + fr] - sp[m * W * TL::F + fl]);
            if (first_only_m) {
                auto cellM = [&](int cg) { return sp[(5 + m) * TL::QS + TL::pidx(l, cg - gl0 + LAP)]; };
                double a = 0.0, b = 0.0;
                closure_m(cellM, n_dom, 0, lo, hi, a);
                closure_m(cellM, n_dom, 1, lo, hi, b);
                val = scale * (b - a);
            } else if (last_only_p) {
                val = scale * nfix_lds[l * 5 + m];
            }
            if (STG) {
                rv[m] = acc_sum(duo[r][m], val);   // the complete RHS, staged below
            } else if (ACC) {
                dp[m * vstride] = acc_sum(duo[r][m], val);
            } else {
                dp[m * vstride] = val;
            }
        }
        if (STG) {
            if (sa.krk == 1) {
#pragma unroll
                for (int m = 0; m < 5; ++m) rv[m] = rk3_stage_value<1>(fnv[r][m], 0.0, rv[m], sdt);
            } else if (sa.krk == 2) {
#pragma unroll
                for (int m = 0; m < 5; ++m) rv[m] = rk3_stage_value<2>(fnv[r][m], fvv[r][m], rv[m], sdt);
            } else {
#pragma unroll
                for (int m = 0; m < 5; ++m) rv[m] = rk3_stage_value<3>(fnv[r][m], fvv[r][m], rv[m], sdt);
            }
            // the new f's primitives for the next stage's RHS (and OCFD_bc on the planes)
            if (sa.pd.ptr) {
                const CellPrims p = cell_primitives(rv, k.cv);
                sa.pd.ptr[fidx(sa.pd.pitch, k.nyg, cx, cy, cz)] = p.d;
                sa.pu.ptr[fidx(sa.pu.pitch, k.nyg, cx, cy, cz)] = p.u;
                sa.pv.ptr[fidx(sa.pv.pitch, k.nyg, cx, cy, cz)] = p.v;
                sa.pw.ptr[fidx(sa.pw.pitch, k.nyg, cx, cy, cz)] = p.w;
                sa.pT.ptr[fidx(sa.pT.pitch, k.nyg, cx, cy, cz)] = p.T;
                // the stage kernel's expressions for the sound speed, pressure and CFL speed
                const double c = sqrt(p.T) / k.ama;
                if (sa.pcc.ptr) sa.pcc.ptr[fidx(sa.pcc.pitch, k.nyg, cx, cy, cz)] = c;
                if (sa.pP.ptr) sa.pP.ptr[fidx(sa.pP.pitch, k.nyg, cx, cy, cz)] = p.e / p.d * (k.gamma - 1.0);
                if (sa.cfl_part) cfl_max = fmax(cfl_max, fmax(fabs(p.u) + c, fmax(fabs(p.v) + c, fabs(p.w) + c)));
                const int ix = cx - LAP, iy = cy - LAP, iz = cz - LAP;
                if (sa.halo) {
                    // the ghost fill's copies of this cell: ghost g < LAP of a direction holds
                    // interior n + g, ghost n + LAP + g holds interior LAP + g (ghosted
                    // coordinates), one direction at a time (faces only); with n < 2 LAP a
                    // cell feeds both sides
                    auto put = [&](int gx, int gy, int gz) {
                        sa.pd.ptr[fidx(sa.pd.pitch, k.nyg, gx, gy, gz)] = p.d;
                        sa.pu.ptr[fidx(sa.pu.pitch, k.nyg, gx, gy, gz)] = p.u;
                        sa.pv.ptr[fidx(sa.pv.pitch, k.nyg, gx, gy, gz)] = p.v;
                        sa.pw.ptr[fidx(sa.pw.pitch, k.nyg, gx, gy, gz)] = p.w;
                        sa.pT.ptr[fidx(sa.pT.pitch, k.nyg, gx, gy, gz)] = p.T;
                        if (sa.pcc.ptr) sa.pcc.ptr[fidx(sa.pcc.pitch, k.nyg, gx, gy, gz)] = c;
                    };
                    if (ix < LAP) put(cx + k.nx, cy, cz);
                    if (ix >= k.nx - LAP) put(cx - k.nx, cy, cz);
                    if (iy < LAP) put(cx, cy + k.ny, cz);
                    if (iy >= k.ny - LAP) put(cx, cy - k.ny, cz);
                    if (iz < LAP) put(cx, cy, cz + k.nz);
                    if (iz >= k.nz - LAP) put(cx, cy, cz - k.nz);
                }
                if (sa.planes && (ix == 0 || ix == k.nx - 1 || iy == 0 || iy == k.ny - 1 || iz == 0 || iz == k.nz - 1))
                    plane_conserved(p, k.cv, rv);
            }
            const size_t fo = (size_t)(cx - LAP) +
                              (size_t)sa.f.pitch * ((size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP));
            const size_t fstride = (size_t)sa.f.pitch * k.ny * k.nz;
#pragma unroll
            for (int m = 0; m < 5; ++m) sa.f.ptr[fo + m * fstride] = rv[m];
        }
    }
    return cfl_max;
}

// *out = the maximum of v over the work-group's NTH threads and 0 (every thread calls it;
// red: NTH doubles of LDS nobody reads meanwhile).  Through LDS rather than lane shuffles:
// a work-group of 255 threads has an inactive lane whose registers a shuffle would read.
// Wavefront 0 is full (NTH >= 64) and finishes the reduction; fmax is exact and ignores
// NaN, so the order does not matter.
template <int NTH>
__device__ __forceinline__ void tile_wg_max(double v, double *red, double *out)
{
    static_assert(NTH >= 64, "the reduction needs a full first wavefront");
    red[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        double b = 0.0;
        for (int i = threadIdx.x; i < NTH; i += 64) b = fmax(b, red[i]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) b = fmax(b, __shfl_xor(b, o));
        if (threadIdx.x == 0) *out = b;
    }
}

// Grid form: one work-group per tile.  The grid is 1-D with 8 * per work-groups and
// work-group b runs tile (b % 8) * per + b / 8 (x tiles fastest): under the round-robin
// dispatch of work-groups to the eight XCDs each XCD gets a contiguous range of tiles, so
// neighbouring tiles, which share halo lines and the cache lines of their edge columns,
// meet in the same L2.
template <int DIR, int L, int W, int K, bool FIELDS, bool ACC, bool MIX, bool BND, bool STG = false>
__global__ void __launch_bounds__(320)
k_rhs_tile(KParams k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_metric_fields mf,
           ws_soa du, Box box, int gx, int gy, int ntiles, StageArgs sa)
{
    using TL = Tile<DIR, L, W, K>;
    __shared__ double sp[10 * TL::QS + TL::PAD];
    const int per = (ntiles + 7) / 8;
    const int tile = ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
    if (tile >= ntiles) return;
    const int r = tile / gx;
    const TileGeo g = tile_geo<DIR, L, W, K>(box, tile - r * gx, r % gy, r / gy);
    TileLoads<DIR, L, W, K> ld;
    tile_load<DIR, L, W, K>(k.nyg, d, u, v, w, cc, g, ld, k.off32);
    const double cmax = tile_compute<DIR, L, W, K, FIELDS, ACC, MIX, BND, STG>(k, mf, du, g, ld, sp, sa, [] {});
    if (STG && sa.cfl_part) {
        __syncthreads();   // phase 3 has read its faces from sp
        tile_wg_max<TL::NT>(cmax, sp, sa.cfl_part + tile);
    }
}

// Persistent form for one box: the resident work-groups each walk a contiguous run of
// the tiles in launch order (x tiles fastest), so neighbouring tiles, which share halo
// lines, run on the same work-group and XCD.  With PREFETCH the next tile's loads are
// issued as soon as this tile's split has consumed the load registers.
template <int DIR, int L, int W, int K, bool ACC, bool MIX, bool PREFETCH>
__global__ void __launch_bounds__(320)
k_rhs_persist(KParams k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du, Box box,
              int gx, int gy, int ntiles, int chunk)
{
    using TL = Tile<DIR, L, W, K>;
    __shared__ double sp[10 * TL::QS + TL::PAD];
    const int first = blockIdx.x * chunk;
    const int last = min(first + chunk, ntiles);
    if (first >= last) return;
    const ws_metric_fields nomf{};
    auto geo = [&](int i) {
        const int r = i / gx;
        return tile_geo<DIR, L, W, K>(box, i - r * gx, r % gy, r / gy);
    };
    TileLoads<DIR, L, W, K> ld;
    TileGeo g = geo(first);
    if (PREFETCH) tile_load<DIR, L, W, K>(k.nyg, d, u, v, w, cc, g, ld, k.off32);
    for (int i = first; i < last; ++i) {
        if (!PREFETCH) tile_load<DIR, L, W, K>(k.nyg, d, u, v, w, cc, g, ld, k.off32);
        const bool more = i + 1 < last;
        const TileGeo gn = more ? geo(i + 1) : g;
        auto next = [&] {
            if (PREFETCH && more) tile_load<DIR, L, W, K>(k.nyg, d, u, v, w, cc, gn, ld, k.off32);
        };
        tile_compute<DIR, L, W, K, false, ACC, MIX, false, false>(k, nomf, du, g, ld, sp, StageArgs{}, next);
        __syncthreads();   // the next tile reuses the LDS buffer
        g = gn;
    }
}

// Batched form: the tiles of one direction over every block of a batch, in one
// persistent launch.  Work-group g walks the contiguous range [g*chunk, (g+1)*chunk)
// of the tile list, so neighbouring tiles (which share halo lines) run on the same
// work-group, hence the same XCD and L2.  Each block brings its own fields, extents
// and spacing; the physics constants and metrics are the batch's.
template <int DIR, int L, int W, int K, bool ACC, bool MIX, bool PREFETCH>
__global__ void __launch_bounds__(320)
k_rhs_batch(KParams k, const BatchBlock *blocks, const int4 *tiles, int ntiles, int chunk)
{
    using TL = Tile<DIR, L, W, K>;
    __shared__ double sp[10 * TL::QS + TL::PAD];
    const int first = blockIdx.x * chunk;
    const int last = min(first + chunk, ntiles);
    if (first >= last) return;
    const ws_metric_fields nomf{};
    auto box_of = [](const BatchBlock &b) {
        Box box;
        box.s[0] = box.s[1] = box.s[2] = LAP;
        box.e[0] = b.n[0] + LAP; box.e[1] = b.n[1] + LAP; box.e[2] = b.n[2] + LAP;
        return box;
    };
    // tile entries are uniform across the work-group: keep them (and the geometry
    // derived from them) in scalar registers
    auto tile_at = [&](int i) {
        const int4 v = tiles[i];
        return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                         __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
    };
    TileLoads<DIR, L, W, K> ld;
    int4 tl = tile_at(first);
    TileGeo g = tile_geo<DIR, L, W, K>(box_of(blocks[tl.x]), tl.y, tl.z, tl.w);
    if (PREFETCH) {
        const BatchBlock &b = blocks[tl.x];
        tile_load<DIR, L, W, K>(b.n[1] + 2 * LAP, b.d, b.u, b.v, b.w, b.cc, g, ld, k.off32);
    }
    for (int i = first; i < last; ++i) {
        const BatchBlock &b = blocks[tl.x];
        KParams kb = k;
        kb.nx = b.n[0]; kb.ny = b.n[1]; kb.nz = b.n[2];
        kb.nyg = b.n[1] + 2 * LAP; kb.nzg = b.n[2] + 2 * LAP;
        kb.rh[DIR] = b.rh[DIR];
        if (!PREFETCH) tile_load<DIR, L, W, K>(kb.nyg, b.d, b.u, b.v, b.w, b.cc, g, ld, k.off32);
        const bool more = i + 1 < last;
        const int4 tn = more ? tile_at(i + 1) : tl;
        const TileGeo gn = tile_geo<DIR, L, W, K>(box_of(blocks[tn.x]), tn.y, tn.z, tn.w);
        // with PREFETCH the next tile's loads are issued as soon as this tile's split has
        // consumed the load registers, and complete during its reconstruction
        auto next = [&] {
            if (PREFETCH && more) {
                const BatchBlock &bn = blocks[tn.x];
                tile_load<DIR, L, W, K>(bn.n[1] + 2 * LAP, bn.d, bn.u, bn.v, bn.w, bn.cc, gn, ld, k.off32);
            }
        };
        tile_compute<DIR, L, W, K, false, ACC, MIX, false, false>(kb, nomf, b.du, g, ld, sp, StageArgs{}, next);
        __syncthreads();   // the next tile reuses the LDS buffer
        tl = tn;
        g = gn;
    }
}

// Batched form with one tile per work-group (grid = tile count): no loop-carried
// state, so it needs no more registers than the single-box kernel.
template <int DIR, int L, int W, int K, bool ACC, bool MIX, bool STG = false>
__global__ void __launch_bounds__(320)
k_rhs_batch_grid(KParams k, const BatchBlock *blocks, const int4 *tiles, int ntiles, BatchStage bs)
{
    using TL = Tile<DIR, L, W, K>;
    __shared__ double sp[10 * TL::QS + TL::PAD];
    // XCD-contiguous ranges of the tile list, as in k_rhs_tile
    const int per = (ntiles + 7) / 8;
    const int tile = ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
    if (tile >= ntiles) return;
    const int4 tl = tiles[tile];
    const BatchBlock &b = blocks[tl.x];
    KParams kb = k;
    kb.nx = b.n[0]; kb.ny = b.n[1]; kb.nz = b.n[2];
    kb.nyg = b.n[1] + 2 * LAP; kb.nzg = b.n[2] + 2 * LAP;
    kb.rh[DIR] = b.rh[DIR];
    Box box;
    box.s[0] = box.s[1] = box.s[2] = LAP;
    box.e[0] = b.n[0] + LAP; box.e[1] = b.n[1] + LAP; box.e[2] = b.n[2] + LAP;
    const TileGeo g = tile_geo<DIR, L, W, K>(box, tl.y, tl.z, tl.w);
    TileLoads<DIR, L, W, K> ld;
    tile_load<DIR, L, W, K>(kb.nyg, b.d, b.u, b.v, b.w, b.cc, g, ld, k.off32);
    const ws_metric_fields nomf{};
    StageArgs sa{};
    if (STG) {
        sa = StageArgs{bs.swapped ? b.fb : b.fa, bs.swapped ? b.fa : b.fb, bs.dt_dev, bs.dt, bs.krk};
        if (bs.prims) {
            sa.planes = bs.planes;
            sa.pd = b.od; sa.pu = b.ou; sa.pv = b.ov; sa.pw = b.ow; sa.pT = b.oT;
            if (bs.cc) sa.pcc = b.occ;
            if (bs.pressure) sa.pP = b.P;
            sa.cfl_part = bs.cfl_part;
            sa.halo = bs.halo;
        }
    }
    const double cmax = tile_compute<DIR, L, W, K, false, ACC, MIX, false, STG>(kb, nomf, b.du, g, ld, sp, sa, [] {});
    if (STG && sa.cfl_part) {
        __syncthreads();   // phase 3 has read its faces from sp
        tile_wg_max<TL::NT>(cmax, sp, sa.cfl_part + tile);
    }
}

// work-groups of a persistent kernel that fit on the device at once
template <typename Kern>
int resident_of(Kern kern, int threads)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0) != hipSuccess || per <= 0) per = 1;
    return cus * per;
}

template <int DIR, int L, int W, int K, bool ACC, bool MIX, bool PF>
void launch_persist(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du,
                    const Box &box, int gx, int gy, int gz, hipStream_t s)
{
    using TL = Tile<DIR, L, W, K>;
    static const int resident = resident_of(k_rhs_persist<DIR, L, W, K, ACC, MIX, PF>, TL::NT);
    const int n = gx * gy * gz;
    if (n <= 0) return;
    int grid = resident < n ? resident : n;
    const int chunk = (n + grid - 1) / grid;
    grid = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_rhs_persist<DIR, L, W, K, ACC, MIX, PF>), dim3(grid), dim3(TL::NT), 0, s, k, d, u, v, w,
                       cc, du, box, gx, gy, n, chunk);
}

template <int DIR, int L, int W, int K, bool FIELDS, bool ACC, bool MIX>
void launch_dir(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                const ws_metric_fields &mf, ws_soa du, const Box &box, int sched, hipStream_t s)
{
    using TL = Tile<DIR, L, W, K>;
    const int nx = box.e[0] - box.s[0], ny = box.e[1] - box.s[1], nz = box.e[2] - box.s[2];
    // y/z tiles start at the aligned column at or below the box start
    const int nxa = box.e[0] - box.s[0] / TL::XALIGN * TL::XALIGN;
    dim3 grid;
    if (DIR == 0) grid = dim3((nx + L - 1) / L, (ny + W - 1) / W, nz);
    else if (DIR == 1) grid = dim3((nxa + W - 1) / W, (ny + L - 1) / L, nz);
    else grid = dim3((nxa + W - 1) / W, (nz + L - 1) / L, ny);
    const int ntiles = (int)(grid.x * grid.y * grid.z);
    const dim3 grid1(8 * ((ntiles + 7) / 8));
    if (k.bnd_any[DIR]) {
        // one-sided closures in this direction: the grid schedule with the closure code
        hipLaunchKernelGGL((k_rhs_tile<DIR, L, W, K, FIELDS, ACC, MIX, true>), grid1, dim3(TL::NT), 0, s, k, d, u, v,
                           w, cc, mf, du, box, (int)grid.x, (int)grid.y, ntiles, StageArgs{});
    } else if (FIELDS || sched == WS_SCHED_GRID) {
        hipLaunchKernelGGL((k_rhs_tile<DIR, L, W, K, FIELDS, ACC, MIX, false>), grid1, dim3(TL::NT), 0, s, k, d, u,
                           v, w, cc, mf, du, box, (int)grid.x, (int)grid.y, ntiles, StageArgs{});
    } else if (sched == WS_SCHED_PREFETCH) {
        launch_persist<DIR, L, W, K, ACC, MIX, true>(k, d, u, v, w, cc, du, box, grid.x, grid.y, grid.z, s);
    } else {
        launch_persist<DIR, L, W, K, ACC, MIX, false>(k, d, u, v, w, cc, du, box, grid.x, grid.y, grid.z, s);
    }
}

// Tile grid of a grid-schedule launch over `box` (x tiles fastest).
template <int DIR, int L, int W, int K>
dim3 tile_grid(const Box &box)
{
    using TL = Tile<DIR, L, W, K>;
    const int nx = box.e[0] - box.s[0], ny = box.e[1] - box.s[1], nz = box.e[2] - box.s[2];
    const int nxa = box.e[0] - box.s[0] / TL::XALIGN * TL::XALIGN;
    if (DIR == 0) return dim3((nx + L - 1) / L, (ny + W - 1) / W, nz);
    if (DIR == 1) return dim3((nxa + W - 1) / W, (ny + L - 1) / L, nz);
    return dim3((nxa + W - 1) / W, (nz + L - 1) / L, ny);
}

// The last launch of an RK3 stage that applies the stage itself (grid schedule, constant
// metrics, no closures).
template <int DIR, int L, int W, int K, bool MIX>
void launch_dir_stage(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du,
                      const Box &box, const StageArgs &sa, hipStream_t s)
{
    using TL = Tile<DIR, L, W, K>;
    const dim3 grid = tile_grid<DIR, L, W, K>(box);
    const int ntiles = (int)(grid.x * grid.y * grid.z);
    const ws_metric_fields nomf{};
    hipLaunchKernelGGL((k_rhs_tile<DIR, L, W, K, false, true, MIX, false, true>), dim3(8 * ((ntiles + 7) / 8)),
                       dim3(TL::NT), 0, s, k, d, u, v, w, cc, nomf, du, box, (int)grid.x, (int)grid.y, ntiles, sa);
}

// The first launch writes du (or accumulates into it when the caller asks), the other two
// read-modify-write it.  Which direction pays for that read shifts the time differently
// per precision, so the order is a tuning choice (WS_ORDER_*).  With a stage (sa) the last
// launch applies it instead of writing du.
template <int L0, int W0, int L1, int W1, int L2, int W2, int K, bool FIELDS, bool MIX>
void launch_all(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                const ws_metric_fields &mf, ws_soa du, const Box &box, int accumulate, int sched, int order,
                const StageArgs *sa, hipStream_t s)
{
    if (order == WS_ORDER_ZYX) {
        if (accumulate) launch_dir<2, L2, W2, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
        else launch_dir<2, L2, W2, K, FIELDS, false, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
        launch_dir<1, L1, W1, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
        if constexpr (!FIELDS) {
            if (sa) {
                launch_dir_stage<0, L0, W0, K, MIX>(k, d, u, v, w, cc, du, box, *sa, s);
                return;
            }
        }
        launch_dir<0, L0, W0, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
        return;
    }
    if (accumulate) launch_dir<0, L0, W0, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
    else launch_dir<0, L0, W0, K, FIELDS, false, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
    launch_dir<1, L1, W1, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
    if constexpr (!FIELDS) {
        if (sa) {
            launch_dir_stage<2, L2, W2, K, MIX>(k, d, u, v, w, cc, du, box, *sa, s);
            return;
        }
    }
    launch_dir<2, L2, W2, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, sched, s);
}

template <int S, bool FIELDS, bool MIX>
void launch_shape_n(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                    const ws_metric_fields &mf, ws_soa du, const Box &box, int acc, int sched, int order,
                    const StageArgs *sa, hipStream_t s)
{
    constexpr TileShapeDef D = kTileShapes[S];
    launch_all<D.L0, D.W0, D.L1, D.W1, D.L2, D.W2, D.K, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, sched,
                                                                    order, sa, s);
}

template <bool FIELDS, bool MIX>
void launch_shape(int shape, const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                  const ws_metric_fields &mf, ws_soa du, const Box &box, int acc, int sched, int order,
                  const StageArgs *sa, hipStream_t s)
{
    switch (shape) {
    case 0: launch_shape_n<0, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, sched, order, sa, s); break;
    case 1: launch_shape_n<1, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, sched, order, sa, s); break;
    case 2: launch_shape_n<2, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, sched, order, sa, s); break;
    case 3: launch_shape_n<3, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, sched, order, sa, s); break;
    default: launch_shape_n<4, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, sched, order, sa, s); break;
    }
}

// tiles of the stage-applying launch (the last direction of `order`) of tile shape S
template <int S>
long stage_tiles_n(const Box &box, int order)
{
    constexpr TileShapeDef D = kTileShapes[S];
    const dim3 g = order == WS_ORDER_ZYX ? tile_grid<0, D.L0, D.W0, D.K>(box) : tile_grid<2, D.L2, D.W2, D.K>(box);
    return (long)g.x * g.y * g.z;
}

// ---- batched launches ------------------------------------------------------------------

template <int DIR, int L, int W, int K, bool ACC, bool MIX, bool PF>
void launch_batch_dir(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    const int n = bt.ntiles[DIR];
    if (n <= 0) return;
    if constexpr (DIR == 2 && ACC) {
        if (bt.stage) {
            // the last launch applies the RK3 stage (grid schedule only: the caller checks)
            hipLaunchKernelGGL((k_rhs_batch_grid<DIR, L, W, K, true, MIX, true>), dim3(8 * ((n + 7) / 8)),
                               dim3(Tile<DIR, L, W, K>::NT), 0, s, k, bt.blocks, bt.tiles[DIR], n, *bt.stage);
            return;
        }
    }
    if (bt.schedule == WS_SCHED_GRID) {
        hipLaunchKernelGGL((k_rhs_batch_grid<DIR, L, W, K, ACC, MIX>), dim3(8 * ((n + 7) / 8)),
                           dim3(Tile<DIR, L, W, K>::NT), 0, s, k, bt.blocks, bt.tiles[DIR], n, BatchStage{});
        return;
    }
    static const int resident = resident_of(k_rhs_batch<DIR, L, W, K, ACC, MIX, PF>, Tile<DIR, L, W, K>::NT);
    int grid = resident < n ? resident : n;
    const int chunk = (n + grid - 1) / grid;
    grid = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_rhs_batch<DIR, L, W, K, ACC, MIX, PF>), dim3(grid), dim3(Tile<DIR, L, W, K>::NT), 0, s,
                       k, bt.blocks, bt.tiles[DIR], n, chunk);
}

// one tile per work-group over direction DIR's tile list (the grid schedule), STG: this
// launch applies the RK3 stage bt.stage
template <int DIR, int L, int W, int K, bool ACC, bool MIX, bool STG>
void launch_batch_grid(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    const int n = bt.ntiles[DIR];
    if (n <= 0) return;
    hipLaunchKernelGGL((k_rhs_batch_grid<DIR, L, W, K, ACC, MIX, STG>), dim3(8 * ((n + 7) / 8)),
                       dim3(Tile<DIR, L, W, K>::NT), 0, s, k, bt.blocks, bt.tiles[DIR], n,
                       STG ? *bt.stage : BatchStage{});
}

template <int S, bool MIX, bool PF>
void launch_batch_shape_n(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    constexpr TileShapeDef D = kTileShapes[S];
    if (bt.stage && bt.stage_order == WS_ORDER_ZYX) {
        // z writes du, y accumulates, x adds its terms and applies the stage
        launch_batch_grid<2, D.L2, D.W2, D.K, false, MIX, false>(k, bt, s);
        launch_batch_grid<1, D.L1, D.W1, D.K, true, MIX, false>(k, bt, s);
        launch_batch_grid<0, D.L0, D.W0, D.K, true, MIX, true>(k, bt, s);
        return;
    }
    launch_batch_dir<0, D.L0, D.W0, D.K, false, MIX, PF>(k, bt, s);
    launch_batch_dir<1, D.L1, D.W1, D.K, true, MIX, PF>(k, bt, s);
    launch_batch_dir<2, D.L2, D.W2, D.K, true, MIX, PF>(k, bt, s);
}

template <bool MIX, bool PF>
void launch_batch_pf(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    switch (bt.shape) {
    case 0: launch_batch_shape_n<0, MIX, PF>(k, bt, s); break;
    case 1: launch_batch_shape_n<1, MIX, PF>(k, bt, s); break;
    case 2: launch_batch_shape_n<2, MIX, PF>(k, bt, s); break;
    case 3: launch_batch_shape_n<3, MIX, PF>(k, bt, s); break;
    default: launch_batch_shape_n<4, MIX, PF>(k, bt, s); break;
    }
}

template <bool MIX>
int launch_rhs_tile(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                    const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s,
                    const StageArgs *sa = nullptr)
{
    Box box;
    box.s[0] = (int)job.start.x; box.s[1] = (int)job.start.y; box.s[2] = (int)job.start.z;
    box.e[0] = (int)job.end.x; box.e[1] = (int)job.end.y; box.e[2] = (int)job.end.z;
    ws_metric_fields mf{};
    const int shape = sa && c.stage_shape >= 0 ? c.stage_shape : c.tile_shape;
    const int order = sa ? c.stage_order : c.dir_order;
    KParams k = c.kp;
    k.off32 = prim_off32(k, d, u, v, w, cc) ? 1 : 0;
    if (m) {
        mf = *m;
        launch_shape<true, MIX>(shape, k, d, u, v, w, cc, mf, du, box, accumulate, c.schedule, order, nullptr, s);
    } else {
        launch_shape<false, MIX>(shape, k, d, u, v, w, cc, mf, du, box, accumulate, c.schedule, order, sa, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, MIX ? "mixed-precision RHS launch" : "tiled RHS launch");
    return WS_OK;
}

template <bool MIX>
int launch_rhs_batch_tile(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    if (bt.schedule == WS_SCHED_PREFETCH) launch_batch_pf<MIX, true>(k, bt, s);
    else launch_batch_pf<MIX, false>(k, bt, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, MIX ? "batched mixed-precision RHS launch" : "batched RHS launch");
    return WS_OK;
}

}  // namespace

}  // namespace ws
