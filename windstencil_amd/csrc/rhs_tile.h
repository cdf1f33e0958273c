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
// Phase 3 turns face fluxes into cell updates.
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
    // threads write the same point of consecutive lines), so their line stride is
    // padded to NL + 2 (= 2 mod 4 doubles): those writes spread over the banks and
    // every line still starts 16-byte aligned
    static constexpr int LS = DIR == 0 ? NL : NL + 2;
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
    // face buffer index of face f of line l (line-major; F = L + 1 is odd for the even
    // tile lengths used, so the y/z flux difference, which reads one face of
    // consecutive lines per lane, touches distinct banks)
    __device__ static int fidx(int l, int f) { return l * F + f; }
};

// One tile (tile coordinates bx, by, bz within the box) of one direction, with the
// block's LDS buffer sp (10 * QS + PAD doubles).  Every thread of the block calls it.
template <int DIR, int L, int W, int K, bool FIELDS, bool ACC, bool MIX>
__device__ __forceinline__ void rhs_tile_body(const KParams &k, const ws_field &d, const ws_field &u,
                                              const ws_field &v, const ws_field &w, const ws_field &cc,
                                              const ws_metric_fields &mf, const ws_soa &du, const Box &box,
                                              int bx, int by, int bz, double *sp)
{
    using TL = Tile<DIR, L, W, K>;
    const int t = threadIdx.x;

    // tile origin: x0 always tiles x; the line axis is x (DIR 0), y (DIR 1) or z (DIR 2)
    int x0, y0, z0;
    if (DIR == 0) {
        x0 = box.s[0] + bx * L;
        y0 = box.s[1] + by * W;
        z0 = box.s[2] + bz;
    } else if (DIR == 1) {
        x0 = box.s[0] / TL::XALIGN * TL::XALIGN + bx * W;
        y0 = box.s[1] + by * L;
        z0 = box.s[2] + bz;
    } else {
        x0 = box.s[0] / TL::XALIGN * TL::XALIGN + bx * W;
        z0 = box.s[2] + by * L;
        y0 = box.s[1] + bz;
    }
    // cells of the tile that lie in the box, along the line and across lines
    // (lines lo_across .. lim_across-1; only the first aligned y/z tile has lo_across > 0)
    const int lim_line = DIR == 0 ? box.e[0] - x0 : DIR == 1 ? box.e[1] - y0 : box.e[2] - z0;
    const int lim_across = DIR == 0 ? box.e[1] - y0 : box.e[0] - x0;
    const int lo_across = DIR == 0 ? 0 : box.s[0] - x0;

    const double cA1 = k.A[3 * DIR + 0], cA2 = k.A[3 * DIR + 1], cA3 = k.A[3 * DIR + 2];
    const double cSn = k.Anorm[DIR];
    const double cK1 = k.Ak[3 * DIR + 0], cK2 = k.Ak[3 * DIR + 1], cK3 = k.Ak[3 * DIR + 2];

    // ---- phase 1: split every point of the tile ----------------------------------------
    // all loads of this thread's points are issued before any split arithmetic
    double ld[TL::ROUNDS][5];
    int lidx[TL::ROUNDS];
    bool lok[TL::ROUNDS];
#pragma unroll
    for (int r = 0; r < TL::ROUNDS; ++r) {
        const int pt = t + r * TL::NT;
        int l, p;
        if (DIR == 0) { p = pt % TL::NL; l = pt / TL::NL; }     // x fastest
        else          { l = pt % W;      p = pt / W; }          // x fastest
        // the line needs points -LAP .. lim_line+LAP-1 of lines lo_across .. lim_across-1
        lok[r] = pt < TL::NPTS && l >= lo_across && l < lim_across && p < lim_line + 2 * LAP;
        lidx[r] = TL::pidx(l, p);
        if (lok[r]) {
            int x, y, z;
            if (DIR == 0)      { x = x0 - LAP + p; y = y0 + l; z = z0; }
            else if (DIR == 1) { x = x0 + l; y = y0 - LAP + p; z = z0; }
            else               { x = x0 + l; y = y0; z = z0 - LAP + p; }
            const int nyg = k.nyg;
            ld[r][0] = d.ptr[fidx(d.pitch, nyg, x, y, z)];
            ld[r][1] = u.ptr[fidx(u.pitch, nyg, x, y, z)];
            ld[r][2] = v.ptr[fidx(v.pitch, nyg, x, y, z)];
            ld[r][3] = w.ptr[fidx(w.pitch, nyg, x, y, z)];
            ld[r][4] = cc.ptr[fidx(cc.pitch, nyg, x, y, z)];
        }
    }
#pragma unroll
    for (int r = 0; r < TL::ROUNDS; ++r) {
        if (!lok[r]) continue;
        const double dd = ld[r][0], uu = ld[r][1], vv = ld[r][2], ww = ld[r][3], c = ld[r][4];
        double fp[5], fm[5];
        if (FIELDS) {
            const int pt = t + r * TL::NT;
            int l, p, x, y, z;
            if (DIR == 0) { p = pt % TL::NL; l = pt / TL::NL; }
            else          { l = pt % W;      p = pt / W; }
            if (DIR == 0)      { x = x0 - LAP + p; y = y0 + l; z = z0; }
            else if (DIR == 1) { x = x0 + l; y = y0 - LAP + p; z = z0; }
            else               { x = x0 + l; y = y0; z = z0 - LAP + p; }
            const int nyg = k.nyg;
            const ws_field &f1 = mf.A[3 * DIR + 0], &f2 = mf.A[3 * DIR + 1], &f3 = mf.A[3 * DIR + 2];
            const double A1 = f1.ptr[fidx(f1.pitch, nyg, x, y, z)];
            const double A2 = f2.ptr[fidx(f2.pitch, nyg, x, y, z)];
            const double A3 = f3.ptr[fidx(f3.pitch, nyg, x, y, z)];
            const double sn = sqrt(fma(A3, A3, fma(A2, A2, A1 * A1)));
            const double r = 1.0 / sn;
            split_fast(k, dd, uu, vv, ww, c, A1, A2, A3, sn, A1 * r, A2 * r, A3 * r, fp, fm);
        } else {
            split_fast(k, dd, uu, vv, ww, c, cA1, cA2, cA3, cSn, cK1, cK2, cK3, fp, fm);
        }
        const int li = lidx[r];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            sp[m * TL::QS + li] = fp[m];
            sp[(5 + m) * TL::QS + li] = fm[m];
        }
    }
    __syncthreads();

    // ---- phase 3 prologue: accumulating launches read du; issue those loads now so
    // they complete while this thread reconstructs its faces
    constexpr int CPT = (W * L + TL::NT - 1) / TL::NT;    // cells per thread (upper bound)
    const size_t vstride = (size_t)du.pitch * k.ny * k.nz;
    double duo[ACC ? CPT : 1][5];
    if (ACC) {
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
            const int c = t + r * TL::NT;
            int l, i;
            if (DIR == 0) { i = c % L; l = c / L; }
            else          { l = c % W; i = c / W; }
            if (c >= W * L || l < lo_across || l >= lim_across || i >= lim_line) continue;
            int cx, cy, cz;
            if (DIR == 0)      { cx = x0 + i; cy = y0 + l; cz = z0; }
            else if (DIR == 1) { cx = x0 + l; cy = y0 + i; cz = z0; }
            else               { cx = x0 + l; cy = y0; cz = z0 + i; }
            const double *dp = du.ptr + (size_t)(cx - LAP) +
                               (size_t)du.pitch * ((size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP));
#pragma unroll
            for (int m = 0; m < 5; ++m) duo[r][m] = dp[m * vstride];
        }
    }

    // ---- phase 2: K faces of one line per thread, all five variables --------------------
    const int line = t / TL::T, seg = t % TL::T;
    const int f0 = seg * K;                               // first face of this thread
    const bool active = (t < TL::NT) && (line >= lo_across) && (line < lim_across) && (f0 <= lim_line);
    double Fv[5][K];
    if (active) {
        const int base = TL::pidx(line, f0);
#pragma unroll 1
        for (int m = 0; m < 5; ++m) {
            const double *qp = sp + m * TL::QS + base;
            const double *qm = sp + (5 + m) * TL::QS + base;
            // a stored face f (f <= lim_line) reads only points f..f+7, all written in
            // phase 1; faces past lim_line read stale LDS and are not stored
            if (MIX) {
                // FP32 reconstruction relative to FP64 base points; the merge is FP64
                double bP[K], bM[K];
                float cP[K], cM[K];
                weno_k_mixed<false, K>([&](int j) { return qp[j]; }, bP, cP);
                weno_k_mixed<true, K>([&](int j) { return qm[j]; }, bM, cM);
#pragma unroll
                for (int j = 0; j < K; ++j) Fv[m][j] = (bP[j] + bM[j]) + (double)(cP[j] + cM[j]);
            } else {
                double nP[K], dP[K], nM[K], dM[K];
                weno_k<false, K>([&](int j) { return qp[j]; }, nP, dP);
                weno_k<true, K>([&](int j) { return qm[j]; }, nM, dM);
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const double den = dP[j] * dM[j];
                    Fv[m][j] = fma(nP[j], dM[j], nM[j] * dP[j]) * recip_nr(den);
                }
            }
        }
    }
    __syncthreads();
    // face fluxes reuse the split buffer
    if (active) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (f0 + j > lim_line || f0 + j >= TL::F) break;
#pragma unroll
            for (int m = 0; m < 5; ++m) sp[m * W * TL::F + TL::fidx(line, f0 + j)] = Fv[m][j];
        }
    }
    __syncthreads();

    // ---- phase 3: flux difference of each cell ---------------------------------------------
    const double scale0 = -k.rh[DIR];
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
        const int c = t + r * TL::NT;
        if (c >= W * L) break;
        int l, i;
        if (DIR == 0) { i = c % L; l = c / L; }
        else          { l = c % W; i = c / W; }
        if (l < lo_across || l >= lim_across || i >= lim_line) continue;
        int cx, cy, cz;
        if (DIR == 0)      { cx = x0 + i; cy = y0 + l; cz = z0; }
        else if (DIR == 1) { cx = x0 + l; cy = y0 + i; cz = z0; }
        else               { cx = x0 + l; cy = y0; cz = z0 + i; }
        const double aj = FIELDS ? mf.ajac.ptr[fidx(mf.ajac.pitch, k.nyg, cx, cy, cz)] : k.ajac;
        const double scale = aj * scale0;
        double *dp = du.ptr + (size_t)(cx - LAP) +
                     (size_t)du.pitch * ((size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP));
        const int fl = TL::fidx(l, i), fr = TL::fidx(l, i + 1);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double val = scale * (sp[m * W * TL::F + fr] - sp[m * W * TL::F + fl]);
            if (ACC) dp[m * vstride] = duo[r][m] + val;
            else dp[m * vstride] = val;
        }
    }
}

template <int DIR, int L, int W, int K, bool FIELDS, bool ACC, bool MIX>
__global__ void __launch_bounds__(320)
k_rhs_tile(KParams k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_metric_fields mf,
           ws_soa du, Box box)
{
    using TL = Tile<DIR, L, W, K>;
    __shared__ double sp[10 * TL::QS + TL::PAD];
    rhs_tile_body<DIR, L, W, K, FIELDS, ACC, MIX>(k, d, u, v, w, cc, mf, du, box, blockIdx.x, blockIdx.y,
                                                  blockIdx.z, sp);
}

// Batched form: the tiles of one direction over every block of a batch, in one
// persistent launch.  Work-group g walks the contiguous range [g*chunk, (g+1)*chunk)
// of the tile list, so neighbouring tiles (which share halo lines) run on the same
// work-group, hence the same XCD and L2.  Each block brings its own fields, extents
// and spacing; the physics constants and metrics are the batch's.
template <int DIR, int L, int W, int K, bool ACC, bool MIX>
__global__ void __launch_bounds__(320)
k_rhs_batch(KParams k, const BatchBlock *blocks, const int4 *tiles, int ntiles, int chunk)
{
    using TL = Tile<DIR, L, W, K>;
    __shared__ double sp[10 * TL::QS + TL::PAD];
    const int first = blockIdx.x * chunk;
    const int last = min(first + chunk, ntiles);
    const ws_metric_fields nomf{};
    for (int i = first; i < last; ++i) {
        const int4 tl = tiles[i];
        const BatchBlock &b = blocks[tl.x];
        KParams kb = k;
        kb.nx = b.n[0]; kb.ny = b.n[1]; kb.nz = b.n[2];
        kb.nyg = b.n[1] + 2 * LAP; kb.nzg = b.n[2] + 2 * LAP;
        kb.rh[DIR] = b.rh[DIR];
        Box box;
        box.s[0] = box.s[1] = box.s[2] = LAP;
        box.e[0] = b.n[0] + LAP; box.e[1] = b.n[1] + LAP; box.e[2] = b.n[2] + LAP;
        rhs_tile_body<DIR, L, W, K, false, ACC, MIX>(kb, b.d, b.u, b.v, b.w, b.cc, nomf, b.du, box, tl.y, tl.z,
                                                     tl.w, sp);
        __syncthreads();   // the next tile reuses the LDS buffer
    }
}

template <int DIR, int L, int W, int K, bool FIELDS, bool ACC, bool MIX>
void launch_dir(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                const ws_metric_fields &mf, ws_soa du, const Box &box, hipStream_t s)
{
    using TL = Tile<DIR, L, W, K>;
    const int nx = box.e[0] - box.s[0], ny = box.e[1] - box.s[1], nz = box.e[2] - box.s[2];
    // y/z tiles start at the aligned column at or below the box start
    const int nxa = box.e[0] - box.s[0] / TL::XALIGN * TL::XALIGN;
    dim3 grid;
    if (DIR == 0) grid = dim3((nx + L - 1) / L, (ny + W - 1) / W, nz);
    else if (DIR == 1) grid = dim3((nxa + W - 1) / W, (ny + L - 1) / L, nz);
    else grid = dim3((nxa + W - 1) / W, (nz + L - 1) / L, ny);
    hipLaunchKernelGGL((k_rhs_tile<DIR, L, W, K, FIELDS, ACC, MIX>), grid, dim3(TL::NT), 0, s, k, d, u, v, w, cc, mf,
                       du, box);
}

template <int L0, int W0, int L1, int W1, int K, bool FIELDS, bool MIX>
void launch_all(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                const ws_metric_fields &mf, ws_soa du, const Box &box, int accumulate, hipStream_t s)
{
    if (accumulate) launch_dir<0, L0, W0, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, s);
    else launch_dir<0, L0, W0, K, FIELDS, false, MIX>(k, d, u, v, w, cc, mf, du, box, s);
    launch_dir<1, L1, W1, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, s);
    launch_dir<2, L1, W1, K, FIELDS, true, MIX>(k, d, u, v, w, cc, mf, du, box, s);
}

template <int S, bool FIELDS, bool MIX>
void launch_shape_n(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                    const ws_metric_fields &mf, ws_soa du, const Box &box, int acc, hipStream_t s)
{
    constexpr TileShapeDef D = kTileShapes[S];
    launch_all<D.L0, D.W0, D.L1, D.W1, D.K, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, s);
}

template <bool FIELDS, bool MIX>
void launch_shape(int shape, const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                  const ws_metric_fields &mf, ws_soa du, const Box &box, int acc, hipStream_t s)
{
    switch (shape) {
    case 0: launch_shape_n<0, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, s); break;
    case 1: launch_shape_n<1, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, s); break;
    default: launch_shape_n<2, FIELDS, MIX>(k, d, u, v, w, cc, mf, du, box, acc, s); break;
    }
}

// ---- batched launches ------------------------------------------------------------------

// work-groups of k_rhs_batch that fit on the device at once (the persistent grid)
template <int DIR, int L, int W, int K, bool ACC, bool MIX>
int resident_groups()
{
    static const int n = [] {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rhs_batch<DIR, L, W, K, ACC, MIX>,
                                                         Tile<DIR, L, W, K>::NT, 0) != hipSuccess || per <= 0)
            per = 1;
        return cus * per;
    }();
    return n;
}

template <int DIR, int L, int W, int K, bool ACC, bool MIX>
void launch_batch_dir(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    const int n = bt.ntiles[DIR];
    if (n <= 0) return;
    int grid = resident_groups<DIR, L, W, K, ACC, MIX>();
    if (grid > n) grid = n;
    const int chunk = (n + grid - 1) / grid;
    grid = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_rhs_batch<DIR, L, W, K, ACC, MIX>), dim3(grid), dim3(Tile<DIR, L, W, K>::NT), 0, s, k,
                       bt.blocks, bt.tiles[DIR], n, chunk);
}

template <int S, bool MIX>
void launch_batch_shape_n(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    constexpr TileShapeDef D = kTileShapes[S];
    launch_batch_dir<0, D.L0, D.W0, D.K, false, MIX>(k, bt, s);
    launch_batch_dir<1, D.L1, D.W1, D.K, true, MIX>(k, bt, s);
    launch_batch_dir<2, D.L1, D.W1, D.K, true, MIX>(k, bt, s);
}

template <bool MIX>
int launch_rhs_tile(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                    const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s)
{
    Box box;
    box.s[0] = (int)job.start.x; box.s[1] = (int)job.start.y; box.s[2] = (int)job.start.z;
    box.e[0] = (int)job.end.x; box.e[1] = (int)job.end.y; box.e[2] = (int)job.end.z;
    ws_metric_fields mf{};
    const int shape = c.tile_shape;
    if (m) {
        mf = *m;
        launch_shape<true, MIX>(shape, c.kp, d, u, v, w, cc, mf, du, box, accumulate, s);
    } else {
        launch_shape<false, MIX>(shape, c.kp, d, u, v, w, cc, mf, du, box, accumulate, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, MIX ? "mixed-precision RHS launch" : "tiled RHS launch");
    return WS_OK;
}

template <bool MIX>
int launch_rhs_batch_tile(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    switch (bt.shape) {
    case 0: launch_batch_shape_n<0, MIX>(k, bt, s); break;
    case 1: launch_batch_shape_n<1, MIX>(k, bt, s); break;
    default: launch_batch_shape_n<2, MIX>(k, bt, s); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, MIX ? "batched mixed-precision RHS launch" : "batched RHS launch");
    return WS_OK;
}

}  // namespace

}  // namespace ws
