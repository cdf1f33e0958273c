// k_rhs_tiled.hip -- LDS-tiled inviscid RHS, FP64 (the fast path).
//
// One launch per direction.  A 256-thread block owns a tile of one plane:
//   x-direction: 15 cells along x (16 faces) x 16 rows along y, at one z;
//   y-direction: 15 columns along x x 16 cells along y (17 faces), at one z;
//   z-direction: 15 columns along x x 16 cells along z (17 faces), at one y.
// Phase 1 splits every point of the tile's lines, halo included, into LDS
// (10 doubles per point).  Phase 2 gives each thread one face: it reconstructs
// all five variables with the positive and the negative rule and merges each
// pair over one division.  Phase 3 turns face fluxes into cell updates.
// Tile widths divide the benchmark grids (600 = 40*15, 256 = 16*16, 128 = 8*16);
// other sizes leave a partial last tile, masked.
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

namespace {

struct Box {
    int s[3];   // start (ghosted coordinates)
    int e[3];   // end
};

constexpr int NT = 256;          // threads per block
constexpr int TA = 15;           // cells along the line (x) / columns across (y, z)
constexpr int TL = 16;           // rows (x) / cells along the line (y, z)
constexpr int QS = 384;          // LDS entries per quantity (>= 23*16 and 24*16)

template <int DIR, bool FIELDS, bool ACC>
__global__ void __launch_bounds__(NT) k_rhs_tile(KParams k, ws_field d, ws_field u, ws_field v, ws_field w,
                                                 ws_field cc, ws_metric_fields mf, ws_soa du, Box box)
{
    __shared__ double sp[10 * QS];
    const int t = threadIdx.x;

    // tile origin
    int x0, y0, z0;
    if (DIR == 0) {
        x0 = box.s[0] + blockIdx.x * TA;
        y0 = box.s[1] + blockIdx.y * TL;
        z0 = box.s[2] + blockIdx.z;
    } else if (DIR == 1) {
        x0 = box.s[0] + blockIdx.x * TA;
        y0 = box.s[1] + blockIdx.y * TL;
        z0 = box.s[2] + blockIdx.z;
    } else {
        x0 = box.s[0] + blockIdx.x * TA;
        z0 = box.s[2] + blockIdx.y * TL;
        y0 = box.s[1] + blockIdx.z;
    }

    // constant-metric split coefficients of this direction
    const double cA1 = k.A[3 * DIR + 0], cA2 = k.A[3 * DIR + 1], cA3 = k.A[3 * DIR + 2];
    const double cSn = k.Anorm[DIR];
    const double cK1 = k.Ak[3 * DIR + 0], cK2 = k.Ak[3 * DIR + 1], cK3 = k.Ak[3 * DIR + 2];

    // ---- phase 1: split every point of the tile ---------------------------------
    constexpr int NPTS = DIR == 0 ? (TA + 2 * LAP) * TL : TA * (TL + 2 * LAP);
    for (int pt = t; pt < NPTS; pt += NT) {
        int x, y, z, li;
        bool ok;
        if (DIR == 0) {
            const int px = pt % (TA + 2 * LAP), r = pt / (TA + 2 * LAP);
            x = x0 - LAP + px; y = y0 + r; z = z0;
            li = r * 24 + px;
            ok = (y < box.e[1]) && (x < box.e[0] + LAP);
        } else {
            const int a = pt % TA, p = pt / TA;
            x = x0 + a;
            if (DIR == 1) { y = y0 - LAP + p; z = z0; ok = (y < box.e[1] + LAP); }
            else          { z = z0 - LAP + p; y = y0; ok = (z < box.e[2] + LAP); }
            ok = ok && (x < box.e[0]);
            li = p * 16 + a;
        }
        if (ok) {
            const int nyg = k.nyg;
            const double dd = d.ptr[fidx(d.pitch, nyg, x, y, z)];
            const double uu = u.ptr[fidx(u.pitch, nyg, x, y, z)];
            const double vv = v.ptr[fidx(v.pitch, nyg, x, y, z)];
            const double ww = w.ptr[fidx(w.pitch, nyg, x, y, z)];
            const double c = cc.ptr[fidx(cc.pitch, nyg, x, y, z)];
            double fp[5], fm[5];
            if (FIELDS) {
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
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                sp[m * QS + li] = fp[m];
                sp[(5 + m) * QS + li] = fm[m];
            }
        }
    }
    __syncthreads();

    // ---- phase 2: one face per thread, all five variables --------------------------
    double F[5];
    bool face_ok;
    int base, stride;
    if (DIR == 0) {
        const int f = t % 16, r = t / 16;
        face_ok = (y0 + r < box.e[1]) && (x0 + f <= box.e[0]);
        base = r * 24 + f;
        stride = 1;
    } else {
        const int a = t % TA, f = t / TA;      // f in 0..16 for t < 255
        const int lim = DIR == 1 ? box.e[1] - y0 : box.e[2] - z0;
        face_ok = (t < TA * (TL + 1)) && (x0 + a < box.e[0]) && (f <= lim);
        base = f * 16 + a;
        stride = 16;
    }
    if (face_ok) {
#pragma unroll 1
        for (int m = 0; m < 5; ++m) {
            const double *qp = sp + m * QS + base;
            const double *qm = sp + (5 + m) * QS + base;
            double nP, dP, nM, dM;
            weno_fast<false>([&](int j) { return qp[j * stride]; }, nP, dP);
            weno_fast<true>([&](int j) { return qm[j * stride]; }, nM, dM);
            F[m] = fma(nP, dM, nM * dP) / (12.0 * (dP * dM));
        }
    }
    __syncthreads();
    // face fluxes reuse the split buffer: face[m][t]
    if (face_ok) {
#pragma unroll
        for (int m = 0; m < 5; ++m) sp[m * NT + t] = F[m];
    }
    __syncthreads();

    // ---- phase 3: flux difference of each cell ---------------------------------------
    int cx, cy, cz, fl, fr;
    bool cell_ok;
    if (DIR == 0) {
        const int c = t % 16, r = t / 16;
        cx = x0 + c; cy = y0 + r; cz = z0;
        cell_ok = (c < TA) && (cx < box.e[0]) && (cy < box.e[1]);
        fl = r * 16 + c;
        fr = fl + 1;
    } else {
        const int a = t % TA, c = t / TA;      // c in 0..15 for t < 240
        cx = x0 + a;
        if (DIR == 1) { cy = y0 + c; cz = z0; cell_ok = (cy < box.e[1]); }
        else          { cz = z0 + c; cy = y0; cell_ok = (cz < box.e[2]); }
        cell_ok = cell_ok && (t < TA * TL) && (cx < box.e[0]);
        fl = c * TA + a;
        fr = fl + TA;
    }
    if (cell_ok) {
        const double aj = FIELDS ? mf.ajac.ptr[fidx(mf.ajac.pitch, k.nyg, cx, cy, cz)] : k.ajac;
        const double scale = -aj * k.rh[DIR];
        const size_t vstride = (size_t)du.pitch * k.ny * k.nz;
        double *dp = du.ptr + (size_t)(cx - LAP) + (size_t)du.pitch * ((size_t)(cy - LAP) + (size_t)k.ny * (size_t)(cz - LAP));
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double val = scale * (sp[m * NT + fr] - sp[m * NT + fl]);
            if (ACC) dp[m * vstride] += val;
            else dp[m * vstride] = val;
        }
    }
}

template <int DIR, bool FIELDS, bool ACC>
void launch_dir(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                const ws_metric_fields &mf, ws_soa du, const Box &box, hipStream_t s)
{
    const int nx = box.e[0] - box.s[0], ny = box.e[1] - box.s[1], nz = box.e[2] - box.s[2];
    dim3 grid;
    if (DIR == 0) grid = dim3((nx + TA - 1) / TA, (ny + TL - 1) / TL, nz);
    else if (DIR == 1) grid = dim3((nx + TA - 1) / TA, (ny + TL - 1) / TL, nz);
    else grid = dim3((nx + TA - 1) / TA, (nz + TL - 1) / TL, ny);
    hipLaunchKernelGGL((k_rhs_tile<DIR, FIELDS, ACC>), grid, dim3(NT), 0, s, k, d, u, v, w, cc, mf, du, box);
}

template <bool FIELDS>
void launch_all(const KParams &k, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                const ws_metric_fields &mf, ws_soa du, const Box &box, int accumulate, hipStream_t s)
{
    if (accumulate) launch_dir<0, FIELDS, true>(k, d, u, v, w, cc, mf, du, box, s);
    else launch_dir<0, FIELDS, false>(k, d, u, v, w, cc, mf, du, box, s);
    launch_dir<1, FIELDS, true>(k, d, u, v, w, cc, mf, du, box, s);
    launch_dir<2, FIELDS, true>(k, d, u, v, w, cc, mf, du, box, s);
}

}  // namespace

int launch_rhs_tiled(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                     const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s)
{
    Box box;
    box.s[0] = (int)job.start.x; box.s[1] = (int)job.start.y; box.s[2] = (int)job.start.z;
    box.e[0] = (int)job.end.x; box.e[1] = (int)job.end.y; box.e[2] = (int)job.end.z;
    ws_metric_fields mf{};
    if (m) {
        mf = *m;
        launch_all<true>(c.kp, d, u, v, w, cc, mf, du, box, accumulate, s);
    } else {
        launch_all<false>(c.kp, d, u, v, w, cc, mf, du, box, accumulate, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "tiled RHS launch");
    return WS_OK;
}

}  // namespace ws
