// k_aux.hip -- memory-bound helpers around the RHS: ghost fill, sound speed,
// primitive/conserved conversion, SSP-RK3 stage updates, box fill, CFL step.
// Their arithmetic follows the reference's expressions with contraction off, so
// the solver state stays bit-identical to the CPU restatement between RHS calls.
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

namespace {

constexpr int NT = 256;

struct FieldList {
    ws_field f[8];
};

// Periodic LAP-deep face halo of the periodic directions.  Linear index over the face
// slabs of one field:
// x slabs (2*LAP*ny*nz), y slabs (2*LAP*nx*nz), z slabs (2*LAP*nx*ny).
// Ghost points read only interior points, so the three directions need no ordering.
__device__ __forceinline__ void ghost_point(const ws_field &f, int nx, int ny, int nz, int nyg, bool px, bool py,
                                            bool pz, long i)
{
    // slabs of non-periodic directions are left to the caller
    const long nxs = px ? 2L * LAP * ny * nz : 0, nys = py ? 2L * LAP * nx * nz : 0;
    const long nzs = pz ? 2L * LAP * nx * ny : 0;
    int dx, dy, dz, sx, sy, sz;
    if (i < nxs) {
        const int g = (int)(i % (2 * LAP));            // 0..3 left ghosts, 4..7 right ghosts
        const int y = (int)((i / (2 * LAP)) % ny);
        const int z = (int)(i / (2L * LAP * ny));
        dy = sy = y + LAP; dz = sz = z + LAP;
        if (g < LAP) { dx = g;                sx = nx + g; }           // [0,4) <- [nx, nx+4)
        else         { dx = nx + LAP + g - LAP; sx = g; }               // [nx+4, nx+8) <- [4, 8)
    } else if ((i -= nxs) < nys) {
        const int x = (int)(i % nx);
        const int g = (int)((i / nx) % (2 * LAP));
        const int z = (int)(i / (2L * LAP * nx));
        dx = sx = x + LAP; dz = sz = z + LAP;
        if (g < LAP) { dy = g;                sy = ny + g; }
        else         { dy = ny + g;           sy = g; }
    } else if ((i -= nys) < nzs) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const int g = (int)(i / ((long)nx * ny));
        dx = sx = x + LAP; dy = sy = y + LAP;
        if (g < LAP) { dz = g;                sz = nz + g; }
        else         { dz = nz + g;           sz = g; }
    } else {
        return;
    }
    f.ptr[fidx(f.pitch, nyg, dx, dy, dz)] = f.ptr[fidx(f.pitch, nyg, sx, sy, sz)];
}

__global__ void __launch_bounds__(NT) k_ghost_fill(FieldList fl, int nx, int ny, int nz, int nyg, int px, int py,
                                                   int pz)
{
    ghost_point(fl.f[blockIdx.y], nx, ny, nz, nyg, px != 0, py != 0, pz != 0,
                (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// to = from over a ghosted box of nxg x rows points (rows = nyg * nzg), one field per
// blockIdx.y; each field's pitch is at least nxg
__global__ void __launch_bounds__(NT) k_copy_fields(FieldList from, FieldList to, int nxg, long rows)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= (long)nxg * rows) return;
    const long row = i / nxg;
    const int x = (int)(i - row * nxg);
    const ws_field &a = from.f[blockIdx.y], &b = to.f[blockIdx.y];
    b.ptr[(size_t)x + (size_t)b.pitch * (size_t)row] = a.ptr[(size_t)x + (size_t)a.pitch * (size_t)row];
}

__global__ void __launch_bounds__(NT) k_sound_speed(ws_field T, ws_field cc, int sx, int sy, int sz, int nx,
                                                    int ny, int nz, int nyg, double ama)
{
#pragma clang fp contract(off)
    const long total = (long)nx * ny * nz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = sx + (int)(i % nx);
    const int y = sy + (int)((i / nx) % ny);
    const int z = sz + (int)(i / ((long)nx * ny));
    cc.ptr[fidx(cc.pitch, nyg, x, y, z)] = sqrt(T.ptr[fidx(T.pitch, nyg, x, y, z)]) / ama;
}

// Sound speed over a whole ghosted field treated as one flat array (pitch padding
// included, which no kernel reads): cc = sqrt(T)/Ma, two points per thread.
__global__ void __launch_bounds__(NT) k_sound_flat(const double2 *T, double2 *cc, long n2, double ama)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= n2) return;
    const double2 t = T[i];
    cc[i] = make_double2(sqrt(t.x) / ama, sqrt(t.y) / ama);
}

__global__ void __launch_bounds__(NT) k_cons_to_pri(KParams k, ws_soa f, ws_field d, ws_field u, ws_field v,
                                                    ws_field w, ws_field T, ws_field P)
{
#pragma clang fp contract(off)
    const long total = (long)k.nx * k.ny * k.nz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = (int)(i % k.nx);
    const int y = (int)((i / k.nx) % k.ny);
    const int z = (int)(i / ((long)k.nx * k.ny));
    const double d1 = f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 0)];
    const double u1 = f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 1)] / d1;
    const double v1 = f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 2)] / d1;
    const double w1 = f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 3)] / d1;
    double T1 = f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 4)];
    T1 = T1 - 0.5 * d1 * (u1 * u1 + v1 * v1 + w1 * w1);
    const int gx = x + LAP, gy = y + LAP, gz = z + LAP;
    d.ptr[fidx(d.pitch, k.nyg, gx, gy, gz)] = d1;
    u.ptr[fidx(u.pitch, k.nyg, gx, gy, gz)] = u1;
    v.ptr[fidx(v.pitch, k.nyg, gx, gy, gz)] = v1;
    w.ptr[fidx(w.pitch, k.nyg, gx, gy, gz)] = w1;
    T.ptr[fidx(T.pitch, k.nyg, gx, gy, gz)] = T1 / (d1 * k.cv);
    if (P.ptr) {
        T1 = T1 / d1;
        P.ptr[fidx(P.pitch, k.nyg, gx, gy, gz)] = T1 * (k.gamma - 1.0);
    }
}

__global__ void __launch_bounds__(NT) k_pri_to_cons(KParams k, ws_soa f, ws_field d, ws_field u, ws_field v,
                                                    ws_field w, ws_field T, int sx, int sy, int sz, int bx,
                                                    int by, int bz)
{
#pragma clang fp contract(off)
    const long total = (long)bx * by * bz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = sx + (int)(i % bx);
    const int y = sy + (int)((i / bx) % by);
    const int z = sz + (int)(i / ((long)bx * by));
    const int gx = x + LAP, gy = y + LAP, gz = z + LAP;
    const double dd = d.ptr[fidx(d.pitch, k.nyg, gx, gy, gz)];
    const double uu = u.ptr[fidx(u.pitch, k.nyg, gx, gy, gz)];
    const double vv = v.ptr[fidx(v.pitch, k.nyg, gx, gy, gz)];
    const double ww = w.ptr[fidx(w.pitch, k.nyg, gx, gy, gz)];
    const double TT = T.ptr[fidx(T.pitch, k.nyg, gx, gy, gz)];
    f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 0)] = dd;
    f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 1)] = uu * dd;
    f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 2)] = vv * dd;
    f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 3)] = ww * dd;
    f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, 4)] = dd * ((uu * uu + vv * vv + ww * ww) * 0.5 + k.cv * TT);
}

template <int KRK>
__global__ void __launch_bounds__(NT) k_time_advance(KParams k, ws_soa f, ws_soa fn, ws_soa du, ws_soa fl,
                                                     double dt_val, const double *dt_dev)
{
#pragma clang fp contract(off)
    const long total = (long)k.nx * k.ny * k.nz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double dt = dt_dev ? *dt_dev : dt_val;
    const int x = (int)(i % k.nx);
    const int y = (int)((i / k.nx) % k.ny);
    const int z = (int)(i / ((long)k.nx * k.ny));
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const size_t a = sidx(f.pitch, k.ny, k.nz, x, y, z, m);
        const double fnv = fn.ptr[sidx(fn.pitch, k.ny, k.nz, x, y, z, m)];
        const double duv = du.ptr[sidx(du.pitch, k.ny, k.nz, x, y, z, m)];
        const double r = rk3_stage_value<KRK>(fnv, KRK == 1 ? 0.0 : f.ptr[a], duv, dt);
        f.ptr[a] = r;
        if (KRK == 3 && fl.ptr)
            fl.ptr[(size_t)(x + LAP) + (size_t)fl.pitch * ((size_t)(y + LAP) + (size_t)(k.ny + 2 * LAP) *
                   ((size_t)(z + LAP) + (size_t)(k.nz + 2 * LAP) * m))] = r;
    }
}

// One RK3 stage of a solver step, fused: the stage update of f (as k_time_advance), the
// primitives of the new f (as k_cons_to_pri), with PLANES the re-derivation of f on the six
// boundary planes from those primitives (as k_pri_to_cons over the planes), and with
// SOUND the sound speed of the interior (as k_sound_speed; a ghost fill then copies it).
// Same expressions with contraction off as the separate kernels, so the state is
// bit-identical; f is read once and the primitives are not read back.  KRK = 0: the last
// RHS launch already wrote the stage's f; this reads it, writes the primitives, and
// writes f back only on the boundary planes it re-derives.  CFLP (the last stage of a
// step): each work-group also writes the largest max(|u|,|v|,|w|) + c of its cells'
// new primitives to cfl_part[blockIdx.x], k_cfl_partial's quantity for the next step.
template <int KRK, bool PLANES, bool SOUND, bool CFLP>
__device__ __forceinline__ double stage_cell(const KParams &k, ws_soa f, ws_soa fn, ws_soa du, ws_soa fl,
                                             double dt_val, const double *dt_dev, ws_field d, ws_field u, ws_field v,
                                             ws_field w, ws_field T, ws_field P, ws_field cc, long i)
{
#pragma clang fp contract(off)
    const double dt = (KRK != 0 && dt_dev) ? *dt_dev : dt_val;
    const int x = (int)(i % k.nx);
    const int y = (int)((i / k.nx) % k.ny);
    const int z = (int)(i / ((long)k.nx * k.ny));
    double r[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        if (KRK == 0) {
            r[m] = f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, m)];
            continue;
        }
        const double fnv = fn.ptr[sidx(fn.pitch, k.ny, k.nz, x, y, z, m)];
        const double duv = du.ptr[sidx(du.pitch, k.ny, k.nz, x, y, z, m)];
        r[m] = rk3_stage_value<KRK == 0 ? 1 : KRK>(fnv, KRK == 1 ? 0.0 : f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, m)],
                                                   duv, dt);
        if (KRK == 3 && fl.ptr)
            fl.ptr[(size_t)(x + LAP) + (size_t)fl.pitch * ((size_t)(y + LAP) + (size_t)(k.ny + 2 * LAP) *
                   ((size_t)(z + LAP) + (size_t)(k.nz + 2 * LAP) * m))] = r[m];
    }
    const CellPrims p = cell_primitives(r, k.cv);
    const int gx = x + LAP, gy = y + LAP, gz = z + LAP;
    d.ptr[fidx(d.pitch, k.nyg, gx, gy, gz)] = p.d;
    u.ptr[fidx(u.pitch, k.nyg, gx, gy, gz)] = p.u;
    v.ptr[fidx(v.pitch, k.nyg, gx, gy, gz)] = p.v;
    w.ptr[fidx(w.pitch, k.nyg, gx, gy, gz)] = p.w;
    T.ptr[fidx(T.pitch, k.nyg, gx, gy, gz)] = p.T;
    if (P.ptr) P.ptr[fidx(P.pitch, k.nyg, gx, gy, gz)] = p.e / p.d * (k.gamma - 1.0);
    const double c = sqrt(p.T) / k.ama;
    if (SOUND) cc.ptr[fidx(cc.pitch, k.nyg, gx, gy, gz)] = c;
    const bool plane = PLANES && (x == 0 || x == k.nx - 1 || y == 0 || y == k.ny - 1 || z == 0 || z == k.nz - 1);
    if (plane) plane_conserved(p, k.cv, r);
    if (KRK != 0 || plane)
#pragma unroll
        for (int m = 0; m < 5; ++m) f.ptr[sidx(f.pitch, k.ny, k.nz, x, y, z, m)] = r[m];
    // k_cfl_partial's expression on the primitives just written
    return CFLP ? fmax(fabs(p.u) + c, fmax(fabs(p.v) + c, fabs(p.w) + c)) : 0.0;
}

// *out = the work-group's maximum of mx (every thread calls it); fmax is exact and ignores
// NaN, so the order of the reduction does not matter
__device__ __forceinline__ void wg_max(double mx, double *out)
{
    __shared__ double wmax[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
#pragma unroll
        for (int q = 1; q < NT / 64; ++q) b = fmax(b, wmax[q]);
        *out = b;
    }
}

template <int KRK, bool PLANES, bool SOUND, bool CFLP>
__global__ void __launch_bounds__(NT) k_stage(KParams k, ws_soa f, ws_soa fn, ws_soa du, ws_soa fl, double dt_val,
                                              const double *dt_dev, ws_field d, ws_field u, ws_field v, ws_field w,
                                              ws_field T, ws_field P, ws_field cc, double *cfl_part)
{
    const long total = (long)k.nx * k.ny * k.nz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double mx = 0.0;
    if (i < total) mx = stage_cell<KRK, PLANES, SOUND, CFLP>(k, f, fn, du, fl, dt_val, dt_dev, d, u, v, w, T, P, cc, i);
    if (CFLP) wg_max(mx, cfl_part + blockIdx.x);
}

__global__ void __launch_bounds__(NT) k_fill(double value, double *ptr, uint32_t pitch, uint32_t sx, uint32_t sy,
                                             uint32_t sz)
{
    const long total = (long)sx * sy * sz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t x = (uint32_t)(i % sx), y = (uint32_t)((i / sx) % sy), z = (uint32_t)(i / ((long)sx * sy));
    ptr[(size_t)x + (size_t)pitch * ((size_t)y + (size_t)z * sy)] = value;
}

// CFL: per-block maxima of max(|u|,|v|,|w|) + c over the interior, then one block
// reduces them and writes dt.  No atomics, so the result is reproducible.
__global__ void __launch_bounds__(NT) k_cfl_partial(KParams k, ws_field u, ws_field v, ws_field w, ws_field T,
                                                    double *partial)
{
    __shared__ double red[NT];
    const long total = (long)k.nx * k.ny * k.nz;
    double mx = 0.0;
    for (long i = (long)blockIdx.x * NT + threadIdx.x; i < total; i += (long)gridDim.x * NT) {
        const int x = (int)(i % k.nx) + LAP;
        const int y = (int)((i / k.nx) % k.ny) + LAP;
        const int z = (int)(i / ((long)k.nx * k.ny)) + LAP;
        const double c = sqrt(T.ptr[fidx(T.pitch, k.nyg, x, y, z)]) / k.ama;
        const double a = fabs(u.ptr[fidx(u.pitch, k.nyg, x, y, z)]) + c;
        const double b = fabs(v.ptr[fidx(v.pitch, k.nyg, x, y, z)]) + c;
        const double e = fabs(w.ptr[fidx(w.pitch, k.nyg, x, y, z)]) + c;
        mx = fmax(mx, fmax(a, fmax(b, e)));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(NT) k_cfl_final(const double *partial, int n, double cfl, double hmin,
                                                  double *dt_out)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    double mx = 0.0;
#pragma unroll 8
    for (int i = threadIdx.x; i < n; i += NT) mx = fmax(mx, partial[i]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *dt_out = cfl * hmin / red[0];
}

__global__ void __launch_bounds__(NT) k_min_reduce(const double *v, int n, double *out)
{
    __shared__ double red[NT];
    double mn = v[0];
    for (int i = threadIdx.x; i < n; i += NT) mn = fmin(mn, v[i]);
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---- batched helpers: every block of a batch in one launch ------------------------------
// Each block owns a contiguous range of the grid's work-groups (AuxBlock); a work-group
// finds its block by bisection over the ascending range starts (uniform, scalar loads) and
// then does what the single-block kernel's work-group of the same index in that block does.

template <bool GHOST>
__device__ __forceinline__ const AuxBlock &aux_block(const AuxBlock *tab, int nb, long wg)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((GHOST ? tab[mid].ghost_wg0 : tab[mid].cell_wg0) <= wg) lo = mid;
        else hi = mid - 1;
    }
    return tab[lo];
}

__global__ void __launch_bounds__(NT) k_ghost_fill_batch(const AuxBlock *tab, int nb, int set)
{
    const long wg = blockIdx.x;
    const AuxBlock &b = aux_block<true>(tab, nb, wg);
    ghost_point(b.g[set][blockIdx.y], b.n[0], b.n[1], b.n[2], b.nyg, true, true, true,
                (wg - b.ghost_wg0) * NT + threadIdx.x);
}

// the sound speed of every block's interior (k_sound_flat's expression)
__global__ void __launch_bounds__(NT) k_sound_batch(const AuxBlock *tab, int nb, double ama)
{
#pragma clang fp contract(off)
    const long wg = blockIdx.x;
    const AuxBlock &b = aux_block<false>(tab, nb, wg);
    const int nx = b.n[0], ny = b.n[1];
    const long i = (wg - b.cell_wg0) * NT + threadIdx.x;
    if (i >= (long)nx * ny * b.n[2]) return;
    const int x = (int)(i % nx) + LAP, y = (int)((i / nx) % ny) + LAP, z = (int)(i / ((long)nx * ny)) + LAP;
    const ws_field &T = b.g[0][4], &cc = b.g[0][5];
    cc.ptr[fidx(cc.pitch, b.nyg, x, y, z)] = sqrt(T.ptr[fidx(T.pitch, b.nyg, x, y, z)]) / ama;
}

// k_stage<0, PLANES, SOUND, CFLP> of every block: work-group cell_wg0 + q of a block covers
// the block's cells [256 q, 256 q + 256), as work-group q of the single-block launch does,
// so the CFL partials of a block are the single-block kernel's
template <bool PLANES, bool SOUND, bool CFLP>
__global__ void __launch_bounds__(NT) k_stage_batch(KParams k, const AuxBlock *tab, int nb, int swapped,
                                                    double *cfl_part)
{
    const long wg = blockIdx.x;
    const AuxBlock &b = aux_block<false>(tab, nb, wg);
    KParams kb = k;
    kb.nx = b.n[0]; kb.ny = b.n[1]; kb.nz = b.n[2]; kb.nyg = b.nyg;
    const long total = (long)kb.nx * kb.ny * kb.nz;
    const long i = (wg - b.cell_wg0) * NT + threadIdx.x;
    const ws_soa f = swapped ? b.fb : b.fa;
    double mx = 0.0;
    if (i < total)
        mx = stage_cell<0, PLANES, SOUND, CFLP>(kb, f, f, f, ws_soa{}, 0.0, nullptr, b.g[0][0], b.g[0][1], b.g[0][2],
                                                b.g[0][3], b.g[0][4], b.P, b.g[0][5], i);
    if (CFLP) wg_max(mx, cfl_part + wg);
}

// one work-group per block: k_cfl_final over the block's partials
__global__ void __launch_bounds__(NT) k_cfl_final_batch(const AuxBlock *tab, const double *part, double cfl,
                                                        double *dts, const int2 *ranges)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    const AuxBlock &b = tab[blockIdx.x];
    const long lo = ranges ? (long)ranges[blockIdx.x].x : b.cell_wg0;
    const long hi = ranges ? (long)ranges[blockIdx.x].y : b.cell_wg1;
    double mx = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += NT) mx = fmax(mx, part[i]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dts[blockIdx.x] = cfl * b.hmin / red[0];
}

inline unsigned blocks_for(long n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

int launch_ghost_fill(const Ctx &c, const ws_field *f, int n, hipStream_t s)
{
    const KParams &k = c.kp;
    const int px = c.cfg.periodic[0], py = c.cfg.periodic[1], pz = c.cfg.periodic[2];
    const long per = 2L * LAP * ((px ? (long)k.ny * k.nz : 0) + (py ? (long)k.nx * k.nz : 0) +
                                 (pz ? (long)k.nx * k.ny : 0));
    if (per == 0) return WS_OK;
    for (int base = 0; base < n; base += 8) {
        FieldList fl{};
        const int cnt = n - base < 8 ? n - base : 8;
        for (int i = 0; i < cnt; ++i) fl.f[i] = f[base + i];
        hipLaunchKernelGGL(k_ghost_fill, dim3(blocks_for(per), cnt), dim3(NT), 0, s, fl, k.nx, k.ny, k.nz, k.nyg,
                           px, py, pz);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "ghost fill launch");
}

int launch_copy_fields(const Ctx &c, const ws_field *from, const ws_field *to, int n, hipStream_t s)
{
    const KParams &k = c.kp;
    const int nxg = k.nx + 2 * LAP;
    const long rows = (long)k.nyg * k.nzg;
    if (n < 1 || n > 8) return set_error(WS_ERR_INVALID, "copy of 1 to 8 fields");
    FieldList a{}, b{};
    for (int q = 0; q < n; ++q) {
        if (from[q].pitch < (uint32_t)nxg || to[q].pitch < (uint32_t)nxg)
            return set_error(WS_ERR_INVALID, "field copy: pitch below the ghosted row");
        a.f[q] = from[q];
        b.f[q] = to[q];
    }
    hipLaunchKernelGGL(k_copy_fields, dim3(blocks_for((long)nxg * rows), n), dim3(NT), 0, s, a, b, nxg, rows);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "field copy launch");
}

int launch_prepare(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field T, ws_field cc,
                   hipStream_t s)
{
    const KParams &k = c.kp;
    const ws_field g[5] = {d, u, v, w, T};
    int r = launch_ghost_fill(c, g, 5, s);
    if (r) return r;
    const ws_job box{{0, 0, 0}, {(uint32_t)(k.nx + 2 * LAP), (uint32_t)(k.ny + 2 * LAP), (uint32_t)(k.nz + 2 * LAP)}};
    const long n = (long)T.pitch * (k.ny + 2 * LAP) * (k.nz + 2 * LAP);
    // the flat form needs one layout and 16-byte alignment; otherwise the box kernel
    if (T.pitch != cc.pitch || n % 2 || ((uintptr_t)T.ptr | (uintptr_t)cc.ptr) % 16)
        return launch_sound_speed(c, T, cc, box, s);
    hipLaunchKernelGGL(k_sound_flat, dim3(blocks_for(n / 2)), dim3(NT), 0, s, (const double2 *)T.ptr,
                       (double2 *)cc.ptr, n / 2, k.ama);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "sound speed launch");
}

int launch_sound_speed(const Ctx &c, ws_field T, ws_field cc, ws_job job, hipStream_t s)
{
    const int bx = (int)(job.end.x - job.start.x), by = (int)(job.end.y - job.start.y),
              bz = (int)(job.end.z - job.start.z);
    hipLaunchKernelGGL(k_sound_speed, dim3(blocks_for((long)bx * by * bz)), dim3(NT), 0, s, T, cc, (int)job.start.x,
                       (int)job.start.y, (int)job.start.z, bx, by, bz, c.kp.nyg, c.kp.ama);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "sound speed launch");
}

int launch_cons_to_pri(const Ctx &c, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w, ws_field T,
                       ws_field P, hipStream_t s)
{
    const KParams &k = c.kp;
    hipLaunchKernelGGL(k_cons_to_pri, dim3(blocks_for((long)k.nx * k.ny * k.nz)), dim3(NT), 0, s, k, f, d, u, v, w,
                       T, P);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "cons_to_pri launch");
}

int launch_pri_to_cons(const Ctx &c, ws_soa f, ws_field d, ws_field u, ws_field v, ws_field w, ws_field T,
                       ws_job job, hipStream_t s)
{
    const int bx = (int)(job.end.x - job.start.x), by = (int)(job.end.y - job.start.y),
              bz = (int)(job.end.z - job.start.z);
    hipLaunchKernelGGL(k_pri_to_cons, dim3(blocks_for((long)bx * by * bz)), dim3(NT), 0, s, c.kp, f, d, u, v, w, T,
                       (int)job.start.x, (int)job.start.y, (int)job.start.z, bx, by, bz);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "pri_to_cons launch");
}

int launch_time_advance(const Ctx &c, int krk, ws_soa f, ws_soa fn, ws_soa du, ws_soa fl, double dt,
                        const double *dt_dev, hipStream_t s)
{
    const KParams &k = c.kp;
    const dim3 g(blocks_for((long)k.nx * k.ny * k.nz));
    if (krk == 1) hipLaunchKernelGGL(k_time_advance<1>, g, dim3(NT), 0, s, k, f, fn, du, fl, dt, dt_dev);
    else if (krk == 2) hipLaunchKernelGGL(k_time_advance<2>, g, dim3(NT), 0, s, k, f, fn, du, fl, dt, dt_dev);
    else hipLaunchKernelGGL(k_time_advance<3>, g, dim3(NT), 0, s, k, f, fn, du, fl, dt, dt_dev);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "time advance launch");
}

int launch_stage(const Ctx &c, int krk, const ws_state &st, ws_soa fl, double dt, const double *dt_dev, bool planes,
                 bool sound, hipStream_t s, double *cfl_part)
{
    const KParams &k = c.kp;
    const dim3 g(blocks_for((long)k.nx * k.ny * k.nz));
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, g, dim3(NT), 0, s, k, st.f, st.fn, st.du, fl, dt, dt_dev, st.d, st.u, st.v, st.w,
                           st.T, st.P, st.cc, cfl_part);
    };
    // krk 0: the primitives (and planes, sound speed) of an f the last RHS launch updated
    if (krk < 0 || krk > 3) return set_error(WS_ERR_INVALID, "krk must be 0, 1, 2 or 3");
    if (cfl_part) {
        // the last stage of a step: krk 3, or 0 after the RHS applied stage 3
        if ((krk != 0 && krk != 3) || sound)
            return set_error(WS_ERR_INVALID, "the CFL maxima come from the last stage of a step");
        if (krk == 0) planes ? go(k_stage<0, true, false, true>) : go(k_stage<0, false, false, true>);
        else planes ? go(k_stage<3, true, false, true>) : go(k_stage<3, false, false, true>);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? WS_OK : hip_error(e, "stage launch");
    }
    const int sel = (krk == 0 ? 12 : (krk - 1) * 4) + (planes ? 2 : 0) + (sound ? 1 : 0);
    switch (sel) {
    case 12: go(k_stage<0, false, false, false>); break;
    case 13: go(k_stage<0, false, true, false>); break;
    case 14: go(k_stage<0, true, false, false>); break;
    case 15: go(k_stage<0, true, true, false>); break;
    case 0: go(k_stage<1, false, false, false>); break;
    case 1: go(k_stage<1, false, true, false>); break;
    case 2: go(k_stage<1, true, false, false>); break;
    case 3: go(k_stage<1, true, true, false>); break;
    case 4: go(k_stage<2, false, false, false>); break;
    case 5: go(k_stage<2, false, true, false>); break;
    case 6: go(k_stage<2, true, false, false>); break;
    case 7: go(k_stage<2, true, true, false>); break;
    case 8: go(k_stage<3, false, false, false>); break;
    case 9: go(k_stage<3, false, true, false>); break;
    case 10: go(k_stage<3, true, false, false>); break;
    case 11: go(k_stage<3, true, true, false>); break;
    default: return set_error(WS_ERR_INVALID, "krk must be 0, 1, 2 or 3");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "stage launch");
}

int launch_fill(double value, double *ptr, uint32_t pitch, uint32_t sx, uint32_t sy, uint32_t sz, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(blocks_for((long)sx * sy * sz)), dim3(NT), 0, s, value, ptr, pitch, sx, sy, sz);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "fill launch");
}

int launch_cfl(Ctx &c, const ws_state &st, double cfl, double *dt_dev, hipStream_t s)
{
    const KParams &k = c.kp;
    if (!c.cfl_partial) {
        c.cfl_blocks = 1024;
        hipError_t e = hipMalloc(&c.cfl_partial, sizeof(double) * c.cfl_blocks);
        if (e != hipSuccess) return hip_error(e, "cfl scratch allocation");
    }
    const double hmin = fmin(k.h[0], fmin(k.h[1], k.h[2]));
    hipLaunchKernelGGL(k_cfl_partial, dim3(c.cfl_blocks), dim3(NT), 0, s, k, st.u, st.v, st.w, st.T, c.cfl_partial);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(NT), 0, s, (const double *)c.cfl_partial, c.cfl_blocks, cfl, hmin,
                       dt_dev);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "cfl launch");
}

int launch_cfl_final(const Ctx &c, const double *partial, int n, double cfl, double *dt_dev, hipStream_t s)
{
    const KParams &k = c.kp;
    const double hmin = fmin(k.h[0], fmin(k.h[1], k.h[2]));
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(NT), 0, s, partial, n, cfl, hmin, dt_dev);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "cfl launch");
}

long stage_blocks(const Ctx &c) { return blocks_for((long)c.kp.nx * c.kp.ny * c.kp.nz); }

int launch_min_reduce(const double *v, int n, double *out, hipStream_t s)
{
    if (n <= 0) return set_error(WS_ERR_INVALID, "empty reduction");
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(NT), 0, s, v, n, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "min reduce launch");
}

int launch_ghost_fill_batch(const AuxBlock *tab, int nb, long ghost_wgs, int set, hipStream_t s)
{
    if (set < 0 || set > 1) return set_error(WS_ERR_INVALID, "field set must be 0 or 1");
    if (ghost_wgs <= 0 || nb <= 0) return WS_OK;
    hipLaunchKernelGGL(k_ghost_fill_batch, dim3((unsigned)ghost_wgs, 6), dim3(NT), 0, s, tab, nb, set);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "batched ghost fill launch");
}

int launch_prepare_batch(const KParams &k, const AuxBlock *tab, int nb, long ghost_wgs, long cell_wgs,
                         hipStream_t s)
{
    if (nb <= 0 || cell_wgs <= 0) return WS_OK;
    hipLaunchKernelGGL(k_sound_batch, dim3((unsigned)cell_wgs), dim3(NT), 0, s, tab, nb, k.ama);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "batched sound speed launch");
    return launch_ghost_fill_batch(tab, nb, ghost_wgs, 0, s);
}

int launch_stage_batch(const KParams &k, const AuxBlock *tab, int nb, long cell_wgs, int swapped, bool planes,
                       bool sound, double *cfl_part, hipStream_t s)
{
    if (nb <= 0 || cell_wgs <= 0) return WS_OK;
    const dim3 g((unsigned)cell_wgs);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, g, dim3(NT), 0, s, k, tab, nb, swapped, cfl_part); };
    const int sel = (planes ? 4 : 0) + (sound ? 2 : 0) + (cfl_part ? 1 : 0);
    switch (sel) {
    case 0: go(k_stage_batch<false, false, false>); break;
    case 1: go(k_stage_batch<false, false, true>); break;
    case 2: go(k_stage_batch<false, true, false>); break;
    case 3: go(k_stage_batch<false, true, true>); break;
    case 4: go(k_stage_batch<true, false, false>); break;
    case 5: go(k_stage_batch<true, false, true>); break;
    case 6: go(k_stage_batch<true, true, false>); break;
    default: go(k_stage_batch<true, true, true>); break;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "batched stage launch");
}

int launch_cfl_batch(const AuxBlock *tab, int nb, const double *cfl_part, double cfl, double *dts, double *dt_dev,
                     hipStream_t s, const int2 *ranges)
{
    if (nb <= 0) return set_error(WS_ERR_INVALID, "empty batch");
    hipLaunchKernelGGL(k_cfl_final_batch, dim3(nb), dim3(NT), 0, s, tab, cfl_part, cfl, dts, ranges);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "batched cfl launch");
    return launch_min_reduce(dts, nb, dt_dev, s);
}

}  // namespace ws
