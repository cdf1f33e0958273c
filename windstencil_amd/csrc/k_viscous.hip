// k_viscous.hip -- viscous terms of the RHS (the reference's du_vis_Jacobian3d path).
//
// The tiled kernel family with constant metrics runs one fused kernel, k_vis_fused (below),
// whose fluxes never leave the chip.  The anchor family and per-cell metric fields run two
// kernels around a ghost fill, each marching column tiles along z:
//   k_vis_flux: per interior cell, Sutherland viscosity, the twelve central-difference
//               gradients (u, v, w, T along x, y, z), the stress tensor and heat flux, and
//               the four viscous fluxes of each of the three directions (twelve ghosted
//               scratch fields, interior written);
//   ghost fill of the twelve flux fields (periodic faces);
//   k_vis_div:  per interior cell, du[1..4] += Ajac * (central derivative of the direction's
//               fluxes), directions x, y, z in turn.
// The reference runs 12 derivative launches, three flux launches, three exchanges, twelve
// more derivative launches and twelve update launches per RHS, with the gradients and the
// fluxes' derivatives round-tripping through HBM; here the gradients never leave
// registers.
//
// EXACT kernels evaluate every expression in the reference's order with contraction off
// and divide as the reference does ((stencil / 60) / h): bit-identical to the CPU
// restatement; they serve the anchor kernel family.  The fast kernels (the tiled family)
// multiply the stencil sum by 1 / (60 h) (1 / (840 h) for the eighth order): two correctly
// rounded divisions per derivative become one multiplication, which moves a derivative by
// a few units in the last place.
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

namespace {

// A work-group owns a TX x TY column tile of up to ZC consecutive planes and marches it
// along z.  Each thread keeps the z-stencil of its column in registers, so every plane is
// read from HBM once per tile; the x and y stencils of a plane come from an LDS copy of
// the plane's tile and its halo (double-buffered: one barrier per plane).  The loads of
// plane z + 1 are issued before plane z is computed.
constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr int ZC = 32;

struct VisParams {
    int nx, ny, nz, nyg;
    double h[3];
    double rh[3];            // 1 / h
    double rch[3];           // 1 / (60 h) or 1 / (840 h): the fast kernels' derivative scale
    double amu_c0, tsb, vis_c;
    double metric[10];       // constant Akx..Asz, Ajac
    int bnd[6];              // closed faces (D0_bound): x-, x+, y-, y+, z-, z+
    int per[3];              // periodic directions
};

// the twelve flux fields (flux m of direction d at f[4*d + m]), one ghosted layout
struct EvFields {
    double *f[12];
    uint32_t pitch;
};

__device__ __forceinline__ double cd_stencil6(const double *s)
{
#pragma clang fp contract(off)
    return (45.0 * (s[4] - s[2]) - 9.0 * (s[5] - s[1]) + (s[6] - s[0])) / 60.0;
}

__device__ __forceinline__ double cd_stencil8(const double *s)
{
#pragma clang fp contract(off)
    return (672.0 * (s[5] - s[3]) - 168.0 * (s[6] - s[2]) + 32.0 * (s[7] - s[1]) - 3.0 * (s[8] - s[0])) / 840.0;
}

__device__ __forceinline__ double cd_sum6(const double *s)
{
    return 45.0 * (s[4] - s[2]) - 9.0 * (s[5] - s[1]) + (s[6] - s[0]);
}

__device__ __forceinline__ double cd_sum8(const double *s)
{
    return 672.0 * (s[5] - s[3]) - 168.0 * (s[6] - s[2]) + 32.0 * (s[7] - s[1]) - 3.0 * (s[8] - s[0]);
}

// central derivative along direction d of the stencil values s[0 .. ORDER]
template <int ORDER, bool EXACT>
__device__ __forceinline__ double cd_dir(const double *s, const VisParams &p, int d)
{
    if constexpr (EXACT) {
#pragma clang fp contract(off)
        const double tmp = ORDER == 8 ? cd_stencil8(s) : cd_stencil6(s);
        return tmp / p.h[d];
    } else {
        return (ORDER == 8 ? cd_sum8(s) : cd_sum6(s)) * p.rch[d];
    }
}

// The forms of the four cells next to a closed face (the reference's D0_bound,
// OCFD_D0bound_scheme_kernel, src/boundary/OCFD_bound_Scheme.cpp:82-263, consulted by
// OCFD_CD6_kernel / OCFD_CD8_kernel before the central form, src/kernels/OCFD_Schemes.cpp:106-160):
// at distance 0 a one-sided first-order difference, at 1 a second-order, at 2 a fourth-order
// and at 3 a sixth-order central difference, the low face checked first; the reference's
// expression order, contraction off.  s[q] = f(c - R + q); none reads a point past the face.
// True and the undivided derivative in tmp when a form covers cell c of n.
template <int R>
__device__ __forceinline__ bool cd_closure(const double *s0, int c, int n, bool lo, bool hi, double &tmp)
{
#pragma clang fp contract(off)
    const double *s = s0 + R;
    if (lo && c <= 3) {
        if (c == 0) tmp = (s[1] - s[0]);
        else if (c == 1) tmp = (s[1] - s[-1]) * 0.5;
        else if (c == 2) tmp = (s[-2] - 8.0 * s[-1] + 8.0 * s[1] - s[2]) / 12.0;
        else tmp = (s[3] - s[-3] - 9.0 * (s[2] - s[-2]) + 45.0 * (s[1] - s[-1])) / 60.0;
        return true;
    }
    if (hi && c >= n - 4) {
        if (c == n - 1) tmp = (s[0] - s[-1]);
        else if (c == n - 2) tmp = (s[1] - s[-1]) * 0.5;
        else if (c == n - 3) tmp = (s[-2] - 8.0 * s[-1] + 8.0 * s[1] - s[2]) / 12.0;
        else tmp = (s[3] - s[-3] - 9.0 * (s[2] - s[-2]) + 45.0 * (s[1] - s[-1])) / 60.0;
        return true;
    }
    return false;
}

// Derivative along direction d at interior index c: a closure form next to a closed face
// (BND), else the central form.  The exact kernels divide the form by h as the reference
// does (put_d0_kernel), the fast ones multiply by 1/h.
template <int ORDER, bool EXACT, bool BND>
__device__ __forceinline__ double cd_at(const double *s, const VisParams &p, int d, int c)
{
    if constexpr (BND) {
        const int n = d == 0 ? p.nx : d == 1 ? p.ny : p.nz;
        double tmp;
        if (cd_closure<ORDER / 2>(s, c, n, p.bnd[2 * d] != 0, p.bnd[2 * d + 1] != 0, tmp)) {
            if constexpr (EXACT) {
#pragma clang fp contract(off)
                return tmp / p.h[d];
            } else {
                return tmp * p.rh[d];
            }
        }
    }
    return cd_dir<ORDER, EXACT>(s, p, d);
}

// Plane tile with an R-deep halo in x and y (no corners): tile coordinates (px, py),
// px in [0, TX + 2R), py in [0, TY + 2R); the cells are px, py in [R, R + T).
template <int R>
struct Plane {
    static constexpr int PX = TX + 2 * R, PY = TY + 2 * R;
    static constexpr int NH = 2 * R * (TX + TY);       // halo points
    static constexpr int HPT = (NH + NT - 1) / NT;     // halo points per thread
    __device__ static bool xhalo(int h) { return h < 2 * R * TY; }
    __device__ static void halo(int h, int &px, int &py)
    {
        if (xhalo(h)) {
            const int side = h / (R * TY), r = h % (R * TY);
            py = R + r / R;
            px = side ? R + TX + r % R : r % R;
        } else {
            h -= 2 * R * TY;
            const int side = h / (R * TX), r = h % (R * TX);
            px = R + r % TX;
            py = side ? R + TY + r / TX : r / TX;
        }
    }
};

// Work-group -> tile.  The grid has 8 * per work-groups; work-group b runs tile
// (b % 8) * per + b / 8, so the tiles one XCD receives under round-robin dispatch form a
// contiguous range (neighbouring rows of one z-chunk), whose halos share that XCD's L2.
struct March {
    int x0, y0, z0, z1;      // tile origin (interior coordinates) and plane range
    int tx, ty;              // this thread's column in the tile
    bool tile_ok;            // the tile exists
    bool act;                // the column is inside the box: its cells are computed
    bool in;                 // the column is inside the ghosted box: its values are loaded
                             // (columns past the box are the ghosts of the cells near it)
};

__device__ __forceinline__ March march_tile(const VisParams &p)
{
    const int gx = (p.nx + TX - 1) / TX, gy = (p.ny + TY - 1) / TY, gz = (p.nz + ZC - 1) / ZC;
    const int ntiles = gx * gy * gz;
    const int per = (ntiles + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b % 8) * per + b / 8;
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    March m;
    m.x0 = bx * TX; m.y0 = by * TY;
    m.z0 = bz * ZC; m.z1 = min(m.z0 + ZC, p.nz);
    m.tx = (int)threadIdx.x % TX; m.ty = (int)threadIdx.x / TX;
    m.tile_ok = tile < ntiles;
    m.act = m.x0 + m.tx < p.nx && m.y0 + m.ty < p.ny;
    m.in = m.x0 + m.tx < p.nx + LAP && m.y0 + m.ty < p.ny + LAP;
    return m;
}

inline unsigned march_blocks(int nx, int ny, int nz)
{
    const long ntiles = (long)((nx + TX - 1) / TX) * ((ny + TY - 1) / TY) * ((nz + ZC - 1) / ZC);
    return (unsigned)(8 * ((ntiles + 7) / 8));
}

// This thread's halo points of the plane tile: tile coordinates and whether the point
// lies inside the ghosted box (points outside are only read by the stencils of cells
// outside the box)
template <int R>
struct HaloPts {
    int px[Plane<R>::HPT], py[Plane<R>::HPT];
    bool ok[Plane<R>::HPT];
};

template <int R>
__device__ __forceinline__ HaloPts<R> halo_points(const March &t, const VisParams &p)
{
    HaloPts<R> hp;
#pragma unroll
    for (int s = 0; s < Plane<R>::HPT; ++s) {
        const int h = (int)threadIdx.x + s * NT;
        hp.px[s] = hp.py[s] = 0;
        hp.ok[s] = false;
        if (h < Plane<R>::NH) {
            Plane<R>::halo(h, hp.px[s], hp.py[s]);
            const int gx = t.x0 + LAP - R + hp.px[s], gy = t.y0 + LAP - R + hp.py[s];
            hp.ok[s] = gx < p.nx + 2 * LAP && gy < p.nyg;
        }
    }
    return hp;
}

// z-column of a ghosted field through (x, y): element of ghosted plane zg at col.p[zg * col.s]
struct Col {
    const double *p;
    size_t s;
};

__device__ __forceinline__ Col column(const double *ptr, uint32_t pitch, int nyg, int x, int y)
{
    return Col{ptr + fidx(pitch, nyg, x, y, 0), (size_t)pitch * (size_t)nyg};
}

template <int ORDER, bool FIELDS, bool EXACT, bool BND>
__global__ void __launch_bounds__(NT) k_vis_flux(VisParams p, ws_field u, ws_field v, ws_field w, ws_field T,
                                                 ws_metric_fields mf, EvFields ev)
{
#pragma clang fp contract(off)
    constexpr int R = ORDER / 2, Q = 2 * R + 1;
    using PL = Plane<R>;
    __shared__ double lds[2][4][PL::PY][PL::PX];
    const March t = march_tile(p);
    if (!t.tile_ok) return;
    const int nyg = p.nyg;
    const int x = t.x0 + t.tx + LAP, y = t.y0 + t.ty + LAP;
    const ws_field fl[4] = {u, v, w, T};
    const HaloPts<R> hp = halo_points<R>(t, p);
    // slot q of column f holds ghosted plane zg - R + q of the current plane zg
    Col k[4];
    double c[4][Q];
    const int zb = t.z0 + LAP, ze = t.z1 + LAP;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        k[f] = column(fl[f].ptr, fl[f].pitch, nyg, x, y);
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) c[f][q] = t.in ? k[f].p[(size_t)(zb - R + q) * k[f].s] : 0.0;
    }
    // plane zg + 1: its new column slot and its halo points
    double nq[4], hv[PL::HPT][4];
    auto fetch = [&](int zg) {
#pragma unroll
        for (int f = 0; f < 4; ++f) nq[f] = t.in ? k[f].p[(size_t)(zg + R) * k[f].s] : 0.0;
#pragma unroll
        for (int s = 0; s < PL::HPT; ++s) {
            const int gx = t.x0 + LAP - R + hp.px[s], gy = t.y0 + LAP - R + hp.py[s];
#pragma unroll
            for (int f = 0; f < 4; ++f) hv[s][f] = hp.ok[s] ? fl[f].ptr[fidx(fl[f].pitch, nyg, gx, gy, zg)] : 0.0;
        }
    };
    fetch(zb);
    size_t o = fidx(ev.pitch, nyg, x, y, zb);             // output offset of plane zg
    const size_t oplane = (size_t)ev.pitch * nyg;
    for (int zg = zb; zg < ze; ++zg, o += oplane) {
        const int b = zg & 1;
#pragma unroll
        for (int f = 0; f < 4; ++f) c[f][Q - 1] = nq[f];
#pragma unroll
        for (int s = 0; s < PL::HPT; ++s) {
            if ((int)threadIdx.x + s * NT >= PL::NH) break;
#pragma unroll
            for (int f = 0; f < 4; ++f) lds[b][f][hp.py[s]][hp.px[s]] = hv[s][f];
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) lds[b][f][R + t.ty][R + t.tx] = c[f][R];
        if (zg + 1 < ze) fetch(zg + 1);
        __syncthreads();
        if (t.act) {
            auto met = [&](int m) -> double {
                if (FIELDS) {
                    const ws_field &g = m < 9 ? mf.A[m] : mf.ajac;
                    return g.ptr[fidx(g.pitch, nyg, x, y, zg)];
                }
                return p.metric[m];
            };
            // gradients along the computational directions (k, i, s = x, y, z)
            const int cx = x - LAP, cy = y - LAP, cz = zg - LAP;
            auto dx = [&](int f) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = lds[b][f][R + t.ty][t.tx + q];
                return cd_at<ORDER, EXACT, BND>(s, p, 0, cx);
            };
            auto dy = [&](int f) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = lds[b][f][t.ty + q][R + t.tx];
                return cd_at<ORDER, EXACT, BND>(s, p, 1, cy);
            };
            const double uk = dx(0), vk = dx(1), wk = dx(2), Tk = dx(3);
            const double ui = dy(0), vi = dy(1), wi = dy(2), Ti = dy(3);
            const double us = cd_at<ORDER, EXACT, BND>(c[0], p, 2, cz), vs = cd_at<ORDER, EXACT, BND>(c[1], p, 2, cz);
            const double ws = cd_at<ORDER, EXACT, BND>(c[2], p, 2, cz), Ts = cd_at<ORDER, EXACT, BND>(c[3], p, 2, cz);
            const double Akx = met(0), Aky = met(1), Akz = met(2), Aix = met(3), Aiy = met(4), Aiz = met(5);
            const double Asx = met(6), Asy = met(7), Asz = met(8), aj1 = met(9);
            // viscosity (get_Amu_kernal) and the stress tensor (vis_flux_ker)
            const double tc = c[3][R];
            const double amu = p.amu_c0 * sqrt(tc * tc * tc) / (p.tsb + tc);
            const double ux = uk * Akx + ui * Aix + us * Asx, vx = vk * Akx + vi * Aix + vs * Asx;
            const double wx = wk * Akx + wi * Aix + ws * Asx;
            const double uy = uk * Aky + ui * Aiy + us * Asy, vy = vk * Aky + vi * Aiy + vs * Asy;
            const double wy = wk * Aky + wi * Aiy + ws * Asy;
            const double uz = uk * Akz + ui * Aiz + us * Asz, vz = vk * Akz + vi * Aiz + vs * Asz;
            const double wz = wk * Akz + wi * Aiz + ws * Asz;
            const double div = ux + vy + wz;
            const double s11 = (2.0 * ux - 2.0 / 3.0 * div) * amu;
            const double s22 = (2.0 * vy - 2.0 / 3.0 * div) * amu;
            const double s33 = (2.0 * wz - 2.0 / 3.0 * div) * amu;
            const double s12 = (uy + vx) * amu, s13 = (uz + wx) * amu, s23 = (vz + wy) * amu;
            const double amuk = amu * p.vis_c;
            const double Tx = Tk * Akx + Ti * Aix + Ts * Asx;
            const double Ty = Tk * Aky + Ti * Aiy + Ts * Asy;
            const double Tz = Tk * Akz + Ti * Aiz + Ts * Asz;
            const double uc = c[0][R], vc = c[1][R], wc = c[2][R];
            const double E1 = uc * s11 + vc * s12 + wc * s13 + amuk * Tx;
            const double E2 = uc * s12 + vc * s22 + wc * s23 + amuk * Ty;
            const double E3 = uc * s13 + vc * s23 + wc * s33 + amuk * Tz;
            const double Am[9] = {Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double akx = Am[3 * d] * aj1, aky = Am[3 * d + 1] * aj1, akz = Am[3 * d + 2] * aj1;
                ev.f[4 * d + 0][o] = akx * s11 + aky * s12 + akz * s13;
                ev.f[4 * d + 1][o] = akx * s12 + aky * s22 + akz * s23;
                ev.f[4 * d + 2][o] = akx * s13 + aky * s23 + akz * s33;
                ev.f[4 * d + 3][o] = akx * E1 + aky * E2 + akz * E3;
            }
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int q = 0; q < Q - 1; ++q) c[f][q] = c[f][q + 1];
    }
}

template <int ORDER, bool FIELDS, bool EXACT, bool BND>
__global__ void __launch_bounds__(NT) k_vis_div(VisParams p, EvFields ev, ws_field ajac, ws_soa du)
{
#pragma clang fp contract(off)
    constexpr int R = ORDER / 2, Q = 2 * R + 1;
    using PL = Plane<R>;
    // x-direction fluxes with their x-halo, y-direction fluxes with their y-halo
    __shared__ double lx[2][4][TY][PL::PX];
    __shared__ double ly[2][4][PL::PY][TX];
    const March t = march_tile(p);
    if (!t.tile_ok) return;
    const int nyg = p.nyg;
    const int x = t.x0 + t.tx + LAP, y = t.y0 + t.ty + LAP;
    const HaloPts<R> hp = halo_points<R>(t, p);
    const size_t oplane = (size_t)ev.pitch * nyg;
    const size_t o0 = fidx(ev.pitch, nyg, x, y, 0);       // this column's offset in plane 0
    // z-columns of the four z-direction fluxes (slot q: ghosted plane zg - R + q)
    double cz[4][Q];
    const int zb = t.z0 + LAP, ze = t.z1 + LAP;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) cz[m][q] = t.in ? ev.f[8 + m][o0 + (size_t)(zb - R + q) * oplane] : 0.0;
    const size_t vstride = (size_t)du.pitch * p.ny * p.nz;
    const size_t dplane = (size_t)du.pitch * p.ny;
    double *dp = du.ptr + (size_t)(x - LAP) + (size_t)du.pitch * ((size_t)(y - LAP) + (size_t)p.ny * (size_t)t.z0);
    // plane zg + 1: new z-column slots, the column's x and y fluxes, du, the halo points
    double nz_[4], nx_[4], ny_[4], dn[4], hv[PL::HPT][4];
    auto fetch = [&](int zg, const double *q) {
        const size_t oz = o0 + (size_t)zg * oplane;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            nz_[m] = t.in ? ev.f[8 + m][oz + (size_t)R * oplane] : 0.0;
            nx_[m] = t.in ? ev.f[m][oz] : 0.0;
            ny_[m] = t.in ? ev.f[4 + m][oz] : 0.0;
            dn[m] = t.act ? q[(m + 1) * vstride] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < PL::HPT; ++s) {
            const int h = (int)threadIdx.x + s * NT;
            const int gx = t.x0 + LAP - R + hp.px[s], gy = t.y0 + LAP - R + hp.py[s];
            const size_t oh = fidx(ev.pitch, nyg, gx, gy, zg);
            const int d = PL::xhalo(h) ? 0 : 1;
#pragma unroll
            for (int m = 0; m < 4; ++m) hv[s][m] = hp.ok[s] ? ev.f[4 * d + m][oh] : 0.0;
        }
    };
    fetch(zb, dp);
    for (int zg = zb; zg < ze; ++zg, dp += dplane) {
        const int b = zg & 1;
        double acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            cz[m][Q - 1] = nz_[m];
            lx[b][m][t.ty][R + t.tx] = nx_[m];
            ly[b][m][R + t.ty][t.tx] = ny_[m];
            acc[m] = dn[m];
        }
#pragma unroll
        for (int s = 0; s < PL::HPT; ++s) {
            const int h = (int)threadIdx.x + s * NT;
            if (h >= PL::NH) break;
            if (PL::xhalo(h)) {
#pragma unroll
                for (int m = 0; m < 4; ++m) lx[b][m][hp.py[s] - R][hp.px[s]] = hv[s][m];
            } else {
#pragma unroll
                for (int m = 0; m < 4; ++m) ly[b][m][hp.py[s]][hp.px[s] - R] = hv[s][m];
            }
        }
        if (zg + 1 < ze) fetch(zg + 1, dp + dplane);
        __syncthreads();
        if (t.act) {
            const double aj = FIELDS ? ajac.ptr[fidx(ajac.pitch, nyg, x, y, zg)] : p.metric[9];
            const int cx = x - LAP, cy = y - LAP, ck = zg - LAP;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = lx[b][m][t.ty][t.tx + q];
                acc[m] = acc[m] + aj * cd_at<ORDER, EXACT, BND>(s, p, 0, cx);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = ly[b][m][t.ty + q][t.tx];
                acc[m] = acc[m] + aj * cd_at<ORDER, EXACT, BND>(s, p, 1, cy);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) acc[m] = acc[m] + aj * cd_at<ORDER, EXACT, BND>(cz[m], p, 2, ck);
#pragma unroll
            for (int m = 0; m < 4; ++m) dp[(m + 1) * vstride] = acc[m];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int q = 0; q < Q - 1; ++q) cz[m][q] = cz[m][q + 1];
    }
}

// ---- fused form (fast family, constant metrics) ------------------------------------------
//
// One kernel for the whole viscous update: no flux field goes through memory.  A work-group
// of FX x FY = 32 x 16 threads owns the flux points of a (FX - 2R) x (FY - 2R) cell tile plus
// an R-deep ring around it and marches a chunk of planes along z.  At each flux plane every
// thread computes the twelve gradients at its point (x and y stencils from an LDS copy of
// the plane with a further R-deep border, the z stencil from its column's register window)
// and the fluxes its point contributes: x fluxes on the tile's rows, y fluxes on its
// columns, all twelve inside the tile.  The x and y fluxes go to LDS and the tile's threads
// take their derivatives at once; the z fluxes stay in a register ring, so a cell's update
// completes R planes later: du += Ajac * (dx F_x + dy F_y) + Ajac * dz F_z.  Data points are
// addressed in interior coordinates: a periodic direction wraps them onto the interior (the
// flux at a point past the face is the one of its periodic image, bit for bit), a closed one
// clamps them (the closure forms read no point past the face), so neither the fields' ghost
// layers nor any scratch is touched.  Algorithmic traffic 96 B per cell: u, v, w, T in and
// du[1..4] read and written.
constexpr int FX = 32, FY = 16, FNT = FX * FY;

template <int R>
struct Fused {
    static constexpr int TX = FX - 2 * R, TY = FY - 2 * R;     // cells of the tile
    static constexpr int PX = FX + 2 * R, PY = FY + 2 * R;     // plane copy: flux points + border
    static constexpr int NBX = 2 * R * FY;                     // border points left and right
    static constexpr int NB = NBX + 2 * R * FX;                // and below and above (no corners)
    static_assert(NB <= FNT, "one border point per thread");
    // border point h -> plane-copy coordinates
    __device__ static void border(int h, int &px, int &py)
    {
        if (h < NBX) {
            const int side = h / (R * FY), r = h % (R * FY);
            py = R + r / R;
            px = side ? R + FX + r % R : r % R;
        } else {
            h -= NBX;
            const int side = h / (R * FX), r = h % (R * FX);
            px = R + r % FX;
            py = side ? R + FY + r / FX : r / FX;
        }
    }
};

// a coordinate of the data onto the interior [0, n): periodic images, or clamped
__device__ __forceinline__ int wrap_coord(int c, int n, bool periodic)
{
    if (periodic) {
        c %= n;
        return c < 0 ? c + n : c;
    }
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// DIAG: the constant metrics have zero off-diagonal terms (a Cartesian box), so each metric
// product keeps only its diagonal term; the values are the general form's (a zero product
// adds nothing to a sum or an FMA), with about 10 % fewer instructions.
template <int ORDER, bool BND, bool DIAG>
__global__ void __launch_bounds__(FNT) k_vis_fused(VisParams p, ws_field u, ws_field v, ws_field w, ws_field T,
                                                   ws_soa du, int zc)
{
    // contraction stays on here (the fast family's arithmetic need not follow the reference's
    // order): every thread runs the same instructions for its flux point, so the flux at a
    // point past a periodic face still equals the one at its image bit for bit
    constexpr int R = ORDER / 2, Q = 2 * R + 1;
    using FZ = Fused<R>;
    __shared__ double pl[4][FZ::PY][FZ::PX];      // u, v, w, T of the plane: flux points and border
    __shared__ double fxa[4][FZ::TY][FX];          // x fluxes on the tile's rows
    __shared__ double fya[4][FY][FZ::TX];          // y fluxes on the tile's columns
    __shared__ double fzr[Q][4][FZ::TY * FZ::TX];  // z fluxes of the tile, ring of Q planes
    __shared__ double rpl[R + 1][4][FZ::TY * FZ::TX];   // Ajac (dx F_x + dy F_y), ring of R + 1 planes
    const int gx = (p.nx + FZ::TX - 1) / FZ::TX, gy = (p.ny + FZ::TY - 1) / FZ::TY, gz = (p.nz + zc - 1) / zc;
    const int ntiles = gx * gy * gz;
    const int per = (ntiles + 7) / 8;
    const int tile = ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
    if (tile >= ntiles) return;
    const int x0 = (tile % gx) * FZ::TX, y0 = ((tile / gx) % gy) * FZ::TY;
    const int z0 = (tile / (gx * gy)) * zc, z1 = min(z0 + zc, p.nz);
    const int tx = (int)threadIdx.x % FX, ty = (int)threadIdx.x / FX;
    const bool perx = p.per[0] != 0, pery = p.per[1] != 0, perz = p.per[2] != 0;
    // this thread's flux point and (inside the tile) cell
    const int fx = x0 - R + tx, fy = y0 - R + ty;
    const bool xrow = ty >= R && ty < FY - R, ycol = tx >= R && tx < FX - R;   // x / y fluxes needed
    const bool cell = xrow && ycol && fx < p.nx && fy < p.ny;
    const int nyg = p.nyg;
    const ws_field fl[4] = {u, v, w, T};
    auto at = [&](const ws_field &f, int x, int y, int z) {
        return f.ptr[fidx(f.pitch, nyg, x + LAP, y + LAP, z + LAP)];
    };
    // the column of this flux point, and this thread's border point of the plane copy
    const int cxw = wrap_coord(fx, p.nx, perx), cyw = wrap_coord(fy, p.ny, pery);
    int bpx = 0, bpy = 0, bxw = 0, byw = 0;
    const bool bord = (int)threadIdx.x < FZ::NB;
    if (bord) {
        FZ::border((int)threadIdx.x, bpx, bpy);
        bxw = wrap_coord(x0 - 2 * R + bpx, p.nx, perx);
        byw = wrap_coord(y0 - 2 * R + bpy, p.ny, pery);
    }
    // flux planes zf0 .. zf1 - 1 produce the chunk's cells z0 .. z1 - 1
    const int zf0 = z0 - R, zf1 = z1 + R;
    double c[4][Q];                 // the column's u, v, w, T at planes zf - R .. zf + R
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) c[f][q] = at(fl[f], cxw, cyw, wrap_coord(zf0 - R + q, p.nz, perz));
    double nq[4], hb[4];            // next column slot, border point of the next plane
    auto fetch = [&](int zf) {
        const int zc_ = wrap_coord(zf + R, p.nz, perz), zb = wrap_coord(zf, p.nz, perz);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            nq[f] = at(fl[f], cxw, cyw, zc_);
            hb[f] = bord ? at(fl[f], bxw, byw, zb) : 0.0;
        }
    };
    const size_t vstride = (size_t)du.pitch * p.ny * p.nz;
    auto du_at = [&](int z) { return du.ptr + (size_t)fx + (size_t)du.pitch * ((size_t)fy + (size_t)p.ny * (size_t)z); };
    fetch(zf0);
    // flux plane zf keeps its z fluxes in slot (zf - zf0) % Q of fzr and Ajac (dx F_x + dy F_y)
    // in slot (zf - zf0) % (R + 1) of rpl
    const int ci = (ty - R) * FZ::TX + (tx - R);   // this cell's index in the rings (tile threads)
    const double aj = p.metric[9];
    const double Akx = p.metric[0], Aky = p.metric[1], Akz = p.metric[2], Aix = p.metric[3], Aiy = p.metric[4];
    const double Aiz = p.metric[5], Asx = p.metric[6], Asy = p.metric[7], Asz = p.metric[8];
    // One barrier after the plane copy is written and one after the fluxes are: every read of
    // the plane copy precedes the second, every read of the flux arrays the next first one.
    for (int zf = zf0; zf < zf1; ++zf) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            c[f][Q - 1] = nq[f];
            pl[f][ty + R][tx + R] = c[f][R];
            if (bord) pl[f][bpy][bpx] = hb[f];
        }
        // du of the cell this step completes (plane zf - R), used after the second barrier
        const int zo = zf - R;
        const bool out = cell && zo >= z0;
        double duo[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) duo[m] = out ? du_at(zo)[(m + 1) * vstride] : 0.0;
        if (zf + 1 < zf1) fetch(zf + 1);
        __syncthreads();
        // gradients and fluxes at this thread's flux point of plane zf
        {
            auto dx = [&](int f) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = pl[f][ty + R][tx + q];
                return cd_at<ORDER, false, BND>(s, p, 0, fx);
            };
            auto dy = [&](int f) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = pl[f][ty + q][tx + R];
                return cd_at<ORDER, false, BND>(s, p, 1, fy);
            };
            const double uk = dx(0), vk = dx(1), wk = dx(2), Tk = dx(3);
            const double ui = dy(0), vi = dy(1), wi = dy(2), Ti = dy(3);
            const double us = cd_at<ORDER, false, BND>(c[0], p, 2, zf), vs = cd_at<ORDER, false, BND>(c[1], p, 2, zf);
            const double ws = cd_at<ORDER, false, BND>(c[2], p, 2, zf), Ts = cd_at<ORDER, false, BND>(c[3], p, 2, zf);
            const double tc = c[3][R];
            const double amu = p.amu_c0 * sqrt_nr(tc * tc * tc) * recip_nr(p.tsb + tc);
            // physical derivatives from the computational ones (k, i, s)
            auto gx = [&](double gk, double gi, double gs) { return DIAG ? gk * Akx : gk * Akx + gi * Aix + gs * Asx; };
            auto gy = [&](double gk, double gi, double gs) { return DIAG ? gi * Aiy : gk * Aky + gi * Aiy + gs * Asy; };
            auto gz = [&](double gk, double gi, double gs) { return DIAG ? gs * Asz : gk * Akz + gi * Aiz + gs * Asz; };
            const double ux = gx(uk, ui, us), vx = gx(vk, vi, vs), wx = gx(wk, wi, ws);
            const double uy = gy(uk, ui, us), vy = gy(vk, vi, vs), wy = gy(wk, wi, ws);
            const double uz = gz(uk, ui, us), vz = gz(vk, vi, vs), wz = gz(wk, wi, ws);
            const double div = ux + vy + wz;
            const double s11 = (2.0 * ux - 2.0 / 3.0 * div) * amu;
            const double s22 = (2.0 * vy - 2.0 / 3.0 * div) * amu;
            const double s33 = (2.0 * wz - 2.0 / 3.0 * div) * amu;
            const double s12 = (uy + vx) * amu, s13 = (uz + wx) * amu, s23 = (vz + wy) * amu;
            const double amuk = amu * p.vis_c;
            const double Tx = gx(Tk, Ti, Ts), Ty = gy(Tk, Ti, Ts), Tz = gz(Tk, Ti, Ts);
            const double uc = c[0][R], vc = c[1][R], wc = c[2][R];
            const double E1 = uc * s11 + vc * s12 + wc * s13 + amuk * Tx;
            const double E2 = uc * s12 + vc * s22 + wc * s23 + amuk * Ty;
            const double E3 = uc * s13 + vc * s23 + wc * s33 + amuk * Tz;
            auto flux = [&](int d, double out[4]) {
                if constexpr (DIAG) {
                    // row d of the metrics keeps its diagonal term A[4d] only
                    const double a = p.metric[4 * d] * aj;
                    out[0] = a * (d == 0 ? s11 : d == 1 ? s12 : s13);
                    out[1] = a * (d == 0 ? s12 : d == 1 ? s22 : s23);
                    out[2] = a * (d == 0 ? s13 : d == 1 ? s23 : s33);
                    out[3] = a * (d == 0 ? E1 : d == 1 ? E2 : E3);
                } else {
                    const double akx = p.metric[3 * d] * aj, aky = p.metric[3 * d + 1] * aj;
                    const double akz = p.metric[3 * d + 2] * aj;
                    out[0] = akx * s11 + aky * s12 + akz * s13;
                    out[1] = akx * s12 + aky * s22 + akz * s23;
                    out[2] = akx * s13 + aky * s23 + akz * s33;
                    out[3] = akx * E1 + aky * E2 + akz * E3;
                }
            };
            double F[4];
            if (xrow) {
                flux(0, F);
#pragma unroll
                for (int m = 0; m < 4; ++m) fxa[m][ty - R][tx] = F[m];
            }
            if (ycol) {
                flux(1, F);
#pragma unroll
                for (int m = 0; m < 4; ++m) fya[m][ty][tx - R] = F[m];
            }
            if (xrow && ycol) {
                flux(2, F);
                const int slot = (zf - zf0) % Q;
#pragma unroll
                for (int m = 0; m < 4; ++m) fzr[slot][m][ci] = F[m];
            }
        }
        __syncthreads();
        if (xrow && ycol) {
            // x and y derivatives of the plane's fluxes at this cell
            const int ps = (zf - zf0) % (R + 1);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = fxa[m][ty - R][tx - R + q];
                const double ddx = cd_at<ORDER, false, BND>(s, p, 0, fx);
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = fya[m][ty - R + q][tx - R];
                const double ddy = cd_at<ORDER, false, BND>(s, p, 1, fy);
                rpl[ps][m][ci] = aj * ddx + aj * ddy;
            }
            // the cell of plane zf - R is complete: its z fluxes span planes zf - 2R .. zf (the
            // ring slots the thread wrote itself, so no barrier is needed for them)
            if (out) {
                double *dp = du_at(zo);
                const int s0 = (zf - zf0 + 1) % Q;      // slot of flux plane zf - 2R
                const int po = (zf - zf0 + 1) % (R + 1);  // slot of plane zf - R
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    double sz[Q];
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        const int sl = s0 + q;
                        sz[q] = fzr[sl >= Q ? sl - Q : sl][m][ci];
                    }
                    const double ddz = cd_at<ORDER, false, BND>(sz, p, 2, zo);
                    dp[(m + 1) * vstride] = (duo[m] + rpl[po][m][ci]) + aj * ddz;
                }
            }
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int q = 0; q < Q - 1; ++q) c[f][q] = c[f][q + 1];
    }
}

inline unsigned fused_blocks(int nx, int ny, int nz, int order, int zc)
{
    const int R = order / 2;
    const long ntiles = (long)((nx + FX - 2 * R - 1) / (FX - 2 * R)) * ((ny + FY - 2 * R - 1) / (FY - 2 * R)) *
                        ((nz + zc - 1) / zc);
    return (unsigned)(8 * ((ntiles + 7) / 8));
}

template <int ORDER, bool FIELDS, bool EXACT, bool BND>
void launch_order_b(const VisParams &vp, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields &mf,
                    const EvFields &ev, ws_soa du, const Ctx &c, hipStream_t s, int &r)
{
    const unsigned g = march_blocks(vp.nx, vp.ny, vp.nz);
    hipLaunchKernelGGL((k_vis_flux<ORDER, FIELDS, EXACT, BND>), dim3(g), dim3(NT), 0, s, vp, u, v, w, T, mf, ev);
    // periodic directions only: the closure forms read no flux past a closed face
    ws_field fl[12];
    for (int q = 0; q < 12; ++q) fl[q] = ws_field{ev.f[q], ev.pitch, 0};
    if ((r = launch_ghost_fill(c, fl, 12, s))) return;
    hipLaunchKernelGGL((k_vis_div<ORDER, FIELDS, EXACT, BND>), dim3(g), dim3(NT), 0, s, vp, ev, mf.ajac, du);
}

template <int ORDER, bool FIELDS, bool EXACT>
void launch_order(const VisParams &vp, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields &mf,
                  const EvFields &ev, ws_soa du, const Ctx &c, hipStream_t s, int &r)
{
    bool bnd = false;
    for (int q = 0; q < 6; ++q) bnd = bnd || vp.bnd[q];
    if (bnd) launch_order_b<ORDER, FIELDS, EXACT, true>(vp, u, v, w, T, mf, ev, du, c, s, r);
    else launch_order_b<ORDER, FIELDS, EXACT, false>(vp, u, v, w, T, mf, ev, du, c, s, r);
}

template <int ORDER, bool EXACT>
void launch_fields(const VisParams &vp, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields *m,
                   const EvFields &ev, ws_soa du, const Ctx &c, hipStream_t s, int &r)
{
    ws_metric_fields mf{};
    if (m) {
        mf = *m;
        launch_order<ORDER, true, EXACT>(vp, u, v, w, T, mf, ev, du, c, s, r);
    } else {
        launch_order<ORDER, false, EXACT>(vp, u, v, w, T, mf, ev, du, c, s, r);
    }
}

}  // namespace

int launch_viscous(Ctx &c, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields *m, ws_soa du,
                   hipStream_t s)
{
    const ws_config &cfg = c.cfg;
    if (cfg.viscous != 6 && cfg.viscous != 8) return set_error(WS_ERR_INVALID, "viscous terms are off (cfg.viscous)");
    VisParams vp{};
    vp.nx = cfg.nx; vp.ny = cfg.ny; vp.nz = cfg.nz; vp.nyg = cfg.ny + 2 * LAP;
    vp.h[0] = cfg.hx; vp.h[1] = cfg.hy; vp.h[2] = cfg.hz;
    const double den = cfg.viscous == 8 ? 840.0 : 60.0;
    for (int d = 0; d < 3; ++d) {
        vp.rch[d] = 1.0 / (den * vp.h[d]);
        vp.rh[d] = 1.0 / vp.h[d];
    }
    for (int q = 0; q < 6; ++q) vp.bnd[q] = cfg.d0_bound[q] ? 1 : 0;
    for (int q = 0; q < 3; ++q) vp.per[q] = cfg.periodic[q] ? 1 : 0;
    // constants as the reference's opencfd_para_init forms them (src/runtime/OCFD_init.cpp:379-396)
    vp.tsb = 110.4 / cfg.ref_t;
    vp.amu_c0 = 1.0 / cfg.re * (1.0 + 110.4 / cfg.ref_t);
    const double cp = 1.0 / ((cfg.gamma - 1.0) * cfg.ama * cfg.ama);
    vp.vis_c = cp / cfg.pr;
    for (int q = 0; q < 10; ++q) vp.metric[q] = cfg.metric[q];
    // the anchor kernel family keeps the reference's arithmetic bit for bit
    const bool exact = cfg.kernel == WS_KERNEL_ANCHOR;
    bool bnd = false;
    for (int q = 0; q < 6; ++q) bnd = bnd || vp.bnd[q];
    if (!exact && !m) {
        // the fast family with constant metrics: one fused kernel, no scratch
        const int zc = c.vis_zchunk;
        const unsigned g = fused_blocks(cfg.nx, cfg.ny, cfg.nz, cfg.viscous, zc);
        const double *A = cfg.metric;
        const bool diag = A[1] == 0.0 && A[2] == 0.0 && A[3] == 0.0 && A[5] == 0.0 && A[6] == 0.0 && A[7] == 0.0;
        auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(g), dim3(FNT), 0, s, vp, u, v, w, T, du, zc); };
        const int sel = (cfg.viscous == 8 ? 4 : 0) + (bnd ? 2 : 0) + (diag ? 1 : 0);
        switch (sel) {
        case 0: go(k_vis_fused<6, false, false>); break;
        case 1: go(k_vis_fused<6, false, true>); break;
        case 2: go(k_vis_fused<6, true, false>); break;
        case 3: go(k_vis_fused<6, true, true>); break;
        case 4: go(k_vis_fused<8, false, false>); break;
        case 5: go(k_vis_fused<8, false, true>); break;
        case 6: go(k_vis_fused<8, true, false>); break;
        default: go(k_vis_fused<8, true, true>); break;
        }
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? WS_OK : hip_error(e, "fused viscous launch");
    }
    // the two-kernel form: twelve ghosted flux fields of scratch, allocated at the first call
    const uint32_t pitch = (uint32_t)((cfg.nx + 2 * LAP + 31) / 32 * 32);
    const size_t per = (size_t)pitch * (cfg.ny + 2 * LAP) * (cfg.nz + 2 * LAP);
    if (!c.vis_scratch || c.vis_scratch_elems < 12 * per) {
        if (c.vis_scratch) (void)hipFree(c.vis_scratch);
        c.vis_scratch = nullptr;
        c.vis_scratch_elems = 0;
        hipError_t e = hipMalloc(&c.vis_scratch, 12 * per * sizeof(double));
        if (e != hipSuccess) { c.vis_scratch = nullptr; return set_error(WS_ERR_NOMEM, "viscous scratch allocation"); }
        c.vis_scratch_elems = 12 * per;
    }
    EvFields ev;
    for (int q = 0; q < 12; ++q) ev.f[q] = c.vis_scratch + q * per;
    ev.pitch = pitch;
    int r = WS_OK;
    if (cfg.viscous == 8) {
        if (exact) launch_fields<8, true>(vp, u, v, w, T, m, ev, du, c, s, r);
        else launch_fields<8, false>(vp, u, v, w, T, m, ev, du, c, s, r);
    } else {
        if (exact) launch_fields<6, true>(vp, u, v, w, T, m, ev, du, c, s, r);
        else launch_fields<6, false>(vp, u, v, w, T, m, ev, du, c, s, r);
    }
    if (r) return r;
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "viscous launch");
}

}  // namespace ws
