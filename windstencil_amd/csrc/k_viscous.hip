// k_viscous.hip -- viscous terms of the RHS (the reference's du_vis_Jacobian3d path).
//
// Two kernels around a ghost fill, each marching column tiles along z:
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
    double rch[3];           // 1 / (60 h) or 1 / (840 h): the fast kernels' derivative scale
    double amu_c0, tsb, vis_c;
    double metric[10];       // constant Akx..Asz, Ajac
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

template <int ORDER, bool FIELDS, bool EXACT>
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
            auto dx = [&](int f) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = lds[b][f][R + t.ty][t.tx + q];
                return cd_dir<ORDER, EXACT>(s, p, 0);
            };
            auto dy = [&](int f) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = lds[b][f][t.ty + q][R + t.tx];
                return cd_dir<ORDER, EXACT>(s, p, 1);
            };
            const double uk = dx(0), vk = dx(1), wk = dx(2), Tk = dx(3);
            const double ui = dy(0), vi = dy(1), wi = dy(2), Ti = dy(3);
            const double us = cd_dir<ORDER, EXACT>(c[0], p, 2), vs = cd_dir<ORDER, EXACT>(c[1], p, 2);
            const double ws = cd_dir<ORDER, EXACT>(c[2], p, 2), Ts = cd_dir<ORDER, EXACT>(c[3], p, 2);
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

template <int ORDER, bool FIELDS, bool EXACT>
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
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = lx[b][m][t.ty][t.tx + q];
                acc[m] = acc[m] + aj * cd_dir<ORDER, EXACT>(s, p, 0);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                double s[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] = ly[b][m][t.ty + q][t.tx];
                acc[m] = acc[m] + aj * cd_dir<ORDER, EXACT>(s, p, 1);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) acc[m] = acc[m] + aj * cd_dir<ORDER, EXACT>(cz[m], p, 2);
#pragma unroll
            for (int m = 0; m < 4; ++m) dp[(m + 1) * vstride] = acc[m];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int q = 0; q < Q - 1; ++q) cz[m][q] = cz[m][q + 1];
    }
}

template <int ORDER, bool FIELDS, bool EXACT>
void launch_order(const VisParams &vp, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields &mf,
                  const EvFields &ev, ws_soa du, const Ctx &c, hipStream_t s, int &r)
{
    const unsigned g = march_blocks(vp.nx, vp.ny, vp.nz);
    hipLaunchKernelGGL((k_vis_flux<ORDER, FIELDS, EXACT>), dim3(g), dim3(NT), 0, s, vp, u, v, w, T, mf, ev);
    ws_field fl[12];
    for (int q = 0; q < 12; ++q) fl[q] = ws_field{ev.f[q], ev.pitch, 0};
    if ((r = launch_ghost_fill(c, fl, 12, s))) return;
    hipLaunchKernelGGL((k_vis_div<ORDER, FIELDS, EXACT>), dim3(g), dim3(NT), 0, s, vp, ev, mf.ajac, du);
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
    // scratch: twelve ghosted flux fields, allocated at the first call (grow-only)
    const uint32_t pitch = (uint32_t)((cfg.nx + 2 * LAP + 31) / 32 * 32);
    const size_t per = (size_t)pitch * (cfg.ny + 2 * LAP) * (cfg.nz + 2 * LAP);
    if (!c.vis_scratch || c.vis_scratch_elems < 12 * per) {
        if (c.vis_scratch) hipFree(c.vis_scratch);
        c.vis_scratch = nullptr;
        c.vis_scratch_elems = 0;
        hipError_t e = hipMalloc(&c.vis_scratch, 12 * per * sizeof(double));
        if (e != hipSuccess) { c.vis_scratch = nullptr; return set_error(WS_ERR_NOMEM, "viscous scratch allocation"); }
        c.vis_scratch_elems = 12 * per;
    }
    EvFields ev;
    for (int q = 0; q < 12; ++q) ev.f[q] = c.vis_scratch + q * per;
    ev.pitch = pitch;
    VisParams vp{};
    vp.nx = cfg.nx; vp.ny = cfg.ny; vp.nz = cfg.nz; vp.nyg = cfg.ny + 2 * LAP;
    vp.h[0] = cfg.hx; vp.h[1] = cfg.hy; vp.h[2] = cfg.hz;
    const double den = cfg.viscous == 8 ? 840.0 : 60.0;
    for (int d = 0; d < 3; ++d) vp.rch[d] = 1.0 / (den * vp.h[d]);
    // constants as the reference's opencfd_para_init forms them (src/runtime/OCFD_init.cpp:379-396)
    vp.tsb = 110.4 / cfg.ref_t;
    vp.amu_c0 = 1.0 / cfg.re * (1.0 + 110.4 / cfg.ref_t);
    const double cp = 1.0 / ((cfg.gamma - 1.0) * cfg.ama * cfg.ama);
    vp.vis_c = cp / cfg.pr;
    for (int q = 0; q < 10; ++q) vp.metric[q] = cfg.metric[q];
    // the anchor kernel family keeps the reference's arithmetic bit for bit
    const bool exact = cfg.kernel == WS_KERNEL_ANCHOR;
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
