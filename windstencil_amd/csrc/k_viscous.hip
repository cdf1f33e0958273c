// k_viscous.hip -- viscous terms of the RHS (the reference's du_vis_Jacobian3d path).
//
// Two memory-bound kernels around a ghost fill:
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
// registers.  Expressions follow the reference's order with contraction off, so the result
// is bit-identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

namespace {

constexpr int NT = 256;

struct VisParams {
    int nx, ny, nz, nyg;
    double rh[3];            // unused by the kernels (division by h follows the reference)
    double h[3];
    double amu_c0, tsb, vis_c;
    double metric[10];       // constant Akx..Asz, Ajac
};

struct EvFields {
    ws_field f[12];          // flux m of direction d at f[4*d + m]
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

// central derivative along dir of a ghosted field at ghosted point (x, y, z), divided by h
template <int ORDER>
__device__ __forceinline__ double cd_at(const ws_field &f, int nyg, int dir, int x, int y, int z, double h)
{
#pragma clang fp contract(off)
    constexpr int R = ORDER / 2;
    double s[2 * R + 1];
    const size_t stride = dir == 0 ? 1 : dir == 1 ? (size_t)f.pitch : (size_t)f.pitch * nyg;
    const double *c = f.ptr + fidx(f.pitch, nyg, x, y, z);
#pragma unroll
    for (int q = -R; q <= R; ++q) s[q + R] = c[(long)q * (long)stride];
    const double tmp = ORDER == 8 ? cd_stencil8(s) : cd_stencil6(s);
    return tmp / h;
}

template <int ORDER, bool FIELDS>
__global__ void __launch_bounds__(NT) k_vis_flux(VisParams p, ws_field u, ws_field v, ws_field w, ws_field T,
                                                 ws_metric_fields mf, EvFields ev)
{
#pragma clang fp contract(off)
    const long total = (long)p.nx * p.ny * p.nz;
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= total) return;
    const int x = (int)(i % p.nx) + LAP, y = (int)((i / p.nx) % p.ny) + LAP, z = (int)(i / ((long)p.nx * p.ny)) + LAP;
    const int nyg = p.nyg;
    auto met = [&](int m) -> double {
        if (FIELDS) {
            const ws_field &f = m < 9 ? mf.A[m] : mf.ajac;
            return f.ptr[fidx(f.pitch, nyg, x, y, z)];
        }
        return p.metric[m];
    };
    // gradients along the computational directions (k, i, s = x, y, z)
    const double uk = cd_at<ORDER>(u, nyg, 0, x, y, z, p.h[0]), vk = cd_at<ORDER>(v, nyg, 0, x, y, z, p.h[0]);
    const double wk = cd_at<ORDER>(w, nyg, 0, x, y, z, p.h[0]), Tk = cd_at<ORDER>(T, nyg, 0, x, y, z, p.h[0]);
    const double ui = cd_at<ORDER>(u, nyg, 1, x, y, z, p.h[1]), vi = cd_at<ORDER>(v, nyg, 1, x, y, z, p.h[1]);
    const double wi = cd_at<ORDER>(w, nyg, 1, x, y, z, p.h[1]), Ti = cd_at<ORDER>(T, nyg, 1, x, y, z, p.h[1]);
    const double us = cd_at<ORDER>(u, nyg, 2, x, y, z, p.h[2]), vs = cd_at<ORDER>(v, nyg, 2, x, y, z, p.h[2]);
    const double ws = cd_at<ORDER>(w, nyg, 2, x, y, z, p.h[2]), Ts = cd_at<ORDER>(T, nyg, 2, x, y, z, p.h[2]);
    const double Akx = met(0), Aky = met(1), Akz = met(2), Aix = met(3), Aiy = met(4), Aiz = met(5);
    const double Asx = met(6), Asy = met(7), Asz = met(8), aj1 = met(9);
    // viscosity (get_Amu_kernal) and the stress tensor (vis_flux_ker)
    const double tc = T.ptr[fidx(T.pitch, nyg, x, y, z)];
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
    const size_t gc = fidx(u.pitch, nyg, x, y, z);
    const double uc = u.ptr[gc], vc = v.ptr[fidx(v.pitch, nyg, x, y, z)], wc = w.ptr[fidx(w.pitch, nyg, x, y, z)];
    const double E1 = uc * s11 + vc * s12 + wc * s13 + amuk * Tx;
    const double E2 = uc * s12 + vc * s22 + wc * s23 + amuk * Ty;
    const double E3 = uc * s13 + vc * s23 + wc * s33 + amuk * Tz;
    const double Am[9] = {Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double akx = Am[3 * d] * aj1, aky = Am[3 * d + 1] * aj1, akz = Am[3 * d + 2] * aj1;
        const double e[4] = {akx * s11 + aky * s12 + akz * s13, akx * s12 + aky * s22 + akz * s23,
                             akx * s13 + aky * s23 + akz * s33, akx * E1 + aky * E2 + akz * E3};
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const ws_field &o = ev.f[4 * d + m];
            o.ptr[fidx(o.pitch, nyg, x, y, z)] = e[m];
        }
    }
}

template <int ORDER, bool FIELDS>
__global__ void __launch_bounds__(NT) k_vis_div(VisParams p, EvFields ev, ws_field ajac, ws_soa du)
{
#pragma clang fp contract(off)
    const long total = (long)p.nx * p.ny * p.nz;
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= total) return;
    const int ix = (int)(i % p.nx), iy = (int)((i / p.nx) % p.ny), iz = (int)(i / ((long)p.nx * p.ny));
    const int x = ix + LAP, y = iy + LAP, z = iz + LAP;
    const double aj = FIELDS ? ajac.ptr[fidx(ajac.pitch, p.nyg, x, y, z)] : p.metric[9];
    double *dp = du.ptr + (size_t)ix + (size_t)du.pitch * ((size_t)iy + (size_t)p.ny * (size_t)iz);
    const size_t vstride = (size_t)du.pitch * p.ny * p.nz;
    double acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = dp[(m + 1) * vstride];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = acc[m] + aj * cd_at<ORDER>(ev.f[4 * d + m], p.nyg, d, x, y, z, p.h[d]);
#pragma unroll
    for (int m = 0; m < 4; ++m) dp[(m + 1) * vstride] = acc[m];
}

inline unsigned blocks_for(long n) { return (unsigned)((n + NT - 1) / NT); }

template <int ORDER, bool FIELDS>
void launch_order(const VisParams &vp, ws_field u, ws_field v, ws_field w, ws_field T, const ws_metric_fields &mf,
                  const EvFields &ev, ws_soa du, const Ctx &c, hipStream_t s, int &r)
{
    const unsigned g = blocks_for((long)vp.nx * vp.ny * vp.nz);
    hipLaunchKernelGGL((k_vis_flux<ORDER, FIELDS>), dim3(g), dim3(NT), 0, s, vp, u, v, w, T, mf, ev);
    if ((r = launch_ghost_fill(c, ev.f, 12, s))) return;
    hipLaunchKernelGGL((k_vis_div<ORDER, FIELDS>), dim3(g), dim3(NT), 0, s, vp, ev, mf.ajac, du);
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
    for (int q = 0; q < 12; ++q) ev.f[q] = ws_field{c.vis_scratch + q * per, pitch, 0};
    VisParams vp{};
    vp.nx = cfg.nx; vp.ny = cfg.ny; vp.nz = cfg.nz; vp.nyg = cfg.ny + 2 * LAP;
    vp.h[0] = cfg.hx; vp.h[1] = cfg.hy; vp.h[2] = cfg.hz;
    vp.rh[0] = 1.0 / cfg.hx; vp.rh[1] = 1.0 / cfg.hy; vp.rh[2] = 1.0 / cfg.hz;
    // constants as the reference's opencfd_para_init forms them (src/runtime/OCFD_init.cpp:379-396)
    vp.tsb = 110.4 / cfg.ref_t;
    vp.amu_c0 = 1.0 / cfg.re * (1.0 + 110.4 / cfg.ref_t);
    const double cp = 1.0 / ((cfg.gamma - 1.0) * cfg.ama * cfg.ama);
    vp.vis_c = cp / cfg.pr;
    for (int q = 0; q < 10; ++q) vp.metric[q] = cfg.metric[q];
    ws_metric_fields mf{};
    if (m) mf = *m;
    int r = WS_OK;
    if (cfg.viscous == 8) {
        if (m) launch_order<8, true>(vp, u, v, w, T, mf, ev, du, c, s, r);
        else launch_order<8, false>(vp, u, v, w, T, mf, ev, du, c, s, r);
    } else {
        if (m) launch_order<6, true>(vp, u, v, w, T, mf, ev, du, c, s, r);
        else launch_order<6, false>(vp, u, v, w, T, mf, ev, du, c, s, r);
    }
    if (r) return r;
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WS_OK : hip_error(e, "viscous launch");
}

}  // namespace ws
