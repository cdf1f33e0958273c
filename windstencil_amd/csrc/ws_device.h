// ws_device.h -- device-side building blocks of the inviscid RHS (FP64).
//
// Two forms of each scalar kernel live here:
//   *_ref   the reference's operation order with contraction disabled, so that a
//           kernel built from them reproduces the CPU restatement bit for bit;
//   *_fast  the form the tiled kernels use: WENO smoothness written through
//           forward differences, product weights through prefix/suffix
//           products, and the division deferred so that the positive and
//           negative reconstructions of a face share one.
#pragma once

#include <hip/hip_runtime.h>

#include "ws_internal.h"

namespace ws {

__device__ __forceinline__ size_t fidx(uint32_t pitch, int nyg, int x, int y, int z)
{
    return (size_t)x + (size_t)pitch * ((size_t)y + (size_t)nyg * (size_t)z);
}

__device__ __forceinline__ size_t sidx(uint32_t pitch, int ny, int nz, int x, int y, int z, int m)
{
    return (size_t)x + (size_t)pitch * ((size_t)y + (size_t)ny * ((size_t)z + (size_t)nz * (size_t)m));
}

// ---------------------------------------------------------------------------
// Reference-order forms (bitwise anchor)
// ---------------------------------------------------------------------------

// WENO7-SYMBO positive reconstruction, OCFD_Schemes.cpp:306-368; s[0..7] = f(i-4..i+3),
// result at face i-1/2.  `R` is an index map so that the negative rule reuses the text:
// the negative reconstruction (OCFD_Schemes.cpp:491-553) is the positive rule on the
// reversed window, which is exactly how the reference spells it index by index.
template <bool REV>
__device__ __forceinline__ double weno_ref(const double *w)
{
#pragma clang fp contract(off)
    auto s = [&](int k) -> double { return REV ? w[7 - k] : w[k]; };
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, S3 = 0.0, S4 = 0.0, t, t1;
    t =  -2.0*s(0) +  9.0*s(1) - 18.0*s(2) + 11.0*s(3); S0 += 720.0*t*t;
    t =       s(1) -  6.0*s(2) +  3.0*s(3) +  2.0*s(4); S1 += 720.0*t*t;
    t =  -2.0*s(2) -  3.0*s(3) +  6.0*s(4) -      s(5); S2 += 720.0*t*t;
    t = -11.0*s(3) + 18.0*s(4) -  9.0*s(5) +  2.0*s(6); S3 += 720.0*t*t;
    t = -26.0*s(4) + 57.0*s(5) - 42.0*s(6) + 11.0*s(7); S4 += 720.0*t*t;

    t = -6.0*s(0) + 24.0*s(1) - 30.0*s(2) + 12.0*s(3); S0 += 780.0*t*t;
    t =                          6.0*s(2) - 12.0*s(3) +  6.0*s(4); S1 += 780.0*t*t;
    t =                          6.0*s(3) - 12.0*s(4) +  6.0*s(5); S2 += 780.0*t*t;
    t = 12.0*s(3) - 30.0*s(4) + 24.0*s(5) -  6.0*s(6); S3 += 780.0*t*t;
    t = 18.0*s(4) - 48.0*s(5) + 42.0*s(6) - 12.0*s(7); S4 += 780.0*t*t;

    t = -6.0*s(0) + 18.0*(s(1) - s(2)) + 6.0*s(3); S0 += 781.0*t*t;
    t = -6.0*s(1) + 18.0*(s(2) - s(3)) + 6.0*s(4); S1 += 781.0*t*t;
    t = -6.0*s(2) + 18.0*(s(3) - s(4)) + 6.0*s(5); S2 += 781.0*t*t;
    t = -6.0*s(3) + 18.0*(s(4) - s(5)) + 6.0*s(6); S3 += 781.0*t*t;
    t = -6.0*s(4) + 18.0*(s(5) - s(6)) + 6.0*s(7); S4 += 781.0*t*t;

    t  = fmax(S0, S1);
    t1 = fmax(S2, S3);
    t  = fmax(t, t1);
    S4 = fmax(t, S4);
    {
        double t2, t3;
        t  = (0.0401954833730)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S3)*(2.592e-6+S4);
        t1 = (0.2493800006710)*(2.592e-6+S0)*(2.592e-6+S2)*(2.592e-6+S3)*(2.592e-6+S4);
        t2 = (0.4802686256260)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S3)*(2.592e-6+S4);
        t3 = (0.2009775476730)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S4);
        S4 = (0.0291783426580)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S3);
        S3 = t3; S2 = t2; S1 = t1; S0 = t;
    }
    t1 = 0.0;
    t = -3.0*s(0) + 13.0*s(1) - 23.0*s(2) + 25.0*s(3); t1 += S0*t;
    t =      s(1) -  5.0*s(2) + 13.0*s(3) +  3.0*s(4); t1 += S1*t;
    t =     -s(2) +  7.0*s(3) +  7.0*s(4) -      s(5); t1 += S2*t;
    t =  3.0*s(3) + 13.0*s(4) -  5.0*s(5) +      s(6); t1 += S3*t;
    t = 25.0*s(4) - 23.0*s(5) + 13.0*s(6) -  3.0*s(7); t1 += S4*t;
    t1 /= (12.0*(S0+S1+S2+S3+S4));
    return t1;
}

// Steger-Warming split, canonical form (OCFD_NS_Jacobian3d.cpp:69-127), no contraction.
__device__ __forceinline__ void split_ref(const KParams &k, double d, double u, double v, double w,
                                          double cc, double A1, double A2, double A3,
                                          double fp[5], double fm[5])
{
#pragma clang fp contract(off)
    double ss = sqrt(A1*A1 + A2*A2 + A3*A3);
    double Ak1 = A1 / ss, Ak2 = A2 / ss, Ak3 = A3 / ss;
    double vs = A1*u + A2*v + A3*w;
    double E1 = vs, E2 = vs - cc*ss, E3 = vs + cc*ss;
    double eps2 = k.eps2;
    double E1P = (E1 + sqrt(E1*E1 + eps2)) * 0.50;
    double E2P = (E2 + sqrt(E2*E2 + eps2)) * 0.50;
    double E3P = (E3 + sqrt(E3*E3 + eps2)) * 0.50;
    double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    double tmp0 = d / (2.0*k.gamma);
    double uc1 = u - cc*Ak1, uc2 = u + cc*Ak1;
    double vc1 = v - cc*Ak2, vc2 = v + cc*Ak2;
    double wc1 = w - cc*Ak3, wc2 = w + cc*Ak3;
    double vvc1 = (uc1*uc1 + vc1*vc1 + wc1*wc1) * 0.50;
    double vvc2 = (uc2*uc2 + vc2*vc2 + wc2*wc2) * 0.50;
    double vv = (k.gamma - 1.0) * (u*u + v*v + w*w);
    double W2 = k.c3*cc*cc;
    const double C1 = k.c1;
    fp[0] = tmp0 * (C1*E1P + E2P + E3P);
    fp[1] = tmp0 * (C1*E1P*u + E2P*uc1 + E3P*uc2);
    fp[2] = tmp0 * (C1*E1P*v + E2P*vc1 + E3P*vc2);
    fp[3] = tmp0 * (C1*E1P*w + E2P*wc1 + E3P*wc2);
    fp[4] = tmp0 * (E1P*vv + E2P*vvc1 + E3P*vvc2 + W2*(E2P + E3P));
    fm[0] = tmp0 * (C1*E1M + E2M + E3M);
    fm[1] = tmp0 * (C1*E1M*u + E2M*uc1 + E3M*uc2);
    fm[2] = tmp0 * (C1*E1M*v + E2M*vc1 + E3M*vc2);
    fm[3] = tmp0 * (C1*E1M*w + E2M*wc1 + E3M*wc2);
    fm[4] = tmp0 * (E1M*vv + E2M*vvc1 + E3M*vvc2 + W2*(E2M + E3M));
}

// SSP-RK3 stage value of one conserved component (src/solver/OCFD_time.cpp:15-65) in the
// reference's order with contraction off: fnv the step's initial state, fv the stage's
// current f (unused in stage 1), duv the RHS.  Shared by the stage kernels and the RHS
// launch that applies the stage itself, so both give the same bits.
template <int KRK>
__device__ __forceinline__ double rk3_stage_value(double fnv, double fv, double duv, double dt)
{
#pragma clang fp contract(off)
    if (KRK == 1) return fnv + dt * duv;
    if (KRK == 2) {
        const double t1 = 3.0 / 4.0, t2 = 1.0 / 4.0;
        return t1 * fnv + t2 * (fv + dt * duv);
    }
    const double t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
    return t1 * fnv + t2 * (fv + dt * duv);
}

// du + this direction's flux difference, without contraction: a launch that adds to du and
// one that applies the RK3 stage to the sum then hold the same bits whatever the compiler
// fuses in either.
__device__ __forceinline__ double acc_sum(double duv, double val)
{
#pragma clang fp contract(off)
    return duv + val;
}

// Primitives of one cell's conserved values r (get_duvwT, src/kernels/commen_kernel.cpp:78-110)
// with contraction off: e is the internal energy per volume, T = e / (d Cv).  Shared by the
// stage kernel and the RHS launch that derives the primitives itself.
struct CellPrims {
    double d, u, v, w, e, T;
};

__device__ __forceinline__ CellPrims cell_primitives(const double r[5], double cv)
{
#pragma clang fp contract(off)
    CellPrims p;
    p.d = r[0];
    p.u = r[1] / p.d;
    p.v = r[2] / p.d;
    p.w = r[3] / p.d;
    p.e = r[4] - 0.5 * p.d * (p.u * p.u + p.v * p.v + p.w * p.w);
    p.T = p.e / (p.d * cv);
    return p;
}

// Conserved values re-derived from the primitives on a boundary plane (OCFD_bc,
// src/boundary/OCFD_boundary.cpp:18-92), contraction off.
__device__ __forceinline__ void plane_conserved(const CellPrims &p, double cv, double r[5])
{
#pragma clang fp contract(off)
    r[0] = p.d;
    r[1] = p.u * p.d;
    r[2] = p.v * p.d;
    r[3] = p.w * p.d;
    r[4] = p.d * ((p.u * p.u + p.v * p.v + p.w * p.w) * 0.5 + cv * p.T);
}

// ---------------------------------------------------------------------------
// Fast forms (tiled kernels)
// ---------------------------------------------------------------------------

// Nonlinear weights are homogeneous of degree four in (eps + S_k), so the indicators may be
// scaled freely.  The FP64 form scales them by 1/(720*39): the second-difference term
// (780*36/720 = 39 in units of the 720 term) gets coefficient 1, the first-derivative
// term 1/39 (folded into the candidates' difference coefficients as 1/sqrt(39)), the
// third-difference term 39.05/39, and eps becomes 2.592e-6/(720*39).  The FP32 form keeps
// the 1/720 scale (eps 3.6e-9): at 1/(720*39) the product of four (eps + S) factors of a
// smooth window would fall below FP32's normal range.
constexpr double kD0 = 0.0401954833730;
constexpr double kD1 = 0.2493800006710;
constexpr double kD2 = 0.4802686256260;
constexpr double kD3 = 0.2009775476730;
constexpr double kD4 = 0.0291783426580;
constexpr double kEps39 = 2.592e-6 / (720.0 * 39.0);
constexpr double kC39 = 39.05 / 39.0;
constexpr double kA39 = 0.16012815380508713;   // 1/sqrt(39)

// Reciprocal: hardware estimate plus two Newton steps (error within a couple of ulp;
// the argument is a positive normal number wherever it is used).
__device__ __forceinline__ double recip_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Square root of a positive normal number: reciprocal-square-root estimate refined
// with one coupled Newton step on (sqrt, 1/(2 sqrt)) and a final residual correction.
__device__ __forceinline__ double sqrt_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, x);
    return fma(d, h, g);
}

// K consecutive faces from K+7 points s[0..K+6] (face j uses s[j..j+7]).  The forward
// differences, the scaled first differences, the squared second differences and the
// squared third differences (with eps added) are formed once per point and shared by the
// K faces.  The linear weights and the 1/12 of the candidates are folded into the
// candidate coefficients, and the weighted sums are evaluated in Horner form over the
// prefix products of (eps + S_k).  The candidates are taken relative to the face's base
// point w3 and written in Newton form over the shared forward differences, so they cost
// three FMAs each; face j's value is base[j] + num[j] / den[j].
template <bool REV, int K, typename Load>
__device__ __forceinline__ void weno_k(Load ld, double (&num)[K], double (&den)[K], double (&base)[K])
{
    // Every fused multiply-add here is written out: implicit contraction could fuse a shared
    // difference differently at different positions j, and a face computed by two threads
    // (at a tile edge) must come out bit for bit the same in both.
#pragma clang fp contract(off)
    constexpr int NP = K + 7;
    double s[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) s[i] = ld(i);
    double D1[NP - 1], D2[NP - 2], D3[NP - 3], G1[NP - 1], C3[NP - 3], B2[NP - 2];
#pragma unroll
    for (int i = 0; i < NP - 1; ++i) { D1[i] = s[i + 1] - s[i]; G1[i] = (6.0 * kA39) * D1[i]; }
#pragma unroll
    for (int i = 0; i < NP - 2; ++i) { D2[i] = D1[i + 1] - D1[i]; B2[i] = D2[i] * D2[i]; }
#pragma unroll
    for (int i = 0; i < NP - 3; ++i) { D3[i] = D2[i + 1] - D2[i]; C3[i] = fma(kC39 * D3[i], D3[i], kEps39); }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        // window point k of this face, and the anchors of candidate k's differences
        auto w = [&](int k) { return REV ? s[j + 7 - k] : s[j + k]; };
        auto i1 = [&](int k) { return REV ? j + 6 - k : j + k; };
        auto i2 = [&](int k) { return REV ? j + 5 - k : j + k; };
        auto i3 = [&](int k) { return REV ? j + 4 - k : j + k; };
        const double sb = REV ? -kA39 : kA39;
        // first-derivative terms over sqrt(39) (the sign of a term is irrelevant: it is squared)
        const double a0 = fma(11.0 * kA39, D3[i3(0)], fma(sb * 15.0, D2[i2(0)], G1[i1(0)]));
        const double a1 = fma(2.0 * kA39, D3[i3(1)], fma(sb * 9.0, D2[i2(1)], G1[i1(1)]));
        const double a2 = fma(-1.0 * kA39, D3[i3(2)], fma(sb * 3.0, D2[i2(2)], G1[i1(2)]));
        const double a3 = fma(2.0 * kA39, D3[i3(3)], fma(sb * -3.0, D2[i2(3)], G1[i1(3)]));
        const double a4 = fma(11.0 * kA39, D3[i3(4)], fma(sb * -9.0, D2[i2(4)], G1[i1(4)]));
        // second-derivative terms of candidates 0, 3, 4 (1 and 2 are plain second differences)
        const double b0 = REV ? fma(-2.0, D3[j + 4], D2[j + 5]) : fma(2.0, D3[j], D2[j]);
        const double b3 = REV ? D2[j + 2] + D3[j + 1] : D2[j + 3] - D3[j + 3];
        const double b4 = REV ? fma(2.0, D3[j], D2[j + 1]) : fma(-2.0, D3[j + 4], D2[j + 4]);
        const double bb1 = REV ? B2[j + 3] : B2[j + 2];
        const double bb2 = REV ? B2[j + 2] : B2[j + 3];
        // e_k = eps + S_k (eps rides in C3)
        const double e0 = fma(a0, a0, fma(b0, b0, C3[i3(0)]));
        const double e1 = fma(a1, a1, bb1 + C3[i3(1)]);
        const double e2 = fma(a2, a2, bb2 + C3[i3(2)]);
        const double e3 = fma(a3, a3, fma(b3, b3, C3[i3(3)]));
        double e4 = fma(a4, a4, fma(b4, b4, C3[i3(4)]));
        e4 = fmax(fmax(fmax(e0, e1), fmax(e2, e3)), e4);
        // candidates relative to the base point w(3), in Newton form over the forward
        // differences at each candidate's first point (linear weights folded in): each
        // candidate's coefficients sum to 12, so the base drops out
        const double sd = REV ? -1.0 : 1.0;   // sign of odd differences of the reversed window
        auto d1 = [&](int k) { return D1[REV ? j + 6 - k : j + k]; };
        auto d2 = [&](int k) { return D2[REV ? j + 5 - k : j + k]; };
        auto d3 = [&](int k) { return D3[REV ? j + 4 - k : j + k]; };
        const double c0 = fma(sd * kD0 * (13.0 / 12.0), d3(0), fma(kD0 * (4.0 / 3.0), d2(0), (sd * kD0 * 0.5) * d1(0)));
        const double c1 = fma(sd * kD1 * (1.0 / 4.0), d3(1), fma(kD1 * (5.0 / 6.0), d2(1), (sd * kD1 * 0.5) * d1(1)));
        const double c2 = fma(sd * kD2 * (-1.0 / 12.0), d3(2), fma(kD2 * (1.0 / 3.0), d2(2), (sd * kD2 * 0.5) * d1(2)));
        const double c3 = fma(sd * kD3 * (1.0 / 12.0), d3(3), fma(kD3 * (-1.0 / 6.0), d2(3), (sd * kD3 * 0.5) * d1(3)));
        const double c4 = fma(sd * kD4 * (-1.0 / 4.0), d3(4), fma(kD4 * (1.0 / 3.0), d2(4),
                              fma(sd * kD4 * -0.5, d1(4), (sd * kD4) * d1(3))));
        base[j] = w(3);
        const double p01 = e0 * e1;
        double N = fma(c0, e1, c1 * e0);
        double D = fma(kD0, e1, kD1 * e0);
        N = fma(N, e2, c2 * p01);
        D = fma(D, e2, kD2 * p01);
        const double p012 = p01 * e2;
        N = fma(N, e3, c3 * p012);
        D = fma(D, e3, kD3 * p012);
        const double p0123 = p012 * e3;
        num[j] = fma(N, e4, c4 * p0123);
        den[j] = fma(D, e4, kD4 * p0123);
    }
}

// One face of weno_k from the 8-point window q[0..7]: base + num / den is the face value.
template <bool REV>
__device__ __forceinline__ void face64(const double *q, double &num, double &den, double &base)
{
    double n[1], d[1], b[1];
    weno_k<REV, 1>([&](int i) { return q[i]; }, n, d, b);
    num = n[0];
    den = d[0];
    base = b[0];
}

// Mixed-precision form of weno_k: the window stays FP64 in LDS; its first differences
// are formed in FP64 and rounded to FP32, and everything after that (second and third
// differences, smoothness, weights, candidates) runs in FP32 on values taken relative to
// the face's base point w3.  The face value is base[j] + corr[j]: base in FP64, the FP32
// correction carries an error relative to the local variation of the window, not to the
// size of the flux.
//
// Range check: the weights are products of four (eps + S_k) factors, homogeneous of degree
// eight in the window's differences, and the numerator of degree nine.  D is at least
// d_4 eps^4 (about 5e-36 with eps = 3.6e-9), so it never falls below FP32's normal range,
// but it overflows once the window's first differences reach a few hundred (a strong
// discontinuity of the split flux); N may overflow a little earlier.  Bit j of `bad` is set
// unless |N| + D <= 2^125 (so that D, 1/D and N are normal numbers): the caller then
// recomputes that face in FP64 (weno_k), so FP32 weights are used only where they are
// representable.  Below that limit the FP32 correction's error stays relative to the
// window's variation (no catastrophic cancellation: the indicators are sums of squares
// and the weights a convex combination), so no other check is needed.
template <bool REV, int K, typename Load>
__device__ __forceinline__ void weno_k_mixed(Load ld, double (&base)[K], float (&corr)[K], unsigned &bad)
{
    // explicit FMAs only, as in weno_k: the same face must round the same at every position j
#pragma clang fp contract(off)
    bad = 0;
    constexpr int NP = K + 7;
    constexpr float fD0 = (float)kD0, fD1 = (float)kD1, fD2 = (float)kD2, fD3 = (float)kD3, fD4 = (float)kD4;
    // second and third differences carry sqrt(39): the 780-term of the indicators (39 in units
    // of the 720-term) becomes a plain square, and the other terms take 1/sqrt(39) in their
    // constant coefficients
    constexpr float fR = 6.2449979983983983f;          // sqrt(39)
    constexpr float fIR = 0.16012815380508713f;        // 1/sqrt(39)
    constexpr float fC = 39.05f / 39.0f, fEps = 3.6e-9f;
    double s[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) s[i] = ld(i);
    float D1[NP - 1], D2[NP - 2], D3[NP - 3], G1[NP - 1], C3[NP - 3], B2[NP - 2];
#pragma unroll
    for (int i = 0; i < NP - 1; ++i) { D1[i] = (float)(s[i + 1] - s[i]); G1[i] = 6.0f * D1[i]; }
#pragma unroll
    for (int i = 0; i < NP - 2; ++i) { D2[i] = (D1[i + 1] - D1[i]) * fR; B2[i] = D2[i] * D2[i]; }
#pragma unroll
    for (int i = 0; i < NP - 3; ++i) { D3[i] = D2[i + 1] - D2[i]; C3[i] = fmaf(fC * D3[i], D3[i], fEps); }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        auto i1 = [&](int k) { return REV ? j + 6 - k : j + k; };
        auto i2 = [&](int k) { return REV ? j + 5 - k : j + k; };
        auto i3 = [&](int k) { return REV ? j + 4 - k : j + k; };
        const float sb = REV ? -fIR : fIR;
        const float a0 = fmaf(11.0f * fIR, D3[i3(0)], fmaf(sb * 15.0f, D2[i2(0)], G1[i1(0)]));
        const float a1 = fmaf(2.0f * fIR, D3[i3(1)], fmaf(sb * 9.0f, D2[i2(1)], G1[i1(1)]));
        const float a2 = fmaf(-1.0f * fIR, D3[i3(2)], fmaf(sb * 3.0f, D2[i2(2)], G1[i1(2)]));
        const float a3 = fmaf(2.0f * fIR, D3[i3(3)], fmaf(sb * -3.0f, D2[i2(3)], G1[i1(3)]));
        const float a4 = fmaf(11.0f * fIR, D3[i3(4)], fmaf(sb * -9.0f, D2[i2(4)], G1[i1(4)]));
        const float b0 = REV ? fmaf(-2.0f, D3[j + 4], D2[j + 5]) : fmaf(2.0f, D3[j], D2[j]);
        const float b3 = REV ? D2[j + 2] + D3[j + 1] : D2[j + 3] - D3[j + 3];
        const float b4 = REV ? fmaf(2.0f, D3[j], D2[j + 1]) : fmaf(-2.0f, D3[j + 4], D2[j + 4]);
        const float bb1 = REV ? B2[j + 3] : B2[j + 2];
        const float bb2 = REV ? B2[j + 2] : B2[j + 3];
        // e_k = eps + S_k (eps rides in C3)
        const float e0 = fmaf(a0, a0, fmaf(b0, b0, C3[i3(0)]));
        const float e1 = fmaf(a1, a1, bb1 + C3[i3(1)]);
        const float e2 = fmaf(a2, a2, bb2 + C3[i3(2)]);
        const float e3 = fmaf(a3, a3, fmaf(b3, b3, C3[i3(3)]));
        float e4 = fmaf(a4, a4, fmaf(b4, b4, C3[i3(4)]));
        e4 = fmaxf(fmaxf(fmaxf(e0, e1), fmaxf(e2, e3)), e4);
        // candidates relative to the base point w3 in Newton form over the forward
        // differences at each candidate's first point (D2 and D3 carry sqrt(39)); linear
        // weights folded in
        const float sd = REV ? -1.0f : 1.0f;
        auto d1 = [&](int k) { return D1[REV ? j + 6 - k : j + k]; };
        auto d2 = [&](int k) { return D2[REV ? j + 5 - k : j + k]; };
        auto d3 = [&](int k) { return D3[REV ? j + 4 - k : j + k]; };
        const float c0 = fmaf(sd * fD0 * (13.0f / 12.0f) * fIR, d3(0), fmaf(fD0 * (4.0f / 3.0f) * fIR, d2(0), (sd * fD0 * 0.5f) * d1(0)));
        const float c1 = fmaf(sd * fD1 * (1.0f / 4.0f) * fIR, d3(1), fmaf(fD1 * (5.0f / 6.0f) * fIR, d2(1), (sd * fD1 * 0.5f) * d1(1)));
        const float c2 = fmaf(sd * fD2 * (-1.0f / 12.0f) * fIR, d3(2), fmaf(fD2 * (1.0f / 3.0f) * fIR, d2(2), (sd * fD2 * 0.5f) * d1(2)));
        const float c3 = fmaf(sd * fD3 * (1.0f / 12.0f) * fIR, d3(3), fmaf(fD3 * (-1.0f / 6.0f) * fIR, d2(3), (sd * fD3 * 0.5f) * d1(3)));
        const float c4 = fmaf(sd * fD4 * (-1.0f / 4.0f) * fIR, d3(4), fmaf(fD4 * (1.0f / 3.0f) * fIR, d2(4),
                              fmaf(sd * fD4 * -0.5f, d1(4), (sd * fD4) * d1(3))));
        const float p01 = e0 * e1;
        float N = fmaf(c0, e1, c1 * e0);
        float D = fmaf(fD0, e1, fD1 * e0);
        N = fmaf(N, e2, c2 * p01);
        D = fmaf(D, e2, fD2 * p01);
        const float p012 = p01 * e2;
        N = fmaf(N, e3, c3 * p012);
        D = fmaf(D, e3, fD3 * p012);
        const float p0123 = p012 * e3;
        N = fmaf(N, e4, c4 * p0123);
        D = fmaf(D, e4, fD4 * p0123);
        corr[j] = N * __builtin_amdgcn_rcpf(D);
        base[j] = REV ? s[j + 4] : s[j + 3];
        // D >= eps^4 > 2^-126 always, so D and 1/D are normal and N finite when |N| + D <=
        // 2^125 (false for NaN; a finite |N| above 2^125 is sent to FP64 as well)
        const bool ok = fabsf(N) + D <= 4.25352958e+37f;
        bad |= ok ? 0u : (1u << j);
    }
}

// ---- one-sided closures at non-periodic domain faces (the reference's D0_bound) ------
// Reduced WENO5 reconstructions on 5-point windows s0..s4 without linear weights
// (src/boundary/OCFD_bound_Scheme.cpp:540-719, src/kernels/OCFD_Schemes.cpp:1110-1201), in
// the reference's operation order with contraction off.  The full P variant builds its
// first two weights from S1 and S2 before they are assigned (both still 0); that is the
// reference's behaviour and is kept.
__device__ __forceinline__ double bs_sign1(double x) { return x >= 0 ? 1.0 : -1.0; }

__device__ __forceinline__ double bs_minmod2(double a, double b)
{
#pragma clang fp contract(off)
    return 0.5 * (bs_sign1(a) + bs_sign1(b)) * fmin(fabs(a), fabs(b));
}

__device__ __forceinline__ double w5_p(double s0, double s1, double s2, double s3, double s4)
{
#pragma clang fp contract(off)
    const double ep = 1e-6;
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, tmp;
    tmp = s2 - 2.0 * s3 + s4; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s2 - 4.0 * s3 + s4; S0 += 3 * tmp * tmp;
    const double a0 = ((12.0 * ep + S1) * (12.0 * ep + S1)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    const double q03 = (2.0 * s2 + 5.0 * s3 - s4);
    tmp = s1 - 2.0 * s2 + s3; S1 += 13 * tmp * tmp;
    tmp = s1 - s3; S1 += 3 * tmp * tmp;
    const double a1 = ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    const double q13 = (-s1 + 5.0 * s2 + 2.0 * s3);
    tmp = s0 - 2.0 * s1 + s2; S2 += 13 * tmp * tmp;
    tmp = s0 - 4.0 * s1 + 3.0 * s2; S2 += 3 * tmp * tmp;
    const double a2 = ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S1) * (12.0 * ep + S1));
    const double q23 = (2.0 * s0 - 7.0 * s1 + 11.0 * s2);
    return (a0 * q03 + a1 * q13 + a2 * q23) / (6.0 * (a0 + a1 + a2));
}

__device__ __forceinline__ double w5_m(double s0, double s1, double s2, double s3, double s4)
{
#pragma clang fp contract(off)
    const double ep = 1e-6;
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, tmp;
    tmp = s2 - 2.0 * s1 + s0; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s2 - 4.0 * s1 + s0; S0 += 3 * tmp * tmp;
    const double a0 = 1.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    const double q03 = (2.0 * s2 + 5.0 * s1 - s0);
    tmp = s3 - 2.0 * s2 + s1; S1 += 13 * tmp * tmp;
    tmp = s3 - s1; S1 += 3 * tmp * tmp;
    const double a1 = 1.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    const double q13 = (-s3 + 5.0 * s2 + 2.0 * s1);
    tmp = s4 - 2.0 * s3 + s2; S2 += 13 * tmp * tmp;
    tmp = s4 - 4.0 * s3 + 3.0 * s2; S2 += 3 * tmp * tmp;
    const double a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    const double q23 = (2.0 * s4 - 7.0 * s3 + 11.0 * s2);
    return (a0 * q03 + a1 * q13 + a2 * q23) / (6.0 * (a0 + a1 + a2));
}

// single candidate k (weight a_k = 1/(12 ep + S_k)^2 cancels); q is the candidate, S its smoothness
__device__ __forceinline__ double w5_one(double S, double q)
{
#pragma clang fp contract(off)
    const double ep = 1e-6;
    const double a = 1.0 / ((12.0 * ep + S) * (12.0 * ep + S));
    return a * q / (6.0 * a);
}

__device__ __forceinline__ double w5_two(double Sa, double qa, double Sb, double qb)
{
#pragma clang fp contract(off)
    const double ep = 1e-6;
    const double a = 1.0 / ((12.0 * ep + Sa) * (12.0 * ep + Sa));
    const double b = 1.0 / ((12.0 * ep + Sb) * (12.0 * ep + Sb));
    return (a * qa + b * qb) / (6.0 * (a + b));
}

// smoothness of the three WENO5 candidates, in the reference's expression order
__device__ __forceinline__ double w5_S0p(double s2, double s3, double s4)
{
#pragma clang fp contract(off)
    double S = 0.0, tmp;
    tmp = s2 - 2.0 * s3 + s4; S += 13 * tmp * tmp;
    tmp = 3.0 * s2 - 4.0 * s3 + s4; S += 3 * tmp * tmp;
    return S;
}
__device__ __forceinline__ double w5_S1p(double s1, double s2, double s3)
{
#pragma clang fp contract(off)
    double S = 0.0, tmp;
    tmp = s1 - 2.0 * s2 + s3; S += 13 * tmp * tmp;
    tmp = s1 - s3; S += 3 * tmp * tmp;
    return S;
}
__device__ __forceinline__ double w5_S2p(double s0, double s1, double s2)
{
#pragma clang fp contract(off)
    double S = 0.0, tmp;
    tmp = s0 - 2.0 * s1 + s2; S += 13 * tmp * tmp;
    tmp = s0 - 4.0 * s1 + 3.0 * s2; S += 3 * tmp * tmp;
    return S;
}

// P face f (between cells f-1 and f) of a line of n cells; cell(c) gives the split value
// of cell c.  Returns true and the face value when a closure covers the face
// (OCFD_bound_scheme_kernel_p, OCFD_bound_Scheme.cpp:722-807; 5-point windows on cells
// f-3 .. f+1).  hi: the high face's form, 0 none, 1 D0_bound, 2 the fused path's rim kernel
// (split_Jac3d_Stager_Warming_ker_boundary, src/solver/OCFD_NS_Jacobian3d.cpp:1063-1085),
// which extrapolates the last face linearly, 1.5 f(n-1) - 0.5 f(n-2), where D0_bound
// limits the slope with minmod (the same value in exact arithmetic, other roundings).
template <typename Cell>
__device__ __forceinline__ bool closure_p(Cell cell, int n, int f, bool lo, int hi, double &out)
{
#pragma clang fp contract(off)
    if (lo && f <= 3) {
        if (f == 0) {
            out = cell(0);
        } else if (f == 1) {            // lift: candidate 0 on cells f-1 .. f+1
            const double s2 = cell(f - 1), s3 = cell(f), s4 = cell(f + 1);
            out = w5_one(w5_S0p(s2, s3, s4), (2.0 * s2 + 5.0 * s3 - s4));
        } else if (f == 2) {            // lift_plus: candidates 0 and 1
            const double s1 = cell(f - 2), s2 = cell(f - 1), s3 = cell(f), s4 = cell(f + 1);
            out = w5_two(w5_S0p(s2, s3, s4), (2.0 * s2 + 5.0 * s3 - s4), w5_S1p(s1, s2, s3),
                         (-s1 + 5.0 * s2 + 2.0 * s3));
        } else {
            out = w5_p(cell(f - 3), cell(f - 2), cell(f - 1), cell(f), cell(f + 1));
        }
        return true;
    }
    if (hi && f >= n - 2) {
        if (f == n) {
            const double a = cell(n - 1), b = cell(n - 2);
            out = hi == 2 ? 1.5 * a - 0.5 * b : a + 0.5 * bs_minmod2(a - b, a - b);
        } else if (f == n - 1) {        // right_plus: candidates 1 and 2
            const double s0 = cell(f - 3), s1 = cell(f - 2), s2 = cell(f - 1), s3 = cell(f);
            out = w5_two(w5_S1p(s1, s2, s3), (-s1 + 5.0 * s2 + 2.0 * s3), w5_S2p(s0, s1, s2),
                         (2.0 * s0 - 7.0 * s1 + 11.0 * s2));
        } else {
            out = w5_p(cell(f - 3), cell(f - 2), cell(f - 1), cell(f), cell(f + 1));
        }
        return true;
    }
    return false;
}

// M face f (OCFD_bound_scheme_kernel_m, OCFD_bound_Scheme.cpp:808-893; windows on cells
// f-2 .. f+2, mirrored candidates).  With the rim kernel's form (hi == 2,
// OCFD_NS_Jacobian3d.cpp:1075-1085) the last two faces both take 0.5 f(n-1) + 0.5 f(n-2),
// so the last cell's M difference vanishes.
template <typename Cell>
__device__ __forceinline__ bool closure_m(Cell cell, int n, int f, bool lo, int hi, double &out)
{
#pragma clang fp contract(off)
    if (lo && f <= 3) {
        if (f == 0) {                   // lift: candidate 2 on cells f .. f+2
            const double s2 = cell(f), s3 = cell(f + 1), s4 = cell(f + 2);
            double S = 0.0, tmp;
            tmp = s4 - 2.0 * s3 + s2; S += 13 * tmp * tmp;
            tmp = s4 - 4.0 * s3 + 3.0 * s2; S += 3 * tmp * tmp;
            out = w5_one(S, (2.0 * s4 - 7.0 * s3 + 11.0 * s2));
        } else if (f == 1) {            // lift_plus: candidates 1 and 2
            const double s1 = cell(f - 1), s2 = cell(f), s3 = cell(f + 1), s4 = cell(f + 2);
            double S1 = 0.0, S2 = 0.0, tmp;
            tmp = s3 - 2.0 * s2 + s1; S1 += 13 * tmp * tmp;
            tmp = s3 - s1; S1 += 3 * tmp * tmp;
            tmp = s4 - 2.0 * s3 + s2; S2 += 13 * tmp * tmp;
            tmp = s4 - 4.0 * s3 + 3.0 * s2; S2 += 3 * tmp * tmp;
            out = w5_two(S1, (-s3 + 5.0 * s2 + 2.0 * s1), S2, (2.0 * s4 - 7.0 * s3 + 11.0 * s2));
        } else {
            out = w5_m(cell(f - 2), cell(f - 1), cell(f), cell(f + 1), cell(f + 2));
        }
        return true;
    }
    if (hi && f >= n - 3) {
        if (f == n) {
            out = hi == 2 ? 0.5 * cell(n - 1) + 0.5 * cell(n - 2) : cell(n - 1);
        } else if (f == n - 1) {
            const double a = cell(n - 1), b = cell(n - 2);
            out = hi == 2 ? 0.5 * a + 0.5 * b : a - 0.5 * bs_minmod2(a - b, a - b);
        } else if (f == n - 2) {        // right_plus: candidates 0 and 1
            const double s0 = cell(f - 2), s1 = cell(f - 1), s2 = cell(f), s3 = cell(f + 1);
            double S0 = 0.0, S1 = 0.0, tmp;
            tmp = s2 - 2.0 * s1 + s0; S0 += 13 * tmp * tmp;
            tmp = 3.0 * s2 - 4.0 * s1 + s0; S0 += 3 * tmp * tmp;
            tmp = s3 - 2.0 * s2 + s1; S1 += 13 * tmp * tmp;
            tmp = s3 - s1; S1 += 3 * tmp * tmp;
            out = w5_two(S0, (2.0 * s2 + 5.0 * s1 - s0), S1, (-s3 + 5.0 * s2 + 2.0 * s1));
        } else {
            out = w5_m(cell(f - 2), cell(f - 1), cell(f), cell(f + 1), cell(f + 2));
        }
        return true;
    }
    return false;
}

// Steger-Warming split in the form the tiled kernels use: A1..A3 the metric terms,
// sn = |A|, k1..k3 = A/|A| supplied by the caller (constants when metrics are uniform).
__device__ __forceinline__ void split_fast(const KParams &k, double d, double u, double v, double w,
                                           double cc, double A1, double A2, double A3, double sn,
                                           double k1, double k2, double k3, double fp[5], double fm[5])
{
    const double vs = fma(A3, w, fma(A2, v, A1 * u));
    const double cs = cc * sn;
    const double E1 = vs, E2 = vs - cs, E3 = vs + cs;
    // eps2 > 0 keeps the radicands positive and normal
    const double E1P = 0.5 * (E1 + sqrt_nr(fma(E1, E1, k.eps2)));
    const double E2P = 0.5 * (E2 + sqrt_nr(fma(E2, E2, k.eps2)));
    const double E3P = 0.5 * (E3 + sqrt_nr(fma(E3, E3, k.eps2)));
    const double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    const double t0 = d * k.half_rcp_gamma;
    const double ck1 = cc * k1, ck2 = cc * k2, ck3 = cc * k3;
    const double uc1 = u - ck1, uc2 = u + ck1;
    const double vc1 = v - ck2, vc2 = v + ck2;
    const double wc1 = w - ck3, wc2 = w + ck3;
    const double h1 = 0.5 * fma(wc1, wc1, fma(vc1, vc1, uc1 * uc1));
    const double h2 = 0.5 * fma(wc2, wc2, fma(vc2, vc2, uc2 * uc2));
    const double vv = k.gm1 * fma(w, w, fma(v, v, u * u));
    const double W2 = k.c3 * cc * cc;
    // positive part
    {
        const double a = t0 * (k.c1 * E1P), b = t0 * E2P, c = t0 * E3P;
        fp[0] = a + b + c;
        fp[1] = fma(c, uc2, fma(b, uc1, a * u));
        fp[2] = fma(c, vc2, fma(b, vc1, a * v));
        fp[3] = fma(c, wc2, fma(b, wc1, a * w));
        fp[4] = t0 * fma(W2, E2P + E3P, fma(E3P, h2, fma(E2P, h1, E1P * vv)));
    }
    {
        const double a = t0 * (k.c1 * E1M), b = t0 * E2M, c = t0 * E3M;
        fm[0] = a + b + c;
        fm[1] = fma(c, uc2, fma(b, uc1, a * u));
        fm[2] = fma(c, vc2, fma(b, vc1, a * v));
        fm[3] = fma(c, wc2, fma(b, wc1, a * w));
        fm[4] = t0 * fma(W2, E2M + E3M, fma(E3M, h2, fma(E2M, h1, E1M * vv)));
    }
}

}  // namespace ws
