/*
 * ws_oracle_cell.h -- cell-level arithmetic of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This header is part of the CPU oracle that the
 * test-suite, smoke() and the CPU-baseline leg of bench.py use as a checker.
 * Nothing in the shipped library includes it.
 *
 * The functions restate, operation for operation and in the same evaluation
 * order, the reference's scalar formulas.  They are written over the type
 * `orc_real` so that the same text compiles twice:
 *   - as C with `orc_real == double` (ws_oracle.c, the oracle proper);
 *   - as C++ with `orc_real` an operation-counting class (opcount.cpp), which
 *     pins the algorithmic FLOP count per cell.
 * Only +, -, *, /, unary minus, sqrt() and fmax() appear, so both builds see
 * exactly the same arithmetic.
 */
#ifndef WS_ORACLE_CELL_H
#define WS_ORACLE_CELL_H

#ifndef ORC_REAL_DEFINED
typedef double orc_real;
#define ORC_SQRT(x) sqrt(x)
#define ORC_FMAX(a, b) fmax((a), (b))
#endif

/* WENO7-SYMBO, positive-flux reconstruction.
 * Restates OCFD_weno7_SYMBO_kernel_P_opt, src/kernels/OCFD_Schemes.cpp:306-368.
 * s[0..7] = f(i-4 .. i+3); the result is the face value at i-1/2. */
static inline orc_real orc_cell_weno7_symbo_p_ex(const orc_real *s, orc_real *parts)
{
    orc_real S0, S1, S2, S3, S4, tmp, tmp1;
    S0 = 0.0; S1 = 0.0; S2 = 0.0; S3 = 0.0; S4 = 0.0;

    /* first-derivative terms, :315-320 */
    tmp =  -2.0*s[0] +  9.0*s[1] - 18.0*s[2] + 11.0*s[3]; S0 += 720.0*tmp*tmp;
    tmp =       s[1] -  6.0*s[2] +  3.0*s[3] +  2.0*s[4]; S1 += 720.0*tmp*tmp;
    tmp =  -2.0*s[2] -  3.0*s[3] +  6.0*s[4] -      s[5]; S2 += 720.0*tmp*tmp;
    tmp = -11.0*s[3] + 18.0*s[4] -  9.0*s[5] +  2.0*s[6]; S3 += 720.0*tmp*tmp;
    tmp = -26.0*s[4] + 57.0*s[5] - 42.0*s[6] + 11.0*s[7]; S4 += 720.0*tmp*tmp;

    /* second-derivative terms, :322-327 */
    tmp = -6.0*s[0] + 24.0*s[1] - 30.0*s[2] + 12.0*s[3]; S0 += 780.0*tmp*tmp;
    tmp =                          6.0*s[2] - 12.0*s[3] +  6.0*s[4]; S1 += 780.0*tmp*tmp;
    tmp =                          6.0*s[3] - 12.0*s[4] +  6.0*s[5]; S2 += 780.0*tmp*tmp;
    tmp = 12.0*s[3] - 30.0*s[4] + 24.0*s[5] -  6.0*s[6]; S3 += 780.0*tmp*tmp;
    tmp = 18.0*s[4] - 48.0*s[5] + 42.0*s[6] - 12.0*s[7]; S4 += 780.0*tmp*tmp;

    /* third-derivative terms, :329-334 */
    tmp = -6.0*s[0] + 18.0*(s[1] - s[2]) + 6.0*s[3]; S0 += 781.0*tmp*tmp;
    tmp = -6.0*s[1] + 18.0*(s[2] - s[3]) + 6.0*s[4]; S1 += 781.0*tmp*tmp;
    tmp = -6.0*s[2] + 18.0*(s[3] - s[4]) + 6.0*s[5]; S2 += 781.0*tmp*tmp;
    tmp = -6.0*s[3] + 18.0*(s[4] - s[5]) + 6.0*s[6]; S3 += 781.0*tmp*tmp;
    tmp = -6.0*s[4] + 18.0*(s[5] - s[6]) + 6.0*s[7]; S4 += 781.0*tmp*tmp;

    /* the downwind indicator is replaced by the largest of all five, :336-339 */
    tmp  = ORC_FMAX(S0, S1);
    tmp1 = ORC_FMAX(S2, S3);
    tmp  = ORC_FMAX(tmp, tmp1);
    S4   = ORC_FMAX(tmp, S4);
    if (parts) { parts[0] = S0; parts[1] = S1; parts[2] = S2; parts[3] = S3; parts[4] = S4; }

    /* product-form nonlinear weights, :342-352 */
    {
        orc_real tmp2, tmp3;
        tmp  = (0.0401954833730)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S3)*(2.592e-6+S4);
        tmp1 = (0.2493800006710)*(2.592e-6+S0)*(2.592e-6+S2)*(2.592e-6+S3)*(2.592e-6+S4);
        tmp2 = (0.4802686256260)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S3)*(2.592e-6+S4);
        tmp3 = (0.2009775476730)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S4);
        S4   = (0.0291783426580)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S3);
        S3 = tmp3; S2 = tmp2; S1 = tmp1; S0 = tmp;
    }
    if (parts) { parts[5] = S0; parts[6] = S1; parts[7] = S2; parts[8] = S3; parts[9] = S4; }

    /* candidate reconstructions and normalisation, :357-365 */
    tmp1 = 0.0;
    tmp = -3.0*s[0] + 13.0*s[1] - 23.0*s[2] + 25.0*s[3]; tmp1 += S0*tmp; if (parts) parts[10] = tmp;
    tmp =      s[1] -  5.0*s[2] + 13.0*s[3] +  3.0*s[4]; tmp1 += S1*tmp; if (parts) parts[11] = tmp;
    tmp =     -s[2] +  7.0*s[3] +  7.0*s[4] -      s[5]; tmp1 += S2*tmp; if (parts) parts[12] = tmp;
    tmp =  3.0*s[3] + 13.0*s[4] -  5.0*s[5] +      s[6]; tmp1 += S3*tmp; if (parts) parts[13] = tmp;
    tmp = 25.0*s[4] - 23.0*s[5] + 13.0*s[6] -  3.0*s[7]; tmp1 += S4*tmp; if (parts) parts[14] = tmp;

    tmp1 /= (12.0*(S0+S1+S2+S3+S4));
    return tmp1;
}

/* parts (optional, 15 entries): [0..4] smoothness indicators after the
 * downwind max, [5..9] unnormalised weights, [10..14] candidate values. */
static inline orc_real orc_cell_weno7_symbo_p(const orc_real *s)
{
    return orc_cell_weno7_symbo_p_ex(s, 0);
}

/* WENO7-SYMBO, negative-flux reconstruction.
 * Restates OCFD_weno7_SYMBO_kernel_M_opt, src/kernels/OCFD_Schemes.cpp:491-553,
 * typed separately (mirrored indices) so that M(s) == P(reverse(s)) is a test,
 * not a definition.  s[0..7] = f(i-4 .. i+3); result at face i-1/2. */
static inline orc_real orc_cell_weno7_symbo_m(const orc_real *s)
{
    orc_real S0, S1, S2, S3, S4, tmp, tmp1;
    S0 = 0.0; S1 = 0.0; S2 = 0.0; S3 = 0.0; S4 = 0.0;

    /* :500-505 */
    tmp =  -2.0*s[7] +  9.0*s[6] - 18.0*s[5] + 11.0*s[4]; S0 += 720.0*tmp*tmp;
    tmp =       s[6] -  6.0*s[5] +  3.0*s[4] +  2.0*s[3]; S1 += 720.0*tmp*tmp;
    tmp =  -2.0*s[5] -  3.0*s[4] +  6.0*s[3] -      s[2]; S2 += 720.0*tmp*tmp;
    tmp = -11.0*s[4] + 18.0*s[3] -  9.0*s[2] +  2.0*s[1]; S3 += 720.0*tmp*tmp;
    tmp = -26.0*s[3] + 57.0*s[2] - 42.0*s[1] + 11.0*s[0]; S4 += 720.0*tmp*tmp;

    /* :507-512 */
    tmp = -6.0*s[7] + 24.0*s[6] - 30.0*s[5] + 12.0*s[4]; S0 += 780.0*tmp*tmp;
    tmp =                          6.0*s[5] - 12.0*s[4] +  6.0*s[3]; S1 += 780.0*tmp*tmp;
    tmp =                          6.0*s[4] - 12.0*s[3] +  6.0*s[2]; S2 += 780.0*tmp*tmp;
    tmp = 12.0*s[4] - 30.0*s[3] + 24.0*s[2] -  6.0*s[1]; S3 += 780.0*tmp*tmp;
    tmp = 18.0*s[3] - 48.0*s[2] + 42.0*s[1] - 12.0*s[0]; S4 += 780.0*tmp*tmp;

    /* :514-519 */
    tmp = -6.0*s[7] + 18.0*(s[6] - s[5]) + 6.0*s[4]; S0 += 781.0*tmp*tmp;
    tmp = -6.0*s[6] + 18.0*(s[5] - s[4]) + 6.0*s[3]; S1 += 781.0*tmp*tmp;
    tmp = -6.0*s[5] + 18.0*(s[4] - s[3]) + 6.0*s[2]; S2 += 781.0*tmp*tmp;
    tmp = -6.0*s[4] + 18.0*(s[3] - s[2]) + 6.0*s[1]; S3 += 781.0*tmp*tmp;
    tmp = -6.0*s[3] + 18.0*(s[2] - s[1]) + 6.0*s[0]; S4 += 781.0*tmp*tmp;

    /* :521-526 */
    tmp  = ORC_FMAX(S0, S1);
    tmp1 = ORC_FMAX(S2, S3);
    tmp  = ORC_FMAX(tmp, tmp1);
    S4   = ORC_FMAX(tmp, S4);

    /* :528-539 */
    {
        orc_real tmp2, tmp3;
        tmp  = (0.0401954833730)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S3)*(2.592e-6+S4);
        tmp1 = (0.2493800006710)*(2.592e-6+S0)*(2.592e-6+S2)*(2.592e-6+S3)*(2.592e-6+S4);
        tmp2 = (0.4802686256260)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S3)*(2.592e-6+S4);
        tmp3 = (0.2009775476730)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S4);
        S4   = (0.0291783426580)*(2.592e-6+S0)*(2.592e-6+S1)*(2.592e-6+S2)*(2.592e-6+S3);
        S3 = tmp3; S2 = tmp2; S1 = tmp1; S0 = tmp;
    }

    /* :543-550 */
    tmp1 = 0.0;
    tmp = -3.0*s[7] + 13.0*s[6] - 23.0*s[5] + 25.0*s[4]; tmp1 += S0*tmp;
    tmp =      s[6] -  5.0*s[5] + 13.0*s[4] +  3.0*s[3]; tmp1 += S1*tmp;
    tmp =     -s[5] +  7.0*s[4] +  7.0*s[3] -      s[2]; tmp1 += S2*tmp;
    tmp =  3.0*s[4] + 13.0*s[3] -  5.0*s[2] +      s[1]; tmp1 += S3*tmp;
    tmp = 25.0*s[3] - 23.0*s[2] + 13.0*s[1] -  3.0*s[0]; tmp1 += S4*tmp;

    tmp1 /= (12.0*(S0+S1+S2+S3+S4));
    return tmp1;
}

/* Steger-Warming flux-vector splitting of one point in one direction.
 * Restates split_Jac3d_Stager_Warming_ker_origin,
 * src/solver/OCFD_NS_Jacobian3d.cpp:69-127 (the canonical form, dividing by
 * |A|).  A1..A3 are the metric terms of the direction (Akx,Aky,Akz for x,
 * Aix.. for y, Asx.. for z).  C1 = 2(gamma-1), C3 = (3-gamma)/(2(gamma-1))
 * as in src/runtime/OCFD_init.cpp:385-386. */
static inline void orc_cell_split_sw(orc_real d, orc_real u, orc_real v, orc_real w, orc_real cc,
                                     orc_real A1, orc_real A2, orc_real A3,
                                     double gamma, double epsl_sw, double C1, double C3,
                                     orc_real fp[5], orc_real fm[5])
{
    orc_real ss, Ak1, Ak2, Ak3, vs, E1, E2, E3, E1P, E2P, E3P, E1M, E2M, E3M;
    orc_real tmp0, uc1, uc2, vc1, vc2, wc1, wc2, vvc1, vvc2, vv, W2;

    ss  = ORC_SQRT(A1*A1 + A2*A2 + A3*A3);
    Ak1 = A1 / ss;
    Ak2 = A2 / ss;
    Ak3 = A3 / ss;

    vs = A1*u + A2*v + A3*w;

    E1 = vs;
    E2 = vs - cc*ss;
    E3 = vs + cc*ss;

    E1P = (E1 + ORC_SQRT(E1*E1 + epsl_sw*epsl_sw)) * 0.50;
    E2P = (E2 + ORC_SQRT(E2*E2 + epsl_sw*epsl_sw)) * 0.50;
    E3P = (E3 + ORC_SQRT(E3*E3 + epsl_sw*epsl_sw)) * 0.50;

    E1M = E1 - E1P;
    E2M = E2 - E2P;
    E3M = E3 - E3P;

    tmp0 = d / (2.0*gamma);
    uc1 = u - cc*Ak1;
    uc2 = u + cc*Ak1;
    vc1 = v - cc*Ak2;
    vc2 = v + cc*Ak2;
    wc1 = w - cc*Ak3;
    wc2 = w + cc*Ak3;
    vvc1 = (uc1*uc1 + vc1*vc1 + wc1*wc1) * 0.50;
    vvc2 = (uc2*uc2 + vc2*vc2 + wc2*wc2) * 0.50;
    vv = (gamma - 1.0) * (u*u + v*v + w*w);
    W2 = C3*cc*cc;

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

/* One flux-difference update, as put_du_p_kernel / put_du_m_kernel do it:
 * du += -Ajac*(h_r - h_l)/h   (src/kernels/OCFD_Schemes.cpp:233,273). */
static inline orc_real orc_cell_flux_diff(orc_real du, orc_real ajac, orc_real h_r, orc_real h_l, double h)
{
    return du + -ajac*(h_r - h_l)/h;
}

/* Conserved -> primitive, src/kernels/commen_kernel.cpp:84-109.
 * out: d, u, v, w, T, P. */
static inline void orc_cell_cons_to_pri(orc_real f0, orc_real f1, orc_real f2, orc_real f3, orc_real f4,
                                        double Cv, double gamma, orc_real out[6])
{
    orc_real d1, u1, v1, w1, T1;
    d1 = f0;
    u1 = f1 / d1;
    v1 = f2 / d1;
    w1 = f3 / d1;
    T1 = f4;
    T1 = T1 - 0.5*d1*(u1*u1 + v1*v1 + w1*w1);
    out[0] = d1; out[1] = u1; out[2] = v1; out[3] = w1;
    out[4] = T1 / (d1*Cv);
    T1 = T1 / d1;
    out[5] = T1 * (gamma - 1.0);
}

/* Primitive -> conserved, src/kernels/commen_kernel.cpp:32-66. */
static inline void orc_cell_pri_to_cons(orc_real d, orc_real u, orc_real v, orc_real w, orc_real T,
                                        double Cv, orc_real f[5])
{
    f[0] = d;
    f[1] = u*d;
    f[2] = v*d;
    f[3] = w*d;
    f[4] = d*((u*u + v*v + w*w)*0.5 + Cv*T);
}

/* Sound speed, src/kernels/commen_kernel.cpp:145-154. */
static inline orc_real orc_cell_sound_speed(orc_real T, double ama)
{
    return ORC_SQRT(T) / ama;
}

/* ---- viscous terms ---------------------------------------------------------------------- */

/* Central first derivative of one stencil, not yet divided by h: CD6 (s[0..6] centred on s[3],
 * src/kernels/OCFD_Schemes.cpp:99-105) and CD8 (s[0..8] centred on s[4], :130-138). */
static inline orc_real orc_cell_cd6(const orc_real *s)
{
    return (45.0*(s[4] - s[2]) - 9.0*(s[5] - s[1]) + (s[6] - s[0])) / 60.0;
}

static inline orc_real orc_cell_cd8(const orc_real *s)
{
    return (672.0*(s[5] - s[3]) - 168.0*(s[6] - s[2]) + 32.0*(s[7] - s[1]) - 3.0*(s[8] - s[0])) / 840.0;
}

/* Sutherland viscosity of one cell (get_Amu_kernal, src/kernels/commen_kernel.cpp:123-141);
 * c0 = (1 + 110.4/Ref_T)/Re and tsb = 110.4/Ref_T are run constants. */
static inline orc_real orc_cell_amu(orc_real t, double c0, double tsb)
{
    return c0 * ORC_SQRT(t*t*t) / (tsb + t);
}

/* The four viscous fluxes of direction dir at one cell (vis_flux_ker,
 * src/solver/OCFD_NS_Jacobian3d.cpp:477-620): g[4*d + var] the cell's gradients along
 * computational direction d (var = u, v, w, T), A[0..8] its metric terms Akx, Aky, Akz, Aix,
 * Aiy, Aiz, Asx, Asy, Asz, aj1 its Ajac, a1 its viscosity, vis_c = Cp/Pr.  The stress and heat
 * flux are formed again for every direction, as the reference does. */
static inline void orc_cell_vis_flux(const orc_real g[12], orc_real u, orc_real v, orc_real w, orc_real a1,
                                     double vis_c, const orc_real A[9], orc_real aj1, int dir, orc_real ev[4])
{
    orc_real ux, vx, wx, uy, vy, wy, uz, vz, wz, div;
    orc_real s11, s12, s13, s22, s23, s33, amuk, Tx, Ty, Tz, E1, E2, E3, akx, aky, akz;
    ux = g[0] * A[0] + g[4] * A[3] + g[8] * A[6];
    vx = g[1] * A[0] + g[5] * A[3] + g[9] * A[6];
    wx = g[2] * A[0] + g[6] * A[3] + g[10] * A[6];
    uy = g[0] * A[1] + g[4] * A[4] + g[8] * A[7];
    vy = g[1] * A[1] + g[5] * A[4] + g[9] * A[7];
    wy = g[2] * A[1] + g[6] * A[4] + g[10] * A[7];
    uz = g[0] * A[2] + g[4] * A[5] + g[8] * A[8];
    vz = g[1] * A[2] + g[5] * A[5] + g[9] * A[8];
    wz = g[2] * A[2] + g[6] * A[5] + g[10] * A[8];
    div = ux + vy + wz;
    s11 = (2.0 * ux - 2.0 / 3.0 * div) * a1;
    s22 = (2.0 * vy - 2.0 / 3.0 * div) * a1;
    s33 = (2.0 * wz - 2.0 / 3.0 * div) * a1;
    s12 = (uy + vx) * a1;
    s13 = (uz + wx) * a1;
    s23 = (vz + wy) * a1;
    amuk = a1 * vis_c;
    Tx = g[3] * A[0] + g[7] * A[3] + g[11] * A[6];
    Ty = g[3] * A[1] + g[7] * A[4] + g[11] * A[7];
    Tz = g[3] * A[2] + g[7] * A[5] + g[11] * A[8];
    E1 = u * s11 + v * s12 + w * s13 + amuk * Tx;
    E2 = u * s12 + v * s22 + w * s23 + amuk * Ty;
    E3 = u * s13 + v * s23 + w * s33 + amuk * Tz;
    akx = A[3 * dir] * aj1;
    aky = A[3 * dir + 1] * aj1;
    akz = A[3 * dir + 2] * aj1;
    ev[0] = akx * s11 + aky * s12 + akz * s13;
    ev[1] = akx * s12 + aky * s22 + akz * s23;
    ev[2] = akx * s13 + aky * s23 + akz * s33;
    ev[3] = akx * E1 + aky * E2 + akz * E3;
}

#endif
