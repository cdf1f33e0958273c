"""Second CPU restatement of the inviscid RHS, vectorised with numpy.

TEST INFRASTRUCTURE ONLY: used by tests/ as a checker, never by the library.

It is deliberately not a transcription of the reference's expressions.  The
WENO7-SYMBO reconstruction (src/kernels/OCFD_Schemes.cpp:306-368) is rewritten
in divided-difference form: every smoothness term is expressed through the
forward differences F1, F2, F3 anchored at the first point of its candidate
stencil, and the product-form weights through prefix/suffix products.  The
flux difference is taken of the summed face flux hP + hM.  Agreement with the
C oracle (which follows the reference expression by expression) therefore
checks both the transcription and the rewrite that the GPU kernels use.

Array layout: ghosted fields are numpy arrays of shape (nz+8, ny+8, nx+8)
(x fastest, LAP = 4); the RHS is (5, nz, ny, nx).
"""
import numpy as np

LAP = 4
EPS_W = 2.592e-6
# linear weights of the five candidates (OCFD_Schemes.cpp:344-348)
D_LIN = (0.0401954833730, 0.2493800006710, 0.4802686256260, 0.2009775476730, 0.0291783426580)

# a_k = 6*F1_k + B1[k]*F2_k + B3[k]*F3_k (the 720-weighted first-derivative term of candidate k)
_A_F2 = (15.0, 9.0, 3.0, -3.0, -9.0)
_A_F3 = (11.0, 2.0, -1.0, 2.0, 11.0)


def weno_p(s):
    """Positive-flux face value from an 8-point window s[0..7] = f(i-4..i+3)
    (any broadcastable arrays); the face is i-1/2."""
    F1 = [s[k + 1] - s[k] for k in range(7)]
    F2 = [F1[k + 1] - F1[k] for k in range(6)]
    F3 = [F2[k + 1] - F2[k] for k in range(5)]
    a = [6.0 * F1[k] + _A_F2[k] * F2[k] + _A_F3[k] * F3[k] for k in range(5)]
    b = [F2[0] + 2.0 * F3[0], F2[2], F2[3], F2[3] - F3[3], F2[4] - 2.0 * F3[4]]
    S = [720.0 * a[k] * a[k] + (780.0 * 36.0) * b[k] * b[k] + (781.0 * 36.0) * F3[k] * F3[k]
         for k in range(5)]
    S[4] = np.maximum(np.maximum(np.maximum(S[0], S[1]), np.maximum(S[2], S[3])), S[4])
    e = [EPS_W + S[k] for k in range(5)]
    p01 = e[0] * e[1]
    p012 = p01 * e[2]
    s34 = e[3] * e[4]
    s234 = e[2] * s34
    w = [D_LIN[0] * (e[1] * s234),
         D_LIN[1] * (e[0] * s234),
         D_LIN[2] * (p01 * s34),
         D_LIN[3] * (p012 * e[4]),
         D_LIN[4] * (p012 * e[3])]
    q = [-3.0 * s[0] + 13.0 * s[1] - 23.0 * s[2] + 25.0 * s[3],
         s[1] - 5.0 * s[2] + 13.0 * s[3] + 3.0 * s[4],
         -s[2] + 7.0 * s[3] + 7.0 * s[4] - s[5],
         3.0 * s[3] + 13.0 * s[4] - 5.0 * s[5] + s[6],
         25.0 * s[4] - 23.0 * s[5] + 13.0 * s[6] - 3.0 * s[7]]
    num = w[0] * q[0] + w[1] * q[1] + w[2] * q[2] + w[3] * q[3] + w[4] * q[4]
    den = 12.0 * (w[0] + w[1] + w[2] + w[3] + w[4])
    return num / den


def weno_m(s):
    """Negative-flux face value: the positive-flux rule on the reversed window."""
    return weno_p([s[7 - k] for k in range(8)])


def split_sw(d, u, v, w, cc, A1, A2, A3, gamma, epsl):
    """Steger-Warming split fluxes (f+, f-) as two lists of 5 arrays
    (src/solver/OCFD_NS_Jacobian3d.cpp:69-127), written with reciprocals."""
    C1 = 2.0 * (gamma - 1.0)
    C3 = (3.0 - gamma) / (2.0 * (gamma - 1.0))
    ss = np.sqrt(A1 * A1 + A2 * A2 + A3 * A3)
    r = 1.0 / ss
    k1, k2, k3 = A1 * r, A2 * r, A3 * r
    vs = A1 * u + A2 * v + A3 * w
    E = [vs, vs - cc * ss, vs + cc * ss]
    EP = [0.5 * (x + np.sqrt(x * x + epsl * epsl)) for x in E]
    EM = [x - y for x, y in zip(E, EP)]
    t0 = d * (0.5 / gamma)
    uc1, uc2 = u - cc * k1, u + cc * k1
    vc1, vc2 = v - cc * k2, v + cc * k2
    wc1, wc2 = w - cc * k3, w + cc * k3
    h1 = 0.5 * (uc1 * uc1 + vc1 * vc1 + wc1 * wc1)
    h2 = 0.5 * (uc2 * uc2 + vc2 * vc2 + wc2 * wc2)
    vv = (gamma - 1.0) * (u * u + v * v + w * w)
    W2 = C3 * cc * cc

    def comps(e1, e2, e3):
        return [t0 * (C1 * e1 + e2 + e3),
                t0 * (C1 * e1 * u + e2 * uc1 + e3 * uc2),
                t0 * (C1 * e1 * v + e2 * vc1 + e3 * vc2),
                t0 * (C1 * e1 * w + e2 * wc1 + e3 * wc2),
                t0 * (e1 * vv + e2 * h1 + e3 * h2 + W2 * (e2 + e3))]

    return comps(*EP), comps(*EM)


def _axis_slices(axis, lo, hi):
    sl = [slice(LAP, -LAP)] * 3
    sl[axis] = slice(lo, hi)
    return tuple(sl)


def inviscid_rhs(prims, metrics, ajac, h, gamma, epsl):
    """RHS from ghosted primitives.

    prims   : (d, u, v, w, cc) ghosted arrays
    metrics : 9 ghosted arrays or scalars (Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz)
    ajac    : ghosted array or scalar
    h       : (hx, hy, hz)
    """
    d = prims[0]
    nzg, nyg, nxg = d.shape
    n = (nzg - 2 * LAP, nyg - 2 * LAP, nxg - 2 * LAP)   # (nz, ny, nx)
    du = np.zeros((5,) + n)
    interior = (slice(LAP, -LAP),) * 3

    def take(a, sl):
        return a[sl] if isinstance(a, np.ndarray) and a.ndim == 3 else a

    for dirn in range(3):                 # 0 = x, 1 = y, 2 = z
        axis = 2 - dirn                   # numpy axis of the direction
        ext = _axis_slices(axis, 0, None)  # full ghosted extent along the direction
        P = [take(x, ext) for x in prims]
        A = [take(metrics[3 * dirn + c], ext) for c in range(3)]
        fp, fm = split_sw(*P, *A, gamma, epsl)
        nd = n[axis]
        aj = take(ajac, interior)
        for m in range(5):
            def window(F, j):
                sl = [slice(None)] * 3
                sl[axis] = slice(j, j + nd + 1)
                return F[tuple(sl)]
            hp = weno_p([window(fp[m], j) for j in range(8)])
            hm = weno_m([window(fm[m], j) for j in range(8)])
            face = hp + hm
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[axis] = slice(0, nd)
            hi[axis] = slice(1, nd + 1)
            du[m] += -aj * (face[tuple(hi)] - face[tuple(lo)]) / h[dirn]
    return du


def ghost_fill_periodic(a):
    """Periodic LAP-deep face halo in x, then y, then z (in place)."""
    L = LAP
    a[L:-L, L:-L, -L:] = a[L:-L, L:-L, L:2 * L]
    a[L:-L, L:-L, :L] = a[L:-L, L:-L, -2 * L:-L]
    a[L:-L, -L:, L:-L] = a[L:-L, L:2 * L, L:-L]
    a[L:-L, :L, L:-L] = a[L:-L, -2 * L:-L, L:-L]
    a[-L:, L:-L, L:-L] = a[L:2 * L, L:-L, L:-L]
    a[:L, L:-L, L:-L] = a[-2 * L:-L, L:-L, L:-L]
    return a


def euler_flux(d, u, v, w, p, A1, A2, A3, gamma):
    """Physical Euler flux A.F(U) with total energy E = p/(gamma-1) + d|u|^2/2
    (independent of the splitting; used to check f+ + f- = A.F)."""
    V = A1 * u + A2 * v + A3 * w
    E = p / (gamma - 1.0) + 0.5 * d * (u * u + v * v + w * w)
    return [d * V, d * u * V + A1 * p, d * v * V + A2 * p, d * w * V + A3 * p, (E + p) * V]


# ---- viscous terms -----------------------------------------------------------------------
# Central first derivatives as antisymmetric weight tables (CD6: OCFD_Schemes.cpp:99-105,
# CD8: :130-138): d f/dx_i ~ sum_q c_q (f_{i+q} - f_{i-q}) / h.
_CD = {6: (45.0 / 60.0, -9.0 / 60.0, 1.0 / 60.0),
       8: (672.0 / 840.0, -168.0 / 840.0, 32.0 / 840.0, -3.0 / 840.0)}


def _dper(a, axis, h, order):
    """Periodic central derivative of an interior array (nz, ny, nx) along axis (2 = x)."""
    out = np.zeros_like(a)
    for q, c in enumerate(_CD[order], start=1):
        out += c * (np.roll(a, -q, axis=axis) - np.roll(a, q, axis=axis))
    return out / h


def viscous_rhs(prims_int, metrics_int, ajac_int, h, gamma, ama, re, pr, ref_t, order=6):
    """Viscous terms (du components 1..4) of periodic interior fields: u, v, w, T arrays
    (nz, ny, nx) in prims_int, the nine metric terms and the Jacobian as interior arrays.
    Written with tensor algebra: grad[i][j] = sum_d A_d[j] * d(f_i)/dxi_d, the stress
    tau = mu (grad + grad^T - 2/3 div I), the heat flux k grad T with k = mu Cp/Pr, and for
    each direction d the flux F_d = J (A_d . tau, A_d . (tau u + k grad T)), whose
    periodic central derivative, times J, is added."""
    u, v, w, T = prims_int
    vel = (u, v, w)
    axes = (2, 1, 0)                     # computational x, y, z directions of the arrays
    tsb = 110.4 / ref_t
    mu = (1.0 + 110.4 / ref_t) / re * np.sqrt(T ** 3) / (tsb + T)
    kappa = mu * (1.0 / ((gamma - 1.0) * ama * ama)) / pr
    A = [[metrics_int[3 * d + j] for j in range(3)] for d in range(3)]     # A[d][j]
    dvel = [[_dper(f, axes[d], h[d], order) for d in range(3)] for f in vel]  # d f_i / d xi_d
    dT = [_dper(T, axes[d], h[d], order) for d in range(3)]
    grad = [[sum(A[d][j] * dvel[i][d] for d in range(3)) for j in range(3)] for i in range(3)]
    gT = [sum(A[d][j] * dT[d] for d in range(3)) for j in range(3)]
    div = grad[0][0] + grad[1][1] + grad[2][2]
    tau = [[mu * (grad[i][j] + grad[j][i] - (2.0 / 3.0 * div if i == j else 0.0)) for j in range(3)]
           for i in range(3)]
    heat = [sum(vel[i] * tau[i][j] for i in range(3)) + kappa * gT[j] for j in range(3)]
    du = np.zeros((5,) + u.shape)
    for d in range(3):
        a = [A[d][j] * ajac_int for j in range(3)]
        flux = [sum(a[j] * tau[i][j] for j in range(3)) for i in range(3)] + [sum(a[j] * heat[j] for j in range(3))]
        for m in range(4):
            du[m + 1] += ajac_int * _dper(flux[m], axes[d], h[d], order)
    return du
