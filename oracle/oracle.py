"""ctypes front-end of the C oracle (libws_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py use it as the checker.  The library never imports it.

Arrays: ghosted fields are float64 C-contiguous (nz+8, ny+8, nx+8); the
interior SoA state / RHS is (5, nz, ny, nx).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libws_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("gamma", ctypes.c_double), ("ama", ctypes.c_double), ("epsl_sw", ctypes.c_double),
                ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("hz", ctypes.c_double),
                ("bc_planes", ctypes.c_int), ("nthreads", ctypes.c_int), ("viscous", ctypes.c_int),
                ("re", ctypes.c_double), ("pr", ctypes.c_double), ("ref_t", ctypes.c_double),
                ("periodic", ctypes.c_int * 3), ("d0_bound", ctypes.c_int * 6), ("non_ref", ctypes.c_int * 6)]


def build():
    """Compile the oracle (gcc) in place."""
    subprocess.run(["make", "-s", "-C", _HERE, "libws_oracle.so", "opcount"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_weno7_symbo_p.restype = ctypes.c_double
        L.orc_weno7_symbo_p.argtypes = [_dp]
        L.orc_weno7_symbo_m.restype = ctypes.c_double
        L.orc_weno7_symbo_m.argtypes = [_dp]
        L.orc_weno7_symbo_p_parts.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_split_sw.argtypes = [ctypes.c_double] * 10 + [_dp, _dp]
        L.orc_cfl_dt.restype = ctypes.c_double
        for fn in (L.orc_closure_p, L.orc_closure_m):
            fn.restype = ctypes.c_int
            fn.argtypes = [_dp, ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_cd6.restype = ctypes.c_double
        L.orc_cd6.argtypes = [_dp]
        L.orc_cd8.restype = ctypes.c_double
        L.orc_cd8.argtypes = [_dp]
        L.orc_cd_closure.restype = ctypes.c_int
        L.orc_cd_closure.argtypes = [_dp, ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def params(nx, ny, nz, gamma=1.4, ama=1.0, epsl_sw=1e-3, h=None, bc_planes=1, nthreads=0, viscous=0,
           re=1000.0, pr=0.72, ref_t=288.15, periodic=(1, 1, 1), d0_bound=(0,) * 6, non_ref=(0,) * 6):
    if h is None:
        h = (1.0 / (nx - 1.0), 1.0 / (ny - 1.0), 1.0 / (nz - 1.0))
    return OrcParams(nx, ny, nz, gamma, ama, epsl_sw, h[0], h[1], h[2], bc_planes, nthreads, viscous, re, pr, ref_t,
                     (ctypes.c_int * 3)(*periodic), (ctypes.c_int * 6)(*d0_bound), (ctypes.c_int * 6)(*non_ref))


def weno_p(s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().orc_weno7_symbo_p(_ptr(s))


def weno_m(s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().orc_weno7_symbo_m(_ptr(s))


def weno_p_parts(s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    S, w, q = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().orc_weno7_symbo_p_parts(_ptr(s), _ptr(S), _ptr(w), _ptr(q))
    return S, w, q


def split_sw(d, u, v, w, cc, A1, A2, A3, gamma=1.4, epsl=1e-3):
    fp, fm = np.zeros(5), np.zeros(5)
    lib().orc_split_sw(d, u, v, w, cc, A1, A2, A3, gamma, epsl, _ptr(fp), _ptr(fm))
    return fp, fm


def ghosted_zeros(p):
    return np.zeros((p.nz + 8, p.ny + 8, p.nx + 8))


def metrics_cartesian(p, scale=(1.0, 1.0, 1.0)):
    """Akx = sx, Aiy = sy, Asz = sz, all other metric terms 0 (ghosted arrays)."""
    one = [np.full((p.nz + 8, p.ny + 8, p.nx + 8), s) for s in scale]
    zero = ghosted_zeros(p)
    return [one[0], zero, zero, zero, one[1], zero, zero, zero, one[2]]


def _metric_ptrs(metrics):
    arr = (_dp * 9)(*[_ptr(m) for m in metrics])
    return arr


def inviscid_rhs(p, d, u, v, w, cc, metrics, ajac, du=None):
    """RHS from ghosted primitives (accumulated into du, zero-initialised here if None)."""
    if du is None:
        du = np.zeros((5, p.nz, p.ny, p.nx))
    lib().orc_inviscid_rhs(ctypes.byref(p), _ptr(d), _ptr(u), _ptr(v), _ptr(w), _ptr(cc),
                           _metric_ptrs(metrics), _ptr(ajac), _ptr(du))
    return du


def ghost_fill_periodic(p, fld):
    lib().orc_ghost_fill_periodic(ctypes.byref(p), _ptr(fld))
    return fld


def sound_speed(p, T):
    cc = np.zeros_like(T)
    lib().orc_sound_speed(ctypes.byref(p), _ptr(T), _ptr(cc))
    return cc


def cons_to_pri(p, f):
    out = [ghosted_zeros(p) for _ in range(6)]
    lib().orc_cons_to_pri(ctypes.byref(p), _ptr(f), *[_ptr(o) for o in out])
    return out


def pri_to_cons(p, d, u, v, w, T):
    f = np.zeros((5, p.nz, p.ny, p.nx))
    s = (ctypes.c_int * 3)(0, 0, 0)
    e = (ctypes.c_int * 3)(p.nx, p.ny, p.nz)
    lib().orc_pri_to_cons_box(ctypes.byref(p), _ptr(f), _ptr(d), _ptr(u), _ptr(v), _ptr(w), _ptr(T), s, e)
    return f


def du_comput(p, d, u, v, w, T, metrics=None, ajac=None):
    """Full RHS of a stage from primitives: ghost fill (in place), sound speed, RHS."""
    if metrics is None:
        metrics = metrics_cartesian(p)
    if ajac is None:
        ajac = np.ones_like(d)
    cc = ghosted_zeros(p)
    du = np.zeros((5, p.nz, p.ny, p.nx))
    lib().orc_du_comput(ctypes.byref(p), _ptr(d), _ptr(u), _ptr(v), _ptr(w), _ptr(T), _ptr(cc),
                        _metric_ptrs(metrics), _ptr(ajac), _ptr(du))
    return du, cc


def cd(s):
    """Central first-derivative stencil (CD6 for 7 points, CD8 for 9), before 1/h."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().orc_cd6(_ptr(s)) if len(s) == 7 else lib().orc_cd8(_ptr(s))


def cd_closure(s, c, n, lo=1, hi=1):
    """Closure form of the viscous derivative at cell c of n (s: the 2r+1 values f(c-r..c+r),
    r = 3 or 4), before 1/h; None where the central form applies."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = ctypes.c_double()
    ok = lib().orc_cd_closure(_ptr(s), (len(s) - 1) // 2, c, n, lo, hi, ctypes.byref(out))
    return out.value if ok else None


def closure(side, line, f, lo=1, hi=1):
    """Closure face value of face f on a line given as an array of n+8 values (4 ghosts
    each side), or None if no closure covers the face.  side: 'p' or 'm'."""
    line = np.ascontiguousarray(line, dtype=np.float64)
    n = len(line) - 8
    out = ctypes.c_double()
    fn = lib().orc_closure_p if side == "p" else lib().orc_closure_m
    ok = fn(ctypes.cast(line.ctypes.data + 4 * 8, _dp), n, f, lo, hi, ctypes.byref(out))
    return out.value if ok else None


def amu(p, T, re, ref_t):
    out = np.zeros((p.nz, p.ny, p.nx))
    lib().orc_amu(ctypes.byref(p), ctypes.c_double(re), ctypes.c_double(ref_t), _ptr(T), _ptr(out))
    return out


def comput_jacobian(p, x, y, z, jacbound=(0, 0, 0)):
    """Metric terms from ghosted mesh coordinates (their face ghosts as the caller's
    exchange left them): [Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz, Ajac], ghosted,
    each initialised to 1.0 where the computation leaves a value untouched (edges and
    corners), as the reference's Init_Jacobian3d does."""
    out = [np.ones_like(x) for _ in range(10)]
    xyz = (_dp * 3)(_ptr(x), _ptr(y), _ptr(z))
    outp = (_dp * 10)(*[_ptr(a) for a in out])
    lib().orc_comput_jacobian(ctypes.byref(p), xyz, (ctypes.c_int * 3)(*jacbound), outp)
    return out


def viscous_rhs(p, u, v, w, T, re, pr, ref_t, order=6, metrics=None, ajac=None, du=None):
    """Viscous terms added to du[1..4] (a zero du unless given) from ghosted u, v, w, T
    whose periodic ghosts are filled."""
    if metrics is None:
        metrics = metrics_cartesian(p)
    if ajac is None:
        ajac = np.ones_like(u)
    if du is None:
        du = np.zeros((5, p.nz, p.ny, p.nx))
    lib().orc_viscous_rhs(ctypes.byref(p), ctypes.c_double(re), ctypes.c_double(pr), ctypes.c_double(ref_t),
                          ctypes.c_int(order), _ptr(u), _ptr(v), _ptr(w), _ptr(T), _metric_ptrs(metrics),
                          _ptr(ajac), _ptr(du))
    return du


def time_advance(p, krk, f, fn, du, dt):
    lib().orc_time_advance(ctypes.byref(p), ctypes.c_int(krk), _ptr(f), _ptr(fn), _ptr(du), ctypes.c_double(dt))
    return f


def cfl_dt(p, cfl, u, v, w, T):
    return lib().orc_cfl_dt(ctypes.byref(p), ctypes.c_double(cfl), _ptr(u), _ptr(v), _ptr(w), _ptr(T))


def rk3_run(p, nsteps, dt, f, d, u, v, w, T):
    """In-place N SSP-RK3 steps (Cartesian metrics); dt <= 0 means CFL = -dt."""
    lib().orc_rk3_run(ctypes.byref(p), ctypes.c_int(nsteps), ctypes.c_double(dt),
                      _ptr(f), _ptr(d), _ptr(u), _ptr(v), _ptr(w), _ptr(T))
    return f


def opcount():
    """Algorithmic FLOP counts from the op-counting build (dict)."""
    import json
    exe = os.path.join(_HERE, "opcount")
    if not os.path.exists(exe):
        build()
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)
