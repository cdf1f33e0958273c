"""GPU-side helpers for the parity tests: drive the HIP library through its C ABI."""
import ctypes
import sys

import numpy as np

import windstencil_amd as W
from windstencil_amd import LAP


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_ctx(nx, ny, nz, kernel=W.WS_KERNEL_TILED, precision=W.WS_PREC_FP64, **kw):
    cfg = W.default_config(nx, ny, nz, kernel=kernel, precision=precision, **kw)
    return W.Context(cfg)


def upload_state(prims, pad=0):
    """TorchState with interior primitives set; f = pri_to_cons is NOT derived here."""
    rho = prims[0]
    nz, ny, nx = rho.shape
    st = W.TorchState(nx, ny, nz, pad=pad)
    st.set_prims(*prims)
    return st


def gpu_du_comput(ctx, st, metrics=None, job=None, accumulate=False, stream=None):
    """Ghost fill d,u,v,w,T; sound speed over the ghosted box; RHS over `job` (default:
    the interior) into st.du, plus the viscous terms when the context has them (as
    ws_du_comput).  Returns du (numpy, interior)."""
    import torch
    L = W.lib()
    s = W.stream_handle(stream)
    nx, ny, nz = st.n
    fl = (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T)
    W.check(L.ws_ghost_fill_periodic(ctx.handle, fl, 5, s), "ghost fill")
    W.check(L.ws_sound_speed(ctx.handle, st.ws.T, st.ws.cc,
                             W.job((0, 0, 0), (nx + 2 * LAP, ny + 2 * LAP, nz + 2 * LAP)), s), "sound speed")
    j = job if job is not None else W.interior_job(nx, ny, nz)
    mp = ctypes.byref(metrics) if metrics is not None else None
    W.check(L.ws_inviscid_rhs(ctx.handle, st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.cc, mp, st.ws.du, j,
                              1 if accumulate else 0, s), "inviscid rhs")
    if ctx.cfg.viscous:
        # du_comput adds the viscous terms after the inviscid RHS
        W.check(L.ws_viscous_rhs(ctx.handle, st.ws.u, st.ws.v, st.ws.w, st.ws.T, mp, st.ws.du, s), "viscous rhs")
    torch.cuda.synchronize()
    return st.interior("du").cpu().numpy()


def oracle_du(orc, prims, metrics=None, ajac=None, eps=1e-3):
    """The oracle's du_comput on the same interior primitives."""
    from tests.helpers import ghosted
    nz, ny, nx = prims[0].shape
    p = orc.params(nx, ny, nz, epsl_sw=eps)
    g = [ghosted(a) for a in prims]
    du, cc = orc.du_comput(p, *g, metrics=metrics, ajac=ajac)
    return du
