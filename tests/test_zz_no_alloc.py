"""ws_ana_state_noalloc (the analysis pass over the library's fixed device array) against
ws_ana_state (which allocates its records per call): identical reports on two calls in a
row, the second on a smaller box, and with planted NaN / negative cells and metric fields."""
import ctypes

import pytest
import torch

import windstencil_amd as W
from windstencil_amd import LAP
from tests.gpu_util import log, make_ctx, upload_state
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def _report(fn, ctx, st, metrics=None, dt=0.0):
    rep = W.WsAnaReport()
    mp = ctypes.byref(metrics) if metrics is not None else None
    W.check(fn(ctx.handle, ctypes.byref(st.ws), mp, dt, ctypes.byref(rep), None), fn.__name__)
    return rep


def _fields(rep):
    return (rep.nan_T, rep.neg_T, rep.neg_d, rep.neg_jac, tuple(rep.first_nan_T), tuple(rep.first_neg_T),
            tuple(rep.first_neg_d), tuple(rep.first_neg_jac), rep.cfl, rep.fatal)


def _box(n, seed):
    nx, ny, nz = n
    prims = random_state(nx, ny, nz, seed, "vortex")
    st = upload_state(prims)
    ctx = make_ctx(nx, ny, nz)
    W.check(W.lib().ws_ghost_fill_periodic(ctx.handle, (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T),
                                           5, None), "ghost fill")
    torch.cuda.synchronize()
    return ctx, st


def _both(ctx, st, metrics=None, dt=0.0):
    L = W.lib()
    a = _fields(_report(L.ws_ana_state, ctx, st, metrics, dt))
    b = _fields(_report(L.ws_ana_state_noalloc, ctx, st, metrics, dt))
    log(f"alloc {a}\nnoalloc {b}")
    assert a == b
    return b


def test_noalloc_matches_allocating_variant_on_two_boxes():
    # first call: 37x21x18 with NaN, negative T and negative density planted
    ctx, st = _box((37, 21, 18), 5)
    st.t["T"][LAP + 5, LAP + 4, LAP + 3] = float("nan")
    st.t["T"][LAP + 9, LAP + 2, 1] = float("nan")
    for k in range(12):
        st.t["T"][LAP + k % 18, LAP + (3 * k) % 21, LAP + (7 * k) % 37] = -0.5
    st.t["d"][LAP + 1, LAP + 2, LAP + 30] = -1.0
    torch.cuda.synchronize()
    r1 = _both(ctx, st, dt=1e-3)
    assert r1[0] == 2 and r1[1] == 12 and r1[2] == 1 and r1[9] == 1 and r1[8] > 0.0
    # second call, smaller box, clean state: nothing of the first call's records survives
    ctx2, st2 = _box((20, 12, 9), 7)
    r2 = _both(ctx2, st2, dt=2e-3)
    assert r2[:4] == (0, 0, 0, 0) and r2[9] == 0 and r2[4] == (W.WS_ANA_NONE,) * 3
    # the same small box with one negative Ajac in metric fields and a NaN in a ghost cell
    keep = []
    mf = W.WsMetricFields()
    shape = (9 + 2 * LAP, 12 + 2 * LAP, 20 + 2 * LAP)
    for m in range(9):
        a = torch.zeros(shape, dtype=torch.float64, device="cuda")
        if m in (0, 4, 8):
            a.fill_(1.0)
        keep.append(a)
        mf.A[m] = W.field_of(a)
    aj = torch.ones(shape, dtype=torch.float64, device="cuda")
    aj[0, 5, 6] = -2.0
    keep.append(aj)
    mf.ajac = W.field_of(aj)
    st2.t["T"][2, LAP + 3, LAP + 4] = float("nan")
    torch.cuda.synchronize()
    r3 = _both(ctx2, st2, metrics=mf, dt=0.0)
    assert r3[0] == 1 and r3[3] == 1 and r3[7] == (6 - LAP, 5 - LAP, -LAP) and r3[8] == 0.0
