"""Failure detection on the device (ws_ana_state): NaN and negative temperature (the
reference's ana_NAN_and_NT), negative Ajac (ana_Jac), negative density, and the CFL
number of a time step against the oracle's CFL step."""
import ctypes

import numpy as np
import pytest
import torch

import windstencil_amd as W
from windstencil_amd import LAP
from tests.gpu_util import log, make_ctx, upload_state
from tests.helpers import ghosted, random_state

pytestmark = pytest.mark.gpu

NX, NY, NZ = 37, 21, 18


def _ana(ctx, st, metrics=None, dt=0.0):
    rep = W.WsAnaReport()
    mp = ctypes.byref(metrics) if metrics is not None else None
    W.check(W.lib().ws_ana_state(ctx.handle, ctypes.byref(st.ws), mp, dt, ctypes.byref(rep), None), "ana")
    return rep


def _fresh():
    prims = random_state(NX, NY, NZ, 5, "vortex")
    st = upload_state(prims)
    ctx = make_ctx(NX, NY, NZ)
    W.check(W.lib().ws_ghost_fill_periodic(ctx.handle, (W.WsField * 5)(st.ws.d, st.ws.u, st.ws.v, st.ws.w, st.ws.T),
                                           5, None), "ghost fill")
    torch.cuda.synchronize()
    return ctx, st, prims


def test_clean_state_and_cfl_number(orc):
    ctx, st, prims = _fresh()
    rep = _ana(ctx, st, dt=1e-3)
    assert (rep.nan_T, rep.neg_T, rep.neg_d, rep.neg_jac, rep.fatal) == (0, 0, 0, 0, 0)
    assert list(rep.first_nan_T) == [W.WS_ANA_NONE] * 3
    # the oracle's CFL step for CFL 1 is min(h) / max(|u_d| + c): cfl(dt) = dt / that
    p = orc.params(NX, NY, NZ)
    g = [ghosted(a) for a in prims]
    dt1 = orc.cfl_dt(p, 1.0, g[1], g[2], g[3], g[4])
    log(f"cfl {rep.cfl:.15g} oracle {1e-3 / dt1:.15g}")
    assert abs(rep.cfl - 1e-3 / dt1) <= 1e-12 * (1e-3 / dt1)
    assert _ana(ctx, st, dt=0.0).cfl == 0.0


def test_nan_and_negative_temperature_are_found_and_located():
    ctx, st, _ = _fresh()
    T = st.t["T"]
    T[LAP + 5, LAP + 4, LAP + 3] = float("nan")          # interior (x, y, z) = (3, 4, 5)
    T[LAP + 9, LAP + 2, 1] = float("nan")                # a ghost cell, x = -3 (later in memory)
    torch.cuda.synchronize()
    rep = _ana(ctx, st)
    assert rep.nan_T == 2 and rep.fatal == 1
    assert list(rep.first_nan_T) == [3, 4, 5]
    ctx, st, _ = _fresh()
    T = st.t["T"]
    cells = [(LAP + k % NZ, LAP + (3 * k) % NY, LAP + (7 * k) % NX) for k in range(12)]
    for c in cells[:10]:
        T[c] = -0.5
    torch.cuda.synchronize()
    rep = _ana(ctx, st)
    assert rep.neg_T == 10 and rep.fatal == 0                  # the reference aborts above 10
    first = min(cells[:10], key=lambda c: (c[0], c[1], c[2]))
    assert list(rep.first_neg_T) == [first[2] - LAP, first[1] - LAP, first[0] - LAP]
    for c in cells[10:]:
        T[c] = -0.5
    torch.cuda.synchronize()
    rep = _ana(ctx, st)
    assert rep.neg_T == 12 and rep.fatal == 1


def test_negative_density_and_jacobian():
    ctx, st, _ = _fresh()
    st.t["d"][LAP + 1, LAP + 2, LAP + 30] = -1.0
    torch.cuda.synchronize()
    rep = _ana(ctx, st)
    assert rep.neg_d == 1 and list(rep.first_neg_d) == [30, 2, 1] and rep.fatal == 0
    # metric fields with one negative Ajac (in a ghost cell, as ana_Jac scans the ghosted box)
    keep = []
    mf = W.WsMetricFields()
    for m in range(9):
        a = torch.zeros((NZ + 2 * LAP, NY + 2 * LAP, NX + 2 * LAP), dtype=torch.float64, device="cuda")
        if m in (0, 4, 8):
            a.fill_(1.0)
        keep.append(a)
        mf.A[m] = W.field_of(a)
    aj = torch.ones((NZ + 2 * LAP, NY + 2 * LAP, NX + 2 * LAP), dtype=torch.float64, device="cuda")
    aj[0, 5, 6] = -2.0
    keep.append(aj)
    mf.ajac = W.field_of(aj)
    torch.cuda.synchronize()
    rep = _ana(ctx, st, metrics=mf)
    assert rep.neg_jac == 1 and list(rep.first_neg_jac) == [6 - LAP, 5 - LAP, -LAP]
    assert _ana(ctx, st).neg_jac == 0                              # constant Ajac = 1
