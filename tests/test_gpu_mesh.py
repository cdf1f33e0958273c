"""GPU parity of the metric terms computed from mesh coordinates (the reference's
Comput_Jacobian3d) against the oracle, bit for bit on every ghosted array, and the
mesh-file loader end to end."""
import ctypes

import numpy as np
import pytest

import windstencil_amd as W
from windstencil_amd import LAP
from tests.gpu_util import gpu_du_comput, log, make_ctx, upload_state
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def _grids(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(-LAP, nz + LAP), np.arange(-LAP, ny + LAP), np.arange(-LAP, nx + LAP),
                          indexing="ij")
    return i.astype(float), j.astype(float), k.astype(float)


def _mesh(kind, nx, ny, nz, h):
    """Ghosted coordinates: 'cylinder' is periodic in j (angle) with x and r linear in i, k;
    'wavy' perturbs a Cartesian mesh smoothly (non-periodic everywhere)."""
    i, j, k = _grids(nx, ny, nz)
    if kind == "cylinder":
        th = 2 * np.pi * j / ny
        r = 1.0 + 0.5 * k * h[2] + 0.05 * np.sin(2 * np.pi * i / nx)
        return [i * h[0], r * np.cos(th), r * np.sin(th)]
    X, Y, Z = i * h[0], j * h[1], k * h[2]
    return [X + 0.03 * np.sin(3 * Y + 2 * Z), Y + 0.02 * np.sin(4 * X) * np.cos(Z), Z + 0.025 * np.sin(2 * X + Y)]


def _torch_fields(arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def _metric_struct(tens):
    mf = W.WsMetricFields()
    for q in range(9):
        mf.A[q] = W.field_of(tens[q])
    mf.ajac = W.field_of(tens[9])
    return mf


CASES = [("cylinder", (12, 24, 10), (0, 1, 0), (1, 0, 1)),
         ("wavy", (17, 13, 11), (0, 0, 0), (1, 1, 1)),
         ("wavy", (9, 10, 12), (0, 0, 0), (0, 1, 0))]


@pytest.mark.parametrize("kind,shape,periodic,jacbound", CASES)
def test_metrics_bitwise_equal_to_oracle(orc, ws, kind, shape, periodic, jacbound):
    import torch
    nx, ny, nz = shape
    ctx = make_ctx(nx, ny, nz, periodic=list(periodic))
    h = (ctx.cfg.hx, ctx.cfg.hy, ctx.cfg.hz)
    p = orc.params(nx, ny, nz, periodic=periodic)
    xyz = _mesh(kind, nx, ny, nz, h)
    for a in xyz:
        orc.ghost_fill_periodic(p, a)
    ref = orc.comput_jacobian(p, *xyz, jacbound=jacbound)
    coords = _torch_fields(xyz)
    outs = [torch.ones_like(coords[0]) for _ in range(10)]
    cf = (W.WsField * 3)(*[W.field_of(c) for c in coords])
    W.check(W.lib().ws_comput_jacobian(ctx.handle, cf, ctypes.byref(_metric_struct(outs)),
                                       (ctypes.c_int32 * 3)(*jacbound), W.stream_handle()))
    torch.cuda.synchronize()
    for q in range(10):
        got = outs[q].cpu().numpy()
        log(f"  {kind} {shape} field {q}: max |diff| {np.abs(got - ref[q]).max():.3e}")
        assert np.array_equal(got, ref[q]), q


def test_mesh_file_loader_and_rhs_on_its_metrics(orc, ws, tmp_path):
    """ws_load_mesh reads OCFD3d-Mesh.dat, fills the coordinates' ghosts (1.0, then the
    periodic halo) and computes the metrics as the oracle does on the same coordinates;
    the anchor RHS on those metrics equals the oracle's bit for bit."""
    import torch
    nx, ny, nz = 12, 24, 10
    periodic, jacbound = (0, 1, 0), (1, 0, 1)
    ctx = make_ctx(nx, ny, nz, periodic=list(periodic), kernel=W.WS_KERNEL_ANCHOR)
    h = (ctx.cfg.hx, ctx.cfg.hy, ctx.cfg.hz)
    xyz = _mesh("cylinder", nx, ny, nz, h)
    inner = [np.ascontiguousarray(a[LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx]) for a in xyz]
    path = str(tmp_path / "OCFD3d-Mesh.dat").encode()
    W.check(W.lib().ws_mesh_write_host(path, nx, ny, nz, (ctypes.c_void_p * 3)(*[a.ctypes.data for a in inner])))
    coords = [torch.zeros((nz + 2 * LAP, ny + 2 * LAP, nx + 2 * LAP), dtype=torch.float64, device="cuda")
              for _ in range(3)]
    outs = [torch.zeros_like(coords[0]) for _ in range(10)]
    mf = _metric_struct(outs)
    cf = (W.WsField * 3)(*[W.field_of(c) for c in coords])
    W.check(W.lib().ws_load_mesh(ctx.handle, path, cf, ctypes.byref(mf), (ctypes.c_int32 * 3)(*jacbound),
                                 W.stream_handle()))
    torch.cuda.synchronize()
    # the oracle on the same coordinates: ghosts 1.0, interior from the file, periodic halo
    p = orc.params(nx, ny, nz, periodic=periodic)
    cx = []
    for a in inner:
        g = np.ones((nz + 2 * LAP, ny + 2 * LAP, nx + 2 * LAP))
        g[LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx] = a
        orc.ghost_fill_periodic(p, g)
        cx.append(g)
    for q in range(3):
        assert np.array_equal(coords[q].cpu().numpy(), cx[q])
    ref = orc.comput_jacobian(p, *cx, jacbound=jacbound)
    for q in range(10):
        assert np.array_equal(outs[q].cpu().numpy(), ref[q]), q
    # the RHS on these metric terms (periodic in y only: the x and z ghosts of the state are
    # the caller's, here the oracle's ghosted random state)
    prims = random_state(nx, ny, nz, 5, "blobs")
    from tests.helpers import ghosted
    g = [ghosted(a) for a in prims]
    for a in g:
        orc.ghost_fill_periodic(orc.params(nx, ny, nz), a)    # fill every direction once
    du_ref, _ = orc.du_comput(p, *g, metrics=ref[:9], ajac=ref[9])
    st = upload_state(prims)
    for name, a in zip(("d", "u", "v", "w", "T"), g):
        st.t[name].copy_(torch.from_numpy(a))
    got = gpu_du_comput(ctx, st, metrics=mf)
    assert np.array_equal(got, du_ref)
