"""Restart files in the reference's OCFD%08d.dat format, through the C ABI's host entry
points (no GPU).  The expected byte layout is restated here from the reference's writer
(src/io/OCFD_IO.c:153-226 and the per-plane records of src/io/OCFD_IO_mpi.c:605-701):
a Fortran-style header record (int32 12, int32 Istep, float64 tt, int32 12), then rho, u,
v, w, T as one record per z plane framed by int32 NX*NY*8 markers."""
import ctypes
import os

import numpy as np
import pytest

import windstencil_amd as W


def _ptrs(arrs):
    return (ctypes.c_void_p * 5)(*[a.ctypes.data for a in arrs])


def _fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return [np.ascontiguousarray(rng.standard_normal((nz, ny, nx))) for _ in range(5)]


def _expected_bytes(fields, istep, tt):
    nz, ny, nx = fields[0].shape
    out = [np.array([12, istep], dtype=np.int32).tobytes(), np.array([tt], dtype=np.float64).tobytes(),
           np.array([12], dtype=np.int32).tobytes()]
    mark = np.array([nx * ny * 8], dtype=np.int32).tobytes()
    for f in fields:
        for z in range(nz):
            out += [mark, f[z].astype(np.float64).tobytes(), mark]
    return b"".join(out)


def test_restart_name():
    buf = ctypes.create_string_buffer(32)
    W.check(W.lib().ws_restart_name(12, buf, 32))
    assert buf.value == b"OCFD00000012.dat"
    assert W.lib().ws_restart_name(5, buf, 8) == W.WS_ERR_INVALID


def test_write_matches_reference_layout(tmp_path):
    nx, ny, nz = 7, 5, 3
    f = _fields(nx, ny, nz, 1)
    path = str(tmp_path / "OCFD00000042.dat").encode()
    W.check(W.lib().ws_restart_write_host(path, nx, ny, nz, 42, 0.125, _ptrs(f)))
    data = open(path, "rb").read()
    assert len(data) == 20 + 5 * nz * (8 + 8 * nx * ny)
    assert data == _expected_bytes(f, 42, 0.125)


def test_read_of_independently_built_file(tmp_path):
    nx, ny, nz = 9, 4, 6
    f = _fields(nx, ny, nz, 2)
    path = tmp_path / "opencfd.dat"
    path.write_bytes(_expected_bytes(f, 1234, 3.5e-2))
    out = [np.zeros((nz, ny, nx)) for _ in range(5)]
    istep, tt = ctypes.c_int32(), ctypes.c_double()
    W.check(W.lib().ws_restart_read_host(str(path).encode(), nx, ny, nz, ctypes.byref(istep), ctypes.byref(tt),
                                         _ptrs(out)))
    assert istep.value == 1234 and tt.value == 3.5e-2
    for a, b in zip(out, f):
        assert np.array_equal(a, b)


def test_round_trip_and_grid_checks(tmp_path):
    nx, ny, nz = 6, 6, 4
    f = _fields(nx, ny, nz, 3)
    path = str(tmp_path / "r.dat").encode()
    L = W.lib()
    W.check(L.ws_restart_write_host(path, nx, ny, nz, 7, 1.0, _ptrs(f)))
    istep, tt = ctypes.c_int32(), ctypes.c_double()
    out = [np.zeros((nz, ny, nx)) for _ in range(5)]
    W.check(L.ws_restart_read_host(path, nx, ny, nz, ctypes.byref(istep), ctypes.byref(tt), _ptrs(out)))
    assert all(np.array_equal(a, b) for a, b in zip(out, f))
    # another grid: the plane records or the file length do not match
    small = [np.zeros((nz - 1, ny, nx)) for _ in range(5)]
    assert L.ws_restart_read_host(path, nx, ny, nz - 1, ctypes.byref(istep), ctypes.byref(tt),
                                  _ptrs(small)) == W.WS_ERR_INVALID
    assert b"longer" in L.ws_last_error()
    other = [np.zeros((nz, ny, nx + 1)) for _ in range(5)]
    assert L.ws_restart_read_host(path, nx + 1, ny, nz, ctypes.byref(istep), ctypes.byref(tt),
                                  _ptrs(other)) == W.WS_ERR_INVALID
    # truncated file
    data = open(path, "rb").read()
    cut = tmp_path / "cut.dat"
    cut.write_bytes(data[:-5])
    assert L.ws_restart_read_host(str(cut).encode(), nx, ny, nz, ctypes.byref(istep), ctypes.byref(tt),
                                  _ptrs(out)) == W.WS_ERR_INVALID
    # corrupted header marker
    bad = tmp_path / "bad.dat"
    bad.write_bytes(np.array([13], dtype=np.int32).tobytes() + data[4:])
    assert L.ws_restart_read_host(str(bad).encode(), nx, ny, nz, ctypes.byref(istep), ctypes.byref(tt),
                                  _ptrs(out)) == W.WS_ERR_INVALID
    assert L.ws_restart_read_host(str(tmp_path / "missing.dat").encode(), nx, ny, nz, ctypes.byref(istep),
                                  ctypes.byref(tt), _ptrs(out)) == W.WS_ERR_INVALID


def test_jacobi_file_round_trip(tmp_path):
    """OCFD3d-Jacobi.dat: 13 fields of plane records, no header (the reference reads it
    from offset 0, src/solver/OCFD_Comput_Jacobian3d.cpp:138-170)."""
    nx, ny, nz = 5, 4, 3
    rng = np.random.default_rng(9)
    f = [np.ascontiguousarray(rng.standard_normal((nz, ny, nx))) for _ in range(13)]
    path = str(tmp_path / "OCFD3d-Jacobi.dat").encode()
    L = W.lib()
    W.check(L.ws_jacobi_write_host(path, nx, ny, nz, (ctypes.c_void_p * 13)(*[a.ctypes.data for a in f])))
    data = open(path, "rb").read()
    mark = np.array([nx * ny * 8], dtype=np.int32).tobytes()
    assert data == b"".join(mark + a[z].tobytes() + mark for a in f for z in range(nz))
    out = [np.zeros((nz, ny, nx)) for _ in range(13)]
    W.check(L.ws_jacobi_read_host(path, nx, ny, nz, (ctypes.c_void_p * 13)(*[a.ctypes.data for a in out])))
    assert all(np.array_equal(a, b) for a, b in zip(out, f))
    # a restart file is not a Jacobi file of the same grid
    r5 = str(tmp_path / "r.dat").encode()
    W.check(L.ws_restart_write_host(r5, nx, ny, nz, 1, 0.0, _ptrs(f[:5])))
    assert L.ws_jacobi_read_host(r5, nx, ny, nz, (ctypes.c_void_p * 13)(*[a.ctypes.data for a in out])) \
        == W.WS_ERR_INVALID


def test_mesh_file_round_trip(tmp_path):
    """OCFD3d-Mesh.dat: x, y, z as plane records, no header (read with the Jacobi file's
    offsets, src/solver/OCFD_Comput_Jacobian3d.cpp:120-135); the first three fields of a
    Jacobi file have the same bytes."""
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(11)
    f = [np.ascontiguousarray(rng.standard_normal((nz, ny, nx))) for _ in range(3)]
    path = str(tmp_path / "OCFD3d-Mesh.dat").encode()
    L = W.lib()
    W.check(L.ws_mesh_write_host(path, nx, ny, nz, (ctypes.c_void_p * 3)(*[a.ctypes.data for a in f])))
    data = open(path, "rb").read()
    mark = np.array([nx * ny * 8], dtype=np.int32).tobytes()
    assert data == b"".join(mark + a[z].tobytes() + mark for a in f for z in range(nz))
    out = [np.zeros((nz, ny, nx)) for _ in range(3)]
    W.check(L.ws_mesh_read_host(path, nx, ny, nz, (ctypes.c_void_p * 3)(*[a.ctypes.data for a in out])))
    assert all(np.array_equal(a, b) for a, b in zip(out, f))
    # another grid's mesh file is rejected
    assert L.ws_mesh_read_host(path, nx, ny, nz - 1, (ctypes.c_void_p * 3)(*[a.ctypes.data for a in out])) \
        == W.WS_ERR_INVALID
