"""The C-ABI library: loads without a GPU, exports every declared symbol, and
rejects bad configurations on the host before any device call."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}()]|\([^)]*\))*\)\s*;", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_expected_surface():
    names = _declared_functions()
    for must in ("ws_inviscid_rhs", "ws_ghost_fill_periodic", "ws_rk3_step", "du_invis_Jacobian3d_x",
                 "du_invis_Jacobian3d_init", "get_duvwT", "OCFD_time_advance", "cuda_mem_value_init_warp"):
        assert must in names


def test_library_exports_every_declared_symbol(ws):
    L = ws.lib()
    missing = [n for n in sorted(_declared_functions()) if not hasattr(L, n)]
    assert not missing, missing
    # and the Python front-end binds each of them with a signature
    assert set(_declared_functions()) <= set(ws._SIGS)


def test_symbols_are_unmangled_c(ws):
    out = subprocess.run(["nm", "-D", "--defined-only", ws.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in _declared_functions():
        assert n in exported, n


def test_descriptor_layouts_match_the_reference(ws):
    # cudaField / cudaSoA / cudaJobPackage sizes under __HIPCC__ (cuda_commen.h:48-104)
    assert ctypes.sizeof(ws.WsField) == 16
    assert ctypes.sizeof(ws.WsSoa) == 24
    assert ctypes.sizeof(ws.WsJob) == 24
    # ws_ana_report: four int64 counts, four int32[3] cells, the CFL number, the flag
    assert ctypes.sizeof(ws.WsAnaReport) == 96
    assert ws.WsAnaReport.cfl.offset == 80 and ws.WsAnaReport.fatal.offset == 88


def test_host_side_validation_without_gpu(ws):
    L = ws.lib()
    assert L.ws_version().startswith(b"windstencil_amd")
    cfg = ws.default_config(64, 32, 16)
    assert cfg.gamma == 1.4 and cfg.metric[0] == 1.0 and cfg.metric[9] == 1.0
    h = ctypes.c_void_p()
    bad = ws.default_config(3, 32, 16)                  # extent below the halo depth
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    assert b"LAP" in L.ws_last_error()
    bad = ws.default_config(64, 32, 16, d0_bound=(0, 0, 1, 0, 0, 0))   # closure on a periodic direction
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    assert b"non-periodic" in L.ws_last_error()
    # the rim kernel's form (2) exists on high faces only
    bad = ws.default_config(64, 32, 16, periodic=(0, 1, 1), d0_bound=(2, 1, 0, 0, 0, 0))
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    assert b"high face" in L.ws_last_error()
    bad = ws.default_config(64, 32, 16, periodic=(0, 1, 1), d0_bound=(1, 3, 0, 0, 0, 0))
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    # viscous terms: every direction periodic or closed on both faces
    bad = ws.default_config(64, 32, 16, periodic=(1, 0, 1), viscous=6)
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -2
    bad = ws.default_config(64, 32, 16, periodic=(1, 0, 1), d0_bound=(0, 0, 1, 0, 0, 0), viscous=6)
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -2
    assert b"both faces" in L.ws_last_error()
    bad = ws.default_config(64, 32, 16, gamma=1.0)
    assert L.ws_ctx_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    assert L.ws_error_string(-3) == b"HIP runtime error"
    rep = ws.WsAnaReport()
    assert L.ws_ana_state(None, None, None, 0.0, ctypes.byref(rep), None) == -1
    v = ctypes.c_double()
    assert L.ws_probe_valu(ws.WS_PREC_FP64, ws.WS_PROBE_FMA, None, None) == -1


def test_product_path_has_no_oracle_dependency():
    """The library sources and the package never reference the oracle."""
    pkg = os.path.join(ROOT, "windstencil_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f
