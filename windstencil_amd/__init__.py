"""windstencil_amd -- MI355X inviscid-RHS engine (Python front-end).

The compute runs in libwindstencil_amd.so (HIP, gfx950), reached through its C
ABI (include/windstencil_amd.h) with ctypes.  torch supplies device memory and
streams only.  Importing this package does not touch the GPU; `lib()` raises
if the shared library has not been built (python -m windstencil_amd.build).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; WINDSTENCIL_AMD_LIB names another build of the same library
LIB_PATH = os.environ.get("WINDSTENCIL_AMD_LIB") or os.path.join(HERE, "libwindstencil_amd.so")
LAP = 4

WS_OK = 0
WS_ERR_INVALID, WS_ERR_UNSUPPORTED, WS_ERR_HIP, WS_ERR_NOMEM = -1, -2, -3, -4
WS_PREC_FP64, WS_PREC_MIXED = 0, 1
WS_SCHED_GRID, WS_SCHED_PERSISTENT, WS_SCHED_PREFETCH = 0, 1, 2
SCHEDULES = (WS_SCHED_GRID, WS_SCHED_PERSISTENT, WS_SCHED_PREFETCH)
WS_KERNEL_TILED, WS_KERNEL_ANCHOR = 0, 1
WS_TILE_15x16_K1, WS_TILE_30x16_K2, WS_TILE_30x16_A16_K2 = 0, 1, 2
WS_TILE_62x8_K2 = 3
WS_TILE_126x2_K2 = 4
WS_PROBE_FMA, WS_PROBE_MUL, WS_PROBE_ADD = 0, 1, 2
WS_ORDER_XYZ, WS_ORDER_ZYX = 0, 1
ORDERS = (WS_ORDER_XYZ, WS_ORDER_ZYX)
TILE_SHAPES = (WS_TILE_15x16_K1, WS_TILE_30x16_K2, WS_TILE_30x16_A16_K2, WS_TILE_62x8_K2, WS_TILE_126x2_K2)

_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32
_dp = ctypes.POINTER(ctypes.c_double)


class WsField(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("pitch", _u32), ("dummy0", _u32)]


class WsSoa(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("pitch", _u32), ("length_y", _u32), ("length_z", _u32),
                ("dummy0", _u32)]


class WsDim3(ctypes.Structure):
    _fields_ = [("x", _u32), ("y", _u32), ("z", _u32)]


class WsJob(ctypes.Structure):
    _fields_ = [("start", WsDim3), ("end", WsDim3)]


class WsMetricFields(ctypes.Structure):
    _fields_ = [("A", WsField * 9), ("ajac", WsField)]


class WsConfig(ctypes.Structure):
    _fields_ = [("nx", _i32), ("ny", _i32), ("nz", _i32), ("periodic", _i32 * 3),
                ("gamma", ctypes.c_double), ("ama", ctypes.c_double), ("epsl_sw", ctypes.c_double),
                ("dt", ctypes.c_double), ("hx", ctypes.c_double), ("hy", ctypes.c_double),
                ("hz", ctypes.c_double), ("metric", ctypes.c_double * 10), ("precision", _i32),
                ("kernel", _i32), ("bc_planes", _i32), ("viscous", _i32), ("re", ctypes.c_double),
                ("pr", ctypes.c_double), ("ref_t", ctypes.c_double), ("d0_bound", _i32 * 6), ("non_ref", _i32 * 6)]


class WsState(ctypes.Structure):
    _fields_ = [("f", WsSoa), ("fn", WsSoa), ("du", WsSoa),
                ("d", WsField), ("u", WsField), ("v", WsField), ("w", WsField),
                ("T", WsField), ("P", WsField), ("cc", WsField)]


class WsAnaReport(ctypes.Structure):
    _fields_ = [("nan_T", ctypes.c_int64), ("neg_T", ctypes.c_int64), ("neg_d", ctypes.c_int64),
                ("neg_jac", ctypes.c_int64), ("first_nan_T", _i32 * 3), ("first_neg_T", _i32 * 3),
                ("first_neg_d", _i32 * 3), ("first_neg_jac", _i32 * 3), ("cfl", ctypes.c_double),
                ("fatal", _i32)]


WS_ANA_NONE = -2147483648


# every extern "C" symbol of include/*.h, with its ctypes signature
_SIGS = {
    "ws_version": (ctypes.c_char_p, []),
    "ws_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "ws_last_error": (ctypes.c_char_p, []),
    "ws_config_default": (None, [ctypes.POINTER(WsConfig), _i32, _i32, _i32]),
    "ws_ctx_create": (ctypes.c_int, [ctypes.POINTER(WsConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "ws_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ws_ctx_get_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsConfig)]),
    "ws_ctx_set_precision": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_set_kernel": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_set_tile_shape": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_get_tile_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_ctx_set_schedule": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_get_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_ctx_set_direction_order": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_get_direction_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_ctx_set_stage_order": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_get_stage_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_ctx_set_stage_shape": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_ctx_get_stage_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_ctx_set_stage_fusion": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_state_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState)]),
    "ws_state_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState)]),
    "ws_inviscid_rhs": (ctypes.c_int, [ctypes.c_void_p] + [WsField] * 5 +
                        [ctypes.POINTER(WsMetricFields), WsSoa, WsJob, _i32, ctypes.c_void_p]),
    "ws_sound_speed": (ctypes.c_int, [ctypes.c_void_p, WsField, WsField, WsJob, ctypes.c_void_p]),
    "ws_ghost_fill_periodic": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsField), _i32, ctypes.c_void_p]),
    "ws_prepare_rhs_inputs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_void_p]),
    "ws_mesh_write_host": (ctypes.c_int, [ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(ctypes.c_void_p)]),
    "ws_mesh_read_host": (ctypes.c_int, [ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(ctypes.c_void_p)]),
    "ws_comput_jacobian": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsField), ctypes.POINTER(WsMetricFields),
                                          ctypes.POINTER(_i32), ctypes.c_void_p]),
    "ws_load_mesh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(WsField),
                                    ctypes.POINTER(WsMetricFields), ctypes.POINTER(_i32), ctypes.c_void_p]),
    "ws_du_comput": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.POINTER(WsMetricFields),
                                    ctypes.c_void_p]),
    "ws_viscous_rhs": (ctypes.c_int, [ctypes.c_void_p] + [WsField] * 4 + [ctypes.POINTER(WsMetricFields), WsSoa,
                                                                      ctypes.c_void_p]),
    "ws_cons_to_pri": (ctypes.c_int, [ctypes.c_void_p, WsSoa] + [WsField] * 6 + [ctypes.c_void_p]),
    "ws_pri_to_cons": (ctypes.c_int, [ctypes.c_void_p, WsSoa] + [WsField] * 5 + [WsJob, ctypes.c_void_p]),
    "ws_bc_planes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_void_p]),
    "ws_time_advance": (ctypes.c_int, [ctypes.c_void_p, _i32, WsSoa, WsSoa, WsSoa, WsSoa, ctypes.c_double,
                                       ctypes.c_void_p]),
    "ws_mem_value_init": (ctypes.c_int, [ctypes.c_double, ctypes.c_void_p, _u32, _u32, _u32, _u32,
                                         ctypes.c_void_p]),
    "ws_cfl_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_double, ctypes.c_void_p,
                                 ctypes.c_void_p]),
    "ws_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), _i32, ctypes.POINTER(ctypes.c_float),
                                   ctypes.c_void_p]),
    "ws_probe_valu": (ctypes.c_int, [_i32, _i32, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "ws_ana_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_void_p, ctypes.c_double,
                                    ctypes.POINTER(WsAnaReport), ctypes.c_void_p]),
    "ws_ana_state_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_void_p,
                                            ctypes.c_double, ctypes.POINTER(WsAnaReport), ctypes.c_void_p]),
    "ws_rk3_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_void_p, ctypes.c_void_p]),
    "ws_rk3_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), _i32, ctypes.c_double, _i32,
                                  ctypes.c_void_p]),
    "ws_restart_name": (ctypes.c_int, [_i32, ctypes.c_char_p, _i32]),
    "ws_restart_write_host": (ctypes.c_int, [ctypes.c_char_p, _i32, _i32, _i32, _i32, ctypes.c_double,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "ws_restart_read_host": (ctypes.c_int, [ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(_i32),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_void_p)]),
    "ws_save_restart": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), _i32, ctypes.c_double,
                                       ctypes.c_char_p, ctypes.c_void_p]),
    "ws_load_restart": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.c_char_p, ctypes.POINTER(_i32),
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "ws_jacobi_write_host": (ctypes.c_int, [ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(ctypes.c_void_p)]),
    "ws_jacobi_read_host": (ctypes.c_int, [ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(ctypes.c_void_p)]),
    "ws_load_jacobi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(WsMetricFields),
                                      ctypes.POINTER(WsField), ctypes.c_void_p]),
    "ws_batch_create": (ctypes.c_int, [ctypes.POINTER(WsConfig), _i32, ctypes.POINTER(_i32), ctypes.POINTER(WsState),
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "ws_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ws_batch_block": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.POINTER(WsState)]),
    "ws_batch_cells": (ctypes.c_int64, [ctypes.c_void_p]),
    "ws_batch_set_precision": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_set_kernel": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_set_tile_shape": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_get_tile_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_batch_set_schedule": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_set_stage_fusion": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_get_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_batch_set_stage_order": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_get_stage_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_batch_set_stage_shape": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "ws_batch_get_stage_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32)]),
    "ws_batch_inviscid_rhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ws_batch_du_comput": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ws_batch_cfl_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "ws_batch_rk3_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "ws_batch_rk3_run": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.c_double, _i32, ctypes.c_void_p]),
    "ws_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]),
    # reference-named surface (include/windstencil_amd_ref.h)
    "ws_ref_bind": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(WsState), ctypes.POINTER(WsMetricFields),
                                   ctypes.c_void_p]),
    "ws_ref_bind_f_lap": (ctypes.c_int, [ctypes.POINTER(WsSoa)]),
    "ws_ref_set_abort_on_error": (None, [ctypes.c_int]),
    "ws_ref_last_status": (ctypes.c_int, []),
    "du_invis_Jacobian3d_init": (None, [WsJob, ctypes.POINTER(ctypes.c_void_p)]),
    "du_invis_Jacobian3d_x": (None, [WsJob, ctypes.POINTER(WsSoa), ctypes.POINTER(WsSoa),
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "du_invis_Jacobian3d": (ctypes.c_void_p, [ctypes.c_void_p]),
    "du_vis_Jacobian3d": (ctypes.c_void_p, [ctypes.c_void_p]),
    "du_comput": (None, [ctypes.c_int]),
    "get_duvwT": (None, []),
    "OCFD_time_advance": (None, [ctypes.c_int]),
    "OCFD_bc": (None, []),
    "cuda_mem_value_init_warp": (None, [ctypes.c_double, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                        ctypes.c_uint, ctypes.c_uint]),
    "exchange_boundary_xyz_packed_dev": (None, [ctypes.c_void_p, ctypes.POINTER(WsField)]),
}

_lib = None


class WsError(RuntimeError):
    pass


def lib():
    """Load libwindstencil_amd.so; raise if it is missing (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WsError(f"{LIB_PATH} not built: run `python -m windstencil_amd.build`")
        # torch bundles its own HIP runtime under the same soname (libamdhip64.so.7).
        # Load it first so that this library binds to the runtime torch uses; the
        # other order leaves torch with a runtime it cannot initialise.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code, what="windstencil_amd"):
    if code != WS_OK:
        L = lib()
        raise WsError(f"{what}: {L.ws_error_string(code).decode()} ({L.ws_last_error().decode()})")


def default_config(nx, ny, nz, **kw):
    c = WsConfig()
    lib().ws_config_default(ctypes.byref(c), nx, ny, nz)
    for k, v in kw.items():
        if k in ("metric", "periodic", "d0_bound", "non_ref"):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(c, k, v)
    return c


class Context:
    """Owns a ws_ctx."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(lib().ws_ctx_create(ctypes.byref(cfg), ctypes.byref(h)), "ws_ctx_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().ws_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p):
        check(lib().ws_ctx_set_precision(self.handle, p), "ws_ctx_set_precision")

    def set_kernel(self, k):
        check(lib().ws_ctx_set_kernel(self.handle, k), "ws_ctx_set_kernel")

    def set_tile_shape(self, s):
        check(lib().ws_ctx_set_tile_shape(self.handle, s), "ws_ctx_set_tile_shape")

    def set_schedule(self, s):
        check(lib().ws_ctx_set_schedule(self.handle, s), "ws_ctx_set_schedule")

    def set_direction_order(self, o):
        check(lib().ws_ctx_set_direction_order(self.handle, o), "ws_ctx_set_direction_order")

    def set_stage_order(self, o):
        """Direction order of the RHS launches that apply an RK3 stage (WS_ORDER_*)."""
        check(lib().ws_ctx_set_stage_order(self.handle, o), "ws_ctx_set_stage_order")

    def set_stage_shape(self, s):
        """Tile shape of the RHS launches that apply an RK3 stage (-1: the context's)."""
        check(lib().ws_ctx_set_stage_shape(self.handle, s), "ws_ctx_set_stage_shape")

    def set_stage_fusion(self, level):
        """0: separate stage kernel; 1: the last RHS launch applies each stage; 2: it also
        derives the primitives of stages 1 and 2; 3 (default): within ws_rk3_run those of
        stage 3 too, the steps alternating two primitive sets."""
        check(lib().ws_ctx_set_stage_fusion(self.handle, int(level)), "ws_ctx_set_stage_fusion")


# ---- torch-backed buffers ---------------------------------------------------------------------

def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ghosted_tensor(nx, ny, nz, pad=0, device="cuda"):
    """Zeroed ghosted field storage (nz+8, ny+8, pitch), pitch = nx+8+pad."""
    import torch
    return torch.zeros((nz + 2 * LAP, ny + 2 * LAP, nx + 2 * LAP + pad), dtype=torch.float64, device=device)


def soa_tensor(nx, ny, nz, pad=0, device="cuda"):
    """Zeroed interior state storage (5, nz, ny, pitch), pitch = nx+pad."""
    import torch
    return torch.zeros((5, nz, ny, nx + pad), dtype=torch.float64, device=device)


def field_of(t):
    assert t.dtype.itemsize == 8 and t.is_contiguous()
    return WsField(t.data_ptr(), t.shape[-1], 0)


def soa_of(t):
    assert t.dtype.itemsize == 8 and t.is_contiguous()
    return WsSoa(t.data_ptr(), t.shape[-1], t.shape[-2], t.shape[-3], 0)


def job(start, end):
    return WsJob(WsDim3(*start), WsDim3(*end))


def interior_job(nx, ny, nz):
    return job((LAP, LAP, LAP), (nx + LAP, ny + LAP, nz + LAP))


class TorchState:
    """A ws_state whose buffers are torch tensors (optionally with padded pitches)."""

    def __init__(self, nx, ny, nz, pad=0, device="cuda"):
        self.n = (nx, ny, nz)
        self.t = {k: soa_tensor(nx, ny, nz, pad, device) for k in ("f", "fn", "du")}
        for k in ("d", "u", "v", "w", "T", "P", "cc"):
            self.t[k] = ghosted_tensor(nx, ny, nz, pad, device)
        self.ws = WsState()
        self.refresh()

    def refresh(self):
        for k in ("f", "fn", "du"):
            setattr(self.ws, k, soa_of(self.t[k]))
        for k in ("d", "u", "v", "w", "T", "P", "cc"):
            setattr(self.ws, k, field_of(self.t[k]))

    def sync_from_ws(self):
        """After the library swapped f and fn inside ws_state, swap the tensors to match."""
        if self.ws.f.ptr != self.t["f"].data_ptr():
            self.t["f"], self.t["fn"] = self.t["fn"], self.t["f"]
        assert self.ws.f.ptr == self.t["f"].data_ptr()

    def interior(self, key):
        nx, ny, nz = self.n
        a = self.t[key]
        if key in ("f", "fn", "du"):
            return a[:, :, :, :nx]
        return a[LAP:LAP + nz, LAP:LAP + ny, LAP:LAP + nx]

    def set_prims(self, rho, u, v, w, T):
        """Upload interior primitives (numpy (nz, ny, nx)); the conserved state f is
        left alone (ws_pri_to_cons derives it)."""
        import torch
        for k, a in zip(("d", "u", "v", "w", "T"), (rho, u, v, w, T)):
            self.interior(k).copy_(torch.from_numpy(a))


class Batch:
    """A ws_batch over torch-backed block states (one TorchState per block)."""

    def __init__(self, base_cfg, extents, states=None):
        self.extents = [tuple(int(v) for v in e) for e in extents]
        self.states = states if states is not None else [TorchState(*e) for e in self.extents]
        n = len(self.extents)
        ext = (_i32 * (3 * n))(*[v for e in self.extents for v in e])
        arr = (WsState * n)(*[s.ws for s in self.states])
        h = ctypes.c_void_p()
        check(lib().ws_batch_create(ctypes.byref(base_cfg), n, ext, arr, ctypes.byref(h)), "ws_batch_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().ws_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cells(self):
        return int(lib().ws_batch_cells(self.handle))

    def sync_states(self):
        """Pick up the f/fn exchange of RK3 steps in every block's TorchState."""
        for i, s in enumerate(self.states):
            st = WsState()
            check(lib().ws_batch_block(self.handle, i, ctypes.byref(st)), "ws_batch_block")
            s.ws.f, s.ws.fn = st.f, st.fn
            s.sync_from_ws()

