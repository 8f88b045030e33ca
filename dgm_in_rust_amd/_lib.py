"""ctypes binding of libdgm.so (include/dgm.h). Fails loudly if the HIP library is missing."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgm.so")

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)

STATUS = {0: "OK", 1: "NEG_DENSITY", 2: "NEG_PRESSURE", 3: "NONFINITE", 4: "BAD_CONFIG", 5: "BAD_MESH",
          6: "HIP_ERROR", 7: "OUT_OF_MEMORY", 8: "BAD_ARGUMENT", 9: "NO_SPACE"}
NO_SPACE = 9

FP64, FP32_BULK = 0, 1
LIMITER_OFF, LIMITER_SEQUENTIAL, LIMITER_JACOBI = 0, 1, 2

EXPORTED = ["dgm_ctx_create", "dgm_ctx_create_host", "dgm_ctx_destroy", "dgm_ctx_info", "dgm_last_error", "dgm_get_connectivity",
            "dgm_get_geometry", "dgm_get_tables", "dgm_initial_solution", "dgm_project", "dgm_cell_points_xy",
            "dgm_compute_residuals", "dgm_compute_residuals_device_noalloc", "dgm_apply_limiter", "dgm_compute_time_step", "dgm_rk3", "dgm_time_march",
            "dgm_state_upload", "dgm_state_download", "dgm_device_time_march", "dgm_device_steps_async",
            "dgm_synchronize", "dgm_stream", "dgm_check_state", "dgm_device_clock", "dgm_lift_coefficient",
            "dgm_profile_steps", "dgm_troubled_counts", "dgm_get_tuning", "dgm_set_tuning", "dgm_autotune",
            "dgm_batch_create", "dgm_batch_destroy", "dgm_batch_last_error", "dgm_batch_footprint", "dgm_batch_add",
            "dgm_batch_remove", "dgm_batch_steps_async", "dgm_batch_synchronize", "dgm_batch_case_info",
            "dgm_batch_case_state", "dgm_batch_lift", "dgm_batch_get_tuning", "dgm_batch_set_tuning",
            "dgm_batch_autotune", "dgm_batch_poll_async", "dgm_batch_poll_wait", "dgm_batch_device_time",
            "dgm_batch_profile_steps"]


class Mesh(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("vertices_xy", _f64p), ("n_boundary_blocks", C.c_int64),
                ("boundary_block_type", _i32p), ("boundary_block_count", _i64p), ("boundary_nodes", _i64p),
                ("n_element_blocks", C.c_int64), ("element_block_count", _i64p), ("element_nodes", _i64p),
                ("element_xy", _f64p)]


class Params(C.Structure):
    _fields_ = [("order", C.c_int32), ("n_cell_gp", C.c_int32), ("n_edge_gp", C.c_int32), ("hcr", C.c_double),
                ("cfl", C.c_double), ("fs_mach", C.c_double), ("fs_aoa_deg", C.c_double), ("fs_rho", C.c_double),
                ("fs_p", C.c_double), ("fs_hcr", C.c_double), ("precision", C.c_int32), ("limiter", C.c_int32)]


class Info(C.Structure):
    _fields_ = [("n_elements", C.c_int64), ("n_edges", C.c_int64), ("n_internal_edges", C.c_int64),
                ("n_internal_elements", C.c_int64), ("n_boundary_elements", C.c_int64), ("n_basis", C.c_int32),
                ("n_cell_gp", C.c_int32), ("n_edge_gp", C.c_int32), ("order", C.c_int32),
                ("precision", C.c_int32), ("limiter", C.c_int32), ("n_limiter_levels", C.c_int32),
                ("fused_pass", C.c_int32), ("cfl", C.c_double), ("device_bytes", C.c_int64)]


class Error(C.Structure):
    _fields_ = [("status", C.c_int32), ("edge", C.c_int64), ("element", C.c_int64), ("message", C.c_char * 256)]


class MarchStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("time", C.c_double), ("last_dt", C.c_double),
                ("troubled_cells_last_stage", C.c_int64)]


class KernelTimes(C.Structure):
    _fields_ = [("surface_ms", C.c_double), ("element_ms", C.c_double), ("detect_ms", C.c_double),
                ("rebuild_ms", C.c_double), ("apply_ms", C.c_double), ("time_step_ms", C.c_double),
                ("control_ms", C.c_double), ("step_ms", C.c_double), ("steps", C.c_int32)]


class BatchParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("limiter", C.c_int32), ("hcr", C.c_double), ("max_cases", C.c_int32),
                ("pad", C.c_int32), ("capacity", C.c_int64 * 4)]


class CaseParams(C.Structure):
    _fields_ = [("order", C.c_int32), ("final_step", C.c_int32), ("final_time", C.c_double), ("cfl", C.c_double),
                ("fs_mach", C.c_double), ("fs_aoa_deg", C.c_double), ("fs_rho", C.c_double), ("fs_p", C.c_double),
                ("fs_hcr", C.c_double)]


class CaseInfo(C.Structure):
    _fields_ = [("live", C.c_int32), ("active", C.c_int32), ("status", C.c_int32), ("steps", C.c_int32),
                ("time", C.c_double), ("last_dt", C.c_double), ("error_edge", C.c_int64),
                ("troubled_last_stage", C.c_int32), ("order", C.c_int32), ("n_elements", C.c_int64),
                ("footprint_bytes", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("surface_block", C.c_int32), ("element_waves", C.c_int32), ("element_block", C.c_int32),
                ("detect_block", C.c_int32), ("remap", C.c_int32), ("weno_grid", C.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class DgmError(RuntimeError):
    def __init__(self, status, edge=-1, element=-1, message=""):
        self.status, self.edge, self.element = status, edge, element
        super().__init__(f"{STATUS.get(status, status)}: {message} (edge {edge}, element {element})")


_lib = None


def lib():
    """Load libdgm.so. Raises if it has not been built: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python dgm_in_rust_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.dgm_ctx_create.argtypes = [C.POINTER(Mesh), C.POINTER(Params), C.POINTER(vp)]
    L.dgm_ctx_create_host.argtypes = [C.POINTER(Mesh), C.POINTER(Params), C.POINTER(vp)]
    L.dgm_ctx_destroy.argtypes = [vp]
    L.dgm_ctx_destroy.restype = None
    L.dgm_ctx_info.argtypes = [vp, C.POINTER(Info)]
    L.dgm_last_error.argtypes = [vp, C.POINTER(Error)]
    L.dgm_get_connectivity.argtypes = [vp] + [_i64p] * 10
    L.dgm_get_geometry.argtypes = [vp] + [_f64p] * 5
    L.dgm_get_tables.argtypes = [vp] + [_f64p] * 7
    L.dgm_initial_solution.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, _f64p]
    L.dgm_project.argtypes = [vp, _f64p, _f64p]
    L.dgm_cell_points_xy.argtypes = [vp, _f64p]
    L.dgm_compute_residuals.argtypes = [vp, _f64p, _f64p, C.c_int32]
    L.dgm_compute_residuals_device_noalloc.argtypes = [vp, vp, vp]   # device pointers
    L.dgm_apply_limiter.argtypes = [vp, _f64p, _u8p]
    L.dgm_compute_time_step.argtypes = [vp, _f64p, _f64p]
    L.dgm_rk3.argtypes = [vp, _f64p, C.c_double]
    L.dgm_time_march.argtypes = [vp, _f64p, C.c_int32, C.c_double, C.POINTER(MarchStats)]
    L.dgm_state_upload.argtypes = [vp, _f64p]
    L.dgm_state_download.argtypes = [vp, _f64p]
    L.dgm_device_time_march.argtypes = [vp, C.c_int32, C.c_double, C.POINTER(MarchStats)]
    L.dgm_device_steps_async.argtypes = [vp, C.c_int32]
    L.dgm_synchronize.argtypes = [vp]
    L.dgm_stream.argtypes = [vp]
    L.dgm_stream.restype = vp
    L.dgm_check_state.argtypes = [vp]
    L.dgm_device_clock.argtypes = [vp, _f64p, _i32p, _f64p]
    L.dgm_lift_coefficient.argtypes = [vp, C.c_double, _f64p]
    L.dgm_troubled_counts.argtypes = [vp, _i32p]
    L.dgm_profile_steps.argtypes = [vp, C.c_int32, C.POINTER(KernelTimes)]
    L.dgm_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.dgm_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.dgm_autotune.argtypes = [vp, C.c_int32, C.POINTER(Tuning)]
    L.dgm_batch_create.argtypes = [C.POINTER(BatchParams), C.POINTER(vp)]
    L.dgm_batch_destroy.argtypes = [vp]
    L.dgm_batch_destroy.restype = None
    L.dgm_batch_last_error.argtypes = [vp, C.POINTER(Error)]
    L.dgm_batch_footprint.argtypes = [vp, C.c_int32, C.c_int64, C.c_int64, C.c_int64, _i64p]
    L.dgm_batch_add.argtypes = [vp, C.POINTER(Mesh), C.POINTER(CaseParams), _f64p, _i32p]
    L.dgm_batch_remove.argtypes = [vp, C.c_int32]
    L.dgm_batch_steps_async.argtypes = [vp, C.c_int32]
    L.dgm_batch_synchronize.argtypes = [vp]
    L.dgm_batch_case_info.argtypes = [vp, C.c_int32, C.POINTER(CaseInfo)]
    L.dgm_batch_case_state.argtypes = [vp, C.c_int32, _f64p]
    L.dgm_batch_lift.argtypes = [vp, C.c_double, C.c_int32, _f64p]
    L.dgm_batch_get_tuning.argtypes = [vp, C.c_int32, C.POINTER(Tuning)]
    L.dgm_batch_set_tuning.argtypes = [vp, C.c_int32, C.POINTER(Tuning)]
    L.dgm_batch_autotune.argtypes = [vp, C.c_int32, C.POINTER(Tuning)]
    L.dgm_batch_profile_steps.argtypes = [vp, C.c_int32, C.POINTER(KernelTimes)]
    L.dgm_batch_poll_async.argtypes = [vp, C.c_double, C.c_double, _i64p]
    L.dgm_batch_poll_wait.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(CaseInfo), _f64p, _i32p]
    L.dgm_batch_device_time.argtypes = [vp, _f64p]
    _lib = L
    return L


def mesh_struct(m):
    """dgm_mesh view of a contiguous MeshInput (the arrays must outlive the struct)."""
    return Mesh(m.vxy.shape[0], ptr(m.vxy), len(m.eb_type), ptr(m.eb_type, _i32p), ptr(m.eb_count, _i64p),
                ptr(m.eb_nodes, _i64p), len(m.el_count), ptr(m.el_count, _i64p), ptr(m.el_nodes, _i64p),
                ptr(m.elem_xy) if m.elem_xy is not None else None)


def ptr(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)
