"""ctypes binding of libmycel_hip.so (include/mycel_hip.h).

The library is the only implementation of the path: if it is missing this module raises, there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmycel_hip.so")

OK = 0
E_ARGUMENT_NULL = 1
E_ARGUMENT = 2
E_OUT_OF_RANGE = 3
E_INVALID_OPERATION = 4
E_NOT_CONVERGED = 5
E_CANCELED = 6
E_DEVICE = 7
E_OUT_OF_MEMORY = 8

OC_NOALLOC_MAX_ELEMENTS = 16777216

PRECISION_FP64 = 0
PRECISION_MIXED_FP32 = 1
PRECISION_MIXED_BF16 = 2
PRECISION_MIXED_GMG = 3
PRECISION_FP64_ORDERED = 4
MATVEC_AUTO = 0
MATVEC_NODE_GATHER = 1
MATVEC_ELEMENT = 2
MATVEC_REFERENCE_ORDER = 3



class MycelError(RuntimeError):
    """Base of the exceptions mapped from library status codes."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class ArgumentNullError(MycelError, ValueError):
    pass


class ArgumentError(MycelError, ValueError):
    pass


class ArgumentOutOfRangeError(MycelError, IndexError):
    pass


class InvalidOperationError(MycelError):
    pass


class OperationCanceledError(MycelError):
    pass


class DeviceError(MycelError):
    pass


class OutOfMemoryError(MycelError, MemoryError):
    pass


_EXC = {
    E_ARGUMENT_NULL: ArgumentNullError,
    E_ARGUMENT: ArgumentError,
    E_OUT_OF_RANGE: ArgumentOutOfRangeError,
    E_INVALID_OPERATION: InvalidOperationError,
    E_NOT_CONVERGED: InvalidOperationError,
    E_CANCELED: OperationCanceledError,
    E_DEVICE: DeviceError,
    E_OUT_OF_MEMORY: OutOfMemoryError,
}


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("h", C.c_double)]


class OpInfo(C.Structure):
    _fields_ = [("total_elements", C.c_int32), ("total_nodes", C.c_int32), ("total_dofs", C.c_int32),
                ("active_elements", C.c_int32), ("active_nodes", C.c_int32), ("active_fixed_dofs", C.c_int32)]


class PcgResultC(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("final_residual", C.c_double), ("converged", C.c_int32),
                ("message", C.c_int32), ("elapsed_ms", C.c_double), ("outer_corrections", C.c_int32),
                ("fallback_iterations", C.c_int32), ("message_value", C.c_double)]


class SimpOptionsC(C.Structure):
    _fields_ = [("volume_fraction", C.c_double), ("filter_radius_mm", C.c_double), ("penal", C.c_double),
                ("move_limit", C.c_double), ("max_iterations", C.c_int32), ("solver_tol_initial", C.c_double),
                ("solver_tol_final", C.c_double), ("solver_max_iterations", C.c_int32), ("rho_min", C.c_double),
                ("protect_boundary_hinges", C.c_int32), ("start_iteration", C.c_int32),
                ("enforce_min_filter_radius", C.c_int32), ("continuation", C.c_int32), ("precision", C.c_int32),
                ("enable_heaviside", C.c_int32), ("initial_lambda", C.c_double), ("initial_design", C.c_void_p),
                ("initial_design_len", C.c_int32), ("resume_state", C.c_int32), ("resume_max_change", C.c_double),
                ("resume_beta", C.c_double), ("resume_history_len", C.c_int32), ("resume_history", C.c_double * 6),
                ("island_deflation", C.c_double)]


class SnapshotC(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("compliance", C.c_double), ("volume_fraction", C.c_double),
                ("max_change", C.c_double), ("grey_level", C.c_double), ("solver_iterations", C.c_int32),
                ("solver_residual", C.c_double), ("max_displacement", C.c_double), ("lambda_", C.c_double),
                ("solve_ms", C.c_double), ("sensitivity_ms", C.c_double), ("filter_ms", C.c_double),
                ("update_ms", C.c_double), ("total_ms", C.c_double), ("beta_next", C.c_double)]


class SimpResultC(C.Structure):
    _fields_ = [("converged", C.c_int32), ("total_iterations", C.c_int32), ("final_compliance", C.c_double),
                ("final_volume_fraction", C.c_double), ("canceled", C.c_int32)]


class JobDescC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("grid", Grid), ("domain", C.c_void_p), ("fixed_dofs", C.c_void_p),
                ("n_fixed", C.c_int32), ("forces", C.c_void_p), ("patch_nodes", C.c_void_p),
                ("n_patch_nodes", C.c_int32), ("e0", C.c_double), ("nu", C.c_double), ("emin_ratio", C.c_double),
                ("penal", C.c_double), ("density_box", C.c_void_p), ("tol", C.c_double), ("max_iter", C.c_int32),
                ("precision", C.c_int32), ("simp", SimpOptionsC)]


class PackStatsC(C.Structure):
    _fields_ = [("batches", C.c_int64), ("member_batches", C.c_int64), ("launches", C.c_int64),
                ("graph_batches", C.c_int64), ("rebuilds", C.c_int64), ("solves", C.c_int64),
                ("build_ms", C.c_double), ("enqueue_ms", C.c_double), ("wait_ms", C.c_double)]


class JobInfoC(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("compliance", C.c_double),
                ("max_displacement", C.c_double), ("pcg_iterations", C.c_int32), ("simp_iterations", C.c_int32),
                ("converged", C.c_int32), ("footprint_bytes", C.c_double), ("queue_ms", C.c_double),
                ("run_ms", C.c_double), ("latency_ms", C.c_double)]


PROGRESS_CB = C.CFUNCTYPE(None, C.c_int32, C.c_double, C.c_void_p)
SNAPSHOT_CB = C.CFUNCTYPE(C.c_int32, C.POINTER(SnapshotC), C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.c_void_p)

_vp = C.c_void_p
_i = C.c_int32
_d = C.c_double

# name -> (restype, argtypes). Every symbol include/mycel_hip.h declares.
SIGNATURES = {
    "mycel_last_error": (C.c_char_p, []),
    "mycel_version": (_i, []),
    "mycel_pcg_message": (_i, [C.POINTER(PcgResultC), C.c_char_p, _i]),
    "mycel_device_count": (_i, [C.POINTER(_i)]),
    "mycel_set_device": (_i, [_i]),
    "mycel_hex8_ke": (_i, [_d, _d, _vp]),
    "mycel_hex8_bc": (_i, [_d, _vp]),
    "mycel_assemble_forces": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp]),
    "mycel_passive_region": (_i, [C.POINTER(Grid), _vp, _i, _i, _vp]),
    "mycel_op_create": (_i, [C.POINTER(Grid), _vp, _vp, _i, _d, _d, _d, _d, _vp, C.POINTER(_vp)]),
    "mycel_op_destroy": (_i, [_vp]),
    "mycel_op_info_get": (_i, [_vp, C.POINTER(OpInfo)]),
    "mycel_op_update_densities": (_i, [_vp, _vp, _d]),
    "mycel_op_update_densities_d": (_i, [_vp, _vp, _d, _vp]),
    "mycel_op_apply_k": (_i, [_vp, _vp, _vp]),
    "mycel_op_apply_a": (_i, [_vp, _vp, _vp, _i]),
    "mycel_op_apply_k_d": (_i, [_vp, _vp, _vp, _vp]),
    "mycel_op_apply_a_d": (_i, [_vp, _vp, _vp, _i, _vp]),
    "mycel_op_apply_k_f32_d": (_i, [_vp, _vp, _vp, _vp]),
    "mycel_op_get_diag_a": (_i, [_vp, _vp]),
    "mycel_op_get_ke": (_i, [_vp, _vp]),
    "mycel_op_get_constrained": (_i, [_vp, _vp]),
    "mycel_op_get_active_elements": (_i, [_vp, _vp]),
    "mycel_op_stream": (_vp, [_vp]),
    "mycel_pcg_solve": (_i, [_vp, _vp, _vp, _d, _i, _i, _vp, PROGRESS_CB, _vp, C.POINTER(PcgResultC)]),
    "mycel_pcg_solve_d": (_i, [_vp, _vp, _vp, _d, _i, _i, C.POINTER(PcgResultC)]),
    "mycel_precond_apply_d": (_i, [_vp, _i, _vp, _vp, _vp]),
    "mycel_static_analysis": (_i, [_vp, _vp, _vp, _d, _i, _i, _i, C.POINTER(_d), C.POINTER(_d), _vp, _vp,
                                   C.POINTER(PcgResultC)]),
    "mycel_von_mises": (_i, [_vp, _vp, _vp]),
    "mycel_average_element_to_node": (_i, [_vp, _vp, _vp]),
    "mycel_sample_nodal_scalar": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "mycel_filter_create": (_i, [C.POINTER(Grid), _vp, _vp, _i, _d, C.POINTER(_vp)]),
    "mycel_filter_destroy": (_i, [_vp]),
    "mycel_filter_get_s": (_i, [_vp, _vp]),
    "mycel_filter_nnz": (_i, [_vp, C.POINTER(C.c_int64)]),
    "mycel_filter_apply": (_i, [_vp, _vp, _vp]),
    "mycel_filter_apply_adjoint": (_i, [_vp, _vp, _vp]),
    "mycel_sensitivities": (_i, [_vp, _vp, _vp, _vp, _d]),
    "mycel_oc_update": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _i, _d, C.POINTER(_d)]),
    "mycel_oc_update_noalloc_d": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _i, _d, _vp, C.POINTER(_d)]),
    "mycel_simp_options_default": (None, [C.POINTER(SimpOptionsC)]),
    "mycel_op_set_matvec_columns": (_i, [_vp, _i]),
    "mycel_op_set_matvec_form": (_i, [_vp, _i, _i]),
    "mycel_op_set_island_deflation": (_i, [_vp, _d]),
    "mycel_run_quantize": (_i, [_vp, C.c_int64, _vp, C.POINTER(C.c_uint32)]),
    "mycel_run_checksum": (_i, [_vp, C.c_int64, C.POINTER(C.c_uint32)]),
    "mycel_server_create": (_i, [_d, _i, C.POINTER(_vp)]),
    "mycel_server_destroy": (_i, [_vp]),
    "mycel_server_submit": (_i, [_vp, C.POINTER(JobDescC), C.POINTER(C.c_int64)]),
    "mycel_server_poll": (_i, [_vp, C.c_int64, C.POINTER(JobInfoC)]),
    "mycel_server_wait": (_i, [_vp, C.c_int64, _d, C.POINTER(JobInfoC)]),
    "mycel_server_fetch": (_i, [_vp, C.c_int64, _vp, _vp]),
    "mycel_server_release": (_i, [_vp, C.c_int64]),
    "mycel_server_stats": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i), C.POINTER(_i)]),
    "mycel_simp_run": (_i, [_vp, _vp, C.POINTER(SimpOptionsC), _vp, _vp, _i, _vp, _i, _vp, SNAPSHOT_CB, _vp,
                            C.POINTER(SimpResultC), C.POINTER(SnapshotC), _i, _vp, _vp, _vp]),
    "mycel_pcg_solve_batch": (_i, [_vp, _vp, _vp, _i, _d, _i, _i, C.POINTER(PcgResultC)]),
    "mycel_pcg_solve_batch_d": (_i, [_vp, _vp, _vp, _i, _d, _i, _i, C.POINTER(PcgResultC)]),
    "mycel_pcg_solve_packed_d": (_i, [_vp, _vp, _vp, _i, _d, _i, C.POINTER(PcgResultC)]),
    "mycel_server_set_packing": (_i, [_vp, _i]),
    "mycel_server_pack_stats": (_i, [_vp, C.POINTER(PackStatsC)]),
    "mycel_simp_run_cases": (_i, [_vp, _vp, _i, C.POINTER(SimpOptionsC), _vp, _vp, _i, _vp, _i, _vp, SNAPSHOT_CB,
                                  _vp, C.POINTER(SimpResultC), C.POINTER(SnapshotC), _i, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load libmycel_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mycel_amd.build` "
                               "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status):
    if status != OK:
        msg = load().mycel_last_error()
        msg = msg.decode("utf-8", "replace") if msg else f"status {status}"
        raise _EXC.get(status, MycelError)(status, msg)


def ptr(a):
    """Pointer to the data of a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def device_ptr(t):
    """Raw device address of a torch tensor (for the *_d entry points)."""
    return C.c_void_p(t.data_ptr())
