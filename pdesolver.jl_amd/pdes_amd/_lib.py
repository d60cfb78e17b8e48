"""ctypes loader of the gfx950 library libpdes_amd.so (built in-tree).

There is no CPU fallback: if the library is missing or no GPU is usable,
loading raises."""
from __future__ import annotations

import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PDES_AMD_LIB", os.path.join(_HERE, "libpdes_amd.so"))

# A batch runs each operator group from its own host thread on its own stream
# (the served workload runs a stream per lane as well).  HIP maps a process's
# streams onto GPU_MAX_HW_QUEUES hardware queues, 4 unless set, and streams
# that share a queue run one after another: with the 5 groups of a mixed 2D
# batch, 8 queues take the 96-instance batched solve from 87.6 s to 81 s, and
# 16 serve 32 mixed requests 3-4% faster than 8 (several lanes of groups).
# HIP reads the variable when it initialises, which is at the library's first
# call unless the process touched the GPU before; an explicit setting in the
# environment wins (a deployment that exports a lower value should raise it).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
_lib = None

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)

# name: (argtypes)
_SIGS = {
    "pdes_last_error": [],
    "pdes_version": [],
    "pdes_ctx_create": [C.POINTER(abi.MeshDesc), C.POINTER(abi.SBPDesc), C.POINTER(abi.PhysDesc), C.POINTER(_vp)],
    "pdes_ctx_destroy": [_vp],
    "pdes_ctx_numdof": [_vp, _i64p],
    "pdes_set_state": [_vp, _vp],
    "pdes_get_state": [_vp, _vp],
    "pdes_get_residual": [_vp, _vp],
    "pdes_state_ptr": [_vp, C.POINTER(_vp), C.POINTER(_vp)],
    "pdes_sync": [_vp],
    "pdes_eval_residual": [_vp, C.c_double],
    "pdes_residual_norm": [_vp, _dp],
    "pdes_jv_complex_step": [_vp, _vp, _vp],
    "pdes_jv_device": [_vp, _vp, _vp],
    "pdes_time_residual": [_vp, C.c_int, _dp],
    "pdes_copy_bandwidth": [C.c_int64, C.c_int, _dp],
    "pdes_estimate_bytes": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)],
    "pdes_device_memory": [C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "pdes_autotune": [_vp, C.c_int, _i32p],
    "pdes_set_launch_variant": [_vp, C.c_int, C.c_int],
    "pdes_rk4": [_vp, C.c_int, C.c_double, _dp],
    "pdes_rk4_noalloc": [_vp, _vp, C.c_int, C.c_double],
    "pdes_jacobian_setup": [_vp, C.c_int],
    "pdes_jacobian_nnzb": [_vp, _i64p, _i64p],
    "pdes_get_bsr_pattern": [_vp, _vp, _vp],
    "pdes_assemble_jacobian": [_vp],
    "pdes_assemble_jacobian_coloring": [_vp],
    "pdes_get_coloring": [_vp, _vp, _i32p, _vp],
    "pdes_get_bsr_values": [_vp, _vp],
    "pdes_bsr_matvec": [_vp, _vp, _vp],
    "pdes_pc_setup": [_vp, C.c_int],
    "pdes_pc_apply": [_vp, _vp, _vp],
    "pdes_gmres": [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                   _i32p, _dp],
    "pdes_newton_solve": [_vp, C.POINTER(abi.NewtonOpts), C.POINTER(abi.NewtonReport)],
    "pdes_eval_functional": [_vp, C.c_int, _vp, C.c_int, _dp],
    "pdes_functional_deriv": [_vp, C.c_int, _vp, C.c_int, _vp],
    "pdes_functional_deriv_m": [_vp, C.c_int, _vp, C.c_int, _vp],
    "pdes_eval_residual_revm": [_vp, _vp, _vp, _vp, _vp, _vp],
    "pdes_bsr_matvec_transpose": [_vp, _vp, _vp],
    "pdes_eval_homotopy": [_vp, _vp],
    "pdes_set_homotopy_boundary": [_vp, C.c_int],
    "pdes_homotopy_jv": [_vp, _vp, _vp],
    "pdes_assemble_homotopy_jacobian": [_vp, C.c_double],
    "pdes_homotopy_solve": [_vp, C.POINTER(abi.NewtonOpts), C.c_double, C.POINTER(abi.NewtonReport)],
    "pdes_adjoint_solve": [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                           C.c_int, _vp, _i32p, _dp],
    "pdes_batch_create": [C.c_int, _vp, _vp, _vp, C.POINTER(_vp)],
    "pdes_batch_destroy": [_vp],
    "pdes_batch_num_groups": [_vp, C.POINTER(C.c_int)],
    "pdes_batch_graph_stats": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "pdes_batch_set_state": [_vp, C.c_int, _vp],
    "pdes_batch_get_state": [_vp, C.c_int, _vp],
    "pdes_batch_newton_solve": [_vp, C.POINTER(abi.NewtonOpts), _vp],
    "pdes_batch_eval_residual": [_vp, C.c_int, _vp],
    "pdes_batch_advance": [_vp, C.POINTER(abi.NewtonOpts), C.c_int, _vp, _vp, C.POINTER(C.c_int)],
    "pdes_batch_export": [_vp, C.c_int, C.POINTER(abi.BatchInstState), _vp, _vp, _vp],
    "pdes_batch_import": [_vp, C.c_int, C.POINTER(abi.BatchInstState), _vp, _vp, _vp],
}


class PdesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "pdes_last_error" else C.c_int
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().pdes_last_error()
        raise PdesError(status, msg.decode() if msg else "")
    return status


def exported_symbols():
    return list(_SIGS.keys())
