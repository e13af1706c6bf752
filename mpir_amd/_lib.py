"""ctypes bindings of libmpir_hip.so (declarations in include/mpir_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmpir_hip.so"

i32, i64, f64, u64 = C.c_int32, C.c_int64, C.c_double, C.c_uint64
P = C.c_void_p


class GmresStats(C.Structure):
    _fields_ = [
        ("n", i32), ("nnz", i64), ("nnz_l", i64), ("nsupernodes", i32), ("nlevels", i32),
        ("max_front", i32), ("factor_flops", f64), ("positive_pivots", i32), ("ir_iterations", i32),
        ("gmres_iterations", i32), ("precond_applications", i32), ("stop_reason", i32),
        ("final_rel_residual", f64), ("analyze_ms", f64), ("factor_ms", f64), ("solve_ms", f64),
        ("device_bytes", i64),
        ("ruiz_sweeps", i32),
        ("preconditioner", i32), ("ilu_sweeps", i32), ("ilu_nl_residual0", f64), ("ilu_nl_residual", f64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LuirStats(C.Structure):
    _fields_ = [
        ("n", i32), ("nnz", i64), ("lnz", i64), ("num_refinement_steps", i32), ("relative_residual", f64),
        ("time_sym", f64), ("time_num_fact", f64), ("time_solve", f64), ("time_ref", f64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerRequest(C.Structure):
    _fields_ = [
        ("status", i32), ("batch", i32), ("lane", i32), ("ir_max", i32), ("gmres_max", i32),
        ("rel_residual_max", f64), ("queue_ms", f64), ("latency_ms", f64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# name -> (restype, argtypes)
SIGNATURES = {
    "mpir_last_error": (C.c_char_p, []),
    "mpir_version": (C.c_char_p, []),
    "mpir_device_ok": (C.c_int, []),
    "mpir_qdldl_etree_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_etree_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_factor_f64_i32": (i32, [i32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_factor_f32_i32": (i32, [i32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_factor_f64_i64": (i64, [i64, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_factor_f32_i64": (i64, [i64, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_solve_f64_f64_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_f32_f64_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_f32_f32_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_f64_f64_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_solve_f32_f64_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_solve_f32_f32_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_factor_f16_i32": (i32, [i32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_factor_bf16_i32": (i32, [i32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_factor_f16_i64": (i64, [i64, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_factor_bf16_i64": (i64, [i64, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_qdldl_solve_f16_f32_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_f16_f64_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_f16_f32_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_solve_f16_f64_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_solve_bf16_f32_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_bf16_f64_i32": (i32, [i32, P, P, P, P, P]),
    "mpir_qdldl_solve_bf16_f32_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_qdldl_solve_bf16_f64_i64": (i64, [i64, P, P, P, P, P]),
    "mpir_luir_qdldl_f64_i32": (i32, [i32, P, P, P, P, P, P, i32]),
    "mpir_luir_qdldl_f32_i32": (i32, [i32, P, P, P, P, P, P, i32]),
    "mpir_luir_qdldl_f64_i64": (i64, [i64, P, P, P, P, P, P, i32]),
    "mpir_luir_qdldl_f32_i64": (i64, [i64, P, P, P, P, P, P, i32]),
    "mpir_gmres_ldl_create": (C.c_int, [C.c_int, C.c_int, C.c_int, P]),
    "mpir_gmres_ldl_destroy": (None, [P]),
    "mpir_gmres_ldl_set_max_ir": (C.c_int, [P, i64]),
    "mpir_gmres_ldl_set_max_gmres": (C.c_int, [P, i64]),
    "mpir_gmres_ldl_set_tol": (C.c_int, [P, f64]),
    "mpir_gmres_ldl_set_ordering": (C.c_int, [P, C.c_int]),
    "mpir_gmres_ldl_set_scaling": (C.c_int, [P, C.c_int]),
    "mpir_gmres_ldl_set_schedule": (C.c_int, [P, C.c_int]),
    "mpir_gmres_ldl_set_factor_schedule": (C.c_int, [P, C.c_int]),
    "mpir_gmres_ldl_set_amalgamation": (C.c_int, [P, C.c_int]),
    "mpir_gmres_ldl_set_panel_split": (C.c_int, [P, C.c_int]),
    "mpir_gmres_ldl_factor_schedule_info": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                                  C.POINTER(C.c_int64)]),
    "mpir_gmres_ldl_set_preconditioner": (C.c_int, [P, C.c_int, i32, i32]),
    "mpir_gmres_ldl_ilu_nnz": (C.c_int, [P, P]),
    "mpir_gmres_ldl_export_ilu": (C.c_int, [P, P, P, P, P, P]),
    "mpir_gmres_ldl_log": (i64, [P, C.c_char_p, i64]),
    "mpir_read_mtx_upper": (C.c_int, [C.c_char_p, P, P, P, P, P, P]),
    "mpir_gmres_ldl_analyze": (C.c_int, [P, i32, P, P]),
    "mpir_gmres_ldl_factor": (C.c_int, [P, P]),
    "mpir_gmres_ldl_factor_dev": (C.c_int, [P, P]),
    "mpir_gmres_ldl_compute": (C.c_int, [P, i32, P, P, P]),
    "mpir_gmres_ldl_solve": (C.c_int, [P, P, P, i32, i64]),
    "mpir_gmres_ldl_solve_dev": (C.c_int, [P, P, P, i32, i64]),
    "mpir_gmres_ldl_stats": (C.c_int, [P, P]),
    "mpir_gmres_ldl_perm": (C.c_int, [P, P]),
    "mpir_gmres_ldl_export_factor": (C.c_int, [P, P, P, P, P, P]),
    "mpir_gmres_ldl_precond_dev": (C.c_int, [P, P]),
    "mpir_gmres_ldl_precond_dev_noalloc": (C.c_int, [P, P]),
    "mpir_gmres_ldl_spmv_dev": (C.c_int, [P, P, P]),
    "mpir_gmres_ldl_precond_multi_dev": (C.c_int, [P, P, i32]),
    "mpir_gmres_ldl_stream": (P, [P]),
    "mpir_gmres_ldl_time": (C.c_int, [P, C.c_int, C.c_int, P]),
    "mpir_gmres_ldl_trace_factor": (C.c_int, [P, P, C.c_int64, C.POINTER(C.c_int64)]),
    "mpir_gmres_ldl_trace_precond": (C.c_int, [P, P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mpir_gmres_ldl_supernode_info": (C.c_int, [P, P, P, P, C.c_int32, C.POINTER(C.c_int32)]),
    "mpir_gmres_ldl_autotune": (C.c_int, [P, C.c_int, P, P]),
    "mpir_gmres_ldl_set_tuning": (C.c_int, [P, i32, i32, i32]),
    "mpir_batch_graph": (C.c_int, [P, C.c_int, P]),
    "mpir_spmv_upper_f64": (C.c_int, [i32, P, P, P, P, P]),
    "mpir_residual_upper_f64": (C.c_int, [i32, P, P, P, P, P, P]),
    "mpir_gen_kkt": (C.c_int, [C.c_int, C.c_int, f64, i64, C.c_int, C.c_int, C.c_int, f64, u64,
                               P, P, P, P, P, P, P, P]),
    "mpir_amd_order": (C.c_int, [i32, P, P, P]),
    "mpir_batch_analyze": (C.c_int, [P, i32, P, P, P, i32]),
    "mpir_batch_factor": (C.c_int, [P, P]),
    "mpir_batch_factor_dev": (C.c_int, [P, P]),
    "mpir_batch_solve": (C.c_int, [P, P, P, P]),
    "mpir_batch_solve_dev": (C.c_int, [P, P, P, P]),
    "mpir_batch_stats": (C.c_int, [P, P, P, P, P, P, P, P]),
    "mpir_pattern_cache_info": (C.c_int, [P, P, P]),
    "mpir_qdldl_solve_plan_builds": (C.c_int, [P]),
    "mpir_batch_plan_info": (C.c_int, [i32, P, P, P, C.c_int, P]),
    "mpir_server_create": (C.c_int, [C.c_int, C.c_int, i64, i32, f64, C.c_int, i32, i32, P]),
    "mpir_server_destroy": (None, [P]),
    "mpir_server_add_pattern": (C.c_int, [P, i32, P, P, P, i32]),
    "mpir_server_add_patterns": (C.c_int, [P, i32, P, P, P, P, P, i32]),
    "mpir_server_run": (C.c_int, [P, i64, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mpir_server_summary": (C.c_int, [P, P, P]),
    "mpir_server_plan_phases": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mpir_server_submit": (C.c_int, [P, i32, P, P, i32, P]),
    "mpir_server_poll": (C.c_int, [P, i64, f64, P, P, i32]),
    "mpir_server_solution": (C.c_int, [P, i64, P]),
    "mpir_server_request_data": (C.c_int, [P, i32, u64, i32, P, P, P]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m mpir_amd.build` "
                "(there is no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    msg = load().mpir_last_error()
    return msg.decode() if msg else ""
