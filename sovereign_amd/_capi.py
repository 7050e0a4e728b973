"""ctypes binding of libsovereign_hip.so (include/sovereign_hip.h).

The library is built in-tree by `make -C sovereign_amd` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there is
no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsovereign_hip.so")

SOV_OK = 0
SOV_SINGULAR = 1
SOV_ERR_ARG = -1
SOV_PREC_FP64 = 0
SOV_PREC_FP32 = 1
SOV_NE_AUTO = 0
SOV_NE_DENSE = 1
SOV_NE_SPARSE = 2

# Declared C entry points (name -> (restype, argtypes)); the CPU test suite
# checks that the built library exports every one of them.
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


class DevInfo(C.Structure):
    _fields_ = [
        ("available", C.c_int32),
        ("name", C.c_char * 256),
        ("arch", C.c_char * 64),
        ("compute_major", C.c_int32),
        ("compute_minor", C.c_int32),
        ("memory_bytes", C.c_uint64),
        ("driver_version", C.c_int32),
        ("multiprocessors", C.c_int32),
        ("reason", C.c_char * 256),
    ]


class IpmOptions(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("feasibility_tol", C.c_double),
        ("optimality_tol", C.c_double),
        ("fraction_to_boundary", C.c_double),
        ("enable_scaling", C.c_int32),
        ("precision", C.c_int32),
        ("mixed_switch_mu", C.c_double),
        ("refine_max", C.c_int32),
        ("refine_tol", C.c_double),
        ("normal_path", C.c_int32),
    ]


class IpmResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("objective", C.c_double),
        ("duality_gap", C.c_double),
        ("primal_residual", C.c_double),
        ("dual_residual", C.c_double),
        ("factorizations", C.c_int32),
        ("fp32_factorizations", C.c_int32),
        ("refine_iterations", C.c_int32),
        ("max_solve_residual", C.c_double),
        ("seconds", C.c_double),
        ("factor_seconds", C.c_double),
        ("solve_seconds", C.c_double),
        ("sparse_factor", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class NeInfo(C.Structure):
    _fields_ = [
        ("sparse", C.c_int32),
        ("supernodes", C.c_int32),
        ("levels", C.c_int32),
        ("max_front", C.c_int32),
        ("nnz_l", C.c_int64),
        ("factor_flops", C.c_double),
        ("front_doubles", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("pcg_iterations", C.c_int32),
        ("fp64_from", C.c_int32),
        ("objective", C.c_double),
        ("duality_gap", C.c_double),
        ("primal_residual", C.c_double),
        ("dual_residual", C.c_double),
    ]


class SbChild(C.Structure):
    _fields_ = [
        ("var", C.c_int32), ("status_down", C.c_int32), ("status_up", C.c_int32),
        ("iters_down", C.c_int32), ("iters_up", C.c_int32), ("pad", C.c_int32),
        ("obj_down", C.c_double), ("obj_up", C.c_double), ("score", C.c_double),
    ]


CANDIDATE_ORDER_FN = C.CFUNCTYPE(None, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class BnbOptions(C.Structure):
    _fields_ = [
        ("max_nodes", C.c_int64), ("strong_branch_candidates", C.c_int32), ("child_solver", C.c_int32),
        ("mip_gap", C.c_double), ("integer_tol", C.c_double), ("time_limit_seconds", C.c_double),
        ("candidate_order", CANDIDATE_ORDER_FN), ("candidate_order_user", C.c_void_p),
        ("branch_rule", C.c_int32), ("reliability_threshold", C.c_int32), ("max_strong_per_node", C.c_int32),
        ("plunging", C.c_int32), ("max_cut_rounds", C.c_int32), ("max_cuts_per_node", C.c_int32),
        ("cmir_at_root", C.c_int32), ("warm_start", C.c_int32), ("warm_mu", C.c_double),
    ]


class BnbResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("pad", C.c_int32), ("objective", C.c_double), ("root_bound", C.c_double),
        ("nodes", C.c_int64), ("lp_solves", C.c_int64), ("seconds", C.c_double),
        ("strong_branch_seconds", C.c_double), ("kernel_ms", C.c_double), ("node_lp_solves", C.c_int64),
        ("strong_lp_solves", C.c_int64), ("strong_lp_used", C.c_int64), ("node_lp_seconds", C.c_double),
        ("max_depth", C.c_int32), ("node_lp_failures", C.c_int32), ("cuts_added", C.c_int32),
        ("cuts_cover", C.c_int32), ("cuts_gomory", C.c_int32), ("cuts_cmir", C.c_int32), ("cut_rounds", C.c_int32),
        ("cut_rejections", C.c_int32), ("root_bound_before_cuts", C.c_double), ("root_rows", C.c_int32),
        ("pad2", C.c_int32), ("lp_iterations", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("pad")}


class LpModelC(C.Structure):
    _fields_ = [
        ("n_vars", C.c_int32),
        ("n_rows", C.c_int32),
        ("col_ptr", _i32p),
        ("row_idx", _i32p),
        ("values", _f64p),
        ("obj", _f64p),
        ("obj_constant", C.c_double),
        ("maximize", C.c_int32),
        ("lb", _f64p),
        ("ub", _f64p),
        ("row_sense", C.c_char_p),
        ("rhs", _f64p),
    ]


class MlpShape(C.Structure):
    _fields_ = [("in_", C.c_int32), ("hidden1", C.c_int32), ("hidden2", C.c_int32), ("pad", C.c_int32)]


class QpOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("precision", C.c_int32), ("tolerance", C.c_double),
                ("polish", C.c_int32), ("pad", C.c_int32)]


class QpResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("ipm_status", C.c_int32), ("iterations", C.c_int32), ("polish_fits", C.c_int32),
        ("fp32_factorizations", C.c_int32), ("fp64_factorizations", C.c_int32), ("refinements", C.c_int32),
        ("pad", C.c_int32), ("objective", C.c_double), ("primal", C.c_double), ("stationarity", C.c_double),
        ("complementarity", C.c_double), ("min_lambda", C.c_double), ("consistency", C.c_double),
        ("ipm_primal", C.c_double), ("ipm_dual", C.c_double), ("ipm_complementarity", C.c_double),
        ("factor_ms", C.c_double), ("solve_ms", C.c_double), ("kernel_ms", C.c_double), ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class CutList(C.Structure):
    _fields_ = [
        ("max_cuts", C.c_int32), ("n_cuts", C.c_int32), ("max_nnz", C.c_int64), ("nnz", C.c_int64),
        ("row_ptr", _i32p), ("col", _i32p), ("val", _f64p), ("sense", C.c_char_p), ("rhs", _f64p),
        ("family", _i32p), ("efficacy", _f64p),
    ]


class QpModelC(C.Structure):
    _fields_ = [
        ("n_vars", C.c_int32), ("n_rows", C.c_int32), ("col_ptr", _i32p), ("row_idx", _i32p), ("values", _f64p),
        ("obj", _f64p), ("obj_constant", C.c_double), ("maximize", C.c_int32), ("pad", C.c_int32),
        ("lb", _f64p), ("ub", _f64p), ("row_sense", C.c_char_p), ("rhs", _f64p), ("q_nnz", C.c_int64),
        ("q_row", _i32p), ("q_col", _i32p), ("q_val", _f64p),
    ]


class QpIpmOptions(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32), ("iterative_refinement", C.c_int32), ("feasibility_tol", C.c_double),
        ("optimality_tol", C.c_double), ("fraction_to_boundary", C.c_double), ("common_step", C.c_int32),
        ("retry_mehrotra_start", C.c_int32),
    ]


class QpIpmResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iterations", C.c_int32), ("objective", C.c_double), ("duality_gap", C.c_double),
        ("primal_residual", C.c_double), ("dual_residual", C.c_double), ("factorizations", C.c_int32),
        ("refinement_steps", C.c_int32), ("mehrotra_start", C.c_int32), ("n_form", C.c_int32),
        ("m_form", C.c_int32), ("pad", C.c_int32), ("regularization", C.c_double), ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


_f32p = C.POINTER(C.c_float)

SIGNATURES = {
    "sov_separate_cuts": (C.c_int, [C.POINTER(LpModelC), C.POINTER(C.c_int8), _f64p, C.c_int32, C.c_double, C.c_int32,
                                    C.c_int32, C.POINTER(CutList)]),
    "sov_cut_validity": (C.c_int, [C.POINTER(LpModelC), C.c_int32, _i32p, _f64p, C.c_char, C.c_double, C.c_int32, _f64p,
                                   C.c_double, C.POINTER(C.c_int32)]),
    "sov_dev_probe": (C.c_int, [C.POINTER(DevInfo)]),
    "sov_gpu_operations": (C.c_uint64, []),
    "sov_gpu_factorizations": (C.c_uint64, []),
    "sov_reset_gpu_operations": (None, []),
    "sov_last_error": (C.c_char_p, []),
    "sov_sync": (C.c_int, []),
    "sov_stream": (_vp, []),
    "sov_set_kernel_timing": (None, [C.c_int32]),
    "sov_kernel_times": (C.c_int, [_f64p]),
    "sov_gemm_shapes": (C.c_int, [_i32p, _f64p]),
    "sov_dense_lu": (C.c_int, [C.c_int64, _f64p, _i32p]),
    "sov_dense_lu_dev_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p]),
    "sov_mat_upload_csc": (C.c_int, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.POINTER(_vp)]),
    "sov_mat_free": (C.c_int, [_vp]),
    "sov_spmv": (C.c_int, [_vp, C.c_int32, _f64p, _f64p]),
    "sov_spmv_dev": (C.c_int, [_vp, C.c_int32, _vp, _vp]),
    "sov_spmv_csc": (C.c_int, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int32]),
    "sov_ne_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sov_ne_create_path": (C.c_int, [_vp, C.c_int32, C.POINTER(_vp)]),
    "sov_ne_get_info": (C.c_int, [_vp, C.POINTER(NeInfo)]),
    "sov_min_degree_order": (C.c_int, [C.c_int32, _i32p, _i32p, C.c_int64, _i32p, _i32p,
                                       C.POINTER(C.c_int64)]),
    "sov_ne_factor": (C.c_int, [_vp, _f64p, C.c_int32]),
    "sov_ne_solve": (C.c_int, [_vp, _f64p]),
    "sov_ne_last_residual": (C.c_double, [_vp]),
    "sov_ne_replaced_pivots": (C.c_int32, [_vp]),
    "sov_ne_free": (C.c_int, [_vp]),
    "sov_ipm_default_options": (None, [C.POINTER(IpmOptions)]),
    "sov_ipm_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sov_ipm_create_path": (C.c_int, [_vp, C.c_int32, C.POINTER(_vp)]),
    "sov_ipm_solve": (C.c_int, [_vp, _f64p, _f64p, C.POINTER(IpmOptions), C.POINTER(IpmResult),
                                _f64p, _f64p, _f64p]),
    "sov_ipm_free": (C.c_int, [_vp]),
    "sov_ipm_solve_form": (C.c_int, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p,
                                     C.POINTER(IpmOptions), C.POINTER(IpmResult), _f64p, _f64p, _f64p]),
    "sov_lp_solve": (C.c_int, [C.POINTER(LpModelC), C.POINTER(IpmOptions), C.POINTER(IpmResult),
                               _f64p, _f64p, _f64p]),
    "sov_batch_create": (C.c_int, [C.POINTER(_vp)]),
    "sov_batch_add_matrix": (C.c_int, [_vp, C.c_int32, C.c_int32, _f64p, _i32p]),
    "sov_batch_solve_warm": (C.c_int, [_vp, C.c_int32, _i32p, _f64p, C.POINTER(C.c_int64), _f64p, _f64p,
                                       C.POINTER(C.c_int64), C.POINTER(IpmOptions), C.POINTER(BatchResult), _f64p, _f64p,
                                       C.POINTER(C.c_int64), _f64p, C.POINTER(C.c_int64), C.c_double]),
    "sov_batch_solve": (C.c_int, [_vp, C.c_int32, _i32p, _f64p, C.POINTER(C.c_int64), _f64p, _f64p,
                                  C.POINTER(C.c_int64), C.POINTER(IpmOptions), C.POINTER(BatchResult), _f64p]),
    "sov_batch_last_kernel_ms": (C.c_double, [_vp]),
    "sov_batch_free": (C.c_int, [_vp]),
    "sov_strong_branch": (C.c_int, [C.POINTER(LpModelC), _f64p, _f64p, C.POINTER(C.c_int8), _f64p, C.c_double,
                                    C.c_int32, C.c_int32, C.POINTER(IpmOptions), C.POINTER(SbChild), _i32p, _i32p,
                                    _f64p]),
    "sov_bnb_default_options": (None, [C.POINTER(BnbOptions)]),
    "sov_bnb_solve": (C.c_int, [C.POINTER(LpModelC), C.POINTER(C.c_int8), C.POINTER(BnbOptions),
                                C.POINTER(IpmOptions), C.POINTER(BnbResult), _f64p]),
    "sov_mlp_create": (C.c_int, [C.POINTER(MlpShape), C.c_uint64, C.POINTER(_vp)]),
    "sov_mlp_param_count": (C.c_int64, [_vp]),
    "sov_mlp_get_params": (C.c_int, [_vp, _f32p]),
    "sov_mlp_set_params": (C.c_int, [_vp, _f32p]),
    "sov_mlp_forward": (C.c_int, [_vp, _f32p, C.c_int32, _f32p]),
    "sov_mlp_gradients": (C.c_int, [_vp, _f32p, _f32p, C.c_int32, _f32p, _f32p]),
    "sov_mlp_train": (C.c_int, [_vp, _f32p, _f32p, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                _f32p]),
    "sov_mlp_free": (C.c_int, [_vp]),
    "sov_qp_default_options": (None, [C.POINTER(QpOptions)]),
    "sov_qp_liswet_solve": (C.c_int, [C.c_int32, _f64p, C.POINTER(QpOptions), C.POINTER(QpResult), _f64p, _f64p]),
    "sov_qp_liswet_polish": (C.c_int, [C.c_int32, _f64p, _f64p, _f64p, C.POINTER(QpResult), _f64p, _f64p]),
    "sov_qp_ipm_default_options": (None, [C.POINTER(QpIpmOptions)]),
    "sov_qp_ipm_solve": (C.c_int, [C.POINTER(QpModelC), C.POINTER(QpIpmOptions), C.POINTER(QpIpmResult), _f64p,
                                   _f64p, _f64p]),
}

_lib = None


def load() -> C.CDLL:
    """Loads the in-tree library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C sovereign_amd` "
            "or __graft_entry__.build(); there is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def pf(a: np.ndarray):
    return a.ctypes.data_as(_f64p) if a is not None else None


def pi(a: np.ndarray):
    return a.ctypes.data_as(_i32p) if a is not None else None


def last_error() -> str:
    return load().sov_last_error().decode()


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
    return rc
