"""Python mirror of the reference's device-path interfaces, over the C ABI.

Names and argument meaning follow the reference:
  device_info()                        cuda::device_info (gpu/include/sovereign/cuda_driver.hpp:14-27)
  dense_lu(a)                          cuda::dense_lu (cuda_driver.hpp:29-34)
  spmv_csc(...)                        cuda::spmv_csc (cuda_driver.hpp:36-40)
  SparseMatrix.multiply / _transpose   SparseMatrixCSC::multiply[_transpose] (sparse/src/sparse_matrix.cpp:91-104)
  NormalEquations(A).factor(d)/solve   NormalEquations (lp/src/interior_point.cpp:129-212)
  InteriorPointSolver(opts).solve(lp)  InteriorPointSolver (lp/include/sovereign/interior_point.hpp:7-36)
  gpu_operations / gpu_factorizations  gpu_spmv.hpp:89-92

Everything runs on the GPU through libsovereign_hip.so; there is no CPU
fallback, and a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi as K
from .problems import FormLp, LpModel

STATUS_NAMES = {0: "Optimal", 1: "Feasible", 2: "Infeasible", 3: "Unbounded", 4: "TimeLimit",
                5: "IterationLimit", 6: "NumericalError", 7: "Error", 8: "NotImplemented"}


def device_info() -> dict:
    lib = K.load()
    info = K.DevInfo()
    lib.sov_dev_probe(C.byref(info))
    return {"available": bool(info.available), "name": info.name.decode(), "arch": info.arch.decode(),
            "compute_major": info.compute_major, "compute_minor": info.compute_minor,
            "memory_bytes": int(info.memory_bytes), "driver_version": info.driver_version,
            "multiprocessors": info.multiprocessors, "reason": info.reason.decode()}


def gpu_available() -> bool:
    return device_info()["available"]


def _require_device():
    info = device_info()
    if not info["available"]:
        raise RuntimeError(f"no usable GPU: {info['reason']}")


def gpu_operations() -> int:
    return int(K.load().sov_gpu_operations())


def gpu_factorizations() -> int:
    return int(K.load().sov_gpu_factorizations())


def reset_gpu_operations() -> None:
    K.load().sov_reset_gpu_operations()


def dense_lu(a: np.ndarray):
    """LU with partial pivoting of a square matrix given as a numpy array
    (row i, column j = a[i, j]). Returns (status, lu, piv) with lu in the same
    orientation (unit-lower L below the diagonal, U on and above) and piv the
    DenseLU permutation: row i of P*A is row piv[i] of A. status 0 ok, 1
    singular (lu is then None)."""
    lib = K.load()
    _require_device()
    n = a.shape[0]
    buf = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    flat = buf.reshape(-1, order="F")
    piv = np.zeros(n, dtype=np.int32)
    rc = K.check(lib.sov_dense_lu(n, K.pf(flat), K.pi(piv)), "sov_dense_lu")
    if rc == K.SOV_SINGULAR:
        return 1, None, None
    return 0, flat.reshape((n, n), order="F"), piv


def spmv_csc(nrows, ncols, col_ptr, row_idx, values, x, transpose=False) -> np.ndarray:
    """y = A x, or A^T x, for a CSC matrix (one-shot; cached per thread)."""
    lib = K.load()
    _require_device()
    cp, ri, va = K.i32(col_ptr), K.i32(row_idx), K.f64(values)
    xv = K.f64(x)
    y = np.zeros(ncols if transpose else nrows)
    K.check(lib.sov_spmv_csc(nrows, ncols, K.pi(cp), K.pi(ri), K.pf(va), K.pf(xv), K.pf(y), 1 if transpose else 0),
            "sov_spmv_csc")
    return y


class SparseMatrix:
    """A CSC matrix resident in HBM (SparseMatrixCSC with its products)."""

    def __init__(self, nrows, ncols, col_ptr, row_idx, values):
        self._lib = K.load()
        _require_device()
        self.nrows, self.ncols = int(nrows), int(ncols)
        self._arrays = (K.i32(col_ptr), K.i32(row_idx), K.f64(values))
        self.handle = C.c_void_p()
        cp, ri, va = self._arrays
        K.check(self._lib.sov_mat_upload_csc(self.nrows, self.ncols, K.pi(cp), K.pi(ri), K.pf(va),
                                             C.byref(self.handle)), "sov_mat_upload_csc")

    def multiply(self, x) -> np.ndarray:
        x = K.f64(x)
        if x.shape[0] != self.ncols:
            raise ValueError("SparseMatrixCSC::multiply dimension mismatch")
        y = np.zeros(self.nrows)
        K.check(self._lib.sov_spmv(self.handle, 0, K.pf(x), K.pf(y)), "sov_spmv")
        return y

    def multiply_transpose(self, x) -> np.ndarray:
        x = K.f64(x)
        if x.shape[0] != self.nrows:
            raise ValueError("SparseMatrixCSC::multiply_transpose dimension mismatch")
        y = np.zeros(self.ncols)
        K.check(self._lib.sov_spmv(self.handle, 1, K.pf(x), K.pf(y)), "sov_spmv")
        return y

    def close(self):
        if self.handle:
            self._lib.sov_mat_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_NE_PATHS = {"auto": K.SOV_NE_AUTO, "dense": K.SOV_NE_DENSE, "sparse": K.SOV_NE_SPARSE}


def min_degree_order(n, col_ptr, row_idx, fill_limit=0):
    """AMDOrdering::compute (amd_ordering.cpp:12-130) on the host: returns
    (perm, iperm, predicted_nnz), or None when fill_limit > 0 is passed."""
    lib = K.load()
    cp, ri = K.i32(col_ptr), K.i32(row_idx)
    perm = np.zeros(n, dtype=np.int32)
    iperm = np.zeros(n, dtype=np.int32)
    pred = C.c_int64(0)
    rc = K.check(lib.sov_min_degree_order(n, K.pi(cp), K.pi(ri), int(fill_limit), K.pi(perm), K.pi(iperm),
                                          C.byref(pred)), "sov_min_degree_order")
    if rc != K.SOV_OK:
        return None
    return perm, iperm, int(pred.value)


class NormalEquations:
    """factor(d) / solve(rhs) of M = A diag(d) A^T + 1e-12 I on the GPU.
    precision: 'fp64' (direct FP64 Cholesky) or 'fp32' (FP32 Cholesky as the
    preconditioner of FP64 conjugate gradients). path: 'auto', 'dense' (blocked
    dense Cholesky) or 'sparse' (minimum-degree ordered multifrontal Cholesky,
    FP64 only), the reference's dense / sparse selector
    (interior_point.cpp:129-175)."""

    def __init__(self, A: SparseMatrix, path: str = "auto"):
        self._lib = K.load()
        self.A = A
        self.handle = C.c_void_p()
        K.check(self._lib.sov_ne_create_path(A.handle, _NE_PATHS[path], C.byref(self.handle)), "sov_ne_create_path")

    def info(self) -> dict:
        """Symbolic figures of the sparse factorization (sparse = 0 for the dense one)."""
        inf = K.NeInfo()
        K.check(self._lib.sov_ne_get_info(self.handle, C.byref(inf)), "sov_ne_get_info")
        return inf.as_dict()

    def factor(self, d, precision: str = "fp64") -> bool:
        d = K.f64(d)
        prec = K.SOV_PREC_FP32 if precision == "fp32" else K.SOV_PREC_FP64
        rc = K.check(self._lib.sov_ne_factor(self.handle, K.pf(d), prec), "sov_ne_factor")
        return rc == K.SOV_OK

    def solve(self, rhs) -> np.ndarray:
        x = K.f64(rhs).copy()
        rc = K.check(self._lib.sov_ne_solve(self.handle, K.pf(x)), "sov_ne_solve")
        if rc != K.SOV_OK:
            raise RuntimeError("solve without a successful factorization")
        return x

    @property
    def last_residual(self) -> float:
        return float(self._lib.sov_ne_last_residual(self.handle))

    @property
    def replaced_pivots(self) -> int:
        return int(self._lib.sov_ne_replaced_pivots(self.handle))

    def close(self):
        if self.handle:
            self._lib.sov_ne_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class InteriorPointOptions:
    """interior_point.hpp:7-24, plus the precision policy of the GPU path."""
    max_iterations: int = 250
    feasibility_tol: float = 1e-9
    optimality_tol: float = 1e-9
    fraction_to_boundary: float = 0.999
    enable_scaling: bool = True
    precision: str = "fp64"  # 'fp64' plain FP64 GPU path, 'mixed' FP32 factors + FP64 refinement
    # mixed: FP64 factorizations once mu < mixed_switch_mu, or earlier once an
    # iteration's FP32-preconditioned CG takes more than 8 iterations
    mixed_switch_mu: float = 1e-6
    refine_max: int = 30
    refine_tol: float = 1e-12
    normal_path: str = "auto"  # 'auto', 'dense' or 'sparse' (NormalEquations)

    def to_c(self) -> K.IpmOptions:
        o = K.IpmOptions()
        o.max_iterations = self.max_iterations
        o.feasibility_tol = self.feasibility_tol
        o.optimality_tol = self.optimality_tol
        o.fraction_to_boundary = self.fraction_to_boundary
        o.enable_scaling = 1 if self.enable_scaling else 0
        o.precision = K.SOV_PREC_FP32 if self.precision == "mixed" else K.SOV_PREC_FP64
        o.mixed_switch_mu = self.mixed_switch_mu
        o.refine_max = self.refine_max
        o.refine_tol = self.refine_tol
        o.normal_path = _NE_PATHS[self.normal_path]
        return o


class InteriorPointSolver:
    """Mehrotra IPM on the GPU. solve(LpModel) -> dict with status, objective,
    residuals, primal values (by variable index) and duals."""

    def __init__(self, options: InteriorPointOptions | None = None):
        self.options = options or InteriorPointOptions()

    def solve(self, model: LpModel) -> dict:
        lib = K.load()
        _require_device()
        st = model.fill_struct(K.LpModelC)
        res = K.IpmResult()
        primal = np.zeros(model.n_vars)
        row_dual = np.zeros(model.n_rows)
        var_dual = np.zeros(model.n_vars)
        opt = self.options.to_c()
        rc = lib.sov_lp_solve(C.byref(st), C.byref(opt), C.byref(res), K.pf(primal), K.pf(row_dual), K.pf(var_dual))
        out = res.as_dict()
        out["status_name"] = STATUS_NAMES.get(res.status, str(res.status))
        out["rc"] = rc
        if rc < 0:
            out["message"] = K.last_error()
        out.update(primal=primal, row_dual=row_dual, var_dual=var_dual)
        return out

    def solve_form(self, lp: FormLp) -> dict:
        """Mehrotra loop on an LP already in IPM form (no scaling, no shift)."""
        lib = K.load()
        _require_device()
        res = K.IpmResult()
        x = np.zeros(lp.n)
        y = np.zeros(lp.m)
        s = np.zeros(lp.n)
        cp, ri, va = K.i32(lp.col_ptr), K.i32(lp.row_idx), K.f64(lp.values)
        b, c = K.f64(lp.b), K.f64(lp.c)
        opt = self.options.to_c()
        rc = lib.sov_ipm_solve_form(lp.m, lp.n, K.pi(cp), K.pi(ri), K.pf(va), K.pf(b), K.pf(c), C.byref(opt),
                                    C.byref(res), K.pf(x), K.pf(y), K.pf(s))
        out = res.as_dict()
        out["status_name"] = STATUS_NAMES.get(res.status, str(res.status))
        out["rc"] = rc
        if rc < 0:
            out["message"] = K.last_error()
        out.update(x=x, y=y, s=s)
        return out


class ResidentIpm:
    """A constraint matrix kept in HBM with its normal-equation plan; solve()
    runs the Mehrotra loop for (b, c). Used to time solves with the inputs
    already resident."""

    def __init__(self, lp: FormLp, normal_path: str = "auto"):
        self._lib = K.load()
        _require_device()
        self.lp = lp
        self.A = SparseMatrix(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values)
        self.handle = C.c_void_p()
        K.check(self._lib.sov_ipm_create_path(self.A.handle, _NE_PATHS[normal_path], C.byref(self.handle)),
                "sov_ipm_create_path")
        self._b, self._c = K.f64(lp.b), K.f64(lp.c)

    def solve(self, options: InteriorPointOptions, want_solution: bool = False) -> dict:
        res = K.IpmResult()
        opt = options.to_c()
        x = np.zeros(self.lp.n) if want_solution else None
        y = np.zeros(self.lp.m) if want_solution else None
        s = np.zeros(self.lp.n) if want_solution else None
        rc = self._lib.sov_ipm_solve(self.handle, K.pf(self._b), K.pf(self._c), C.byref(opt), C.byref(res),
                                     K.pf(x), K.pf(y), K.pf(s))
        K.check(rc, "sov_ipm_solve")
        out = res.as_dict()
        out["status_name"] = STATUS_NAMES.get(res.status, str(res.status))
        if want_solution:
            out.update(x=x, y=y, s=s)
        return out

    def close(self):
        if self.handle:
            self._lib.sov_ipm_free(self.handle)
            self.handle = C.c_void_p()
        self.A.close()
