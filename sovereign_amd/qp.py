"""The structured separable convex QP on the GPU, over the C ABI
(include/sovereign_qp.h).

    min 0.5 |x|^2 - target'x   s.t.  x_r - 2 x_{r+1} + x_{r+2} >= 0,  x free

Mirrors the reference's QP entry point for this shape:
  QpInteriorPointOptions / QpInteriorPointSolver(options).solve(model)
      solver/qp/include/sovereign/qp_interior_point.hpp:7-31
  detect_liswet_structured(model)      solver/qp/src/qp_interior_point.cpp:892-967
  the structured dispatch              qp_interior_point.cpp:3153-3161
  certify_liswet                       qp_interior_point.cpp:1410-1480

Any other convex QP goes to the generic interior point (sov_qp_ipm_solve):
solve_qp_ipm (qp_interior_point.cpp:2848-3116) on build_qp_form's form
(:118-247), from the default starting point and then Mehrotra's
(:3166-3180), its Newton system in Schur-complement form on the GPU.
Everything runs through libsovereign_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _capi as K
from .api import STATUS_NAMES, _require_device

INF = 1e30


@dataclass
class QpModel:
    """A continuous QP in the reference's model terms (types.hpp
    OptimizationModel, restricted to what the structured detector reads):
    minimize 0.5 x'Qx + linear'x over variables with bounds [lb, ub],
    rows sum_j coef_j x_j (sense) rhs. quadratic maps (i, j) -> Q_ij, rows
    are dicts {var index: coefficient}, senses '>=', '<=' or '='."""

    n_vars: int
    linear: np.ndarray
    quadratic: dict
    rows: list
    senses: list
    rhs: np.ndarray
    lb: np.ndarray = field(default=None)
    ub: np.ndarray = field(default=None)
    minimize: bool = True
    constant: float = 0.0

    def __post_init__(self):
        self.linear = np.asarray(self.linear, dtype=np.float64)
        self.rhs = np.asarray(self.rhs, dtype=np.float64)
        if self.lb is None:
            self.lb = np.full(self.n_vars, -INF)
        if self.ub is None:
            self.ub = np.full(self.n_vars, INF)


def liswet_model(target) -> QpModel:
    """The structured QP for `target` as a model: Q = I, linear = -target,
    rows x_r - 2 x_{r+1} + x_{r+2} >= 0, free variables."""
    t = np.asarray(target, dtype=np.float64)
    n = len(t)
    return QpModel(n_vars=n, linear=-t, quadratic={(i, i): 1.0 for i in range(n)},
                   rows=[{r: 1.0, r + 1: -2.0, r + 2: 1.0} for r in range(n - 2)],
                   senses=[">="] * (n - 2), rhs=np.zeros(n - 2))


def detect_liswet(model: QpModel):
    """The target vector when the model has the structured shape, else None
    (the acceptance rules of detect_liswet_structured, :892-967: a
    minimization, n >= 3 free variables, n - 2 rows, Q = I exactly (1e-12),
    row r = x_r - 2 x_{r+1} + x_{r+2} >= 0 with rhs 0 (1e-12))."""
    n = model.n_vars
    if not model.minimize or n < 3 or len(model.rows) != n - 2:
        return None
    if np.any(model.lb > -1e29) or np.any(np.isfinite(model.ub) & (model.ub < 1e29)):
        return None
    diag = np.zeros(n, dtype=bool)
    for (i, j), v in model.quadratic.items():
        if i != j or abs(v - 1.0) > 1e-12:
            return None
        diag[i] = True
    if not diag.all():
        return None
    expected = (1.0, -2.0, 1.0)
    for r, (row, sense) in enumerate(zip(model.rows, model.senses)):
        if sense != ">=" or abs(model.rhs[r]) > 1e-12 or len(row) != 3:
            return None
        for k in range(3):
            if r + k not in row or abs(row[r + k] - expected[k]) > 1e-12:
                return None
    return -model.linear.copy()


@dataclass
class QpInteriorPointOptions:
    """qp_interior_point.hpp:7-18 (the fields the structured path reads),
    plus the precision policy of the GPU path."""
    max_iterations: int = 200
    precision: str = "mixed"  # 'mixed': FP32 factor + FP64 refinement; 'fp64': FP64 throughout
    tolerance: float = 0.0    # 0: the structured tolerances (1e-10; 1e-12 when rows <= 50)
    polish: bool = True

    def to_c(self) -> K.QpOptions:
        o = K.QpOptions()
        o.max_iterations = self.max_iterations
        o.precision = K.SOV_PREC_FP32 if self.precision == "mixed" else K.SOV_PREC_FP64
        o.tolerance = self.tolerance
        o.polish = 1 if self.polish else 0
        return o


def solve_liswet(target, options: QpInteriorPointOptions | None = None) -> dict:
    """Solves the structured QP for `target`; dict with status, objective,
    x, lam (multipliers), the certificate and the interior-point record."""
    lib = K.load()
    _require_device()
    t = K.f64(target)
    n = len(t)
    if n < 3:
        raise ValueError("the structured QP needs at least 3 variables")
    x = np.zeros(n)
    lam = np.zeros(n - 2)
    res = K.QpResult()
    opt = (options or QpInteriorPointOptions()).to_c()
    rc = lib.sov_qp_liswet_solve(n, K.pf(t), C.byref(opt), C.byref(res), K.pf(x), K.pf(lam))
    K.check(rc, "sov_qp_liswet_solve")
    out = res.as_dict()
    out["status_name"] = STATUS_NAMES.get(res.status, str(res.status))
    out.update(x=x, lam=lam)
    return out


def polish_liswet(target, s, z) -> dict:
    """The host polish alone from an interior point (s, z) of the QP for
    `target` (sov_qp_liswet_polish; needs no device): dict with status (0
    certified), objective, x, lam and the certificate."""
    lib = K.load()
    t = K.f64(target)
    n = len(t)
    if n < 3 or len(s) != n - 2 or len(z) != n - 2:
        raise ValueError("target needs n >= 3 entries and s, z n - 2 each")
    x = np.zeros(n)
    lam = np.zeros(n - 2)
    res = K.QpResult()
    K.check(lib.sov_qp_liswet_polish(n, K.pf(t), K.pf(K.f64(s)), K.pf(K.f64(z)), C.byref(res), K.pf(x), K.pf(lam)),
            "sov_qp_liswet_polish")
    out = res.as_dict()
    out.update(x=x, lam=lam)
    return out


@dataclass
class QpIpmOptions:
    """QpInteriorPointOptions (qp_interior_point.hpp:7-18) of the generic path."""
    max_iterations: int = 100
    feasibility_tol: float = 1e-8
    optimality_tol: float = 1e-8
    fraction_to_boundary: float = 0.999
    iterative_refinement: bool = True
    common_step: bool = True
    retry_mehrotra_start: bool = True

    def to_c(self) -> K.QpIpmOptions:
        o = K.QpIpmOptions()
        o.max_iterations = self.max_iterations
        o.feasibility_tol = self.feasibility_tol
        o.optimality_tol = self.optimality_tol
        o.fraction_to_boundary = self.fraction_to_boundary
        o.iterative_refinement = int(self.iterative_refinement)
        o.common_step = int(self.common_step)
        o.retry_mehrotra_start = int(self.retry_mehrotra_start)
        return o


def solve_generic(model: QpModel, options: QpIpmOptions | None = None) -> dict:
    """The generic QP interior point on the GPU. model.quadratic holds terms
    q_ij of 1/2 sum q_ij x_i x_j as the reference stores them: a diagonal
    term is Q_ii, an off-diagonal one adds q_ij / 2 to Q_ij and Q_ji, so the
    entries of a symmetric Q, both (i, j) and (j, i), give that Q."""
    lib = K.load()
    _require_device()
    n, m = model.n_vars, len(model.rows)
    cols = [[] for _ in range(n)]
    for r, row in enumerate(model.rows):
        for j, v in row.items():
            cols[int(j)].append((r, float(v)))
    ptr = np.zeros(n + 1, dtype=np.int32)
    ri, vals = [], []
    for j in range(n):
        for r, v in sorted(cols[j]):
            ri.append(r)
            vals.append(v)
        ptr[j + 1] = len(ri)
    ri = np.array(ri, dtype=np.int32)
    vals = np.array(vals, dtype=np.float64)
    qi = np.array([i for (i, j) in model.quadratic], dtype=np.int32)
    qj = np.array([j for (i, j) in model.quadratic], dtype=np.int32)
    qv = np.array(list(model.quadratic.values()), dtype=np.float64)
    sense = "".join({">=": "G", "<=": "L", "=": "E"}[s] for s in model.senses).encode()
    lb, ub, c, rhs = K.f64(model.lb), K.f64(model.ub), K.f64(model.linear), K.f64(model.rhs)
    st = K.QpModelC()
    st.n_vars, st.n_rows = n, m
    st.col_ptr, st.row_idx, st.values = K.pi(ptr), K.pi(ri), K.pf(vals)
    st.obj, st.obj_constant, st.maximize = K.pf(c), float(getattr(model, "constant", 0.0)), 0 if model.minimize else 1
    st.lb, st.ub, st.row_sense, st.rhs = K.pf(lb), K.pf(ub), sense, K.pf(rhs)
    st.q_nnz, st.q_row, st.q_col, st.q_val = len(qv), K.pi(qi), K.pi(qj), K.pf(qv)
    res = K.QpIpmResult()
    opt = (options or QpIpmOptions()).to_c()
    x, y, z = np.zeros(n), np.zeros(max(m + n, 1)), np.zeros(n)
    K.check(lib.sov_qp_ipm_solve(C.byref(st), C.byref(opt), C.byref(res), K.pf(x), K.pf(y), K.pf(z)),
            "sov_qp_ipm_solve")
    out = res.as_dict()
    out["status_name"] = STATUS_NAMES.get(res.status, str(res.status))
    out.update(x=x, row_dual=y[:res.m_form], var_dual=z)
    return out


class QpInteriorPointSolver:
    """QpInteriorPointSolver (qp_interior_point.hpp:24-31): the structured
    shape goes to the LISWET kernel (status 0 Optimal with a certificate, 6
    NumericalError when the certificate fails), any other convex QP to the
    generic interior point (0 Optimal, 7 Error)."""

    def __init__(self, options: QpInteriorPointOptions | None = None, generic: QpIpmOptions | None = None):
        self.options = options or QpInteriorPointOptions()
        self.generic = generic or QpIpmOptions()

    def solve(self, model: QpModel) -> dict:
        target = detect_liswet(model)
        if target is None:
            return solve_generic(model, generic_options_from_env(self.generic))
        return solve_liswet(target, self.options)


def generic_options_from_env(base: QpIpmOptions) -> QpIpmOptions:
    """The environment switches QpInteriorPointSolver::solve reads
    (qp_interior_point.cpp:3139-3151): SOVEREIGN_QP_MAX_ITERATIONS (at least
    1; malformed text keeps the default) and SOVEREIGN_QP_DISABLE, a comma
    list that may name refinement, common_step and retry."""
    o = QpIpmOptions(**vars(base))
    text = os.environ.get("SOVEREIGN_QP_MAX_ITERATIONS")
    if text is not None:
        try:
            o.max_iterations = max(1, int(text))
        except ValueError:
            pass
    off = "," + os.environ.get("SOVEREIGN_QP_DISABLE", "") + ","
    if ",refinement," in off:
        o.iterative_refinement = False
    if ",common_step," in off:
        o.common_step = False
    if ",retry," in off:
        o.retry_mehrotra_start = False
    return o


def liswet_target(n: int = 20002, seed: int = 20260501, noise_sd: float = 0.01) -> np.ndarray:
    """Synthetic target of the n-variable workload: t_i = i / (n - 1),
    target = sin(3t) + t^2 + e, e ~ Normal(0, noise_sd) (noise_sd is a
    standard deviation)."""
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64) / (n - 1)
    return np.sin(3.0 * t) + t * t + rng.normal(0.0, noise_sd, n)
