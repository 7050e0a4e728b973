"""Strong branching and branch-and-bound over batched child LPs (Python
mirror of include/sovereign_bnb.h and the batch entry points of
include/sovereign_hip.h).

  BatchLpSolver            many small dense LPs in one persistent launch
  strong_branch(...)       pick_strong_branch (milp/src/branch_and_bound.cpp:726-793)
  BranchAndBoundSolver     BranchAndBoundSolver::solve core (branch_and_bound.cpp:906-1523)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as K
from .api import InteriorPointOptions, _require_device
from .problems import LpModel

CHILD_BATCHED = 0
CHILD_PER_LP = 1

# BranchRule (milp/include/sovereign/branch_and_bound.hpp:7-16) and the names
# SOVEREIGN_BRANCH_RULE takes (milp/src/milp_solver.cpp:13-20).
BRANCH_RULES = {"most_fractional": 0, "fractional": 0, "strong": 1, "reliability": 1, "pseudocost": 2,
                "pseudo": 2, "full_strong": 3}


class BatchLpSolver:
    """min c'x, A x = b, 0 <= x <= u for many LPs sharing uploaded matrices."""

    def __init__(self):
        self._lib = K.load()
        _require_device()
        self.handle = C.c_void_p()
        K.check(self._lib.sov_batch_create(C.byref(self.handle)), "sov_batch_create")
        self.shapes = []

    def add_matrix(self, A: np.ndarray) -> int:
        m, n = A.shape
        flat = np.asfortranarray(A, dtype=np.float64).reshape(-1, order="F")
        mid = C.c_int32()
        K.check(self._lib.sov_batch_add_matrix(self.handle, m, n, K.pf(flat), C.byref(mid)), "sov_batch_add_matrix")
        self.shapes.append((m, n))
        return mid.value

    def solve(self, lps, options: InteriorPointOptions | None = None, want_x=False):
        """lps: list of (matrix_id, b, c, u). Returns (results, x list or None)."""
        options = options or InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9)
        ids = np.array([lp[0] for lp in lps], dtype=np.int32)
        b_off, x_off = [], []
        bs, cs, us = [], [], []
        nb = nx = 0
        for mid, b, c, u in lps:
            m, n = self.shapes[mid]
            assert len(b) == m and len(c) == n and len(u) == n
            b_off.append(nb)
            x_off.append(nx)
            nb += m
            nx += n
            bs.append(b)
            cs.append(c)
            us.append(u)
        b_all = K.f64(np.concatenate(bs))
        c_all = K.f64(np.concatenate(cs))
        u_all = K.f64(np.concatenate(us))
        bo = np.array(b_off, dtype=np.int64)
        xo = np.array(x_off, dtype=np.int64)
        res = (K.BatchResult * len(lps))()
        x_all = np.zeros(nx) if want_x else None
        opt = options.to_c()
        K.check(self._lib.sov_batch_solve(self.handle, len(lps), K.pi(ids), K.pf(b_all),
                                          bo.ctypes.data_as(C.POINTER(C.c_int64)), K.pf(c_all), K.pf(u_all),
                                          xo.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(opt), res,
                                          K.pf(x_all)), "sov_batch_solve")
        out = [{f: getattr(r, f) for f, _ in K.BatchResult._fields_ if f != "pad"} for r in res]
        xs = None
        if want_x:
            xs = [x_all[x_off[k]:x_off[k] + self.shapes[ids[k]][1]] for k in range(len(lps))]
        return out, xs

    @property
    def last_kernel_ms(self) -> float:
        return float(self._lib.sov_batch_last_kernel_ms(self.handle))

    def close(self):
        if self.handle:
            self._lib.sov_batch_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def strong_branch(model: LpModel, is_int, x_node, parent_obj, lb=None, ub=None, max_candidates=64,
                  child_solver=CHILD_BATCHED, options: InteriorPointOptions | None = None):
    """Scores the strong-branching candidates of one node. Returns
    (list of candidate dicts, index of the best, batched kernel ms)."""
    lib = K.load()
    _require_device()
    options = options or InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9)
    st = model.fill_struct(K.LpModelC)
    lb = K.f64(model.lb if lb is None else lb)
    ub = K.f64(model.ub if ub is None else ub)
    isi = np.ascontiguousarray(is_int, dtype=np.int8)
    xv = K.f64(x_node)
    out = (K.SbChild * max_candidates)()
    ncand, best = C.c_int32(), C.c_int32()
    kms = C.c_double(0.0)
    opt = options.to_c()
    K.check(lib.sov_strong_branch(C.byref(st), K.pf(lb), K.pf(ub), isi.ctypes.data_as(C.POINTER(C.c_int8)),
                                  K.pf(xv), parent_obj, max_candidates, child_solver, C.byref(opt), out,
                                  C.byref(ncand), C.byref(best), C.byref(kms)), "sov_strong_branch")
    cands = [{f: getattr(out[k], f) for f, _ in K.SbChild._fields_ if f != "pad"} for k in range(ncand.value)]
    return cands, best.value, kms.value


class BranchAndBoundSolver:
    """Best-bound branch-and-bound with plunging and strong branching on the
    GPU (BranchAndBoundOptions, branch_and_bound.hpp:18-59).
    branch_rule: "strong" (reliability branching, the reference's default),
    "full_strong", "pseudocost" or "most_fractional".
    ranker (full_strong): optional callable(x, lb, ub) -> key per variable
    (higher first), e.g. the learned surrogate; None keeps the
    closest-to-1/2 order. Cuts: max_cut_rounds rounds at the root (0: none)
    of at most max_cuts_per_node validated cover and Gomory rounding cuts
    (and cMIR with cmir_at_root), inherited by the whole tree. warm_start:
    batched child LPs start from their parent node LP's first iterate with
    complementarity below warm_mu (the interior-point counterpart of the
    reference's dual simplex warm start from the parent basis)."""

    def __init__(self, strong_branch_candidates=64, child_solver=CHILD_BATCHED, max_nodes=100000,
                 mip_gap=1e-6, integer_tol=1e-6, time_limit_seconds=0.0, ranker=None,
                 options: InteriorPointOptions | None = None, branch_rule="strong", reliability_threshold=4,
                 max_strong_per_node=8, plunging=True, max_cut_rounds=5, max_cuts_per_node=20,
                 cmir_at_root=False, warm_start=True, warm_mu=0.1):
        self.lib = K.load()
        self.bo = K.BnbOptions()
        self.lib.sov_bnb_default_options(C.byref(self.bo))
        if branch_rule not in BRANCH_RULES:
            raise ValueError(f"unknown branch rule {branch_rule!r}")
        self.bo.branch_rule = BRANCH_RULES[branch_rule]
        self.bo.reliability_threshold = reliability_threshold
        self.bo.max_strong_per_node = max_strong_per_node
        self.bo.plunging = int(bool(plunging))
        self.bo.max_cut_rounds = max_cut_rounds
        self.bo.max_cuts_per_node = max_cuts_per_node
        self.bo.cmir_at_root = int(bool(cmir_at_root))
        self.bo.warm_start = int(bool(warm_start))
        self.bo.warm_mu = warm_mu
        self.bo.strong_branch_candidates = strong_branch_candidates
        self.bo.child_solver = child_solver
        self.bo.max_nodes = max_nodes
        self.bo.mip_gap = mip_gap
        self.bo.integer_tol = integer_tol
        self.bo.time_limit_seconds = time_limit_seconds
        self._cb = None
        if ranker is not None:
            def _order(n, x, lb, ub, key, user):
                xs = np.ctypeslib.as_array(x, shape=(n,))
                lbs = np.ctypeslib.as_array(lb, shape=(n,))
                ubs = np.ctypeslib.as_array(ub, shape=(n,))
                kv = np.ctypeslib.as_array(key, shape=(n,))
                kv[:] = ranker(xs.copy(), lbs.copy(), ubs.copy())
            self._cb = K.CANDIDATE_ORDER_FN(_order)
            self.bo.candidate_order = self._cb
        self.options = options or InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9)

    def solve(self, model: LpModel, is_int) -> dict:
        _require_device()
        st = model.fill_struct(K.LpModelC)
        isi = np.ascontiguousarray(is_int, dtype=np.int8)
        res = K.BnbResult()
        x = np.zeros(model.n_vars)
        opt = self.options.to_c()
        rc = self.lib.sov_bnb_solve(C.byref(st), isi.ctypes.data_as(C.POINTER(C.c_int8)), C.byref(self.bo),
                                    C.byref(opt), C.byref(res), K.pf(x))
        K.check(rc, "sov_bnb_solve")
        out = res.as_dict()
        out["x"] = x
        return out
