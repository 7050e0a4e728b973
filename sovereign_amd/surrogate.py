"""Learned branching-score surrogate (Python side of include/sovereign_surrogate.h).

Candidates of a branch-and-bound node are described by 32 features computed
from the model and the node's LP point (fractionality, objective
coefficient, pseudocost-style degradation estimates, column size and
coefficient statistics, activity statistics of the rows the column is in,
node-level counts). The MLP 32 -> 256 -> 256 -> 1 (BF16 MFMA kernels,
FP32 accumulation, Adam) learns the strong-branching score the solver
itself computes (pick_strong_branch, solver/milp/src/branch_and_bound.cpp:
726-793: min(down, up) + 1e-4 max(down, up) of the objective degradations,
infeasible child 1e6), normalised per node by the node's best score. Used as
the ranker of BranchAndBoundSolver, it decides which k candidates get real
(batched) strong branching; the strong-branching scores still pick the
branching variable among them, and full strong branching remains the
reference for its ranking error.

  Mlp                      the GPU network (forward, gradients, train)
  candidate_features(...)  the 32 features of candidates at a node
  node_dataset(...)        (features, targets) of one node from strong branching
  generate_dataset(...)    node datasets over seeded knapsack instances
  SurrogateRanker          ranker callable for BranchAndBoundSolver
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _capi as K
from .api import InteriorPointOptions, InteriorPointSolver, _require_device
from .problems import LpModel, multi_knapsack

N_FEATURES = 32
SHAPE = (N_FEATURES, 256, 256)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_mkp.npz")
_f32p = C.POINTER(C.c_float)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf32(a: np.ndarray):
    return a.ctypes.data_as(_f32p)


class Mlp:
    """MLP in -> h1 -> h2 -> 1 on the GPU."""

    def __init__(self, shape=SHAPE, seed: int = 1):
        self._lib = K.load()
        _require_device()
        self.shape = tuple(shape)
        sh = K.MlpShape(shape[0], shape[1], shape[2], 0)
        self.handle = C.c_void_p()
        K.check(self._lib.sov_mlp_create(C.byref(sh), seed, C.byref(self.handle)), "sov_mlp_create")
        self.n_params = int(self._lib.sov_mlp_param_count(self.handle))

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        K.check(self._lib.sov_mlp_get_params(self.handle, _pf32(p)), "sov_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        assert len(p) == self.n_params
        K.check(self._lib.sov_mlp_set_params(self.handle, _pf32(p)), "sov_mlp_set_params")

    def forward(self, X) -> np.ndarray:
        X = _f32(X).reshape(-1, self.shape[0])
        y = np.zeros(len(X), dtype=np.float32)
        K.check(self._lib.sov_mlp_forward(self.handle, _pf32(X), len(X), _pf32(y)), "sov_mlp_forward")
        return y

    def gradients(self, X, t):
        X = _f32(X).reshape(-1, self.shape[0])
        t = _f32(t)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float(0.0)
        K.check(self._lib.sov_mlp_gradients(self.handle, _pf32(X), _pf32(t), len(X), _pf32(g), C.byref(loss)),
                "sov_mlp_gradients")
        return float(loss.value), g

    def train(self, X, t, epochs=40, batch=256, lr=1e-3, seed=7) -> np.ndarray:
        X = _f32(X).reshape(-1, self.shape[0])
        t = _f32(t)
        hist = np.zeros(epochs, dtype=np.float32)
        K.check(self._lib.sov_mlp_train(self.handle, _pf32(X), _pf32(t), len(X), batch, epochs, lr, seed,
                                        _pf32(hist)), "sov_mlp_train")
        return hist

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, shape=np.array(self.shape), **meta)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        net = cls(tuple(int(v) for v in z["shape"]))
        net.params = z["params"]
        return net

    def close(self):
        if getattr(self, "handle", None):
            self._lib.sov_mlp_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ModelStats:
    """Per-model quantities the features reuse at every node."""

    def __init__(self, model: LpModel):
        self.model = model
        A = model.dense()
        self.A = A
        self.absA = np.abs(A)
        n = model.n_vars
        self.nnz = (A != 0).sum(axis=0)
        col_sum = self.absA.sum(axis=0)
        self.col_mean = col_sum / np.maximum(self.nnz, 1)
        mean_all = max(self.absA[A != 0].mean() if self.nnz.sum() else 1.0, 1e-12)
        self.mean_all = mean_all
        self.col_max = self.absA.max(axis=0) / mean_all
        self.col_min = np.where(A != 0, self.absA, np.inf).min(axis=0)
        self.col_min = np.where(np.isfinite(self.col_min), self.col_min, 0.0) / mean_all
        self.col_std = self.absA.std(axis=0) / mean_all
        c = model.obj if model.maximize else -model.obj  # "higher is better" direction
        self.c = c
        self.cmax = max(np.abs(model.obj).max(), 1e-12)
        ratio = c / np.maximum(self.col_mean, 1e-12)
        self.eff = ratio / max(np.median(np.abs(ratio)), 1e-12)
        self.eff_rank = np.argsort(np.argsort(ratio)) / max(n - 1, 1)
        rhs = np.where(np.abs(model.rhs) > 1e-12, np.abs(model.rhs), 1.0)
        self.rhs_abs = rhs
        self.rel_weight = (self.absA / rhs[:, None]).sum(axis=0)
        self.sense = np.frombuffer(model.row_sense, dtype="S1")


def candidate_features(stats: ModelStats, x, lb, ub, cand) -> np.ndarray:
    """The 32 features of the candidate variables `cand` at a node whose LP
    point is x (bounds lb, ub). Rows: candidates."""
    m = stats.model
    x = np.asarray(x, float)
    lb = np.asarray(lb, float)
    ub = np.asarray(ub, float)
    cand = np.asarray(cand, int)
    n = m.n_vars
    fl = np.floor(x)
    fdown = x - fl
    fup = 1.0 - fdown
    frac = np.minimum(fdown, fup)
    isfrac = frac > 1e-6
    ax = stats.A @ x
    slack = np.where(stats.sense == b"G", ax - m.rhs, m.rhs - ax)
    slack = np.where(stats.sense == b"E", -np.abs(ax - m.rhs), slack) / stats.rhs_abs
    binding = slack < 1e-6
    obj = float(m.obj @ x)
    F = np.zeros((len(cand), N_FEATURES))
    nfrac = max(int(isfrac.sum()), 1)
    # most-fractional rank among the candidates (0 = closest to 1/2)
    closeness = np.abs(frac[cand] - 0.5)
    mf_rank = np.argsort(np.argsort(closeness, kind="stable"), kind="stable") / max(len(cand) - 1, 1)
    for r, j in enumerate(cand):
        rows = np.nonzero(stats.A[:, j])[0]
        s = slack[rows] if len(rows) else np.zeros(1)
        w = (stats.absA[rows, j] / stats.rhs_abs[rows]) if len(rows) else np.zeros(1)
        bnd = binding[rows] if len(rows) else np.zeros(1, bool)
        cj = stats.c[j] / stats.cmax
        F[r] = [
            frac[j], fdown[j], x[j], cj,
            stats.nnz[j] / max(m.n_rows, 1), stats.col_mean[j] / stats.mean_all, stats.col_max[j],
            stats.col_min[j], stats.col_std[j],
            np.tanh(stats.eff[j]), stats.eff_rank[j],
            s.mean(), s.min(), s.max(), s.std(),
            (w * (s < 1e-6)).sum(), w.max(), w.mean(), (w * s).sum(),
            fdown[j] * abs(cj), fup[j] * abs(cj),
            nfrac / n, float(np.mean(x >= ub - 1e-9)), float(np.mean(x <= lb + 1e-9)),
            lb[j], min(ub[j], 1e3), float(np.mean(lb == ub)),
            bnd.mean(), stats.rel_weight[j],
            abs(stats.c[j]) * frac[j] / max(abs(obj), 1e-9) * n,
            mf_rank[r], 1.0,
        ]
    return F


def node_targets(cands) -> np.ndarray:
    """Training targets of one node: sqrt(score / best score) (best = 1);
    the square root spreads the many small scores, which sharpens the
    ranking near the top."""
    s = np.array([c["score"] for c in cands], float)
    top = s.max() if len(s) else 1.0
    return np.sqrt(s / top) if top > 0 else np.zeros_like(s)


def node_inputs(F) -> np.ndarray:
    """Network inputs of one node: each feature standardised over the
    node's candidates (what matters is how a candidate compares with the
    others at its node; node-wide features become 0)."""
    F = np.asarray(F, float)
    return (F - F.mean(axis=0)) / np.maximum(F.std(axis=0), 1e-6)


def node_dataset(model: LpModel, stats: ModelStats, x, obj, lb=None, ub=None, max_candidates=64):
    """Features and targets of one node from (batched) strong branching on
    its candidates; also returns the candidate dicts."""
    from .milp import strong_branch

    lb = model.lb if lb is None else lb
    ub = model.ub if ub is None else ub
    is_int = np.ones(model.n_vars, dtype=np.int8)
    cands, best, _ = strong_branch(model, is_int, x, obj, lb=lb, ub=ub, max_candidates=max_candidates)
    if not cands:
        return np.zeros((0, N_FEATURES)), np.zeros(0), cands
    F = candidate_features(stats, x, lb, ub, [c["var"] for c in cands])
    return F, node_targets(cands), cands


def instance(seed: int, n=150, m=50) -> LpModel:
    """The training/evaluation family: knapsack instances of the benchmark's
    kind (weights U[1,100], capacity half the row sum, correlated profits)
    at a reduced size."""
    return multi_knapsack(n, m, seed=seed)


def generate_dataset(seeds, n=150, m=50, max_candidates=64, log=None):
    """Root-node strong-branching data of seeded knapsack instances:
    X (rows: candidates, raw features), t (targets), node id of each row."""
    Xs, ts, ids = [], [], []
    ipm = InteriorPointSolver(InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9))
    for k, seed in enumerate(seeds):
        model = instance(seed, n, m)
        root = ipm.solve(model)
        if root["status"] != 0:
            continue
        stats = ModelStats(model)
        F, t, _ = node_dataset(model, stats, root["primal"], root["objective"], max_candidates=max_candidates)
        if len(t) < 2:
            continue
        Xs.append(F)
        ts.append(t)
        ids.append(np.full(len(t), seed))
        if log and (k + 1) % 50 == 0:
            log(f"dataset: {k + 1}/{len(seeds)} instances, {sum(len(v) for v in ts)} candidates")
    return np.vstack(Xs), np.concatenate(ts), np.concatenate(ids)


def dataset_inputs(X, ids) -> np.ndarray:
    """node_inputs applied node by node to a dataset."""
    Z = np.zeros_like(np.asarray(X, float))
    for u in np.unique(ids):
        m = ids == u
        Z[m] = node_inputs(X[m])
    return Z


class SurrogateRanker:
    """Ranker for BranchAndBoundSolver: key per variable (higher first) from
    the network; non-fractional variables get -inf."""

    def __init__(self, model: LpModel, net: Mlp):
        self.stats = ModelStats(model)
        self.net = net
        self.calls = 0

    def scores(self, x, lb, ub, cand):
        F = candidate_features(self.stats, x, lb, ub, cand)
        return self.net.forward(node_inputs(F))

    def __call__(self, x, lb, ub):
        self.calls += 1
        frac = np.minimum(x - np.floor(x), np.ceil(x) - x)
        cand = np.nonzero(frac > 1e-6)[0]
        key = np.full(len(x), -np.inf)
        if len(cand):
            key[cand] = self.scores(x, lb, ub, cand)
        return key


def load_trained(model: LpModel, path=WEIGHTS):
    """The committed network trained by sovereign_amd/train_surrogate.py."""
    return SurrogateRanker(model, Mlp.load(path))
