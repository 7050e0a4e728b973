"""Seeded synthetic LP workloads and the general-LP container.

`LpModel` is the array image of the reference's OptimizationModel for LP
(solver/core/include/sovereign/types.hpp:32-59): a CSC constraint matrix,
one sense per row ('L' <=, 'G' >=, 'E' =), variable bounds, and a linear
objective. Generators follow the workload definitions the solver is measured
on; every one takes an explicit seed (default base seed 26119).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

BASE_SEED = 26119
INF = 1e30


@dataclass
class LpModel:
    n_vars: int
    n_rows: int
    col_ptr: np.ndarray
    row_idx: np.ndarray
    values: np.ndarray
    obj: np.ndarray
    lb: np.ndarray
    ub: np.ndarray
    row_sense: bytes
    rhs: np.ndarray
    obj_constant: float = 0.0
    maximize: bool = False
    name: str = ""
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.col_ptr = np.ascontiguousarray(self.col_ptr, dtype=np.int32)
        self.row_idx = np.ascontiguousarray(self.row_idx, dtype=np.int32)
        self.values = np.ascontiguousarray(self.values, dtype=np.float64)
        self.obj = np.ascontiguousarray(self.obj, dtype=np.float64)
        self.lb = np.ascontiguousarray(self.lb, dtype=np.float64)
        self.ub = np.ascontiguousarray(self.ub, dtype=np.float64)
        self.rhs = np.ascontiguousarray(self.rhs, dtype=np.float64)
        if isinstance(self.row_sense, str):
            self.row_sense = self.row_sense.encode()
        assert len(self.row_sense) == self.n_rows
        assert self.col_ptr.shape[0] == self.n_vars + 1

    @property
    def nnz(self) -> int:
        return int(self.col_ptr[-1])

    def fill_struct(self, struct_type):
        """Fills a ctypes struct laid out as sov_lp_model / orc_lp_model."""
        s = struct_type()
        f64p = C.POINTER(C.c_double)
        i32p = C.POINTER(C.c_int32)
        sense = C.create_string_buffer(self.row_sense, len(self.row_sense) + 1)
        self._keep = [sense]
        s.n_vars = self.n_vars
        s.n_rows = self.n_rows
        s.col_ptr = self.col_ptr.ctypes.data_as(i32p)
        s.row_idx = self.row_idx.ctypes.data_as(i32p)
        s.values = self.values.ctypes.data_as(f64p)
        s.obj = self.obj.ctypes.data_as(f64p)
        s.obj_constant = self.obj_constant
        s.maximize = 1 if self.maximize else 0
        s.lb = self.lb.ctypes.data_as(f64p)
        s.ub = self.ub.ctypes.data_as(f64p)
        s.row_sense = C.cast(sense, C.c_char_p)
        s.rhs = self.rhs.ctypes.data_as(f64p)
        return s

    def dense(self) -> np.ndarray:
        A = np.zeros((self.n_rows, self.n_vars))
        for j in range(self.n_vars):
            for p in range(self.col_ptr[j], self.col_ptr[j + 1]):
                A[self.row_idx[p], j] += self.values[p]
        return A


def csc_from_columns(cols_rows, cols_vals, n_rows):
    """CSC arrays from per-column row/value lists."""
    ptr = np.zeros(len(cols_rows) + 1, dtype=np.int32)
    ptr[1:] = np.cumsum([len(r) for r in cols_rows])
    rows = np.concatenate([np.asarray(r, dtype=np.int32) for r in cols_rows]) if cols_rows else np.zeros(0, np.int32)
    vals = np.concatenate([np.asarray(v, dtype=np.float64) for v in cols_vals]) if cols_vals else np.zeros(0)
    return ptr, rows, vals


def transport_balanced(sources: int = 50, sinks: int = 50, seed: int = BASE_SEED) -> LpModel:
    """Balanced transportation LP: supplies uniform int [10, 100], total demand
    equal to total supply split by a uniform Dirichlet, costs uniform int
    [1, 20]; x_ij >= 0, one equality row per source and per sink (one of the
    rows is redundant)."""
    rng = np.random.default_rng(seed)
    supply = rng.integers(10, 101, size=sources).astype(np.float64)
    demand = supply.sum() * rng.dirichlet(np.ones(sinks))
    cost = rng.integers(1, 21, size=(sources, sinks)).astype(np.float64)
    n = sources * sinks
    j = np.arange(n)
    i_src = j // sinks
    i_snk = j % sinks
    rows = np.stack([i_src, sources + i_snk], axis=1).reshape(-1)
    ptr = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    return LpModel(n_vars=n, n_rows=sources + sinks, col_ptr=ptr, row_idx=rows, values=np.ones(2 * n),
                   obj=cost.reshape(-1), lb=np.zeros(n), ub=np.full(n, INF),
                   row_sense=b"E" * (sources + sinks), rhs=np.concatenate([supply, demand]),
                   name=f"transport_{sources}x{sinks}")


def transport_grid(m: int) -> LpModel:
    """Transportation LP with m sources and m sinks, supplies and demands 10,
    cost 1 + 0.04*((7i + 3j) mod 5) (tests/unit/test_correctness_regressions.cpp:74-109)."""
    n = m * m
    j = np.arange(n)
    i_src, i_snk = j // m, j % m
    cost = 1.0 + 0.04 * ((i_src * 7 + i_snk * 3) % 5)
    rows = np.stack([i_src, m + i_snk], axis=1).reshape(-1)
    ptr = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    return LpModel(n_vars=n, n_rows=2 * m, col_ptr=ptr, row_idx=rows, values=np.ones(2 * n), obj=cost,
                   lb=np.zeros(n), ub=np.full(n, INF), row_sense=b"E" * (2 * m), rhs=np.full(2 * m, 10.0),
                   name=f"transport_grid_{m}")


def staircase(products: int, periods: int) -> LpModel:
    """Multi-period production planning with inventory balance rows and one
    capacity row per period (test_correctness_regressions.cpp:158-180)."""
    cols_r, cols_v, obj = [], [], []
    senses, rhs = [], []
    var_index = {}

    def var(name):
        var_index[name] = len(var_index)
        cols_r.append([])
        cols_v.append([])
        return var_index[name]

    def add(row, name, val):
        k = var_index[name]
        cols_r[k].append(row)
        cols_v[k].append(val)

    for t in range(periods):
        cap_terms = []
        for p in range(products):
            mk, st = f"make{p}_{t}", f"stock{p}_{t}"
            var(mk)
            obj.append(1.0 + 0.1 * ((p * 3 + t * 7) % 11))
            var(st)
            obj.append(0.05 + 0.01 * (p % 4))
            r = len(senses)
            add(r, mk, 1.0)
            add(r, st, -1.0)
            if t > 0:
                add(r, f"stock{p}_{t - 1}", 1.0)
            senses.append("E")
            rhs.append(5.0 + float((p * 5 + t * 3) % 9))
            cap_terms.append((mk, 1.0 + 0.1 * (p % 3)))
        r = len(senses)
        for name, v in cap_terms:
            add(r, name, v)
        senses.append("L")
        rhs.append(12.0 * products)
    # rows inside a column must be sorted for the CSC arrays
    for k in range(len(cols_r)):
        order = np.argsort(cols_r[k], kind="stable")
        cols_r[k] = list(np.asarray(cols_r[k])[order])
        cols_v[k] = list(np.asarray(cols_v[k])[order])
    ptr, rows, vals = csc_from_columns(cols_r, cols_v, len(senses))
    n = len(obj)
    return LpModel(n_vars=n, n_rows=len(senses), col_ptr=ptr, row_idx=rows, values=vals, obj=np.array(obj),
                   lb=np.zeros(n), ub=np.full(n, INF), row_sense="".join(senses), rhs=np.array(rhs),
                   name=f"staircase_{products}x{periods}")


def multi_knapsack(n: int = 500, m: int = 200, seed: int = BASE_SEED) -> LpModel:
    """0/1 multidimensional knapsack (maximize): weights uniform int [1, 100],
    capacity floor(0.5 x row sum), profit_j = mean_i w_ij + U[0, 20]; m <= rows,
    binaries as 0 <= x <= 1 (integrality is the branch-and-bound's business)."""
    rng = np.random.default_rng(seed)
    W = rng.integers(1, 101, size=(m, n)).astype(np.float64)
    cap = np.floor(0.5 * W.sum(axis=1))
    profit = W.mean(axis=0) + rng.uniform(0.0, 20.0, size=n)
    ptr = np.arange(0, m * n + 1, m, dtype=np.int32)
    rows = np.tile(np.arange(m, dtype=np.int32), n)
    return LpModel(n_vars=n, n_rows=m, col_ptr=ptr, row_idx=rows, values=W.T.reshape(-1), obj=profit,
                   lb=np.zeros(n), ub=np.ones(n), row_sense=b"L" * m, rhs=cap, maximize=True,
                   name=f"mkp_{n}x{m}")


@dataclass
class FormLp:
    """An LP already in interior-point form: min c'x, A x = b, x >= 0."""
    m: int
    n: int
    col_ptr: np.ndarray
    row_idx: np.ndarray
    values: np.ndarray
    b: np.ndarray
    c: np.ndarray
    x0: np.ndarray = None
    y0: np.ndarray = None
    s0: np.ndarray = None

    def as_model(self) -> LpModel:
        return LpModel(n_vars=self.n, n_rows=self.m, col_ptr=self.col_ptr, row_idx=self.row_idx,
                       values=self.values, obj=self.c, lb=np.zeros(self.n), ub=np.full(self.n, INF),
                       row_sense=b"E" * self.m, rhs=self.b, name=f"random_lp_{self.m}x{self.n}")


def random_sparse_lp(m: int = 8192, n: int = 32768, per_col: int = 6, seed: int = BASE_SEED) -> FormLp:
    """Random sparse feasible LP in equality form: each column has `per_col`
    distinct rows drawn uniformly, values N(0, 1); b = A x0 with
    x0 ~ U(0.1, 2); c = A'y0 + s0 with y0 ~ N(0, 1), s0 ~ U(0.1, 2). (x0, y0,
    s0) is a strictly feasible primal-dual pair, so the LP has an optimum."""
    rng = np.random.default_rng(seed)
    k = min(per_col, m)
    rows = rng.integers(0, m, size=(n, k))
    for _ in range(64):  # redraw columns with repeated rows
        srt = np.sort(rows, axis=1)
        bad = np.any(srt[:, 1:] == srt[:, :-1], axis=1)
        if not bad.any():
            break
        rows[bad] = rng.integers(0, m, size=(int(bad.sum()), k))
    rows = np.sort(rows, axis=1)
    vals = rng.standard_normal((n, k))
    ptr = np.arange(0, n * k + 1, k, dtype=np.int32)
    ri = rows.reshape(-1).astype(np.int32)
    va = vals.reshape(-1)
    x0 = rng.uniform(0.1, 2.0, size=n)
    y0 = rng.standard_normal(m)
    s0 = rng.uniform(0.1, 2.0, size=n)
    cols = np.repeat(np.arange(n), k)
    b = np.zeros(m)
    np.add.at(b, ri, va * x0[cols])
    aty = np.zeros(n)
    np.add.at(aty, cols, va * y0[ri])
    c = aty + s0
    return FormLp(m=m, n=n, col_ptr=ptr, row_idx=ri, values=va, b=b, c=c, x0=x0, y0=y0, s0=s0)
