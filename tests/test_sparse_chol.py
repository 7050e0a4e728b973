"""Sparse Cholesky path of the normal equations, host side (no device):

  * the minimum-degree ordering (sov_min_degree_order) is the reference's
    AMDOrdering::compute to the last index: the same permutation as the
    oracle's restatement (itself checked against the compiled reference in
    test_oracle.py) on staircase, transport, knapsack, random and dense-row
    patterns, and the same verdict under a fill limit;
  * tests/sparse_audit.cpp, compiled with g++ against the symbolic analysis
    the library uses, checks the supernodal structure's invariants and runs
    the GPU kernels' index arithmetic sequentially with every access
    bounds-checked: its solution of M x = b matches a dense solve and the
    oracle's up-looking LDL^T, nnz(L) matches the oracle's, and a forced
    tiny pivot is replaced as the reference replaces it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from sovereign_amd import api
from sovereign_amd import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def normal_pattern(model):
    return O.normal_eq_pattern(model.n_rows, model.n_vars, model.col_ptr, model.row_idx)


def random_pattern(n, per_col, seed):
    rng = np.random.default_rng(seed)
    cols = [[j] + sorted(set(rng.integers(j, n, per_col).tolist()) - {j}) for j in range(n)]
    cp = np.cumsum([0] + [len(c) for c in cols])
    return cp, np.concatenate(cols)


def budget_staircase(products, periods):
    """A staircase plus one row over every make variable: a dense node that the
    ordering must put last."""
    m = P.staircase(products, periods)
    A = np.zeros((m.n_rows + 1, m.n_vars))
    for j in range(m.n_vars):
        for p in range(m.col_ptr[j], m.col_ptr[j + 1]):
            A[m.row_idx[p], j] = m.values[p]
    A[-1, 0::2] = 1.0
    return A


def dense_lower_pattern(A):
    M = (A != 0).astype(float) @ (A != 0).T.astype(float)
    n = M.shape[0]
    cp, ri = [0], []
    for j in range(n):
        rows = [i for i in range(j, n) if M[i, j] != 0 or i == j]
        ri += rows
        cp.append(len(ri))
    return np.array(cp), np.array(ri)


PATTERNS = {
    "staircase_8x40": lambda: normal_pattern(P.staircase(8, 40)),
    "staircase_20x100": lambda: normal_pattern(P.staircase(20, 100)),
    "transport_20x20": lambda: normal_pattern(P.transport_balanced(20, 20)),
    "mkp_60x30": lambda: normal_pattern(P.multi_knapsack(60, 30)),
    "random_300": lambda: random_pattern(300, 3, 7),
    "budget_staircase": lambda: dense_lower_pattern(budget_staircase(6, 30)),
}


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_ordering_matches_reference_rule(name):
    pcp, pri = PATTERNS[name]()
    n = len(pcp) - 1
    perm, iperm, pred = api.min_degree_order(n, pcp, pri)
    ok, operm, oiperm = O.amd_order(n, pcp, pri)
    assert ok == 1
    assert np.array_equal(perm, operm) and np.array_equal(iperm, oiperm)
    assert np.array_equal(np.sort(perm), np.arange(n))
    # under a limit just below the prediction both stop; at the prediction both finish
    assert api.min_degree_order(n, pcp, pri, fill_limit=pred - 1) is None
    assert O.amd_order(n, pcp, pri, fill_limit=pred - 1)[0] == 0
    assert api.min_degree_order(n, pcp, pri, fill_limit=pred) is not None


def test_dense_node_ordered_last():
    pcp, pri = dense_lower_pattern(budget_staircase(6, 30))
    n = len(pcp) - 1
    perm, _, _ = api.min_degree_order(n, pcp, pri)
    assert perm[-1] == n - 1  # the budget row touches every period's makes


@pytest.fixture(scope="module")
def audit(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("sparse") / "sparse_audit")
    csrc = os.path.join(ROOT, "sovereign_amd", "csrc")
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", csrc, "-I", os.path.join(ROOT, "include"),
                    "-o", exe, os.path.join(HERE, "sparse_audit.cpp"), os.path.join(csrc, "sparse_symbolic.cpp")],
                   check=True)
    return exe


def run_audit(exe, pcp, pri, vals, rhs):
    n = len(pcp) - 1
    text = "\n".join([f"{n} {len(pri)}", " ".join(map(str, pcp)), " ".join(map(str, pri)),
                      " ".join(repr(float(v)) for v in vals), " ".join(repr(float(v)) for v in rhs)])
    r = subprocess.run([exe], input=text, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[:2000]
    lines = r.stdout.split("\n")
    head = lines[0].split()
    assert head[0] == "ok"
    stats = dict(zip(["supernodes", "levels", "max_front", "nnz_l", "front_total", "replaced"], map(int, head[1:])))
    return stats, np.array([float(v) for v in lines[1:1 + n]])


def normal_values(model, d, pcp, pri):
    A = np.zeros((model.n_rows, model.n_vars))
    for j in range(model.n_vars):
        for p in range(model.col_ptr[j], model.col_ptr[j + 1]):
            A[model.row_idx[p], j] += model.values[p]
    M = (A * d) @ A.T + 1e-12 * np.eye(model.n_rows)
    vals = np.array([M[pri[p], j] for j in range(model.n_rows) for p in range(pcp[j], pcp[j + 1])])
    return M, vals


@pytest.mark.parametrize("name,model", [("staircase_8x40", P.staircase(8, 40)),
                                        ("staircase_20x100", P.staircase(20, 100)),
                                        ("transport_12x12", P.transport_balanced(12, 12)),  # rank-deficient rows
                                        ("random_lp_200x800", P.random_sparse_lp(200, 800, seed=5).as_model()),
                                        ("mkp_40x20", P.multi_knapsack(40, 20))])
def test_multifrontal_solve_matches_dense_and_oracle(audit, name, model):
    rng = np.random.default_rng(11)
    pcp, pri = normal_pattern(model)
    d = np.exp(rng.uniform(-6, 6, model.n_vars))  # IPM-like spread of D = x / s
    M, vals = normal_values(model, d, pcp, pri)
    b = rng.standard_normal(model.n_rows)
    stats, x = run_audit(audit, pcp, pri, vals, b)
    st, xo, rep, lnnz, _ = O.ldlt_factor_solve(pcp, pri, vals, b, tiny=1e-13)
    assert st == 0
    assert stats["nnz_l"] - model.n_rows == lnnz  # same fill: the postorder is an equivalent ordering
    assert stats["replaced"] == rep
    # backward error of a stable factorization, whatever M's conditioning
    assert np.abs(M @ x - b).max() <= 1e-12 * np.abs(M).max() * np.abs(x).max()
    if np.linalg.cond(M) < 1e10:  # well-posed: the solution itself is pinned
        ref = np.linalg.solve(M, b)
        assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()
        assert np.abs(x - xo).max() <= 1e-8 * np.abs(ref).max()


def test_multifrontal_tiny_pivot_replaced(audit):
    """Two identical rows: the second pivot vanishes, the reference's rule
    replaces it (d = 1e128) and that direction drops out of the solution."""
    A = np.array([[1.0, 2.0, 0.0, 1.0], [1.0, 2.0, 0.0, 1.0], [0.0, 1.0, 3.0, 0.0]])
    M = A @ A.T  # no regularization: the second pivot is rounding noise
    pcp, pri = np.array([0, 3, 5, 6]), np.array([0, 1, 2, 1, 2, 2])
    vals = np.array([M[pri[p], j] for j in range(3) for p in range(pcp[j], pcp[j + 1])])
    b = np.array([1.0, 1.0, 2.0])
    stats, x = run_audit(audit, pcp, pri, vals, b)
    st, xo, rep, _, _ = O.ldlt_factor_solve(pcp, pri, vals, b, tiny=1e-13)
    assert st == 0 and rep == 1 and stats["replaced"] == 1
    assert np.allclose(x, xo, rtol=1e-9, atol=1e-12)
