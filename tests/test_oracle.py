"""Pins the CPU oracle: reference golden vectors, the reference's own
known-answer tests, and what the mathematics fixes (LAPACK partial pivoting,
A D A^T by dense algebra, LP optima by an independent solver, duality bounds).
CPU only."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.optimize import linprog

import golden_models as G
from oracle import oracle as O
from sovereign_amd import problems as P


# ---- golden vectors from the reference build --------------------------------

@pytest.mark.parametrize("key", sorted(G.expected().keys()))
def test_oracle_matches_reference_golden(key):
    """Same status and iteration count, objective to the last bit (both are
    compiled without FMA contraction, so the arithmetic is identical)."""
    exp = G.expected()[key]
    mode = key.split(":")[-1]
    r = O.lp_solve(G.model_for(key), mode)
    assert r["status"] == exp["status"]
    assert r["iterations"] == exp["iterations"]
    assert repr(float(r["objective"])) == exp["objective"]
    assert r["used_sparse"] == exp["used_sparse"]
    np.testing.assert_allclose(G.digest(r["primal"]), exp["primal_digest"], rtol=1e-13, atol=1e-12)


def test_oracle_kernels_match_reference_golden():
    k = G.kernels()
    for n in (7, 40, 150):
        st, lu, piv = O.dense_lu(k[f"lu{n}_a"])
        assert st == 0
        np.testing.assert_array_equal(O.dense_lu_solve(lu, piv, k[f"lu{n}_b"]), k[f"lu{n}_x"])
        np.testing.assert_array_equal(O.dense_lu_solve(lu, piv, k[f"lu{n}_b"], transpose=True), k[f"lu{n}_xt"])
    m, n = 90, 130
    ax = O.spmv(m, n, k["spmv_ptr"], k["spmv_idx"], k["spmv_val"], k["spmv_x"])
    atz = O.spmv(m, n, k["spmv_ptr"], k["spmv_idx"], k["spmv_val"], k["spmv_z"], transpose=True)
    np.testing.assert_array_equal(ax, k["spmv_ax"])
    np.testing.assert_array_equal(atz, k["spmv_atz"])


@pytest.mark.skipif(not O.ref_available(), reason="reference build only exists in the build container")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_matches_reference_fresh_seeds(seed):
    for model in (P.random_sparse_lp(120, 400, 5, seed=seed).as_model(), P.transport_balanced(12, 17, seed=seed)):
        for mode in ("auto", "dense", "sparse"):
            a, r = O.lp_solve(model, mode), O.ref_lp_solve(model, mode)
            assert (a["status"], a["iterations"]) == (r["status"], r["iterations"])
            np.testing.assert_array_equal(a["primal"], r["primal"])
            np.testing.assert_array_equal(a["row_dual"], r["row_dual"])


@pytest.mark.skipif(not O.ref_available(), reason="reference build only exists in the build container")
def test_oracle_pattern_and_ordering_match_reference():
    lp = P.random_sparse_lp(200, 600, 4, seed=7)
    p1 = O.normal_eq_pattern(lp.m, lp.n, lp.col_ptr, lp.row_idx)
    p2 = O.normal_eq_pattern(lp.m, lp.n, lp.col_ptr, lp.row_idx, use_ref=True)
    np.testing.assert_array_equal(p1[0], p2[0])
    np.testing.assert_array_equal(p1[1], p2[1])
    o1 = O.amd_order(lp.m, *p1)
    o2 = O.amd_order(lp.m, *p2, use_ref=True)
    np.testing.assert_array_equal(o1[1], o2[1])


# ---- the reference's own known-answer tests ---------------------------------

def test_dense_lu_two_by_two():
    """tests/unit/test_dense_lu.cpp:7-17"""
    st, lu, piv = O.dense_lu(np.array([[2.0, 1.0], [0.0, 3.0]]))
    assert st == 0
    np.testing.assert_allclose(O.dense_lu_solve(lu, piv, [5.0, 6.0]), [1.5, 2.0], atol=1e-9)


def test_dense_lu_reconstruction_and_singular():
    """cuda_driver.cpp:616-639 (P*A = L*U on a 5x5) and test_gpu_backend.cpp:79-87."""
    n = 5
    a = np.array([[np.sin(1.0 + 3.0 * i + 7.0 * j) + (0.5 if i == j else 0.0) for j in range(n)] for i in range(n)])
    st, lu, piv = O.dense_lu(a)
    assert st == 0
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    np.testing.assert_allclose(L @ U, a[piv], atol=1e-9)
    n = 40
    m = np.array([[np.sin(0.37 * (i * n + j) + 1.0) + (3.0 if (i + 1) % n == j else 0.0) for j in range(n)]
                  for i in range(n)])
    m[:, 5] = 2.0 * m[:, 3]
    assert O.dense_lu(m)[0] == 1


def test_dense_lu_pivots_match_lapack():
    """Partial pivoting with the first largest |a| is LAPACK's getrf rule."""
    rng = np.random.default_rng(11)
    a = rng.standard_normal((60, 60))
    st, lu, piv = O.dense_lu(a)
    lu2, ipiv = sla.lu_factor(a)
    perm = np.arange(60)
    for k, p in enumerate(ipiv):
        perm[[k, p]] = perm[[p, k]]
    np.testing.assert_array_equal(piv, perm)
    np.testing.assert_allclose(lu, lu2, rtol=1e-10, atol=1e-12)


def test_spmv_closed_form():
    """cuda_driver.cpp:640-657: A = [[2,0,.5],[0,4,3]]."""
    ptr, idx, val = [0, 1, 2, 4], [0, 1, 0, 1], [2.0, 4.0, 0.5, 3.0]
    np.testing.assert_allclose(O.spmv(2, 3, ptr, idx, val, [1.0, 2.0, 3.0]), [3.5, 17.0])
    np.testing.assert_allclose(O.spmv(2, 3, ptr, idx, val, [1.0, -1.0], transpose=True), [2.0, -4.0, -2.5])


def test_normal_equation_values_known_answers():
    """test_sparse_symmetric.cpp:94-164: M = [[2,2],[2,5]]; d_1 = 0 skips the column."""
    ptr, idx, val = [0, 2, 3], [0, 1, 1], [1.0, 1.0, 1.0]
    M = O.normal_eq_dense(2, 2, ptr, idx, val, [2.0, 3.0])
    np.testing.assert_allclose(M, [[2 + 1e-12, 2], [2, 5 + 1e-12]], rtol=0, atol=1e-14)
    M0 = O.normal_eq_dense(2, 2, ptr, idx, val, [2.0, 0.0])
    np.testing.assert_allclose(M0, [[2, 2], [2, 2]], atol=1e-11)
    pcp, pri = O.normal_eq_pattern(2, 2, ptr, idx)
    np.testing.assert_array_equal(pcp, [0, 2, 3])
    np.testing.assert_array_equal(pri, [0, 1, 1])


def test_ldlt_known_answers():
    """test_sparse_ldlt.cpp:683-720: breakdown without the tiny-pivot rule,
    one dropped direction with it, and a 2x2 solve."""
    pcp, pri = [0, 2, 3], [0, 1, 1]
    st, *_ = O.ldlt_factor_solve(pcp, pri, [1.0, 2.0, 1.0], [1.0, 1.0])
    assert st == 1
    st, x, *_ = O.ldlt_factor_solve(pcp, pri, [4.0, 2.0, 3.0], [10.0, 7.0])
    assert st == 0
    np.testing.assert_allclose(x, [2.0, 1.0], atol=1e-12)
    assert O.ldlt_factor_solve(pcp, pri, [1.0, 1.0, 1.0], [3.0, 3.0])[0] == 1
    st, x, rep, _, _ = O.ldlt_factor_solve(pcp, pri, [1.0, 1.0, 1.0], [3.0, 3.0], tiny=1e-13)
    assert st == 0 and rep == 1
    assert abs(x[0] + x[1] - 3.0) < 1e-12


def test_ldlt_hub_row_ordered_last():
    """test_sparse_ldlt.cpp:779-813: a hub joined to a chain is ordered last
    and adds no fill."""
    n = 3000
    pcp, pri, vals = [0], [], []
    for j in range(n):
        pri.append(j)
        vals.append(float(n) if j == 0 else 4.0)
        if j == 0:
            pri.extend(range(1, n))
            vals.extend([1.0] * (n - 1))
        elif j + 1 < n:
            pri.append(j + 1)
            vals.append(-1.0)
        pcp.append(len(pri))
    b = np.sin(0.1 * np.arange(n))
    st, x, _, lnnz, perm = O.ldlt_factor_solve(pcp, pri, vals, b)
    assert st == 0 and perm[-1] == 0 and lnnz == (n - 1) + (n - 2)
    # residual against the symmetric matrix
    M = np.zeros((n, n))
    for j in range(n):
        for p in range(pcp[j], pcp[j + 1]):
            M[pri[p], j] = M[j, pri[p]] = vals[p]
    assert np.max(np.abs(M @ x - b)) < 1e-10


def test_ldlt_matches_dense_solve_lcg_matrix():
    """test_sparse_ldlt.cpp:620-681: random 60x150 A (4 per column, LCG seed
    12345), two diagonals, LDL^T against a dense solve."""
    state = [12345]

    def nxt():
        state[0] = (state[0] * 1103515245 + 12345) & 0xFFFFFFFF
        return (state[0] >> 8) & 0xFFFF

    m, n = 60, 150
    ptr, idx, val = [0], [], []
    for _ in range(n):
        rows = []
        while len(rows) < 4:
            r = nxt() % m
            if r not in rows:
                rows.append(r)
        rows.sort()
        for r in rows:
            idx.append(r)
            val.append(0.5 + (nxt() % 1000) / 400.0)
        ptr.append(len(idx))
    A = np.zeros((m, n))
    for j in range(n):
        A[idx[ptr[j]:ptr[j + 1]], j] = val[ptr[j]:ptr[j + 1]]
    pcp, pri = O.normal_eq_pattern(m, n, ptr, idx)
    for rnd in range(2):
        d = np.array([1.0 + (nxt() % 100) / 50.0 if rnd == 0 else 10.0 ** ((nxt() % 13) - 6.0) for _ in range(n)])
        M = A @ np.diag(d) @ A.T + 1e-10 * np.eye(m)
        vals = [M[pri[p], j] - (1e-10 if pri[p] == j else 0.0) for j in range(m) for p in range(pcp[j], pcp[j + 1])]
        b = np.array([(nxt() % 200) / 10.0 - 10.0 for _ in range(m)])
        st, x, *_ = O.ldlt_factor_solve(pcp, pri, vals, b, reg=1e-10)
        assert st == 0
        xd = np.linalg.solve(M, b)
        assert np.max(np.abs(x - xd)) <= 1e-8 * max(1.0, np.max(np.abs(xd)))


# ---- mathematics ------------------------------------------------------------

def test_normal_equations_equal_dense_algebra():
    lp = P.random_sparse_lp(70, 200, 5, seed=3)
    A = lp.as_model().dense()
    d = np.random.default_rng(0).uniform(0.01, 10, size=lp.n)
    M = O.normal_eq_dense(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, d)
    np.testing.assert_allclose(M, A @ np.diag(d) @ A.T + 1e-12 * np.eye(lp.m), rtol=1e-13, atol=1e-12)
    for mode in ("dense", "sparse"):
        ne = O.NormalEquations(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, mode)
        assert ne.factor(d)
        b = np.arange(lp.m, dtype=float)
        np.testing.assert_allclose(M @ ne.solve(b), b, rtol=0, atol=1e-8 * np.max(np.abs(b)))


def test_normal_equations_redundant_row_drops_one_pivot():
    """A balanced transport problem's equality rows sum to zero in pairs
    (supply total = demand total), so A has rank m - 1 and A D A^T is
    singular up to the 1e-12 regularization. With D ~ 1e3 the dependent
    row's pivot falls under 1e-13 |d0|: the LDL^T's tiny-pivot rule
    (sparse_ldlt.cpp:165-173) drops exactly one pivot, and consistent
    right-hand sides (b in range(A)) are still solved: A^T y is unique even
    though y is not. This is the premise of the GPU dropped-pivot test."""
    f = O.ipm_form(P.transport_balanced(20, 20))
    m, n = f["m"], f["n"]
    A = np.zeros((m, n))
    for j in range(n):
        for p in range(f["col_ptr"][j], f["col_ptr"][j + 1]):
            A[f["row_idx"][p], j] = f["values"][p]
    assert np.linalg.matrix_rank(A) == m - 1
    rng = np.random.default_rng(11)
    d = 1e3 * np.exp(rng.uniform(-1, 1, size=n))
    rhs = A @ rng.standard_normal(n)
    ne = O.NormalEquations(m, n, f["col_ptr"], f["row_idx"], f["values"], "sparse")
    assert ne.factor(d)
    assert ne.replaced_pivots == 1
    y = ne.solve(rhs)
    M = A @ np.diag(d) @ A.T + 1e-12 * np.eye(m)
    np.testing.assert_allclose(A.T @ y, np.linalg.lstsq(M, rhs, rcond=None)[0] @ A, rtol=0,
                               atol=1e-7 * np.max(np.abs(A.T @ y)))


def test_spmv_equals_dense_product():
    lp = P.random_sparse_lp(90, 130, 5, seed=9)
    A = lp.as_model().dense()
    x = np.sin(np.arange(lp.n))
    z = 0.5 - (np.arange(lp.m) % 4)
    np.testing.assert_allclose(O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, x), A @ x, atol=1e-13)
    np.testing.assert_allclose(O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, z, True), A.T @ z, atol=1e-13)


@pytest.mark.parametrize("model", [P.transport_balanced(50, 50), G.load_fixture("plan_600x900"),
                                   P.staircase(8, 40)], ids=["transport_c1", "plan_600x900", "staircase"])
def test_ipm_objective_matches_highs(model):
    """Independent check: scipy's HiGHS on the same LP, objective within 1e-8."""
    A = model.dense()
    sense = np.frombuffer(model.row_sense, dtype="S1")
    c = -model.obj if model.maximize else model.obj
    ub_rows = sense == b"L"
    lb_rows = sense == b"G"
    eq = sense == b"E"
    A_ub = np.vstack([A[ub_rows], -A[lb_rows]])
    b_ub = np.concatenate([model.rhs[ub_rows], -model.rhs[lb_rows]])
    bounds = [(lo, None if hi >= 1e29 else hi) for lo, hi in zip(model.lb, model.ub)]
    h = linprog(c, A_ub=A_ub if len(b_ub) else None, b_ub=b_ub if len(b_ub) else None,
                A_eq=A[eq] if eq.any() else None, b_eq=model.rhs[eq] if eq.any() else None, bounds=bounds,
                method="highs")
    assert h.status == 0
    highs = (-h.fun if model.maximize else h.fun) + model.obj_constant
    r = O.lp_solve(model)
    assert r["status"] == 0
    assert abs(r["objective"] - highs) <= 1e-8 * max(1.0, abs(highs))


def test_random_lp_duality_bounds_and_kkt():
    """(x0, y0, s0) of the generator is feasible: b'y0 <= optimum <= c'x0; the
    returned iterate satisfies the KKT conditions to the tolerances."""
    lp = P.random_sparse_lp(150, 600, 6, seed=5)
    r = O.solve_form(lp, enable_scaling=0)
    assert r["status"] == 0
    A = lp.as_model().dense()
    x, y, s = r["x"], r["y"], r["s"]
    assert lp.b @ lp.y0 - 1e-8 <= r["objective"] <= lp.c @ lp.x0 + 1e-8
    assert np.max(np.abs(A @ x - lp.b)) / max(1, np.max(np.abs(lp.b))) < 1e-9
    assert np.max(np.abs(A.T @ y + s - lp.c)) / max(1, np.max(np.abs(lp.c))) < 1e-9
    assert abs(x @ s) / (1 + abs(lp.c @ x)) < 1e-9


def test_ipm_form_restatement_invariants():
    """build_ipm_form: slack per inequality, ub rows, scaling keeps A x = b."""
    rng = np.random.default_rng(4)
    m0, nv = 6, 5
    A0 = rng.uniform(-1, 1, (m0, nv))
    A0[rng.random((m0, nv)) < 0.3] = 0.0
    cols = [list(np.nonzero(A0[:, j])[0]) for j in range(nv)]
    ptr, rows, vals = P.csc_from_columns(cols, [list(A0[c, j]) for j, c in enumerate(cols)], m0)
    lb = np.array([0.0, -1.0, 2.0, 0.0, 0.5])
    ub = np.array([1e30, 3.0, 1e30, 4.0, 1e30])
    model = P.LpModel(n_vars=nv, n_rows=m0, col_ptr=ptr, row_idx=rows, values=vals, obj=rng.standard_normal(nv),
                      lb=lb, ub=ub, row_sense="LGELGE", rhs=rng.standard_normal(m0))
    f = O.ipm_form(model, scaling=True)
    assert f["m"] == m0 + 2  # two finite upper bounds
    assert f["n"] == nv + 4 + 2  # slacks for L, G and the ub rows
    # unscaled x -> scaled: x_scaled = (x - shift) / col_scale satisfies A_s x_s = b_s
    x = np.concatenate([lb + 0.3, np.zeros(f["n"] - nv)])
    As = np.zeros((f["m"], f["n"]))
    for j in range(f["n"]):
        As[f["row_idx"][f["col_ptr"][j]:f["col_ptr"][j + 1]], j] = f["values"][f["col_ptr"][j]:f["col_ptr"][j + 1]]
    A_un = As / f["row_scale"][:, None] / f["col_scale"][None, :]
    b_un = f["b"] / f["row_scale"]
    struct = A_un[:m0, :nv]
    np.testing.assert_allclose(struct, A0, atol=1e-12)
    np.testing.assert_allclose(b_un[:m0], model.rhs - A0 @ lb, atol=1e-12)
