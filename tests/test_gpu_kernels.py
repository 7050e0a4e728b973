"""GPU parity of the device kernels through the C ABI against the oracle:
SpMV (bit-level tolerance 1e-12), dense LU, normal-equation solves
(FP64 direct, FP32 factor + FP64 PCG)."""
import numpy as np
import pytest

import golden_models as G
from oracle import oracle as O
from sovereign_amd import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def test_device_probe(gpu_api):
    info = gpu_api.device_info()
    assert info["arch"].startswith("gfx950")
    assert info["reason"] == ""


def test_spmv_matches_oracle_reference_pattern(gpu_api):
    """test_gpu_backend.cpp:89-116: 90x130 pattern, products within 1e-12."""
    rows, cols = 90, 130
    ptr, idx, val = [0], [], []
    for j in range(cols):
        for i in range(j % 5, rows, 7 + j % 3):
            idx.append(i)
            val.append(np.cos(i + 3 * j))
        ptr.append(len(idx))
    x = np.sin(np.arange(cols))
    z = 0.5 - (np.arange(rows) % 4)
    say("spmv 90x130")
    gpu_api.reset_gpu_operations()
    ax = gpu_api.spmv_csc(rows, cols, ptr, idx, val, x)
    atz = gpu_api.spmv_csc(rows, cols, ptr, idx, val, z, transpose=True)
    assert gpu_api.gpu_operations() == 2
    np.testing.assert_allclose(ax, O.spmv(rows, cols, ptr, idx, val, x), rtol=0, atol=1e-12)
    np.testing.assert_allclose(atz, O.spmv(rows, cols, ptr, idx, val, z, True), rtol=0, atol=1e-12)


def test_spmv_golden_and_large(gpu_api):
    k = G.kernels()
    say("spmv golden 90x130")
    A = gpu_api.SparseMatrix(90, 130, k["spmv_ptr"], k["spmv_idx"], k["spmv_val"])
    np.testing.assert_allclose(A.multiply(k["spmv_x"]), k["spmv_ax"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(A.multiply_transpose(k["spmv_z"]), k["spmv_atz"], rtol=1e-13, atol=1e-13)
    A.close()
    lp = P.random_sparse_lp(8192, 32768, 6, seed=P.BASE_SEED)
    say("spmv 8192x32768")
    A = gpu_api.SparseMatrix(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values)
    x = np.random.default_rng(1).standard_normal(lp.n)
    y = np.random.default_rng(2).standard_normal(lp.m)
    ref_ax = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, x)
    ref_aty = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, y, True)
    np.testing.assert_allclose(A.multiply(x), ref_ax, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(A.multiply_transpose(y), ref_aty, rtol=1e-12, atol=1e-12)
    # a column-free and a row-free edge: zero columns and rows
    A.close()
    ptr = np.array([0, 0, 2, 2], dtype=np.int32)
    B = gpu_api.SparseMatrix(4, 3, ptr, [1, 3], [2.0, -1.0])
    np.testing.assert_array_equal(B.multiply([5.0, 1.0, 7.0]), [0.0, 2.0, 0.0, -1.0])
    np.testing.assert_array_equal(B.multiply_transpose([1.0, 1.0, 1.0, 1.0]), [0.0, 1.0, 0.0])
    B.close()


def lu_test_matrix(n):
    """test_gpu_backend.cpp:31-40"""
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    return np.sin(0.37 * (i * n + j) + 1.0) + np.where((i + 1) % n == j, 3.0, 0.0)


def test_dense_lu_matches_cpu(gpu_api):
    """test_gpu_backend.cpp:65-77: n=150 solve within 1e-9, one GPU factorization."""
    n = 150
    a = lu_test_matrix(n)
    b = 1.0 + (np.arange(n) % 7)
    say("dense lu 150")
    gpu_api.reset_gpu_operations()
    st, lu, piv = gpu_api.dense_lu(a)
    assert st == 0
    assert gpu_api.gpu_factorizations() == 1
    st_c, lu_c, piv_c = O.dense_lu(a)
    x_gpu = O.dense_lu_solve(lu, piv, b)
    x_cpu = O.dense_lu_solve(lu_c, piv_c, b)
    np.testing.assert_allclose(x_gpu, x_cpu, rtol=0, atol=1e-9)
    np.testing.assert_array_equal(piv, piv_c)


@pytest.mark.parametrize("n", [5, 64, 65, 300, 1000])
def test_dense_lu_reconstruction(gpu_api, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    say(f"dense lu reconstruction {n}")
    st, lu, piv = gpu_api.dense_lu(a)
    assert st == 0
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    np.testing.assert_allclose(L @ U, a[piv], atol=1e-10 * n)
    _, _, piv_c = O.dense_lu(a)
    np.testing.assert_array_equal(piv, piv_c)


def test_dense_lu_reports_singular(gpu_api):
    """test_gpu_backend.cpp:79-87"""
    n = 40
    a = lu_test_matrix(n)
    a[:, 5] = 2.0 * a[:, 3]
    say("dense lu singular")
    assert gpu_api.dense_lu(a)[0] == 1
    assert gpu_api.dense_lu(np.zeros((3, 3)))[0] == 1


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_normal_equations_each_tile_shape(gpu_api, monkeypatch, shape, precision):
    """Every MFMA tile shape the autotuner can pick (forced through
    SOVEREIGN_GEMM_SHAPE) gives the oracle's solution; m = 1000 leaves
    partial tiles at every edge."""
    lp = P.random_sparse_lp(1000, 4000, 6, seed=5)
    rng = np.random.default_rng(11)
    d = np.exp(rng.uniform(-3, 3, size=lp.n))
    rhs = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, rng.standard_normal(lp.n))
    ne_c = O.NormalEquations(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, "dense")
    assert ne_c.factor(d)
    at_c = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, ne_c.solve(rhs), True)
    monkeypatch.setenv("SOVEREIGN_GEMM_SHAPE", str(shape))
    say(f"normal equations tile shape {shape} {precision}")
    A = gpu_api.SparseMatrix(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values)
    ne = gpu_api.NormalEquations(A)
    assert ne.factor(d, precision)
    at_g = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, ne.solve(rhs), True)
    ne.close()
    np.testing.assert_allclose(at_g, at_c, rtol=0, atol=1e-7 * max(1.0, np.max(np.abs(at_c))))


def test_autotuner_reports_shapes(gpu_api):
    """After factorizations in both precisions the tuner has chosen a shape
    per product class and timed every candidate."""
    import ctypes as C

    from sovereign_amd import _capi

    shapes = np.full(4, -2, dtype=np.int32)
    times = np.zeros(12)
    _capi.load().sov_gemm_shapes(shapes.ctypes.data_as(C.POINTER(C.c_int32)), times.ctypes.data_as(C.POINTER(C.c_double)))
    assert set(shapes[[0, 2]]) <= {0, 2} and set(shapes[[1, 3]]) <= {0, 1, 2, 3}, shapes
    assert np.all(times > 0), times


@pytest.mark.parametrize("case", ["random_300x1200", "random_1000x4000", "transport_50x50", "plan_600x900"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_normal_equations_solve(gpu_api, case, precision):
    """factor(d) / solve(rhs) against the oracle's NormalEquations on the
    same IPM-form matrix and diagonal. M may be singular (transport's
    redundant row), so the check is on A^T dy, which is unique."""
    if case.startswith("random"):
        m, n = (int(v) for v in case.split("_")[1].split("x"))
        lp = P.random_sparse_lp(m, n, 6, seed=3)
        m_, n_, cp, ri, va = lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values
    else:
        model = P.transport_balanced(50, 50) if case == "transport_50x50" else G.load_fixture(case)
        f = O.ipm_form(model)
        m_, n_, cp, ri, va = f["m"], f["n"], f["col_ptr"], f["row_idx"], f["values"]
    rng = np.random.default_rng(7)
    for scale in (1.0, 1e-4):
        d = np.exp(rng.uniform(-3, 3, size=n_)) * scale
        rhs = O.spmv(m_, n_, cp, ri, va, rng.standard_normal(n_))  # consistent right-hand side
        ne_c = O.NormalEquations(m_, n_, cp, ri, va, "dense" if m_ <= 1200 else "sparse")
        assert ne_c.factor(d)
        x_c = ne_c.solve(rhs)
        say(f"normal equations {case} {precision} scale={scale}")
        A = gpu_api.SparseMatrix(m_, n_, cp, ri, va)
        ne = gpu_api.NormalEquations(A)
        assert ne.factor(d, precision)
        x_g = ne.solve(rhs)
        at_g = O.spmv(m_, n_, cp, ri, va, x_g, True)
        at_c = O.spmv(m_, n_, cp, ri, va, x_c, True)
        np.testing.assert_allclose(at_g, at_c, rtol=0, atol=1e-7 * max(1.0, np.max(np.abs(at_c))))
        # residual of M x = rhs with the FP64 matrix-free product
        Mx = O.spmv(m_, n_, cp, ri, va, d * at_g) + 1e-12 * x_g
        assert np.max(np.abs(Mx - rhs)) <= 1e-8 * max(1.0, np.max(np.abs(rhs)))
        ne.close()
        A.close()


def test_normal_equations_dropped_pivot(gpu_api):
    """A balanced transport problem has one redundant row, so M = A D A^T
    is singular up to the 1e-12 regularization. With D ~ 1e3 that row's
    FP64 pivot falls under the drop threshold 1e-13 |d0|: the factorization
    must replace it (the pivot rule of sparse_ldlt.cpp:165-173), count it,
    and still solve consistent right-hand sides (checked on A^T dy)."""
    model = P.transport_balanced(20, 20)
    f = O.ipm_form(model)
    m_, n_, cp, ri, va = f["m"], f["n"], f["col_ptr"], f["row_idx"], f["values"]
    rng = np.random.default_rng(11)
    d = 1e3 * np.exp(rng.uniform(-1, 1, size=n_))
    rhs = O.spmv(m_, n_, cp, ri, va, rng.standard_normal(n_))
    ne_c = O.NormalEquations(m_, n_, cp, ri, va, "sparse")  # the LDL^T with the pivot rule
    assert ne_c.factor(d)
    assert ne_c.replaced_pivots == 1, ne_c.replaced_pivots
    x_c = ne_c.solve(rhs)
    say("normal equations transport 20x20 dropped pivot")
    A = gpu_api.SparseMatrix(m_, n_, cp, ri, va)
    ne = gpu_api.NormalEquations(A)
    assert ne.factor(d, "fp64")
    assert ne.replaced_pivots >= 1, ne.replaced_pivots
    x_g = ne.solve(rhs)
    at_g = O.spmv(m_, n_, cp, ri, va, x_g, True)
    at_c = O.spmv(m_, n_, cp, ri, va, x_c, True)
    np.testing.assert_allclose(at_g, at_c, rtol=0, atol=1e-7 * max(1.0, np.max(np.abs(at_c))))
    ne.close()
    A.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_solve_graph_replays_match_eager(gpu_api, monkeypatch, precision):
    """The solve sequences replayed as HIP graphs (the triangular solve pair;
    for the FP32 factor, one CG iteration with its solve inside) give
    bit-for-bit the eager launches' result on every replay, with eager work
    (SpMV products) and host synchronisations between the replays, and again
    after a refactorization with another diagonal. m = 2100 leaves a partial
    last super-block."""
    lp = P.random_sparse_lp(2100, 8400, 6, seed=9)
    rng = np.random.default_rng(21)
    A = gpu_api.SparseMatrix(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values)
    ne = gpu_api.NormalEquations(A)
    for refactor in range(2):
        d = np.exp(rng.uniform(-3, 3, size=lp.n))
        assert ne.factor(d, precision)
        for rep in range(4):
            say(f"graph replay {precision} factorization {refactor} replay {rep}")
            rhs = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, rng.standard_normal(lp.n))
            monkeypatch.setenv("SOVEREIGN_GRAPHS", "1")
            x_graph = ne.solve(rhs)
            A.multiply(rng.standard_normal(lp.n))  # eager device work between the replays
            monkeypatch.setenv("SOVEREIGN_GRAPHS", "0")
            x_eager = ne.solve(rhs)
            A.multiply_transpose(rng.standard_normal(lp.m))
            np.testing.assert_array_equal(x_graph, x_eager)
            at = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, x_graph, True)
            Mx = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, d * at) + 1e-12 * x_graph
            assert np.max(np.abs(Mx - rhs)) <= 1e-8 * max(1.0, np.max(np.abs(rhs)))
    ne.close()
    A.close()
