"""GPU parity of the sparse normal-equations path (minimum-degree ordered
multifrontal Cholesky) through the C ABI:

  * factor + solve against the oracle's up-looking LDL^T on staircase and
    random LP patterns with an IPM-like spread of D; the factor's nnz(L)
    equals the oracle's; the dropped-pivot case of the dense path's test
    (a redundant transport row at D ~ 1e3) replaces exactly the pivots the
    oracle replaces;
  * whole interior-point solves of the reference's staircase LPs
    (benchmarks/tools/generate_sparse_scale.py shapes) on the sparse path:
    objectives against the published HiGHS optima
    (benchmarks/reports/sih-online.md:72-75) and the dense GPU path; the
    automatic choice takes the sparse path there and the dense one on a
    random LP whose factor fills in.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sovereign_amd import problems as P

pytestmark = pytest.mark.gpu

# HiGHS optima of the reference's staircase LPs (sih-online.md:72-75)
STAIRCASE_OPT = {(20, 100): 21329.70606060608, (50, 200): 106185.73575757477}


def say(msg):
    print(f"[case] {msg}", flush=True)


def form_of(model):
    f = O.ipm_form(model)
    return f["m"], f["n"], f["col_ptr"], f["row_idx"], f["values"]


@pytest.mark.parametrize("name,model", [("staircase_8x40", P.staircase(8, 40)),
                                        ("staircase_20x100", P.staircase(20, 100)),
                                        ("random_lp_300x1200", P.random_sparse_lp(300, 1200, seed=3).as_model())])
def test_sparse_normal_equations_match_oracle(gpu_api, name, model):
    m, n, cp, ri, va = form_of(model)
    rng = np.random.default_rng(5)
    d = np.exp(rng.uniform(-8, 8, n))
    rhs = O.spmv(m, n, cp, ri, va, rng.standard_normal(n))
    say(f"sparse normal equations {name} m={m}")
    A = gpu_api.SparseMatrix(m, n, cp, ri, va)
    ne = gpu_api.NormalEquations(A, path="sparse")
    info = ne.info()
    assert info["sparse"] == 1
    assert ne.factor(d, "fp64")
    x = ne.solve(rhs)
    pcp, pri = O.normal_eq_pattern(m, n, cp, ri)
    ne_c = O.NormalEquations(m, n, cp, ri, va, "sparse")
    assert ne_c.factor(d)
    xc = ne_c.solve(rhs)
    assert ne.replaced_pivots == ne_c.replaced_pivots
    st, _, _, lnnz, _ = O.ldlt_factor_solve(pcp, pri, np.ones(len(pri)))
    assert info["nnz_l"] - m == lnnz
    # compare through A^T x (the quantity the IPM uses; M may be near-singular)
    at_g = O.spmv(m, n, cp, ri, va, x, True)
    at_c = O.spmv(m, n, cp, ri, va, xc, True)
    np.testing.assert_allclose(at_g, at_c, rtol=0, atol=1e-8 * max(1.0, np.abs(at_c).max()))
    # a second factorization with another D reuses the analysis
    d2 = np.exp(rng.uniform(-4, 4, n))
    assert ne.factor(d2, "fp32")  # FP32 requests factor in FP64 on this path
    x2 = ne.solve(rhs)
    assert ne_c.factor(d2)
    np.testing.assert_allclose(O.spmv(m, n, cp, ri, va, x2, True), O.spmv(m, n, cp, ri, va, ne_c.solve(rhs), True),
                               rtol=0, atol=1e-8 * max(1.0, np.abs(at_c).max()))
    ne.close()
    A.close()


def test_sparse_dropped_pivot(gpu_api):
    model = P.transport_balanced(20, 20)
    m, n, cp, ri, va = form_of(model)
    rng = np.random.default_rng(11)
    d = 1e3 * np.exp(rng.uniform(-1, 1, size=n))
    rhs = O.spmv(m, n, cp, ri, va, rng.standard_normal(n))
    ne_c = O.NormalEquations(m, n, cp, ri, va, "sparse")
    assert ne_c.factor(d) and ne_c.replaced_pivots == 1
    x_c = ne_c.solve(rhs)
    say("sparse normal equations transport 20x20 dropped pivot")
    A = gpu_api.SparseMatrix(m, n, cp, ri, va)
    ne = gpu_api.NormalEquations(A, path="sparse")
    assert ne.factor(d, "fp64")
    assert ne.replaced_pivots == 1
    x_g = ne.solve(rhs)
    at_g = O.spmv(m, n, cp, ri, va, x_g, True)
    at_c = O.spmv(m, n, cp, ri, va, x_c, True)
    np.testing.assert_allclose(at_g, at_c, rtol=0, atol=1e-7 * max(1.0, np.max(np.abs(at_c))))
    ne.close()
    A.close()


@pytest.mark.parametrize("shape", [(20, 100), (50, 200)])
def test_staircase_ipm_sparse_path(gpu_api, shape):
    model = P.staircase(*shape)
    say(f"staircase {shape} ipm sparse")
    r = gpu_api.InteriorPointSolver(gpu_api.InteriorPointOptions(normal_path="sparse")).solve(model)
    assert r["status"] == 0 and r["sparse_factor"] == 1
    opt = STAIRCASE_OPT[shape]
    assert abs(r["objective"] - opt) <= 1e-8 * abs(opt)
    assert r["primal_residual"] <= 1e-8 and r["dual_residual"] <= 1e-8
    # the automatic choice takes the sparse path here
    ra = gpu_api.InteriorPointSolver(gpu_api.InteriorPointOptions(precision="mixed")).solve(model)
    assert ra["sparse_factor"] == 1 and ra["status"] == 0
    assert abs(ra["objective"] - opt) <= 1e-8 * abs(opt)
    if shape == (20, 100):  # the dense GPU path agrees (m = 2100)
        rd = gpu_api.InteriorPointSolver(gpu_api.InteriorPointOptions(normal_path="dense")).solve(model)
        assert rd["sparse_factor"] == 0 and rd["status"] == 0
        assert abs(rd["objective"] - r["objective"]) <= 1e-8 * abs(opt)


def test_auto_path_keeps_dense_for_filling_patterns(gpu_api):
    lp = P.random_sparse_lp(1024, 4096, seed=8)
    say("auto path random lp 1024x4096")
    r = gpu_api.InteriorPointSolver(gpu_api.InteriorPointOptions()).solve_form(lp)
    assert r["status"] == 0 and r["sparse_factor"] == 0
