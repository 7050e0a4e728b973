"""The CPU restatement of the reference's generic QP interior point
(oracle/qp_oracle.py, solve_qp_ipm of solver/qp/src/qp_interior_point.cpp:
2848-3116), pinned three ways:

  * against the reference's own QpInteriorPointSolver on seeded QPs
    (tests/golden/qp_golden.json, written by make_qp_golden.py from
    oracle/_ref): status, iteration count, objective, primal point;
  * against closed forms: a box-constrained separable QP is the clip of
    -c_j / q_j, an equality-constrained QP over free columns is one KKT
    solve, a maximisation is the negated minimisation;
  * against the KKT conditions of the returned point.
"""
import json
import os

import numpy as np
import pytest

from oracle import qp_oracle as Q

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 1e30


def golden_cases():
    with open(os.path.join(HERE, "golden", "qp_golden.json")) as fh:
        return json.load(fh)["cases"]


def dense(case):
    n, m = case["n"], case["m"]
    return (np.array(case["A"]).reshape(m, n), list(case["senses"]), np.array(case["rhs"]), np.array(case["lb"]),
            np.array(case["ub"]), np.array(case["c"]), np.array(case["Q"]).reshape(n, n), case["maximize"],
            case["constant"])


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["name"])
def test_oracle_matches_reference_goldens(case):
    r = Q.qp_solve(*dense(case))
    ref = case["ref"]
    assert r["status"] == ref["status"]
    assert r["iterations"] == ref["iterations"]
    if ref["status"] == 0:
        assert r["objective"] == pytest.approx(ref["objective"], rel=1e-12, abs=1e-12)
        np.testing.assert_allclose(r["x"], ref["x"], rtol=0, atol=1e-10)


def test_box_separable_is_clip():
    rng = np.random.default_rng(7)
    n = 25
    q = rng.uniform(0.5, 3.0, n)
    c = rng.standard_normal(n) * 3
    lb = rng.uniform(-2.0, 0.0, n)
    ub = lb + rng.uniform(0.5, 3.0, n)
    r = Q.qp_solve(np.zeros((0, n)), [], np.zeros(0), lb, ub, c, np.diag(q))
    x = np.clip(-c / q, lb, ub)
    assert r["status"] == 0
    np.testing.assert_allclose(r["x"], x, atol=1e-6)
    assert r["objective"] == pytest.approx(c @ x + 0.5 * x @ (q * x), rel=1e-8)


def test_equality_free_columns_is_one_kkt_solve():
    rng = np.random.default_rng(8)
    n, m = 12, 5
    B = rng.standard_normal((n, n))
    Qm = B.T @ B + 0.5 * np.eye(n)
    A = rng.standard_normal((m, n))
    b = rng.standard_normal(m)
    c = rng.standard_normal(n)
    K = np.block([[Qm, A.T], [A, np.zeros((m, m))]])
    x = np.linalg.solve(K, np.concatenate([-c, b]))[:n]
    r = Q.qp_solve(A, ["E"] * m, b, np.full(n, -INF), np.full(n, INF), c, Qm, constant=1.5)
    assert r["status"] == 0
    np.testing.assert_allclose(r["x"], x, atol=1e-7)
    assert r["objective"] == pytest.approx(1.5 + c @ x + 0.5 * x @ Qm @ x, rel=1e-9)


def test_maximize_is_negated_minimize():
    case = next(c for c in golden_cases() if c["name"] == "maximize_a")
    A, senses, rhs, lb, ub, c, Qm, _, const = dense(case)
    mx = Q.qp_solve(A, senses, rhs, lb, ub, c, Qm, True, const)
    mn = Q.qp_solve(A, senses, rhs, lb, ub, -c, -Qm, False, -const)
    assert mx["status"] == mn["status"] == 0
    assert mx["objective"] == pytest.approx(-mn["objective"], rel=1e-10)
    np.testing.assert_allclose(mx["x"], mn["x"], atol=1e-10)


@pytest.mark.parametrize("name", ["mixed_a", "shifted_a", "free_a"])
def test_kkt_conditions_of_solution(name):
    """Primal feasibility, and for the form's duals: stationarity
    Q x + c - A'y - s = 0 with complementarity x_j s_j ~ 0."""
    case = next(c for c in golden_cases() if c["name"] == name)
    A, senses, rhs, lb, ub, c, Qm, mx, const = dense(case)
    f = Q.build_qp_form(A, senses, rhs, lb, ub, c, Qm, mx, const)
    r = Q.qp_solve(A, senses, rhs, lb, ub, c, Qm, mx, const)
    x = r["x"]
    ax = A @ x
    for k, s in enumerate(senses):
        if s == "L":
            assert ax[k] <= rhs[k] + 1e-7
        elif s == "G":
            assert ax[k] >= rhs[k] - 1e-7
        else:
            assert abs(ax[k] - rhs[k]) <= 1e-7
    assert np.all(x >= lb - 1e-8) and np.all(x <= ub + 1e-8)
    # the form's point: structural shift removed, slacks from the rows
    xf = np.linalg.lstsq(f["A"][:, f["ns"]:], f["b"] - f["A"][:, :f["ns"]] @ (x - f["shift"]), rcond=None)[0] \
        if f["n"] > f["ns"] else np.zeros(0)
    full = np.concatenate([x - f["shift"], xf])
    grad = f["Q"] @ full + f["c"]
    y = r["y"]
    s = grad - f["A"].T @ y
    free = f["free"]
    assert np.max(np.abs(s[free])) < 1e-6 if free.any() else True
    assert np.all(s[~free] > -1e-6)
    assert np.max(np.abs(full[~free] * s[~free])) < 1e-6


def test_plant_dropped_shift_term_fails_golden():
    """Dropping Q·shift from the form's linear term (a plausible slip in
    build_qp_form) moves the shifted case away from the reference's answer."""
    case = next(c for c in golden_cases() if c["name"] == "shifted_a")
    A, senses, rhs, lb, ub, c, Qm, mx, const = dense(case)
    f = Q.build_qp_form(A, senses, rhs, lb, ub, c, Qm, mx, const)
    f["c"] = f["c"].copy()
    f["c"][:f["ns"]] = np.asarray(c)
    r = Q.solve_qp_ipm(f)
    assert abs(r["objective"] - case["ref"]["objective"]) > 1e-4


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libsovref.so")),
                    reason="oracle/_ref is built only where the reference checkout exists")
def test_oracle_matches_reference_on_fresh_seeds():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_qp_golden as G

    from oracle import oracle as O
    for make, seed in [(G.case_mixed, 9001), (G.case_shifted, 9002), (G.case_free, 9003)]:
        p = make(seed)
        args = (p["A"], p["senses"], p["rhs"], p["lb"], p["ub"], p["c"], p["Q"], p["maximize"], p["constant"])
        ref = O.ref_qp_solve(*args)
        mine = Q.qp_solve(*args)
        assert mine["status"] == ref["status"] and mine["iterations"] == ref["iterations"]
        assert mine["objective"] == pytest.approx(ref["objective"], rel=1e-12, abs=1e-12)
