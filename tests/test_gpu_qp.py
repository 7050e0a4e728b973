"""The structured separable QP on the GPU (sov_qp_liswet_solve) against the
reference's solved instances, brute force and the oracle.

Bar: certified optimal (certify_liswet's five checks within 1e-8 x scale),
objective within 1e-12 relative of the reference's, x within 1e-9 of the
reference's x, for both the mixed (FP32 factor + FP64 refinement) and the
plain FP64 paths.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from sovereign_amd import qp

from test_qp_oracle import brute_force, small_cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "data", "liswet_fixtures.npz"))
with open(os.path.join(HERE, "golden", "liswet_golden.json")) as _fh:
    GOLD = json.load(_fh)["cases"]

PRECISIONS = ["mixed", "fp64"]


def run(target, precision="mixed", **kw):
    print(f"qp solve n={len(target)} precision={precision}", flush=True)
    return qp.solve_liswet(target, qp.QpInteriorPointOptions(precision=precision, **kw))


def summary(r):
    keys = ("status", "ipm_status", "iterations", "polish_fits", "fp32_factorizations", "fp64_factorizations",
            "refinements", "objective", "stationarity", "complementarity", "min_lambda", "ipm_primal", "ipm_dual",
            "ipm_complementarity", "kernel_ms")
    return {k: r[k] for k in keys}


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n", [200, 1000, 10000, 20000])
def test_reference_instances(n, precision):
    target, xref = FIX[f"target_{n}"], FIX[f"x_{n}"]
    r = run(target, precision)
    assert r["status"] == 0, summary(r)
    gold = float(GOLD[str(n)]["objective"])
    assert abs(r["objective"] - gold) <= 1e-12 * abs(gold), summary(r)
    assert np.max(np.abs(r["x"] - xref)) <= 1e-9, summary(r)
    cert, lam = O.liswet_certify(target, r["x"])
    assert cert["ok"], cert
    np.testing.assert_allclose(r["lam"], lam, rtol=1e-9, atol=1e-9 * np.max(np.abs(lam)))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_brute_force_small(precision):
    worst = 0.0
    for c in small_cases(24, seed=4242):
        ref = brute_force(c)
        r = run(c, precision)
        assert r["status"] == 0, summary(r)
        worst = max(worst, float(np.max(np.abs(r["x"] - ref)) / max(1.0, np.max(np.abs(ref)))))
    assert worst <= 1e-8


@pytest.mark.parametrize("n", [3, 4, 5])
def test_smallest_shapes(n):
    rng = np.random.default_rng(n)
    c = rng.normal(size=n)
    ref = brute_force(c)
    r = run(c)
    assert r["status"] == 0, summary(r)
    np.testing.assert_allclose(r["x"], ref, atol=1e-9)


def test_closed_forms():
    t = np.linspace(0.0, 1.0, 301)
    convex = 1.0 + 0.5 * t + 3.0 * t * t
    r = run(convex)
    assert r["status"] == 0
    np.testing.assert_allclose(r["x"], convex, atol=1e-9)
    concave = np.sin(3.0 * t)
    r = run(concave)
    assert r["status"] == 0
    np.testing.assert_allclose(r["x"], np.polyval(np.polyfit(t, concave, 1), t), atol=1e-9)


def test_synthetic_workload_matches_oracle():
    target = qp.liswet_target(20002)
    it, xo = O.liswet_ipm(target)
    assert it >= 0
    cert_o, _ = O.liswet_certify(target, xo)
    for precision in PRECISIONS:
        r = run(target, precision)
        assert r["status"] == 0, summary(r)
        assert abs(r["objective"] - cert_o["objective"]) <= 1e-12 * abs(cert_o["objective"]), summary(r)
        assert np.max(np.abs(r["x"] - xo)) <= 1e-9, summary(r)


def test_mixed_path_uses_fp32_factors():
    r = run(FIX["target_10000"], "mixed")
    assert r["fp32_factorizations"] > 0, summary(r)
    r64 = run(FIX["target_10000"], "fp64")
    assert r64["fp32_factorizations"] == 0 and r64["fp64_factorizations"] > 0, summary(r64)


def test_solver_interface():
    t = FIX["target_1000"]
    out = qp.QpInteriorPointSolver().solve(qp.liswet_model(t))
    assert out["status"] == 0 and out["status_name"] == "Optimal"
    assert np.max(np.abs(out["x"] - FIX["x_1000"])) <= 1e-9
    # off the structured shape (bounded columns): the generic interior point,
    # checked against the oracle's restatement of solve_qp_ipm
    from oracle import qp_oracle as QO

    t = FIX["target_200"][:60]
    m = qp.liswet_model(t)
    m.lb = np.zeros(len(t))
    out = qp.QpInteriorPointSolver().solve(m)
    k = len(t) - 2
    A = np.zeros((k, len(t)))
    A[np.arange(k), np.arange(k)], A[np.arange(k), np.arange(k) + 1], A[np.arange(k), np.arange(k) + 2] = 1, -2, 1
    ref = QO.qp_solve(A, ["G"] * k, np.zeros(k), m.lb, np.full(len(t), 1e30), -t, np.eye(len(t)))
    assert out["status"] == ref["status"] == 0
    assert out["objective"] == pytest.approx(ref["objective"], rel=1e-8, abs=1e-10)
    assert np.all(out["x"] >= -1e-9)
