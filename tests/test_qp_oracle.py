"""The structured-QP oracle (oracle/liswet_oracle.cpp) pinned on the CPU.

- Against the reference's own solved instances (tests/golden/data/
  liswet_fixtures.npz from benchmarks/reports/liswet_step6/n{200,1000,10000,
  20000}: target and certified x; objectives from the reference's
  independent high-precision check).
- Against brute-force active-set enumeration on tiny instances (the method
  of the reference's tests/scripts/test_liswet_gi_randomized.py:17-52,
  written here from the KKT conditions): every subset of rows is tried as
  the active set, the equality-constrained projection is solved, and the
  closest KKT point wins.
- Closed forms: a convex target is its own projection; a concave one
  projects onto its least-squares line (all rows active).
- The certificate rejects planted mistakes (a perturbed or mirrored x).
Also the structure detector's acceptance rules (detect_liswet_structured,
solver/qp/src/qp_interior_point.cpp:892-967) on model variants, and the
library's host polish (sov_qp_liswet_polish, which needs no device) from
good, empty, full and perturbed active-set guesses against the same pins.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from sovereign_amd import qp

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "data", "liswet_fixtures.npz"))
with open(os.path.join(HERE, "golden", "liswet_golden.json")) as _fh:
    GOLD = json.load(_fh)["cases"]


def second_difference(k: int) -> np.ndarray:
    a = np.zeros((k, k + 2))
    r = np.arange(k)
    a[r, r] = 1.0
    a[r, r + 1] = -2.0
    a[r, r + 2] = 1.0
    return a


def brute_force(target: np.ndarray) -> np.ndarray:
    """Closest KKT point over all active sets: x = target + A_S' lam_S with
    A_S x = 0, lam_S >= 0, A x >= 0."""
    n = len(target)
    k = n - 2
    a = second_difference(k)
    best = None
    for mask in range(1 << k):
        act = [r for r in range(k) if mask >> r & 1]
        if act:
            aw = a[act]
            # (I  -A_S') (x  )   (target)
            # (A_S  0  ) (lam) = (0     )
            kkt = np.block([[np.eye(n), -aw.T], [aw, np.zeros((len(act), len(act)))]])
            try:
                sol = np.linalg.solve(kkt, np.concatenate([target, np.zeros(len(act))]))
            except np.linalg.LinAlgError:
                continue
            x, lam = sol[:n], sol[n:]
        else:
            x, lam = target.copy(), np.zeros(0)
        if (a @ x).min() < -1e-10 or lam.min(initial=0.0) < -1e-10:
            continue
        dist = float(np.sum((x - target) ** 2))
        if best is None or dist < best[0]:
            best = (dist, x)
    assert best is not None
    return best[1]


def small_cases(count: int, seed: int = 26119):
    rng = np.random.default_rng(seed)
    t = np.linspace(-1.0, 1.0, 8)
    cases = [0.3 + 0.2 * t + 0.4 * t * t, 0.3 + 0.2 * t - 0.4 * t * t]
    while len(cases) < count:
        k = int(rng.integers(5, 11))
        t = np.linspace(-1.0, 1.0, k + 2)
        scale = 10.0 ** rng.uniform(-2.0, 2.0)
        smooth = rng.uniform(-1.0, 1.0) + rng.uniform(-1.0, 1.0) * t
        noise = rng.normal(0.0, rng.uniform(0.0, 0.2), k + 2)
        cases.append(scale * (smooth + 0.3 * t * t + noise))
    return cases


@pytest.mark.parametrize("n", [200, 1000, 10000, 20000])
def test_oracle_reproduces_reference_solutions(n):
    target, xref = FIX[f"target_{n}"], FIX[f"x_{n}"]
    it, x = O.liswet_ipm(target)
    assert it >= 0
    cert, _ = O.liswet_certify(target, x)
    assert cert["ok"], cert
    gold = float(GOLD[str(n)]["objective"])
    assert abs(cert["objective"] - gold) <= 1e-12 * abs(gold)
    assert np.max(np.abs(x - xref)) <= 1e-9


@pytest.mark.parametrize("n", [200, 1000])
def test_reference_solution_certifies_and_mistakes_do_not(n):
    target, xref = FIX[f"target_{n}"], FIX[f"x_{n}"]
    cert, lam = O.liswet_certify(target, xref)
    assert cert["ok"] and cert["min_lambda"] > -1e-8
    assert np.max(np.abs(lam)) == pytest.approx(float(GOLD[str(n)]["lambda_inf"]), rel=1e-9)
    bumped = xref.copy()
    bumped[n // 2] += 1e-6
    assert not O.liswet_certify(target, bumped)[0]["ok"]
    assert not O.liswet_certify(target, xref[::-1].copy() * -1.0)[0]["ok"]


def test_oracle_matches_brute_force_on_small_cases():
    worst = 0.0
    for c in small_cases(40):
        ref = brute_force(c)
        it, x = O.liswet_ipm(c)
        assert it >= 0
        worst = max(worst, float(np.max(np.abs(x - ref)) / max(1.0, np.max(np.abs(ref)))))
    assert worst <= 1e-8


def test_closed_forms():
    t = np.linspace(0.0, 1.0, 50)
    convex = 1.0 + 0.5 * t + 3.0 * t * t
    _, x = O.liswet_ipm(convex)
    np.testing.assert_allclose(x, convex, atol=1e-9)
    concave = np.sin(3.0 * t)  # concave on [0, 1]: the projection is the LS line
    _, x = O.liswet_ipm(concave)
    line = np.polyval(np.polyfit(t, concave, 1), t)
    np.testing.assert_allclose(x, line, atol=1e-9)


def test_synthetic_workload_oracle_certifies():
    target = qp.liswet_target(2002)
    it, x = O.liswet_ipm(target)
    assert it >= 0
    assert O.liswet_certify(target, x)[0]["ok"]


def test_detector_acceptance_rules():
    t = np.linspace(0.0, 1.0, 12) ** 2
    model = qp.liswet_model(t)
    np.testing.assert_array_equal(qp.detect_liswet(model), t)

    def variant(**kw):
        m = qp.liswet_model(t)
        for k, v in kw.items():
            setattr(m, k, v)
        return m

    assert qp.detect_liswet(variant(minimize=False)) is None
    lb = np.full(12, -qp.INF)
    lb[3] = 0.0
    assert qp.detect_liswet(variant(lb=lb)) is None
    quad = {(i, i): 1.0 for i in range(12)}
    quad[(0, 1)] = 0.5
    assert qp.detect_liswet(variant(quadratic=quad)) is None
    quad = {(i, i): 1.0 for i in range(12)}
    quad[(4, 4)] = 2.0
    assert qp.detect_liswet(variant(quadratic=quad)) is None
    senses = [">="] * 10
    senses[2] = "<="
    assert qp.detect_liswet(variant(senses=senses)) is None
    rhs = np.zeros(10)
    rhs[5] = 0.1
    assert qp.detect_liswet(variant(rhs=rhs)) is None
    rows = [dict(r) for r in model.rows]
    rows[1][2] = -1.9
    assert qp.detect_liswet(variant(rows=rows)) is None
    rows = [dict(r) for r in model.rows]
    rows[1][7] = 1.0
    assert qp.detect_liswet(variant(rows=rows)) is None
    assert qp.detect_liswet(variant(rows=model.rows[:-1], senses=[">="] * 9, rhs=np.zeros(9))) is None


def test_synthetic_target_shape():
    t = qp.liswet_target(20002)
    assert t.shape == (20002,)
    u = np.arange(20002) / 20001.0
    resid = t - (np.sin(3 * u) + u * u)
    assert abs(resid.std() - 0.01) < 5e-4 and abs(resid.mean()) < 5e-4
    np.testing.assert_array_equal(t, qp.liswet_target(20002))


# ---- the library's host polish (sov_qp_liswet_polish; no device needed) ----

def _guess_arrays(active):
    """(s, z) whose guess {s < z} is `active`."""
    active = np.asarray(active, dtype=bool)
    return np.where(active, 0.5, 2.0), np.where(active, 2.0, 0.5)


def _true_active(x):
    a = x[:-2] - 2.0 * x[1:-1] + x[2:]
    return np.abs(a) <= 1e-12 * max(1.0, np.max(np.abs(x)))


@pytest.mark.parametrize("n", [200, 1000, 10000, 20000])
@pytest.mark.parametrize("guess", ["all_active", "none_active", "perturbed"])
def test_library_polish_recovers_reference_solution(n, guess):
    target, xref = FIX[f"target_{n}"], FIX[f"x_{n}"]
    truth = _true_active(xref)
    if guess == "all_active":
        act = np.ones(n, dtype=bool)
    elif guess == "none_active":
        act = np.zeros(n, dtype=bool)
    else:
        rng = np.random.default_rng(n)
        act = truth.copy()
        flip = rng.choice(n, size=max(3, n // 50), replace=False)
        act[flip] = ~act[flip]
    s, z = _guess_arrays(act)
    r = qp.polish_liswet(target, s, z)
    assert r["status"] == 0, {k: r[k] for k in ("polish_fits", "stationarity", "min_lambda", "primal")}
    gold = float(GOLD[str(n)]["objective"])
    assert abs(r["objective"] - gold) <= 1e-12 * abs(gold)
    assert np.max(np.abs(r["x"] - xref)) <= 1e-9
    cert, lam = O.liswet_certify(target, r["x"])
    assert cert["ok"]
    np.testing.assert_allclose(r["lam"], lam, rtol=1e-9, atol=1e-9 * np.max(np.abs(lam)))


def test_library_polish_matches_brute_force_from_any_guess():
    rng = np.random.default_rng(7)
    for c in small_cases(30, seed=99):
        ref = brute_force(c)
        act = rng.random(len(c) - 2) < 0.5
        r = qp.polish_liswet(c, *_guess_arrays(act))
        assert r["status"] == 0
        assert np.max(np.abs(r["x"] - ref)) <= 1e-8 * max(1.0, np.max(np.abs(ref)))


def test_library_polish_closed_forms():
    t = np.linspace(0.0, 1.0, 400)
    convex = 1.0 + 0.5 * t + 3.0 * t * t
    r = qp.polish_liswet(convex, *_guess_arrays(np.ones(398, dtype=bool)))
    assert r["status"] == 0
    np.testing.assert_allclose(r["x"], convex, atol=1e-9)
    concave = np.sin(3.0 * t)
    r = qp.polish_liswet(concave, *_guess_arrays(np.zeros(398, dtype=bool)))
    assert r["status"] == 0
    np.testing.assert_allclose(r["x"], np.polyval(np.polyfit(t, concave, 1), t), atol=1e-9)
