"""Writes tests/golden/qp_golden.json: seeded convex QPs and what the
reference's own QpInteriorPointSolver (solver/qp/src/qp_interior_point.cpp,
compiled into oracle/_ref by oracle/Makefile) returns for them: status,
iterations, objective and primal point. The QPs avoid the reference's
special shapes (the structured second-difference model, the projection and
free identity-Hessian fast paths, :3153-3165), so every case goes through its
generic solve_qp_ipm (:2848-3116) with the retry from Mehrotra's start.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_qp_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

INF = 1e30


def psd(rng, n, rank=None, ridge=1e-2):
    B = rng.standard_normal((rank or n, n))
    return B.T @ B / (rank or n) + ridge * np.eye(n)


def feasible_rows(rng, A, x0, senses):
    """rhs so that x0 satisfies every row with some slack on inequalities."""
    ax = A @ x0
    gap = rng.uniform(0.1, 1.0, size=len(senses))
    return np.array([ax[r] + (gap[r] if s == "L" else -gap[r] if s == "G" else 0.0) for r, s in enumerate(senses)])


def case_mixed(seed, n=30, m=15):
    """Bounded and unbounded-above columns, rows of every sense."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.4)
    senses = list(rng.choice(["L", "G", "E"], size=m, p=[0.4, 0.3, 0.3]))
    lb = np.zeros(n)
    ub = np.where(rng.random(n) < 0.5, rng.uniform(2.0, 5.0, n), INF)
    x0 = rng.uniform(0.2, 1.8, n)
    return dict(A=A, senses=senses, rhs=feasible_rows(rng, A, x0, senses), lb=lb, ub=ub,
                c=rng.standard_normal(n), Q=psd(rng, n, rank=n // 2), maximize=False, constant=0.0)


def case_shifted(seed, n=24, m=10):
    """Negative and positive finite lower bounds (the lower-bound shift), a
    constant term, a rank-deficient Q."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    senses = list(rng.choice(["L", "G"], size=m))
    lb = rng.uniform(-3.0, 1.0, n)
    ub = lb + rng.uniform(1.0, 4.0, n)
    x0 = lb + 0.5 * (ub - lb)
    return dict(A=A, senses=senses, rhs=feasible_rows(rng, A, x0, senses), lb=lb, ub=ub,
                c=rng.standard_normal(n), Q=psd(rng, n, rank=n // 3, ridge=0.0), maximize=False, constant=2.5)


def case_free(seed, n=20, m=8):
    """Free columns (lb <= -1e29) under a positive-definite Q, next to
    bounded ones."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    senses = list(rng.choice(["L", "G", "E"], size=m))
    lb = np.where(rng.random(n) < 0.4, -INF, 0.0)
    ub = np.full(n, INF)
    x0 = rng.uniform(0.2, 1.5, n)
    return dict(A=A, senses=senses, rhs=feasible_rows(rng, A, x0, senses), lb=lb, ub=ub,
                c=rng.standard_normal(n), Q=psd(rng, n, ridge=0.1), maximize=False, constant=0.0)


def case_maximize(seed, n=16, m=6):
    """A maximization of a concave objective."""
    rng = np.random.default_rng(seed)
    A = np.abs(rng.standard_normal((m, n)))
    senses = ["L"] * m
    lb = np.zeros(n)
    ub = np.full(n, 10.0)
    return dict(A=A, senses=senses, rhs=np.full(m, 5.0), lb=lb, ub=ub, c=rng.uniform(0.5, 2.0, n),
                Q=-psd(rng, n), maximize=True, constant=-1.0)


def case_lp(seed, n=18, m=9):
    """Q = 0: the generic path on a linear program (independent steps)."""
    rng = np.random.default_rng(seed)
    A = np.abs(rng.standard_normal((m, n))) + 0.1
    senses = ["G"] * m
    lb = np.zeros(n)
    ub = np.full(n, INF)
    return dict(A=A, senses=senses, rhs=rng.uniform(1.0, 3.0, m), lb=lb, ub=ub,
                c=rng.uniform(1.0, 2.0, n), Q=np.zeros((n, n)), maximize=False, constant=0.0)


def case_infeasible(seed, n=6):
    """x_0 + x_1 <= 1 and x_0 + x_1 >= 3: no point, the interior point fails."""
    rng = np.random.default_rng(seed)
    A = np.zeros((2, n))
    A[:, :2] = 1.0
    return dict(A=A, senses=["L", "G"], rhs=np.array([1.0, 3.0]), lb=np.zeros(n), ub=np.full(n, INF),
                c=rng.standard_normal(n), Q=psd(rng, n), maximize=False, constant=0.0)


CASES = [
    ("mixed_a", case_mixed, 101), ("mixed_b", case_mixed, 102), ("mixed_c", case_mixed, 103),
    ("shifted_a", case_shifted, 201), ("shifted_b", case_shifted, 202),
    ("free_a", case_free, 301), ("free_b", case_free, 302),
    ("maximize_a", case_maximize, 401),
    ("lp_a", case_lp, 501),
    ("infeasible_a", case_infeasible, 601),
]


def enc(v):
    return [float(t) for t in np.asarray(v, dtype=float).ravel()]


def main():
    out = []
    for name, make, seed in CASES:
        p = make(seed)
        r = O.ref_qp_solve(p["A"], p["senses"], p["rhs"], p["lb"], p["ub"], p["c"], p["Q"], p["maximize"],
                           p["constant"])
        m, n = p["A"].shape
        out.append({"name": name, "seed": seed, "n": n, "m": m, "A": enc(p["A"]), "senses": "".join(p["senses"]),
                    "rhs": enc(p["rhs"]), "lb": enc(p["lb"]), "ub": enc(p["ub"]), "c": enc(p["c"]), "Q": enc(p["Q"]),
                    "maximize": bool(p["maximize"]), "constant": float(p["constant"]),
                    "ref": {"status": r["status"], "iterations": r["iterations"], "objective": float(r["objective"]),
                            "x": enc(r["x"])}})
        print(f"{name}: status {r['status']} iterations {r['iterations']} objective {r['objective']:.15g}")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qp_golden.json")
    with open(path, "w") as fh:
        json.dump({"source": "reference QpInteriorPointSolver via oracle/_ref", "cases": out}, fh)
        fh.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
