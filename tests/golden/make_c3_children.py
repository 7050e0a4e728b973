"""Writes tests/golden/c3_children.json: the strong-branching children of the
root node of the n=500, m=200 multidimensional knapsack (seed 26119), solved
by the reference's own CPU interior-point solver (oracle/_ref, compiled from
the reference sources by oracle/Makefile) where it is built, else by the
oracle's restatement of it (the two agree bit for bit on the IPM fixtures).

Records the root LP (status, objective, primal to 17 digits), the 64
fractional candidates in the order pick_strong_branch takes them
(solver/milp/src/branch_and_bound.cpp:731-748: closeness of the fractional
part to 1/2, stable), and for each the down (ub = floor x) and up
(lb = ceil x) child LP's status and objective. The GPU test feeds the same
root point to the batched strong branching and compares all 128 children.

Usage: python tests/golden/make_c3_children.py [--workers 8]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sovereign_amd import problems as P  # noqa: E402

N, M, CANDIDATES = 500, 200, 64


def solver():
    return O.ref_lp_solve if O.ref_available() else O.lp_solve


def child(args):
    var, side, xv = args
    model = P.multi_knapsack(N, M, seed=P.BASE_SEED)
    lb, ub = model.lb.copy(), model.ub.copy()
    if side == "down":
        ub[var] = np.floor(xv)
    else:
        lb[var] = np.ceil(xv)
    r = solver()(dataclasses.replace(model, lb=lb, ub=ub, _keep=[]))
    return var, side, int(r["status"]), repr(float(r["objective"])), int(r["iterations"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args()
    model = P.multi_knapsack(N, M, seed=P.BASE_SEED)
    root = solver()(model)
    assert root["status"] == 0
    x = root["primal"]
    frac = np.minimum(x - np.floor(x), np.ceil(x) - x)
    cand = [j for j in range(N) if frac[j] > 1e-6]
    cand.sort(key=lambda j: abs(frac[j] - 0.5))  # list.sort is stable
    cand = cand[:CANDIDATES]
    jobs = [(j, s, float(x[j])) for j in cand for s in ("down", "up")]
    with Pool(args.workers) as pool:
        out = pool.map(child, jobs)
    kids = {f"{v}:{s}": {"status": st, "objective": obj, "iterations": it} for v, s, st, obj, it in out}
    doc = {
        "instance": {"generator": "multi_knapsack", "n": N, "m": M, "seed": P.BASE_SEED},
        "solver": "reference (oracle/_ref)" if O.ref_available() else "oracle restatement",
        "root": {"status": int(root["status"]), "objective": repr(float(root["objective"])),
                 "primal": [repr(float(v)) for v in x]},
        "candidates": [int(j) for j in cand],
        "children": kids,
    }
    with open(os.path.join(HERE, "c3_children.json"), "w") as fh:
        json.dump(doc, fh, indent=0)
    print(f"{len(kids)} children written; root objective {root['objective']!r}")


if __name__ == "__main__":
    main()
