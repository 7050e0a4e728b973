"""Writes tests/golden/cut_golden.json: seeded models, points and the cuts the
reference's own generators return for them (generate_cover_cuts,
generate_mir_cuts, generate_cmir_cuts, solver/cuts/src/cuts.cpp, compiled
into oracle/_ref by oracle/Makefile), and check_cut_validity's verdicts on
planted cuts.

The models mix binary, general-integer and continuous columns, negative
lower bounds, integral and fractional coefficients and all row senses; the
points put binaries near 1 so that covers fire. Two multidimensional
knapsacks at their LP optima (the oracle's interior point) add the
tree search's own workload.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_cut_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sovereign_amd import problems as P  # noqa: E402


def random_model(seed, nv=30, nr=12):
    rng = np.random.default_rng(seed)
    kind = rng.choice([0, 1, 2], size=nv, p=[0.25, 0.25, 0.5]).astype(np.int8)
    lb = np.where(kind == 2, 0.0, np.where(rng.random(nv) < 0.3, -rng.integers(1, 4, nv).astype(float), 0.0))
    ub = np.where(kind == 2, 1.0, lb + rng.integers(2, 9, nv))
    cols_of = [[] for _ in range(nv)]
    rhs = []
    for r in range(nr):
        cols = rng.choice(nv, size=rng.integers(3, 10), replace=False)
        vals = rng.integers(1, 30, len(cols)).astype(float)
        if rng.random() < 0.5:
            vals = vals + rng.choice([0.0, 0.25, 0.5, 0.75], len(cols))
        vals *= np.where(rng.random(len(cols)) < 0.2, -1, 1)
        for j, a in zip(cols, vals):
            cols_of[j].append((r, float(a)))
        rhs.append(0.6 * np.abs(vals).sum() + rng.random())
    senses = "".join(rng.choice(list("LLLGE"), size=nr))
    x = lb + rng.random(nv) * (ub - lb)
    x = np.where(kind == 2, np.clip(rng.random(nv) * 0.6 + 0.5, 0, 1), x)
    cp, ri, vv = [0], [], []
    for j in range(nv):
        for r, a in sorted(cols_of[j]):
            ri.append(r)
            vv.append(a)
        cp.append(len(ri))
    model = P.LpModel(n_vars=nv, n_rows=nr, col_ptr=np.array(cp), row_idx=np.array(ri), values=np.array(vv),
                      obj=np.zeros(nv), lb=lb, ub=ub, row_sense=senses.encode(), rhs=np.array(rhs))
    return model, kind, x


def enc_model(m):
    return {"n_vars": m.n_vars, "n_rows": m.n_rows, "col_ptr": m.col_ptr.tolist(), "row_idx": m.row_idx.tolist(),
            "values": m.values.tolist(), "lb": m.lb.tolist(), "ub": m.ub.tolist(), "row_sense": m.row_sense.decode(),
            "rhs": m.rhs.tolist()}


def enc_cuts(cs):
    return [{"cols": list(c.keys()), "vals": list(c.values()), "rhs": r} for c, r in cs]


def main():
    cases = []
    for seed in range(10, 22):
        m, kind, x = random_model(seed)
        cases.append(("mixed_%d" % seed, m, kind, x))
    for seed in (1, 2):
        m = P.multi_knapsack(60, 8, seed=seed)
        x = O.lp_solve(m)["primal"]
        cases.append(("mkp60x8_%d" % seed, m, np.full(m.n_vars, 2, dtype=np.int8), x))
    out = []
    for name, m, kind, x in cases:
        entry = {"name": name, "model": enc_model(m), "kind": kind.tolist(), "x": [float(v) for v in x], "cuts": {}}
        for fam in (1, 2, 4):
            entry["cuts"][str(fam)] = enc_cuts(O.ref_separate_cuts(m, kind, x, fam))
        # the gate on planted cuts: a valid one, one cutting off a known point, one the box cannot meet
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        point = np.where(kind > 0, np.round(m.lb + rng.random(m.n_vars) * (m.ub - m.lb)), m.lb)
        cols = sorted(rng.choice(m.n_vars, size=4, replace=False).tolist())
        coefs = {int(j): 1.0 for j in cols}
        act = float(sum(point[j] for j in cols))
        box_min = float(sum(m.lb[j] for j in cols))
        # (cut, with the point?): valid; cuts the point off; no box point meets it (no point given); valid >=
        planted = [(coefs, "L", act + 0.5, True), (coefs, "L", act - 1.0, True), (coefs, "L", box_min - 1.0, False),
                   (coefs, "G", act, True)]
        entry["gate"] = []
        for cf, sense, rhs, with_point in planted:
            pts = [point] if with_point else []
            verdict = O.ref_cut_validity(m, cf, sense, rhs, pts)
            entry["gate"].append({"cols": list(cf.keys()), "vals": list(cf.values()), "sense": sense, "rhs": rhs,
                                  "points": [p.tolist() for p in pts], "verdict": verdict})
        out.append(entry)
        print(name, {f: len(v) for f, v in entry["cuts"].items()}, [g["verdict"] for g in entry["gate"]])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cut_golden.json")
    with open(path, "w") as fh:
        json.dump({"source": "reference generate_*_cuts / check_cut_validity via oracle/_ref", "cases": out}, fh)
        fh.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
