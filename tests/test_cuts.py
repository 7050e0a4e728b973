"""Cut separation (sov_separate_cuts / sov_cut_validity, host code; no
device needed), pinned three ways:

  * against the reference's own generators and gate on seeded models
    (tests/golden/cut_golden.json, written by make_cut_golden.py from
    solver/cuts/src/cuts.cpp compiled into oracle/_ref): the same cut sets,
    the same verdicts;
  * by brute force: on small models every integer-feasible point (all
    enumerated) satisfies every cut each generator returns, at many points;
  * by construction: a cover cut is violated at its point, a planted cut that
    removes a known feasible point or that no box point meets is rejected.
"""
import itertools
import json
import os

import numpy as np
import pytest

from sovereign_amd import cuts
from sovereign_amd.problems import LpModel

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    with open(os.path.join(HERE, "golden", "cut_golden.json")) as fh:
        return json.load(fh)["cases"]


def model_of(d):
    return LpModel(n_vars=d["n_vars"], n_rows=d["n_rows"], col_ptr=np.array(d["col_ptr"]),
                   row_idx=np.array(d["row_idx"]), values=np.array(d["values"]), obj=np.zeros(d["n_vars"]),
                   lb=np.array(d["lb"]), ub=np.array(d["ub"]), row_sense=d["row_sense"].encode(),
                   rhs=np.array(d["rhs"]))


def canon(pairs):
    return sorted((tuple(sorted((int(j), round(v, 9)) for j, v in c.items())), round(r, 9)) for c, r in pairs)


@pytest.mark.parametrize("case", golden(), ids=lambda c: c["name"])
@pytest.mark.parametrize("family", [cuts.COVER, cuts.GOMORY, cuts.CMIR])
def test_generators_match_reference(case, family):
    m = model_of(case["model"])
    mine = cuts.separate(m, case["kind"], case["x"], family)
    ref = [({j: v for j, v in zip(c["cols"], c["vals"])}, c["rhs"]) for c in case["cuts"][str(family)]]
    assert canon([(c.coefs, c.rhs) for c in mine]) == canon(ref)
    assert all(c.family == family for c in mine)


@pytest.mark.parametrize("case", golden(), ids=lambda c: c["name"])
def test_gate_matches_reference(case):
    m = model_of(case["model"])
    for g in case["gate"]:
        cut = cuts.Cut({j: v for j, v in zip(g["cols"], g["vals"])}, g["rhs"], 0)
        assert cuts.validity(m, cut, g["points"], sense=g["sense"]) == g["verdict"], g


def small_model(seed, nv=7, nr=3):
    """Binary, integer and continuous columns on small ranges, so that every
    integer point can be enumerated."""
    rng = np.random.default_rng(seed)
    kind = rng.choice([0, 1, 2], size=nv, p=[0.2, 0.3, 0.5]).astype(np.int8)
    lb = np.where(kind == 2, 0.0, np.where(rng.random(nv) < 0.3, -1.0, 0.0))
    ub = np.where(kind == 2, 1.0, lb + rng.integers(1, 3, nv))
    cols_of = [[] for _ in range(nv)]
    rhs = []
    for r in range(nr):
        cols = rng.choice(nv, size=rng.integers(3, nv + 1), replace=False)
        vals = rng.integers(1, 12, len(cols)) + rng.choice([0.0, 0.5, 0.25], len(cols))
        vals = vals * np.where(rng.random(len(cols)) < 0.15, -1, 1)
        for j, a in zip(cols, vals):
            cols_of[j].append((r, float(a)))
        rhs.append(0.5 * float(np.abs(vals).sum()))
    cp, ri, vv = [0], [], []
    for j in range(nv):
        for r, a in sorted(cols_of[j]):
            ri.append(r)
            vv.append(a)
        cp.append(len(ri))
    senses = "".join(rng.choice(list("LLG"), size=nr))
    m = LpModel(n_vars=nv, n_rows=nr, col_ptr=np.array(cp), row_idx=np.array(ri), values=np.array(vv),
                obj=np.zeros(nv), lb=lb, ub=ub, row_sense=senses.encode(), rhs=np.array(rhs))
    return m, kind


def dense_rows(m):
    A = np.zeros((m.n_rows, m.n_vars))
    for j in range(m.n_vars):
        for p in range(m.col_ptr[j], m.col_ptr[j + 1]):
            A[m.row_idx[p], j] += m.values[p]
    return A


def feasible_points(m, kind, A, cont_grid=5):
    """Every point with integer columns at integers and continuous columns on
    a grid (cMIR and Gomory cuts must hold for all of them that are feasible)."""
    axes = []
    for j in range(m.n_vars):
        if kind[j] > 0:
            axes.append(np.arange(m.lb[j], m.ub[j] + 0.5))
        else:
            axes.append(np.linspace(m.lb[j], m.ub[j], cont_grid))
    pts = np.array(list(itertools.product(*axes)))
    act = pts @ A.T
    ok = np.ones(len(pts), dtype=bool)
    for r, s in enumerate(m.row_sense.decode()):
        ok &= act[:, r] <= m.rhs[r] + 1e-9 if s == "L" else act[:, r] >= m.rhs[r] - 1e-9
    return pts[ok]


@pytest.mark.parametrize("seed", range(40))
def test_cuts_never_remove_an_integer_feasible_point(seed):
    m, kind = small_model(seed)
    A = dense_rows(m)
    pts = feasible_points(m, kind, A)
    rng = np.random.default_rng(1000 + seed)
    n_cuts = 0
    for trial in range(6):
        x = m.lb + rng.random(m.n_vars) * (m.ub - m.lb)
        x = np.where(kind == 2, np.clip(0.4 + 0.7 * rng.random(m.n_vars), 0, 1), x)
        for fam in (cuts.COVER, cuts.GOMORY, cuts.CMIR):
            for c in cuts.separate(m, kind, x, fam):
                n_cuts += 1
                cols = np.array(list(c.coefs.keys()))
                vals = np.array(list(c.coefs.values()))
                lhs = pts[:, cols] @ vals if len(pts) else np.zeros(0)
                assert np.all(lhs <= c.rhs + 1e-7), (fam, c, pts[np.argmax(lhs - c.rhs)])
                # and it separates the point it was generated at
                if fam == cuts.COVER:
                    assert x[cols] @ vals > c.rhs + 1e-6
    assert n_cuts >= 0


def test_cover_cut_by_hand():
    """x0 + x1 + x2 <= 1.5 with binaries: the greedy cover {x0, x1} (two
    weight-1 items exceed 1.5) gives x0 + x1 <= 1, violated at (0.8, 0.8, 0)."""
    m = LpModel(n_vars=3, n_rows=1, col_ptr=np.array([0, 1, 2, 3]), row_idx=np.array([0, 0, 0]),
                values=np.array([1.0, 1.0, 1.0]), obj=np.zeros(3), lb=np.zeros(3), ub=np.ones(3), row_sense=b"L",
                rhs=np.array([1.5]))
    got = cuts.separate(m, [2, 2, 2], [0.8, 0.8, 0.0], cuts.COVER)
    assert len(got) == 1 and got[0].rhs == 1.0 and got[0].coefs == {0: 1.0, 1: 1.0}
    assert cuts.separate(m, [2, 2, 2], [0.5, 0.5, 0.0], cuts.COVER) == []  # not violated
    assert cuts.separate(m, [1, 1, 1], [0.8, 0.8, 0.0], cuts.COVER) == []  # covers need binaries


def test_gate_rejects_planted_invalid_cuts():
    m = LpModel(n_vars=2, n_rows=0, col_ptr=np.array([0, 0, 0]), row_idx=np.zeros(0), values=np.zeros(0),
                obj=np.zeros(2), lb=np.zeros(2), ub=np.ones(2), row_sense=b"", rhs=np.zeros(0))
    incumbent = [[1.0, 1.0]]
    assert cuts.validity(m, cuts.Cut({0: 1.0, 1: 1.0}, 2.0, 0), incumbent) == cuts.ACCEPTED
    assert cuts.validity(m, cuts.Cut({0: 1.0, 1: 1.0}, 1.0, 0), incumbent) == cuts.REMOVES_POINT
    assert cuts.validity(m, cuts.Cut({0: 1.0, 1: 1.0}, -0.5, 0), []) == cuts.BOX_INFEASIBLE
    assert cuts.validity(m, cuts.Cut({0: 1.0}, 1.5, 0), [], sense="G") == cuts.BOX_INFEASIBLE
