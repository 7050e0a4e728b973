"""Writes the golden fixtures under tests/golden/ (run in the build container).

1. Converts a few of the reference's own LP fixtures
   (benchmarks/datasets/{scale,gpu}/*.json) into compact CSC arrays
   (data/<name>.npz): data only, read with numpy's safe loader.
2. Runs the reference's own CPU interior-point solver, compiled from its
   sources into oracle/_ref/libsovref.so (oracle/Makefile, target `ref`), on
   those fixtures and on the seeded synthetic workloads, and records status,
   iterations, objective (17 significant digits), residuals and a digest of
   the primal vector in ipm_golden.json. Also records DenseLU solutions and
   SpMV products of the reference on small seeded inputs (kernels_golden.npz).

Usage: python tests/golden/make_golden.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sovereign_amd import problems as P  # noqa: E402

FIXTURES = [
    ("scale/transport_20x20.json", "transport_20x20"),
    ("scale/transport_50x50.json", "transport_50x50"),
    ("gpu/plan_600x900.json", "plan_600x900"),
]


def json_to_model(path: str, name: str) -> P.LpModel:
    with open(path) as fh:
        d = json.load(fh)
    names = [v["name"] for v in d["variables"]]
    index = {n: j for j, n in enumerate(names)}
    nv, nr = len(names), len(d["constraints"])
    cols_r = [[] for _ in range(nv)]
    cols_v = [[] for _ in range(nv)]
    senses, rhs = [], []
    for r, con in enumerate(d["constraints"]):
        for var, val in con["linear"].items():
            cols_r[index[var]].append(r)
            cols_v[index[var]].append(float(val))
        s = con["sense"]
        senses.append("L" if s in ("<=", "le", "Le") else "G" if s in (">=", "ge", "Ge") else "E")
        rhs.append(float(con["rhs"]))
    ptr, rows, vals = P.csc_from_columns(cols_r, cols_v, nr)
    obj = np.zeros(nv)
    for var, val in d["objective"].get("linear", {}).items():
        obj[index[var]] = float(val)
    lb = np.array([float(v.get("lower_bound", 0.0)) for v in d["variables"]])
    ub = np.array([float(v.get("upper_bound", 1e30)) for v in d["variables"]])
    model = P.LpModel(n_vars=nv, n_rows=nr, col_ptr=ptr, row_idx=rows, values=vals, obj=obj, lb=lb, ub=ub,
                      row_sense="".join(senses), rhs=np.array(rhs),
                      obj_constant=float(d["objective"].get("constant", 0.0)),
                      maximize=str(d.get("sense", "minimize")).lower().startswith("max"), name=name)
    model.is_int = np.array([v.get("type", "continuous") in ("binary", "integer") for v in d["variables"]],
                            dtype=np.int8)
    return model


def save_model(model: P.LpModel, path: str):
    np.savez_compressed(path, col_ptr=model.col_ptr, row_idx=model.row_idx, values=model.values, obj=model.obj,
                        lb=model.lb, ub=model.ub, row_sense=np.frombuffer(model.row_sense, dtype=np.uint8),
                        rhs=model.rhs, scalars=np.array([model.n_vars, model.n_rows, int(model.maximize)]),
                        obj_constant=np.array([model.obj_constant]),
                        is_int=getattr(model, "is_int", np.zeros(model.n_vars, dtype=np.int8)))


def load_model(path: str, name: str) -> P.LpModel:
    z = np.load(path, allow_pickle=False)
    nv, nr, mx = (int(v) for v in z["scalars"])
    return P.LpModel(n_vars=nv, n_rows=nr, col_ptr=z["col_ptr"], row_idx=z["row_idx"], values=z["values"],
                     obj=z["obj"], lb=z["lb"], ub=z["ub"], row_sense=z["row_sense"].tobytes(), rhs=z["rhs"],
                     obj_constant=float(z["obj_constant"][0]), maximize=bool(mx), name=name)


def synthetic_models():
    return [
        P.transport_balanced(50, 50, seed=P.BASE_SEED),
        P.transport_balanced(20, 30, seed=P.BASE_SEED + 1),
        P.transport_grid(40),
        P.transport_grid(60),
        P.staircase(8, 40),
        P.random_sparse_lp(300, 1200, 6, seed=P.BASE_SEED).as_model(),
        P.random_sparse_lp(1000, 4000, 6, seed=P.BASE_SEED + 2).as_model(),
    ]


def digest(v: np.ndarray) -> list:
    return [float(np.sum(v)), float(np.sum(v * np.arange(1, len(v) + 1)) / len(v)), float(np.max(np.abs(v)))]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    if not O.ref_available():
        sys.exit("oracle/_ref/libsovref.so is missing: make -C oracle ref")
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
    models = []
    for rel, name in FIXTURES:
        m = json_to_model(os.path.join(args.reference, "benchmarks", "datasets", rel), name)
        save_model(m, os.path.join(HERE, "data", f"{name}.npz"))
        models.append((f"fixture:{name}", m))
    for m in synthetic_models():
        models.append((f"synthetic:{m.name}", m))
    # MILP fixtures of the reference; optimal objectives as published in its
    # benchmarks/reports/EVIDENCE.md:47-50 (OPTIMAL 102 and 133).
    milp = {}
    for rel, name, published in (("miplib/multi_knapsack_18x3.json", "multi_knapsack_18x3", 102.0),
                                 ("miplib/multi_knapsack_24x4.json", "multi_knapsack_24x4", 133.0)):
        m = json_to_model(os.path.join(args.reference, "benchmarks", "datasets", rel), name)
        save_model(m, os.path.join(HERE, "data", f"{name}.npz"))
        milp[name] = {"objective": published, "source": "benchmarks/reports/EVIDENCE.md:47-50"}
    with open(os.path.join(HERE, "milp_golden.json"), "w") as fh:
        json.dump(milp, fh, indent=1, sort_keys=True)
    out = {}
    for key, m in models:
        for mode in ("auto", "dense", "sparse"):
            if mode == "dense" and m.n_rows > 700:
                continue
            r = O.ref_lp_solve(m, mode)
            out[f"{key}:{mode}"] = {
                "status": int(r["status"]), "iterations": int(r["iterations"]),
                "objective": repr(float(r["objective"])), "duality_gap": float(r["duality_gap"]),
                "primal_residual": float(r["primal_residual"]), "dual_residual": float(r["dual_residual"]),
                "used_sparse": int(r["used_sparse"]), "primal_digest": digest(r["primal"]),
            }
            print(key, mode, out[f"{key}:{mode}"]["status"], out[f"{key}:{mode}"]["objective"], flush=True)
    with open(os.path.join(HERE, "ipm_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)

    # DenseLU and SpMV of the reference on seeded inputs.
    rng = np.random.default_rng(P.BASE_SEED)
    kern = {}
    for n in (7, 40, 150):
        a = rng.standard_normal((n, n)) + np.eye(n) * 0.1
        b = rng.standard_normal(n)
        st, x = O.ref_dense_lu_solve(a, b)
        st_t, xt = O.ref_dense_lu_solve(a, b, transpose=True)
        kern[f"lu{n}_a"], kern[f"lu{n}_b"], kern[f"lu{n}_x"], kern[f"lu{n}_xt"] = a, b, x, xt
    lp = P.random_sparse_lp(90, 130, 5, seed=P.BASE_SEED + 3)
    xv = rng.standard_normal(lp.n)
    zv = rng.standard_normal(lp.m)
    kern["spmv_ptr"], kern["spmv_idx"], kern["spmv_val"] = lp.col_ptr, lp.row_idx, lp.values
    kern["spmv_x"], kern["spmv_z"] = xv, zv
    kern["spmv_ax"] = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, xv, use_ref=True)
    kern["spmv_atz"] = O.spmv(lp.m, lp.n, lp.col_ptr, lp.row_idx, lp.values, zv, transpose=True, use_ref=True)
    np.savez_compressed(os.path.join(HERE, "kernels_golden.npz"), **kern)
    print("wrote", os.path.join(HERE, "ipm_golden.json"))


if __name__ == "__main__":
    main()
