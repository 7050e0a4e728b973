"""Loads the golden LP fixtures (compact CSC arrays converted from the
reference's benchmark datasets) and the expected reference results."""
import json
import os

import numpy as np

from sovereign_amd import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load_fixture(name: str) -> P.LpModel:
    z = np.load(os.path.join(GOLDEN, "data", f"{name}.npz"), allow_pickle=False)
    nv, nr, mx = (int(v) for v in z["scalars"])
    model = P.LpModel(n_vars=nv, n_rows=nr, col_ptr=z["col_ptr"], row_idx=z["row_idx"], values=z["values"],
                      obj=z["obj"], lb=z["lb"], ub=z["ub"], row_sense=z["row_sense"].tobytes(), rhs=z["rhs"],
                      obj_constant=float(z["obj_constant"][0]), maximize=bool(mx), name=name)
    model.is_int = z["is_int"] if "is_int" in z.files else np.zeros(nv, dtype=np.int8)
    return model


def milp_expected() -> dict:
    with open(os.path.join(GOLDEN, "milp_golden.json")) as fh:
        return json.load(fh)


def synthetic(name: str) -> P.LpModel:
    table = {
        "transport_50x50": lambda: P.transport_balanced(50, 50, seed=P.BASE_SEED),
        "transport_20x30": lambda: P.transport_balanced(20, 30, seed=P.BASE_SEED + 1),
        "transport_grid_40": lambda: P.transport_grid(40),
        "transport_grid_60": lambda: P.transport_grid(60),
        "staircase_8x40": lambda: P.staircase(8, 40),
        "random_lp_300x1200": lambda: P.random_sparse_lp(300, 1200, 6, seed=P.BASE_SEED).as_model(),
        "random_lp_1000x4000": lambda: P.random_sparse_lp(1000, 4000, 6, seed=P.BASE_SEED + 2).as_model(),
    }
    return table[name]()


def model_for(key: str) -> P.LpModel:
    kind, name = key.split(":")[:2]
    return load_fixture(name) if kind == "fixture" else synthetic(name)


def digest(v: np.ndarray) -> list:
    """Sum, index-weighted mean and max magnitude of a vector (as recorded by
    tests/golden/make_golden.py)."""
    return [float(np.sum(v)), float(np.sum(v * np.arange(1, len(v) + 1)) / len(v)), float(np.max(np.abs(v)))]


def expected() -> dict:
    with open(os.path.join(GOLDEN, "ipm_golden.json")) as fh:
        return json.load(fh)


def kernels() -> dict:
    z = np.load(os.path.join(GOLDEN, "kernels_golden.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}
