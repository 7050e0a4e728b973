"""Writes the structured-QP fixtures (run in the build container).

Copies the reference's own solved instances of the separable convex QP
  min 0.5 |x|^2 - target'x  s.t.  x_r - 2 x_{r+1} + x_{r+2} >= 0
(benchmarks/reports/liswet_step6/n{200,1000,10000,20000}/: input_c.txt is the
target, x.txt the reference's certified solution, independent.json the
objective from the reference's independent high-precision check) into
data/liswet_fixtures.npz (float64 arrays target_<n>, x_<n>) and
liswet_golden.json (objectives as the decimal strings of the reference).
Data only: the text files are read as numbers.

Usage: python tests/golden/make_liswet_fixtures.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = (200, 1000, 10000, 20000)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    base = os.path.join(args.reference, "benchmarks", "reports", "liswet_step6")
    arrays = {}
    golden = {"source": "benchmarks/reports/liswet_step6", "cases": {}}
    for n in SIZES:
        d = os.path.join(base, f"n{n}")
        target = np.loadtxt(os.path.join(d, "input_c.txt"), dtype=np.float64)
        x = np.loadtxt(os.path.join(d, "x.txt"), dtype=np.float64)
        with open(os.path.join(d, "independent.json")) as fh:
            ind = json.load(fh)["independent"]
        assert len(target) == len(x) == n + 2
        arrays[f"target_{n}"] = target
        arrays[f"x_{n}"] = x
        golden["cases"][str(n)] = {
            "variables": n + 2,
            "rows": n,
            "objective": ind["objective_qp"],
            "lambda_inf": ind["lambda_inf"],
        }
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
    np.savez_compressed(os.path.join(HERE, "data", "liswet_fixtures.npz"), **arrays)
    with open(os.path.join(HERE, "liswet_golden.json"), "w") as fh:
        json.dump(golden, fh, indent=1)
        fh.write("\n")
    print("wrote", sorted(arrays))


if __name__ == "__main__":
    main()
