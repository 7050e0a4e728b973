"""Benchmark: interior-point solve time on the m=8192 random sparse LP.

One step = one full Mehrotra IPM solve to a 1e-8 relative gap (and 1e-8
primal/dual residuals) of the LP
  m = 8192 rows, n = 32768 columns, 6 nonzeros per column at uniform rows,
  values N(0,1), b = A x0, c = A'y0 + s0 (seed 26119),
with the matrix already resident in HBM (uploaded and analysed once before the
timed region). `value` is the mixed-precision GPU path (FP32 MFMA
factorizations while mu >= 1e-6, FP64 residuals and PCG refinement); the plain
FP64 GPU path is timed in the same run and reported beside it.

Roofline: the dominant kernel is the MFMA trailing update (SYRK) of the
blocked Cholesky; its algorithmic flops (lower triangle, M(M+1)K per launch)
over its HIP-event-timed duration during the timed region.

CPU baseline: the oracle's restatement of the reference CPU interior-point
path (minimum-degree ordered sparse LDL^T, single thread) on a bounded sample
of the same generator, extrapolated to m = 8192 by the cube of the size ratio
(the factorization dominates and its fill is near-dense on this structure).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "IPM solve s & strong-branch LP solves/s at 1e-8 gap vs plain FP64 GPU path"
MFMA_PEAK = {"f64": 78.6, "f32": 157.3}  # TFLOP/s, MI355X dense matrix peaks


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_solves(ipm, opts, steps, warmup, lib):
    for w in range(warmup):
        r = ipm.solve(opts)
        log(f"warmup {opts.precision}: {r['status_name']} it={r['iterations']} {r['seconds']:.3f}s")
    lib.sov_sync()
    times = np.zeros(7)
    lib.sov_kernel_times(times.ctypes.data_as(C.POINTER(C.c_double)))  # reset
    lib.sov_set_kernel_timing(1)
    out = []
    t0 = time.perf_counter()
    for k in range(steps):
        r = ipm.solve(opts)
        out.append(r)
        log(f"step {k} {opts.precision}: {r['status_name']} it={r['iterations']} obj={r['objective']:.12g} "
            f"{r['seconds']:.3f}s fp32_factors={r['fp32_factorizations']}/{r['factorizations']} "
            f"pcg={r['refine_iterations']}")
    lib.sov_sync()
    wall = time.perf_counter() - t0
    lib.sov_set_kernel_timing(0)
    lib.sov_kernel_times(times.ctypes.data_as(C.POINTER(C.c_double)))
    return out, wall, times


def cpu_baseline(sample_m: int):
    from oracle import oracle as O
    from sovereign_amd import problems as P

    lp = P.random_sparse_lp(sample_m, 4 * sample_m, 6, seed=P.BASE_SEED)
    t0 = time.perf_counter()
    r = O.solve_form(lp, "auto", feasibility_tol=1e-8, optimality_tol=1e-8)
    dt = time.perf_counter() - t0
    factor = (8192 / sample_m) ** 3
    return {
        "value": dt * factor, "unit": "s per IPM solve (extrapolated to m=8192)", "cores": 1, "kind": "port",
        "sample": (f"oracle restatement of the reference CPU IPM (AMD-ordered sparse LDL^T), full solve of the "
                   f"same generator at m={sample_m}, n={4 * sample_m}: {dt:.2f} s, {r['iterations']} iterations, "
                   f"status {r['status']}; x (8192/{sample_m})^3"),
        "host_cpus": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-m", type=int, default=2048)
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--n", type=int, default=32768)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is in DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    from sovereign_amd import _capi, api, problems as P

    lib = _capi.load()
    info = api.device_info()
    log(f"device {info['name']} {info['arch']} available={info['available']} {info['reason']}")
    if not info["available"]:
        sys.exit(f"no usable GPU: {info['reason']}")

    t = time.perf_counter()
    lp = P.random_sparse_lp(args.m, args.n, 6, seed=P.BASE_SEED)
    log(f"generated m={lp.m} n={lp.n} nnz={lp.col_ptr[-1]} in {time.perf_counter() - t:.2f}s")
    t = time.perf_counter()
    ipm = api.ResidentIpm(lp)
    lib.sov_sync()
    setup_s = time.perf_counter() - t
    log(f"resident setup (upload + normal-equation plan) {setup_s:.3f}s")

    tol = dict(feasibility_tol=1e-8, optimality_tol=1e-8)
    mixed = api.InteriorPointOptions(precision="mixed", **tol)
    fp64 = api.InteriorPointOptions(precision="fp64", **tol)
    res_mx, wall_mx, kt_mx = timed_solves(ipm, mixed, args.steps, args.warmup, lib)
    res_64, wall_64, kt_64 = timed_solves(ipm, fp64, args.steps, args.warmup, lib)
    all_optimal = all(r["status"] == 0 for r in res_mx + res_64)
    if not all_optimal:
        log("WARNING: a timed solve did not reach optimality: "
            + str([(r["status_name"], r["duality_gap"]) for r in res_mx + res_64]))
    log(f"kernel ms (mixed run): syrk {kt_mx[0]:.1f} ({int(kt_mx[1])} launches) panel {kt_mx[3]:.1f} "
        f"solve {kt_mx[4]:.1f}; (fp64 run): syrk {kt_64[0]:.1f} ({int(kt_64[1])}) panel {kt_64[3]:.1f} "
        f"solve {kt_64[4]:.1f}")
    sec_mx = wall_mx / args.steps
    sec_64 = wall_64 / args.steps
    obj_gap = abs(res_mx[-1]["objective"] - res_64[-1]["objective"]) / max(1.0, abs(res_64[-1]["objective"]))

    # dominant kernel: the trailing-update SYRK with the most time in the mixed run
    def roof(kt, dtype):
        ms, launches, flops = kt[0], kt[1], kt[2]
        if ms <= 0 or launches <= 0:
            return None
        achieved = flops / (ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK[dtype], "unit": "TFLOP/s",
                "frac": achieved / MFMA_PEAK[dtype], "traffic": None,
                "kernel": f"gemm_update_kernel<{'double' if dtype == 'f64' else 'float'}> (SYRK trailing update)",
                "launches": int(launches), "avg_launch_us": ms * 1e3 / launches,
                "flops_per_launch": flops / launches}

    roof64 = roof(kt_64, "f64")
    line = {
        "metric": METRIC,
        "value": sec_mx,
        "unit": "s per IPM solve (m=8192 LP to 1e-8 gap)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec_mx * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 residuals/refinement + f32 factorizations (mu>=1e-6)",
        "data": "synthetic",
        "config": {"workload": f"random sparse LP m={lp.m} n={lp.n} 6 nnz/col seed {P.BASE_SEED}",
                   "gap_tol": 1e-8, "path": "mixed precision (value); plain FP64 alongside"},
        "all_optimal": all_optimal,
        "fp64_path_s": sec_64,
        "speedup_vs_fp64": sec_64 / sec_mx,
        "objective": res_64[-1]["objective"], "objective_rel_diff_mixed_vs_fp64": obj_gap,
        "iterations": {"mixed": res_mx[-1]["iterations"], "fp64": res_64[-1]["iterations"]},
        "factorizations": {"mixed_fp32": res_mx[-1]["fp32_factorizations"],
                           "mixed_total": res_mx[-1]["factorizations"], "fp64": res_64[-1]["factorizations"]},
        "pcg_iterations_mixed": res_mx[-1]["refine_iterations"],
        "max_solve_residual_mixed": res_mx[-1]["max_solve_residual"],
        "resident_setup_s": setup_s,
        "kernel_ms_fp64_run": {"syrk": kt_64[0], "panel": kt_64[3], "solve": kt_64[4]},
        "kernel_ms_mixed_run": {"syrk": kt_mx[0], "panel": kt_mx[3], "solve": kt_mx[4]},
        "roofline": roof64,
        "roofline_f32_syrk": roof(kt_mx, "f32") if res_mx[-1]["fp32_factorizations"] else None,
    }
    if not args.no_cpu:
        log(f"cpu baseline sample m={args.cpu_sample_m}")
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_m)
    ipm.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
