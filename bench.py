"""Benchmark: interior-point solve time on the m=8192 random sparse LP, and
strong-branching throughput on the n=500, m=200 multidimensional knapsack.

IPM step = one full Mehrotra IPM solve to a 1e-8 relative gap (and 1e-8
primal/dual residuals) of the LP
  m = 8192 rows, n = 32768 columns, 6 nonzeros per column at uniform rows,
  values N(0,1), b = A x0 with x0 ~ U(0.1, 2), c = A'y0 + s0 (seed 26119),
with the matrix already resident in HBM (uploaded and analysed once before the
timed region). `value` is the mixed-precision GPU path (FP32 MFMA
factorizations while mu >= 1e-6, FP64 residuals and PCG refinement); the plain
FP64 GPU path is timed in the same run and reported beside it.

Strong branching: the root LP of the knapsack (500 binaries, 200 <= rows,
seed 26119) is solved, then its 64 most fractional candidates are scored with
2 x 64 = 128 child LPs, once in one persistent batched launch and once with
the plain per-LP GPU interior-point path; both must give the same scores.

Roofline: the dominant MFMA kernel of the measured (mixed) path, the bulk
trailing update (SYRK) of the blocked Cholesky, with its algorithmic flops
(lower triangle, M(M+1)K per launch) over its HIP-event-timed duration (events
on the stream the kernel runs on) during the timed region. Its HBM traffic
per launch comes from rocprofv3 FETCH_SIZE / WRITE_SIZE passes recorded in
profiles/traffic.json (null when absent).

CPU baseline: the oracle's restatement of the reference CPU interior-point
path (minimum-degree ordered sparse LDL^T, one thread): its phases timed at
m = 4096 and scaled to m = 8192 by the per-phase ratios of a full-size run
committed under profiles/ (python bench.py --cpu-phases 4096 8192 makes it).

Separable QP: 20,002 variables, 20,000 second-difference rows, solved on
the mixed and the plain FP64 path (synthetic target and the reference's own
n=20000 instance), with the oracle's one-thread solve as its CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--paths mixed,fp64] [--no-branching] [--no-qp]
       python bench.py --only-qp [--steps K] [--no-cpu]   (the QP leg alone)
       python bench.py --only-branching   (the strong-branching leg alone)
       python bench.py --sb-root-only   (the batched root strong-branching
                                         call alone, warm-up plus three
                                         timed calls: counter profiles)
       python bench.py --copy-probe   (a plain 256 MiB device copy, three
                                       times: the calibration run for the
                                       FETCH_SIZE / WRITE_SIZE counters)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "IPM solve s & strong-branch LP solves/s at 1e-8 gap vs plain FP64 GPU path"
MFMA_PEAK = {"f64": 78.6, "f32": 157.3}  # TFLOP/s, MI355X dense matrix peaks
SHAPES = {0: (64, 64), 1: (128, 64), 2: (128, 128), 3: (32, 32)}


def kernel_name(dtype, shape):
    bm, bn = SHAPES.get(shape, (64, 64))
    t = "double" if dtype == "f64" else "float"
    return f"gemm_update_kernel<{t}, false, true, false, {bm}, {bn}>"


def tuned_shapes(lib):
    """Tile shapes the factorization's autotuner picked, and its timings."""
    sh = np.full(4, -1, dtype=np.int32)
    tm = np.zeros(12)
    lib.sov_gemm_shapes(sh.ctypes.data_as(C.POINTER(C.c_int32)), tm.ctypes.data_as(C.POINTER(C.c_double)))
    names = ["64x64", "128x64", "128x128", "32x32"]
    out = {}
    for p, (dt, off) in enumerate((("f64", 0), ("f32", 6))):
        out[dt] = {"syrk": names[sh[2 * p]] if sh[2 * p] >= 0 else None,
                   "panel": names[sh[2 * p + 1]] if sh[2 * p + 1] >= 0 else None,
                   "syrk_ms": {"64x64": tm[off], "128x128": tm[off + 1]},
                   "panel_ms": {"64x64": tm[off + 2], "128x64": tm[off + 3], "128x128": tm[off + 4],
                                "32x32": tm[off + 5]}}
    return out, sh
NT = 24  # slots of sov_kernel_times
HBM_PEAK = 8000.0  # GB/s, MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_times(lib):
    t = np.zeros(NT)
    lib.sov_kernel_times(t.ctypes.data_as(C.POINTER(C.c_double)))
    return t


def timed_solves(ipm, opts, steps, warmup, lib):
    """Warmup, then `steps` timed solves with HIP events on one launch in
    seven of the bulk SYRK and of the diagonal-block kernel (timing level 2:
    their own launch events, kept off most of the factorization's critical
    path), then one more solve with every kernel class timed, for the
    breakdown. Returns (results, wall of the timed steps, SYRK times of
    the timed steps, per-class times of the extra solve)."""
    for _ in range(warmup):
        r = ipm.solve(opts)
        log(f"warmup {opts.precision}: {r['status_name']} it={r['iterations']} {r['seconds']:.3f}s")
    lib.sov_sync()
    kernel_times(lib)  # reset
    lib.sov_set_kernel_timing(2)
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
    kt = kernel_times(lib)
    lib.sov_set_kernel_timing(1)
    ipm.solve(opts)
    lib.sov_sync()
    lib.sov_set_kernel_timing(0)
    return out, wall, kt, kernel_times(lib)


def kernel_ms(kt):
    return {"syrk_f64": kt[0], "syrk_f32": kt[9], "panel_chain": kt[3], "panel_kernel_f64": kt[12],
            "panel_kernel_f32": kt[15], "strip_update": kt[7], "super_inverse": kt[8], "triangular_solves": kt[4],
            "assembly": kt[5]}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def roofline(kt, dtype, traffic, shape=0):
    """The bulk SYRK of one precision: algorithmic flops M(M+1)K per launch
    over its event-timed launch duration."""
    ms, launches, flops = (kt[0], kt[1], kt[2]) if dtype == "f64" else (kt[9], kt[10], kt[11])
    if ms <= 0 or launches <= 0:
        return None
    achieved = flops / (ms * 1e-3) / 1e12
    name = kernel_name(dtype, shape)
    tr = traffic.get(name)
    return {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK[dtype], "unit": "TFLOP/s",
            "frac": achieved / MFMA_PEAK[dtype],
            "traffic": tr["bytes_per_launch"] if tr else None,
            "kernel": name + " (bulk SYRK trailing update)",
            "launches": int(launches), "avg_launch_us": ms * 1e3 / launches,
            "flops_per_launch": flops / launches,
            "traffic_over_algorithmic_bytes": (tr["bytes_per_launch"] / tr["algorithmic_bytes_per_launch"]
                                               if tr and tr.get("algorithmic_bytes_per_launch") else None),
            "traffic_source": tr.get("source") if tr else None}


def panel_roofline(kt, dtype, traffic):
    """The diagonal-block (panel) kernel: one workgroup factors and inverts a
    128 x 128 block; algorithmic work 2 nb^3 / 3 flop per launch (factor and
    inverse) over its event-timed launch duration."""
    ms, launches, flops = (kt[12], kt[13], kt[14]) if dtype == "f64" else (kt[15], kt[16], kt[17])
    if ms <= 0 or launches <= 0:
        return None
    achieved = flops / (ms * 1e-3) / 1e12
    name = f"potrf_panel_kernel<{'double' if dtype == 'f64' else 'float'}>"
    tr = traffic.get(name)
    return {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK[dtype], "unit": "TFLOP/s",
            "frac": achieved / MFMA_PEAK[dtype], "traffic": tr["bytes_per_launch"] if tr else None,
            "kernel": name + " (diagonal block: factor + inverse, one workgroup, on the critical path)",
            "launches": int(launches), "avg_launch_us": ms * 1e3 / launches, "flops_per_launch": flops / launches,
            "traffic_over_algorithmic_bytes": (tr["bytes_per_launch"] / tr["algorithmic_bytes_per_launch"]
                                               if tr and tr.get("algorithmic_bytes_per_launch") else None),
            "traffic_source": tr.get("source") if tr else None}


def dominant_roofline(kt, traffic, shape_ids):
    """Roofline of the event-timed kernel with the largest time in the timed
    region (the panel kernel and the bulk SYRK of each precision are timed
    there by their own launch events), with the others as secondary."""
    entries = {
        "panel_f32": (kt[15], panel_roofline(kt, "f32", traffic)),
        "panel_f64": (kt[12], panel_roofline(kt, "f64", traffic)),
        "syrk_f32": (kt[9], roofline(kt, "f32", traffic, shape_ids[2])),
        "syrk_f64": (kt[0], roofline(kt, "f64", traffic, shape_ids[0])),
    }
    live = {k: v for k, v in entries.items() if v[1] is not None}
    top = max(live, key=lambda k: live[k][0])
    total = sum(v[0] for v in live.values())
    # one launch in `stride` of each kind carries events: the sampled time
    # times the stride estimates the kind's time in the timed region
    stride = kt[23] if kt[23] >= 1 else 1.0
    main = dict(live[top][1])
    main["timed_region_ms"] = live[top][0] * stride
    main["launch_sampling"] = f"1 in {int(stride)} launches of each kind timed by its own launch events"
    others = []
    for k, (ms, e) in sorted(live.items(), key=lambda kv: -kv[1][0]):
        if k != top:
            e = dict(e)
            e["timed_region_ms"] = ms * stride
            others.append(e)
    main["share_of_event_timed_kernels"] = live[top][0] / total if total > 0 else None
    return main, others


def spmv_roofline(lib, mat, m, n, reps=20):
    """Standalone CSR SpMV (y = A x and y = A' x) on the resident matrix:
    algorithmic bytes 12 nnz + 4 (rows + 1) + 8 (m + n) per product over the
    kernel's event-timed duration."""
    x = np.linspace(0.5, 1.5, n)
    y = np.linspace(-1.0, 1.0, m)
    out = {}
    for tr, vin, vout in ((0, x, np.zeros(m)), (1, y, np.zeros(n))):
        lib.sov_spmv(mat.handle, tr, K_pf(vin), K_pf(vout))  # warm-up
        lib.sov_sync()
        kernel_times(lib)
        lib.sov_set_kernel_timing(1)
        for _ in range(reps):
            lib.sov_spmv(mat.handle, tr, K_pf(vin), K_pf(vout))
        lib.sov_sync()
        kt = kernel_times(lib)
        lib.sov_set_kernel_timing(0)
        ms, launches, nbytes = kt[18], kt[19], kt[20]
        gbps = nbytes / (ms * 1e-3) / 1e9
        out["A'x" if tr else "Ax"] = {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK, "unit": "GB/s",
                                      "frac": gbps / HBM_PEAK, "avg_launch_us": ms * 1e3 / launches,
                                      "bytes_per_launch": nbytes / launches, "launches": int(launches)}
    return out


def solve_roofline(kt, dtype):
    """The blocked triangular solves (GEMV steps): the factor's lower
    triangle read by the forward and the backward pass, m (m + 1) elements
    per solve pair, over the pairs' event-timed duration (one instrumented
    solve, every kernel class timed)."""
    ms, pairs, nbytes = kt[4], kt[21], kt[22]
    if ms <= 0 or pairs <= 0:
        return None
    gbps = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK, "unit": "GB/s", "frac": gbps / HBM_PEAK,
            "solve_pairs": int(pairs), "avg_pair_us": ms * 1e3 / pairs, "bytes_per_pair": nbytes / pairs,
            "note": f"{dtype} path; FP32 and FP64 solves together in the mixed path"}


def K_pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cpu_phases(m, say=None):
    """The oracle's restatement of the reference CPU interior-point path on
    the random LP of size m (4m columns, 6 nonzeros a column, seed 26119),
    phase by phase: the NormalEquations constructor (normal-matrix pattern,
    greedy minimum-degree ordering, symbolic analysis; once per solve), one
    numeric LDL^T factorization and one solve (twice per iteration), one
    thread. Returns seconds per phase."""
    from oracle import oracle as O
    from sovereign_amd import problems as P

    lp = P.random_sparse_lp(m, 4 * m, 6, seed=P.BASE_SEED)
    L = O.lib()
    out = {"m": m}
    t = time.perf_counter()
    h = L.orc_ne_create(lp.m, lp.n, O.pi(O.i32(lp.col_ptr)), O.pi(O.i32(lp.row_idx)), O.pf(O.f64(lp.values)), 0)
    out["create_s"] = time.perf_counter() - t
    if say:
        say(f"cpu phases m={m}: create (pattern, ordering, symbolic) {out['create_s']:.2f}s")
    d = O.f64(lp.x0 / lp.s0)
    t = time.perf_counter()
    L.orc_ne_factor(h, O.pf(d))
    out["factor_s"] = time.perf_counter() - t
    out["factor_nnz"] = int(L.orc_ne_factor_nnz(h))
    if say:
        say(f"cpu phases m={m}: numeric factorization {out['factor_s']:.2f}s, nnz(L) {out['factor_nnz']}")
    r = np.ones(lp.m)
    t = time.perf_counter()
    L.orc_ne_solve(h, O.pf(r))
    out["solve_s"] = time.perf_counter() - t
    L.orc_ne_free(h)
    if say:
        say(f"cpu phases m={m}: solve {out['solve_s']:.3f}s")
    return out


def load_cpu_scaling():
    path = os.path.join(ROOT, "profiles", "cpu_oracle_ipm_phases.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def cpu_baseline(iterations, sample_m=4096):
    """The oracle's CPU path (one thread) on a bounded sample of the
    workload: its phases timed live at m = 4096, scaled to m = 8192 by the
    per-phase ratios of a full-size run on the same kind of host
    (profiles/cpu_oracle_ipm_phases.json, both sizes timed in one run), then
    one solve = the constructor + iterations x (factorization + 2 solves)."""
    ph = cpu_phases(sample_m, say=log)
    sc = load_cpu_scaling()
    if sc is None:
        return {"value": None, "unit": "s per IPM solve (m=8192)", "cores": 1, "kind": "port",
                "sample": f"phases at m={sample_m} only (no full-size scaling profile committed)", "phases": ph}
    r_create = sc["8192"]["create_s"] / sc[str(sample_m)]["create_s"]
    r_factor = sc["8192"]["factor_s"] / sc[str(sample_m)]["factor_s"]
    r_solve = sc["8192"]["solve_s"] / sc[str(sample_m)]["solve_s"]
    create = ph["create_s"] * r_create
    per_it = ph["factor_s"] * r_factor + 2 * ph["solve_s"] * r_solve
    return {
        "value": create + iterations * per_it, "unit": "s per IPM solve (m=8192 LP, 1e-8 gap)", "cores": 1,
        "kind": "port", "nproc": os.cpu_count(),
        "sample": (f"oracle restatement of the reference CPU interior-point path (NormalEquations: greedy minimum-degree "
                   f"ordering + up-looking LDL^T, one thread), phases timed live at m={sample_m} "
                   f"(constructor {ph['create_s']:.1f} s, factorization {ph['factor_s']:.2f} s, solve "
                   f"{ph['solve_s']:.3f} s), scaled to m=8192 by the ratios of a full-size run committed in "
                   f"profiles/cpu_oracle_ipm_phases.json (constructor x{r_create:.2f}, factorization x{r_factor:.2f}); "
                   f"one solve = constructor + {iterations} iterations x (factorization + 2 solves)"),
        "phases_live": ph, "per_iteration_s": per_it, "constructor_s": create,
    }


def branching_leg(lib):
    """Strong branching at the root of the n=500, m=200 knapsack: the 128
    child LPs of 64 candidates batched vs per-LP (same scores), and one
    branch-and-bound node processed three ways (per-LP full strong branching,
    batched full, batched on the surrogate's top 8)."""
    from sovereign_amd import api, milp, problems as P, surrogate as S

    model = P.multi_knapsack(500, 200, seed=P.BASE_SEED)
    is_int = np.ones(model.n_vars, dtype=np.int8)
    root = api.InteriorPointSolver(api.InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9)).solve(model)
    nfrac = int(np.sum(np.abs(root["primal"] - np.round(root["primal"])) > 1e-6))
    out = {"model": "multi_knapsack n=500 m=200 seed 26119 (root node)", "fractional_at_root": nfrac}
    milp.strong_branch(model, is_int, root["primal"], root["objective"], max_candidates=8)  # warm the batch path
    lib.sov_sync()
    t0 = time.perf_counter()
    cb, best_b, kms = milp.strong_branch(model, is_int, root["primal"], root["objective"], max_candidates=64)
    wall_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    cp, best_p, _ = milp.strong_branch(model, is_int, root["primal"], root["objective"], max_candidates=64,
                                       child_solver=milp.CHILD_PER_LP)
    wall_p = time.perf_counter() - t0
    children = 2 * len(cb)
    diff = max([abs(a[f"obj_{s}"] - b[f"obj_{s}"]) / max(1.0, abs(b[f"obj_{s}"]))
                for a, b in zip(cb, cp) for s in ("down", "up")] or [0.0])
    # algorithmic flops of the batched kernel: per child-LP iteration the
    # SYRK n m (m + 1), the Cholesky m^3 / 3 and the GEMVs 12 m n
    mi, ni = model.n_rows, model.n_vars + model.n_rows
    per_it = ni * mi * (mi + 1) + mi ** 3 / 3 + 12 * mi * ni
    iters = sum(c["iters_down"] + c["iters_up"] for c in cb)
    bflops = per_it * iters
    out["roofline_batched_kernel"] = {
        "bound": "mfma", "achieved": bflops / (kms * 1e-3) / 1e12, "peak": MFMA_PEAK["f32"], "unit": "TFLOP/s",
        "frac": bflops / (kms * 1e-3) / 1e12 / MFMA_PEAK["f32"], "kernel": "batch_ipm_kernel (one launch, 128 LPs)",
        "flops_per_child_iteration": per_it, "child_iterations": iters, "kernel_ms": kms,
        "note": "FP32 MFMA SYRK and LDS Cholesky per child; 128 children occupy 128 of 256 CUs"}
    out.update({
        "candidates": len(cb), "child_lps": children,
        "batched_wall_s": wall_b, "batched_kernel_ms": kms, "batched_lp_solves_per_s": children / wall_b,
        "per_lp_wall_s": wall_p, "per_lp_lp_solves_per_s": children / wall_p,
        "speedup_batched_vs_per_lp": wall_p / wall_b,
        "same_choice": best_b == best_p, "max_child_objective_rel_diff": diff,
        "all_children_solved": all(c["status_down"] in (0, 2) and c["status_up"] in (0, 2) for c in cb),
    })
    log(f"strong branching: {children} child LPs batched {wall_b:.3f}s ({kms:.1f} ms kernel), per-LP "
        f"{wall_p:.3f}s, same choice {best_b == best_p}, max rel diff {diff:.2e}")
    # one branch-and-bound node (the root), three ways
    node = {}
    ranker = S.load_trained(model) if os.path.exists(S.WEIGHTS) else None
    setups = [("per_lp_full_64", dict(strong_branch_candidates=64, child_solver=milp.CHILD_PER_LP)),
              ("batched_full_64", dict(strong_branch_candidates=64))]
    if ranker is not None:
        setups.append(("batched_surrogate_top8", dict(strong_branch_candidates=8, ranker=ranker)))
    for name, kw in setups:
        t0 = time.perf_counter()
        r = milp.BranchAndBoundSolver(max_nodes=1, branch_rule="full_strong", **kw).solve(model, is_int)
        node[name] = {"node_wall_s": time.perf_counter() - t0, "strong_branch_s": r["strong_branch_seconds"],
                      "lp_solves": r["lp_solves"]}
        log(f"root node {name}: {node[name]['node_wall_s']:.3f}s (strong branching {r['strong_branch_seconds']:.3f}s, "
            f"{r['lp_solves']} LP solves)")
    out["bnb_root_node"] = node
    if ranker is not None:
        ys = ranker.scores(root["primal"], model.lb, model.ub, [c["var"] for c in cb])
        top8 = set(np.argsort(-ys, kind="stable")[:8])
        out["surrogate_top8_contains_full_sb_choice"] = bool(best_b in top8)
        out["wall_per_node_speedup_surrogate_vs_per_lp_full"] = (
            node["per_lp_full_64"]["node_wall_s"] / node["batched_surrogate_top8"]["node_wall_s"])
    out["bnb_tree"] = tree_leg(model, is_int, ranker)
    if ranker is not None:
        ranker.net.close()
    return out


def sb_root_calls(reps=3):
    """The batched strong-branching call at the knapsack root (64 candidates,
    128 child LPs) alone: one warm-up call, then `reps` timed calls (wall
    seconds and the batched kernel's HIP-event milliseconds)."""
    from sovereign_amd import api, milp, problems as P

    model = P.multi_knapsack(500, 200, seed=P.BASE_SEED)
    is_int = np.ones(model.n_vars, dtype=np.int8)
    root = api.InteriorPointSolver(api.InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9)).solve(model)
    milp.strong_branch(model, is_int, root["primal"], root["objective"], max_candidates=8)
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        cb, _, kms = milp.strong_branch(model, is_int, root["primal"], root["objective"], max_candidates=64)
        out.append({"wall_s": time.perf_counter() - t0, "kernel_ms": kms, "child_lps": 2 * len(cb)})
        log(f"root strong branching: {2 * len(cb)} child LPs {out[-1]['wall_s'] * 1e3:.2f} ms wall, {kms:.2f} ms kernel")
    return out


def tree_leg(model, is_int, ranker, nodes=40, per_lp_nodes=4):
    """Wall seconds per branch-and-bound node over a tree (not only the
    root): the knapsack searched for a fixed node budget with each
    strong-branching setup, the plain per-LP path on a smaller budget (each
    of its nodes costs seconds); the batched full strong branching once with
    children warm-started from their parent's iterate (the default) and once
    cold. Then the same setups solve a small knapsack
    (20 x 4, seed 3) to optimality, where their incumbents must agree."""
    from sovereign_amd import milp

    setups = [("per_lp_full_64", per_lp_nodes, dict(branch_rule="full_strong", strong_branch_candidates=64,
                                                    child_solver=milp.CHILD_PER_LP)),
              ("batched_full_64", nodes, dict(branch_rule="full_strong", strong_branch_candidates=64)),
              ("batched_full_64_cold", nodes, dict(branch_rule="full_strong", strong_branch_candidates=64,
                                                   warm_start=False)),
              ("batched_reliability", nodes, dict(branch_rule="strong"))]
    if ranker is not None:
        setups.append(("batched_surrogate_top8", nodes, dict(branch_rule="full_strong", strong_branch_candidates=8,
                                                             ranker=ranker)))
    out = {"model": "multi_knapsack n=500 m=200 seed 26119", "node_budget": nodes, "per_lp_node_budget": per_lp_nodes}
    for name, budget, kw in setups:
        r = milp.BranchAndBoundSolver(max_nodes=budget, **kw).solve(model, is_int)
        out[name] = {"nodes": r["nodes"], "wall_s": r["seconds"], "wall_s_per_node": r["seconds"] / max(1, r["nodes"]),
                     "lp_solves": r["lp_solves"], "lp_iterations": r["lp_iterations"],
                     "iterations_per_lp": r["lp_iterations"] / max(1, r["lp_solves"]),
                     "root_cuts": r["cuts_added"], "strong_lp_solves": r["strong_lp_solves"],
                     "node_lp_solves": r["node_lp_solves"], "kernel_ms": r["kernel_ms"], "status": r["status"],
                     "incumbent": r["objective"] if r["status"] in (0, 1) else None}
        log(f"tree {name}: {r['nodes']} nodes {r['seconds']:.2f}s ({out[name]['wall_s_per_node'] * 1e3:.1f} ms/node), "
            f"{r['lp_solves']} LP solves ({out[name]['iterations_per_lp']:.2f} iterations each), status {r['status']}")
    base = out["per_lp_full_64"]["wall_s_per_node"]
    for name, _, _ in setups[1:]:
        out[f"per_node_speedup_{name}_vs_per_lp_full"] = base / out[name]["wall_s_per_node"]
    from sovereign_amd import problems as P

    small = P.multi_knapsack(20, 4, seed=3)
    sint = np.ones(small.n_vars, dtype=np.int8)
    objs = {}
    for name, _, kw in setups:
        kw = dict(kw)
        if "ranker" in kw:
            from sovereign_amd import surrogate as S

            kw["ranker"] = S.load_trained(small)
        r = milp.BranchAndBoundSolver(**kw).solve(small, sint)
        objs[name] = {"status": r["status"], "objective": r["objective"], "nodes": r["nodes"], "wall_s": r["seconds"]}
        if "ranker" in kw:
            kw["ranker"].net.close()
    vals = [v["objective"] for v in objs.values()]
    out["small_instance_to_optimality"] = {"model": "multi_knapsack n=20 m=4 seed 3", "runs": objs,
                                           "identical_objectives": max(vals) - min(vals) <= 1e-6 * max(1.0, abs(vals[0]))}
    return out


def qp_leg(steps, with_cpu):
    """The separable convex QP (20,002 variables, 20,000 second-difference
    rows): the synthetic workload (target = sin(3t) + t^2 + Normal(0, 0.01)
    noise, seed 20260501) and the reference's own n=20000 instance
    (tests/golden), each solved `steps` times on the mixed and the plain FP64
    path. Time = the interior-point kernel (HIP events; one launch per solve,
    inputs resident) and the whole call (upload, kernel, host polish and
    certificate). Per-iteration algorithmic traffic: 64 bytes per row."""
    from sovereign_amd import qp

    fx = np.load(os.path.join(ROOT, "tests", "golden", "data", "liswet_fixtures.npz"))
    with open(os.path.join(ROOT, "tests", "golden", "liswet_golden.json")) as fh:
        gold = float(json.load(fh)["cases"]["20000"]["objective"])
    cases = {"synthetic_n20002": qp.liswet_target(20002), "reference_n20000": fx["target_20000"]}
    out = {"model": "separable convex QP, n=20002 vars, m=20000 rows x_i - 2x_{i+1} + x_{i+2} >= 0, Q = I",
           "reference_cpu_gi_s_n20000": 180.9}
    for name, target in cases.items():
        rec = {}
        for prec in ("mixed", "fp64"):
            opts = qp.QpInteriorPointOptions(precision=prec)
            qp.solve_liswet(target, opts)  # warm-up
            runs = [qp.solve_liswet(target, opts) for _ in range(steps)]
            r = runs[-1]
            kms = sorted(x["kernel_ms"] for x in runs)[len(runs) // 2]
            wall = sorted(x["seconds"] for x in runs)[len(runs) // 2]
            bytes_alg = 64.0 * (len(target) - 2) * r["iterations"]
            rec[prec] = {"status": r["status"], "ipm_iterations": r["iterations"], "ipm_status": r["ipm_status"],
                         "fp32_factorizations": r["fp32_factorizations"],
                         "fp64_factorizations": r["fp64_factorizations"], "polish_fits": r["polish_fits"],
                         "objective": r["objective"], "kernel_ms": kms, "call_s": wall,
                         "kernel_algorithmic_GBps": bytes_alg / (kms * 1e-3) / 1e9}
            log(f"qp {name} {prec}: status {r['status']} obj {r['objective']:.15g} kernel {kms:.1f} ms, "
                f"call {wall * 1e3:.1f} ms, {r['iterations']} iterations")
        rec["speedup_mixed_vs_fp64_kernel"] = rec["fp64"]["kernel_ms"] / rec["mixed"]["kernel_ms"]
        if name == "reference_n20000":
            rec["objective_rel_diff_vs_reference"] = abs(rec["fp64"]["objective"] - gold) / abs(gold)
        out[name] = rec
    if with_cpu:
        from oracle import oracle as O

        t0 = time.perf_counter()
        it, xo = O.liswet_ipm(cases["synthetic_n20002"])
        cpu_s = time.perf_counter() - t0
        cert, _ = O.liswet_certify(cases["synthetic_n20002"], xo)
        out["cpu_baseline"] = {"value": cpu_s, "unit": "s per QP solve (synthetic n=20002)", "cores": 1,
                               "kind": "port",
                               "sample": "oracle restatement of the reference structured QP IPM (long double, "
                                         "one thread) plus knot polish, one full solve"}
        out["synthetic_n20002"]["objective_rel_diff_vs_oracle"] = (
            abs(out["synthetic_n20002"]["fp64"]["objective"] - cert["objective"]) / abs(cert["objective"]))
        log(f"qp cpu oracle {cpu_s:.2f}s obj {cert['objective']:.15g}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-branching", action="store_true")
    ap.add_argument("--no-qp", action="store_true")
    ap.add_argument("--only-qp", action="store_true", help="run the separable QP leg alone (profiling)")
    ap.add_argument("--only-branching", action="store_true", help="run the strong-branching leg alone (profiling)")
    ap.add_argument("--sb-root-only", action="store_true",
                    help="only the batched root strong-branching call, three timed calls (counter profiles)")
    ap.add_argument("--paths", default="mixed,fp64")
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--cpu-phases", type=int, nargs="*", default=None,
                    help="time the oracle's CPU phases at these sizes (the full-size scaling profile)")
    args = ap.parse_args()
    if args.cpu_phases is not None:
        rec = {str(m): cpu_phases(m, say=log) for m in (args.cpu_phases or [4096, 8192])}
        rec["nproc"] = os.cpu_count()
        print(json.dumps(rec), flush=True)
        return
    if args.copy_probe:
        import torch

        x = torch.ones(64 * 1024 * 1024, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the multi-GPU (sharded) path is designed but not built.", file=sys.stderr)
        sys.exit(2)
    paths = [p for p in args.paths.split(",") if p]
    if args.only_qp:
        from sovereign_amd import _capi

        _capi.load()
        print(json.dumps({"separable_qp": qp_leg(max(1, args.steps), not args.no_cpu)}), flush=True)
        return
    if args.only_branching:
        from sovereign_amd import _capi

        print(json.dumps({"strong_branching": branching_leg(_capi.load())}), flush=True)
        return
    if args.sb_root_only:
        print(json.dumps({"strong_branching_root": sb_root_calls()}), flush=True)
        return

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
    runs = {}
    for path in paths:
        opts = api.InteriorPointOptions(precision=path, **tol)
        runs[path] = timed_solves(ipm, opts, args.steps, args.warmup, lib)
        log(f"kernel ms ({path}, one instrumented solve): " +
            " ".join(f"{k} {v:.1f}" for k, v in kernel_ms(runs[path][3]).items()))
    spmv = spmv_roofline(lib, ipm.A, lp.m, lp.n)
    log("spmv " + " ".join(f"{k} {v['achieved']:.0f} GB/s ({v['avg_launch_us']:.1f} us)" for k, v in spmv.items()))
    ipm.close()
    # end to end: upload and plan (the reference's NormalEquations constructor
    # runs inside every solve_ipm) plus one solve, per path
    end_to_end = {}
    for path in paths:
        opts = api.InteriorPointOptions(precision=path, **tol)
        lib.sov_sync()
        t = time.perf_counter()
        fresh = api.ResidentIpm(lp)
        r = fresh.solve(opts)
        lib.sov_sync()
        end_to_end[path] = time.perf_counter() - t
        fresh.close()
        log(f"end to end ({path}): setup + solve {end_to_end[path]:.3f}s, status {r['status_name']}")
    main_path = "mixed" if "mixed" in runs else paths[0]
    res, wall, kt, _ = runs[main_path]
    traffic = load_traffic()
    shapes_info, shape_ids = tuned_shapes(lib)
    dominant, secondary = dominant_roofline(kt, traffic, shape_ids)
    line = {
        "metric": METRIC,
        "value": wall / args.steps,
        "unit": "s per IPM solve (m=8192 LP to 1e-8 gap)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32 MFMA factorizations until the f32-preconditioned PCG passes 8 iterations per IPM "
                  "iteration (or mu < 1e-6), f64 residuals/PCG/late factorizations"
                  if main_path == "mixed" else "f64"),
        "data": "synthetic",
        "config": {"workload": f"random sparse LP m={lp.m} n={lp.n} 6 nnz/col seed {P.BASE_SEED}",
                   "gap_tol": 1e-8, "path": f"{main_path} (value)" + (
                       "; plain FP64 alongside" if "fp64" in runs and main_path != "fp64" else "")},
        "all_optimal": all(r["status"] == 0 for rr in runs.values() for r in rr[0]),
        "iterations": {p: rr[0][-1]["iterations"] for p, rr in runs.items()},
        "factorizations": {p: {"fp32": rr[0][-1]["fp32_factorizations"], "total": rr[0][-1]["factorizations"]}
                           for p, rr in runs.items()},
        "resident_setup_s": setup_s,
        "end_to_end_s_incl_setup": end_to_end,
        "kernel_ms_one_instrumented_solve": {p: kernel_ms(rr[3]) for p, rr in runs.items()},
        "roofline": dominant,
        "roofline_secondary": secondary,
        "roofline_spmv": spmv,
        "roofline_triangular_solves": {p: solve_roofline(rr[3], p) for p, rr in runs.items()},
        "autotuned_tiles": shapes_info,
    }
    if "mixed" in runs and "fp64" in runs:
        s_mx = runs["mixed"][1] / args.steps
        s_64 = runs["fp64"][1] / args.steps
        o_mx = runs["mixed"][0][-1]["objective"]
        o_64 = runs["fp64"][0][-1]["objective"]
        line.update({
            "fp64_path_s": s_64, "speedup_vs_fp64": s_64 / s_mx, "objective": o_64,
            "objective_rel_diff_mixed_vs_fp64": abs(o_mx - o_64) / max(1.0, abs(o_64)),
            "pcg_iterations_mixed": runs["mixed"][0][-1]["refine_iterations"],
            "max_solve_residual_mixed": runs["mixed"][0][-1]["max_solve_residual"],
            "roofline_fp64_path": dominant_roofline(runs["fp64"][2], traffic, shape_ids)[0],
        })
    if not args.no_branching:
        line["strong_branching"] = branching_leg(lib)
    if not args.no_qp:
        line["separable_qp"] = qp_leg(max(3, args.steps), not args.no_cpu)
    if not args.no_cpu:
        log("cpu baseline (oracle phases at m=4096, scaled to m=8192)")
        line["cpu_baseline"] = cpu_baseline(line["iterations"][main_path])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
