"""Command line of the solver, as the reference's (solver/cli/main.cpp):

  python -m sovereign_amd.cli solve <model.json|model.mps> [--verify] [--summary-only] [--out <result.json>]
  python -m sovereign_amd.cli convert <model.mps|model.json>      the model as JSON
  python -m sovereign_amd.cli capabilities                        device record
  python -m sovereign_amd.cli version

`solve` prints the result JSON of result_to_json_string (json_io.cpp:190-240)
with the CLI's additions (main.cpp:101-108): requested_device,
gpu_operations, gpu_factorizations, gpu_used, load_seconds and
peak_memory_mb, so a caller that parses the reference's output (the worker's
GPU acceptance test, worker/runner.py:268-276) reads this one unchanged.
SOVEREIGN_DEVICE keeps its meaning: "cuda" (any accelerator request) fails
when no device is usable.

Dispatch follows OptimizationEngine::solve (solver/core/src/engine.cpp):
LP -> the GPU interior point (InteriorPointSolver), MILP -> branch and bound
with the reference's options and environment switches
(solver/milp/src/milp_solver.cpp:31-98), QP -> QpInteriorPointSolver. The
objective of an optimal or feasible result is recomputed from the primal
point on the original model (engine.cpp:152-156).
"""
from __future__ import annotations

import json
import os
import resource
import sys
import time

import numpy as np

from . import api, milp, model_io as IO, qp

VERSION = "sovereign_amd 1.0.0 (LP/MILP/QP interior point and branch and bound on MI355X)"


def _env_int(name, default):
    v = os.environ.get(name)
    return int(v) if v not in (None, "") else default


def _env_flag(name, default):
    v = os.environ.get(name)
    if v is None:
        return default
    return v not in ("0", "false", "off")


def requested_device() -> str:
    """requested_device (solver/gpu/src/gpu_spmv.cpp:29-31)."""
    return os.environ.get("SOVEREIGN_DEVICE", "auto") or "auto"


def _lp_result(model: IO.Model, lp, r: dict) -> IO.Result:
    out = IO.Result(status=IO.STATUS.get(r["status"], "ERROR"), iterations=int(r.get("iterations", 0)),
                    duality_gap=float(r.get("duality_gap", 0.0)), primal_residual=float(r.get("primal_residual", 0.0)),
                    dual_residual=float(r.get("dual_residual", 0.0)), message=r.get("message", "") or "")
    if "primal" in r:
        out.primal = {n: float(v) for n, v in zip(model.var_names, r["primal"])}
        out.dual = {n: float(v) for n, v in zip(model.var_names, r.get("var_dual", np.zeros(model.n_vars)))}
        for i, v in enumerate(r.get("row_dual", [])):
            out.dual[f"row_{i}"] = float(v)
        out.dual_certificate_space = "original_model"
    return out


def solve_lp(model: IO.Model) -> IO.Result:
    lp = model.to_lp_model()
    r = api.InteriorPointSolver(api.InteriorPointOptions()).solve(lp)
    return _lp_result(model, lp, r)


def bnb_options_from_env(model: IO.Model) -> dict:
    """MilpSolver::solve's options (milp_solver.cpp:31-98) for the parts built
    here: branch rule, strong-branching budget, reliability threshold, node
    and time limits, plunging, root cuts (SOVEREIGN_ENABLE_CUTS,
    SOVEREIGN_MAX_CUT_ROUNDS with its size-based default of :46-60,
    SOVEREIGN_MAX_CUTS_PER_NODE, SOVEREIGN_CMIR_MAX_DEPTH >= 0 turns cMIR on
    at the root)."""
    n = model.n_vars
    rule = os.environ.get("SOVEREIGN_BRANCH_RULE", "strong")
    if rule not in milp.BRANCH_RULES:
        rule = "strong"
    scale = n * max(1, len(model.con_names))
    rounds_default = 1 if (scale > 500000 or n > 4000) else 2 if scale > 100000 else 5
    rounds = max(0, _env_int("SOVEREIGN_MAX_CUT_ROUNDS", rounds_default))
    if not _env_flag("SOVEREIGN_ENABLE_CUTS", True):
        rounds = 0
    kw = {"branch_rule": rule,
          "max_cut_rounds": rounds,
          "max_cuts_per_node": max(0, _env_int("SOVEREIGN_MAX_CUTS_PER_NODE", 20)),
          "cmir_at_root": _env_int("SOVEREIGN_CMIR_MAX_DEPTH", -1) >= 0,
          "max_strong_per_node": _env_int("SOVEREIGN_MAX_STRONG_PER_NODE", 2 if n > 1000 else 8),
          "reliability_threshold": _env_int("SOVEREIGN_RELIABILITY_THRESHOLD", 2 if n > 1000 else 4),
          "max_nodes": _env_int("SOVEREIGN_MAX_NODES", 100000),
          "plunging": _env_flag("SOVEREIGN_PLUNGING", True),
          "strong_branch_candidates": 4}
    tl = os.environ.get("SOVEREIGN_TIME_LIMIT")
    if tl:
        kw["time_limit_seconds"] = float(tl)
    return kw


def solve_milp(model: IO.Model) -> IO.Result:
    lp = model.to_lp_model()
    is_int = model.is_int()
    if not is_int.any():
        return solve_lp(model)
    kw = bnb_options_from_env(model)
    if model.n_vars and len(model.con_names) > 272:
        kw["child_solver"] = milp.CHILD_PER_LP  # the batched child form holds at most 272 rows
    r = milp.BranchAndBoundSolver(**kw).solve(lp, is_int)
    st = {0: "OPTIMAL", 1: "FEASIBLE", 2: "INFEASIBLE", 6: "NUMERICAL_ERROR"}.get(r["status"])
    if st is None:
        st = "TIME_LIMIT" if kw.get("time_limit_seconds") and r["seconds"] >= kw["time_limit_seconds"] else "ITERATION_LIMIT"
    out = IO.Result(status=st, nodes=int(r["nodes"]), iterations=int(r["lp_solves"]))
    out.mip = {"node_lp_failures": int(r["node_lp_failures"]), "dropped_subtrees": int(r["node_lp_failures"])}
    if st in ("OPTIMAL", "FEASIBLE"):
        out.primal = {n: float(v) for n, v in zip(model.var_names, r["x"])}
        out.mip.update(has_best_bound=True, best_bound=float(r["objective"]) if st == "OPTIMAL" else float(r["root_bound"]))
    out.message = {"OPTIMAL": "Optimal integer solution found by branch and bound; the tree was exhausted.",
                   "FEASIBLE": "Integer feasible solution found; optimality is not proven (limit or failed node LP).",
                   "INFEASIBLE": "MILP is infeasible: every node was proved infeasible or pruned.",
                   "NUMERICAL_ERROR": "No incumbent, and a node LP failed: infeasibility cannot be certified."
                   }.get(st, "Branch and bound stopped at its limit without an incumbent.")
    out.warnings.append(f"nodes={r['nodes']} lp_solves={r['lp_solves']} branch_rule={kw['branch_rule']}")
    return out


def solve_qp(model: IO.Model) -> IO.Result:
    idx = model.index()
    lin = np.zeros(model.n_vars)
    for n, v in model.obj_linear.items():
        lin[idx[n]] = v
    quad = {}
    Q = model.hessian()
    for i, j in zip(*np.nonzero(Q)):
        quad[(int(i), int(j))] = float(Q[i, j])
    rows = [{idx[n]: v for n, v in lin_r.items()} for lin_r in model.con_linear]
    qm = qp.QpModel(n_vars=model.n_vars, linear=lin, quadratic=quad, rows=rows, senses=list(model.con_senses),
                    rhs=np.array(model.con_rhs), lb=np.array(model.lb), ub=np.array(model.ub),
                    minimize=model.sense == "minimize", constant=model.obj_constant)
    r = qp.QpInteriorPointSolver().solve(qm)
    out = IO.Result(status=IO.STATUS.get(r["status"], "ERROR"), message=r.get("message", "") or "",
                    iterations=int(r.get("iterations", 0)))
    if "x" in r and out.status in ("OPTIMAL", "FEASIBLE"):
        out.primal = {n: float(v) for n, v in zip(model.var_names, r["x"])}
        if "var_dual" in r:
            out.dual = {n: float(v) for n, v in zip(model.var_names, r["var_dual"])}
            for i, v in enumerate(r.get("row_dual", [])):
                out.dual[f"row_{i}"] = float(v)
    for k in ("duality_gap", "primal_residual", "dual_residual"):
        if k in r:
            setattr(out, k, float(r[k]))
    return out


def solve_model(model: IO.Model) -> IO.Result:
    t0 = time.perf_counter()
    if model.problem_type == "LP":
        r = solve_lp(model)
    elif model.problem_type == "MILP":
        r = solve_milp(model)
    else:
        r = solve_qp(model)
    if r.status in ("OPTIMAL", "FEASIBLE") and r.primal:
        r.objective_value = model.objective_of(np.array([r.primal[n] for n in model.var_names]))
    r.runtime_seconds = time.perf_counter() - t0
    return r


def peak_memory_mb() -> float:
    return resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024.0


def cmd_solve(args, out=sys.stdout) -> int:
    if not args:
        print("Usage: solve <model.json|model.mps> [--verify] [--summary-only] [--out <result.json>]", file=sys.stderr)
        return 2
    path, verify, summary, out_path = args[0], False, False, None
    k = 1
    while k < len(args):
        a = args[k]
        if a == "--verify":
            verify = True
        elif a == "--summary-only":
            summary = True
        elif a == "--out" and k + 1 < len(args):
            k += 1
            out_path = args[k]
        else:
            print(f"Unknown argument: {a}", file=sys.stderr)
            return 2
        k += 1
    try:
        device = requested_device()
        info = api.device_info()
        if device == "cuda" and not info["available"]:
            raise RuntimeError("CUDA requested but unavailable: " + info["reason"])
        api.reset_gpu_operations()
        t = time.perf_counter()
        model = IO.load_model_from_file(path)
        load_s = time.perf_counter() - t
        result = solve_model(model)
        payload = IO.result_to_json(result, include_primal=not summary)
        ops = api.gpu_operations()
        payload.update(requested_device=device, gpu_operations=ops, gpu_factorizations=api.gpu_factorizations(),
                       gpu_used=ops > 0, load_seconds=load_s, peak_memory_mb=peak_memory_mb())
        text = json.dumps(payload, indent=2)
        print(text, file=out)
        if out_path:
            with open(out_path, "w") as fh:
                fh.write(text + "\n")
        if verify:
            print("\nverification:\n" + json.dumps(IO.verify(model, result), indent=2), file=out)
        return 1 if result.status == "ERROR" else 0
    except Exception as e:  # the reference prints "error: ..." and exits 1
        print(f"error: {e}", file=sys.stderr)
        return 1


def main(argv=None, out=sys.stdout) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("Usage:\n  solve <model.json|model.mps> [--verify] [--summary-only] [--out <result.json>]\n  version",
              file=sys.stderr)
        return 2
    cmd = argv[0]
    if cmd == "capabilities":
        info = api.device_info()
        caps = {"version": "1.2.0", "cuda_available": bool(info["available"]), "cuda_backend": "hip",
                "cuda_reason": info["reason"],
                "gpu_acceleration": "dense-lu,sparse-matrix-vector,normal-equations,batched-child-lp"}
        if info.get("name"):
            caps["gpu_name"] = info["name"]
            caps["gpu_arch"] = info.get("arch", "")
            caps["gpu_memory_mb"] = int(info.get("memory_bytes", 0)) >> 20
        print(json.dumps(caps), file=out)
        return 0
    if cmd == "version":
        print(VERSION, file=out)
        return 0
    if cmd == "convert" and len(argv) == 2:
        try:
            print(json.dumps(IO.model_to_json(IO.load_model_from_file(argv[1])), indent=2), file=out)
            return 0
        except Exception as e:
            print(str(e), file=sys.stderr)
            return 1
    if cmd == "solve":
        return cmd_solve(argv[1:], out=out)
    print("Usage:\n  solve <model.json|model.mps> [--verify] [--summary-only] [--out <result.json>]\n  version",
          file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main())
