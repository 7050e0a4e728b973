"""The command line end to end on the GPU, as the reference CLI's callers use
it (solver/cli/main.cpp; the worker's acceptance test, worker/runner.py:
268-276): status, objective against the published optima (netlib AFIRO,
benchmarks/datasets/catalog.json, benchmarks/reports/EVIDENCE.md:47-52),
the verifier's verdict, and the GPU accounting fields."""
import io
import json
import os

import pytest

from sovereign_amd import cli

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "data", "models")

CASES = [
    ("afiro.mps", -464.75314286),
    ("afiro.json", -464.75314286),
    ("kuhn_degeneracy.json", -1.25),
    ("illconditioned.json", 3.0e7),
    ("weak_lp_relaxation.json", 100.0),
    ("knapsack6.json", 34.0),
    ("multi_knapsack_18x3.json", 102.0),
    ("set_partition_14x9.json", 6.0),
]


def run_cli(args):
    buf = io.StringIO()
    rc = cli.main(args, out=buf)
    text = buf.getvalue()
    head, _, tail = text.partition("\nverification:\n")
    return rc, json.loads(head), (json.loads(tail) if tail else None)


@pytest.mark.parametrize("name,expected", CASES)
def test_cli_solve_matches_published_optimum(gpu_api, monkeypatch, name, expected):
    monkeypatch.setenv("SOVEREIGN_DEVICE", "cuda")
    print(f"[case] cli solve {name}", flush=True)
    rc, res, ver = run_cli(["solve", os.path.join(DATA, name), "--verify"])
    print(f"[case]   {res['status']} {res['objective_value']} ops {res['gpu_operations']} "
          f"nodes {res['nodes']} {res['runtime_seconds']:.3f}s", flush=True)
    assert rc == 0
    assert res["status"] == "OPTIMAL" and res["optimality_proven"]
    assert res["objective_value"] == pytest.approx(expected, rel=1e-8, abs=1e-9)
    assert ver["is_valid"], ver["issues"]
    # the worker accepts a GPU run only with these fields
    assert res["requested_device"] == "cuda" and res["gpu_used"] is True
    assert isinstance(res["gpu_operations"], int) and res["gpu_operations"] > 0


def test_cli_summary_out_and_convert(gpu_api, tmp_path):
    out = tmp_path / "r.json"
    rc, res, _ = run_cli(["solve", os.path.join(DATA, "afiro.mps"), "--summary-only", "--out", str(out)])
    assert rc == 0 and res["primal"] is None and res["primal_variables"] == 32
    assert json.loads(out.read_text())["status"] == "OPTIMAL"
    buf = io.StringIO()
    assert cli.main(["convert", os.path.join(DATA, "afiro.mps")], out=buf) == 0
    conv = json.loads(buf.getvalue())
    assert conv["problem_type"] == "LP" and len(conv["variables"]) == 32
    buf = io.StringIO()
    assert cli.main(["capabilities"], out=buf) == 0
    assert json.loads(buf.getvalue())["cuda_available"] is True
    assert cli.main(["solve", os.path.join(DATA, "missing.json")], out=io.StringIO()) == 1
