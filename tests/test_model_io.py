"""Model and result I/O in the reference's formats (CPU): the MPS reader
against the reference's own JSON conversion of the same model, the MPS
features the reference reader defines (mps_io.cpp:71-249), JSON round trips,
the verifier's checks, and the LP fixtures solved by the oracle to the
published optima (netlib AFIRO -4.6475314286E+02; catalog.json's
expectations for the robustness models)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from sovereign_amd import model_io as IO

DATA = os.path.join(os.path.dirname(__file__), "golden", "data", "models")


def path(name):
    return os.path.join(DATA, name)


def rows_as_dicts(m):
    return {n: (s, r, lin) for n, s, r, lin in zip(m.con_names, m.con_senses, m.con_rhs, m.con_linear)}


def test_afiro_mps_matches_reference_json_conversion():
    a = IO.load_mps(path("afiro.mps"))
    b = IO.load_json(path("afiro.json"))
    assert a.problem_type == b.problem_type == "LP"
    assert a.sense == b.sense
    assert a.var_names == b.var_names
    assert a.lb == b.lb and a.ub == b.ub
    assert a.obj_linear == b.obj_linear and a.obj_constant == b.obj_constant
    ra, rb = rows_as_dicts(a), rows_as_dicts(b)
    assert ra.keys() == rb.keys()
    for k in ra:
        assert ra[k][0] == rb[k][0] and ra[k][1] == pytest.approx(rb[k][1], abs=0)
        assert ra[k][2] == pytest.approx(rb[k][2], abs=0)


MPS_FEATURES = """NAME          FEATURES
OBJSENSE
    MAX
ROWS
 N  COST
 L  LIM1
 G  LIM2
 E  MYEQN
 E  MYEQ2
 L  RNG1
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         1.0
    MARKER                 'MARKER'                 'INTEND'
    X2        COST         2.0   LIM1         1.0
    X2        MYEQN       -1.0   RNG1         1.0
    X3        COST        -1.0   MYEQN        1.0
    X3        MYEQ2        1.0
    X4        COST         1.5   MYEQ2        1.0
    X5        LIM2         3.0
RHS
    RHS       COST        -7.5   LIM1         4.0
    RHS       LIM2         1.0   MYEQN        7.0
    RHS       MYEQ2        2.0   RNG1         5.0
    RHS2      LIM1        99.0
RANGES
    RNG       MYEQN        3.0   MYEQ2       -2.0
    RNG       RNG1         4.0
BOUNDS
 UP BND       X2           4.0
 MI BND       X3
 UP BND       X4          -1.0
 FR BND       X5
 UP BND2      X1           9.0
ENDATA
"""


def test_mps_features():
    m = IO.load_mps_string(MPS_FEATURES)
    assert m.sense == "maximize" and m.problem_type == "MILP"
    j = m.index()
    assert m.var_types[j["X1"]] == "integer" and m.ub[j["X1"]] == 1.0  # integer without bounds: [0, 1]
    assert (m.lb[j["X2"]], m.ub[j["X2"]]) == (0.0, 4.0)
    assert m.lb[j["X3"]] == -1e30 and m.ub[j["X3"]] == 1e30  # MI
    assert (m.lb[j["X4"]], m.ub[j["X4"]]) == (-1e30, -1.0)  # negative UP without a prior bound
    assert (m.lb[j["X5"]], m.ub[j["X5"]]) == (-1e30, 1e30)  # FR
    assert m.obj_constant == 7.5  # minus the objective row's RHS
    assert m.obj_linear == {"X1": 1.0, "X2": 2.0, "X3": -1.0, "X4": 1.5}
    rows = rows_as_dicts(m)
    assert rows["LIM1"][:2] == ("<=", 4.0)  # the first RHS set wins
    # E row with range +3: [7, 10]; E row with range -2: [0, 2]; L row with range 4: [1, 5]
    assert rows["MYEQN"][:2] == ("<=", 10.0) and rows["MYEQN_lo"][:2] == (">=", 7.0)
    assert rows["MYEQ2"][:2] == ("<=", 2.0) and rows["MYEQ2_lo"][:2] == (">=", 0.0)
    assert rows["RNG1"][:2] == ("<=", 5.0) and rows["RNG1_lo"][:2] == (">=", 1.0)
    assert rows["LIM2"][2] == {"X1": 1.0, "X5": 3.0}
    # free-format records read the same model
    free = IO.load_mps_string(MPS_FEATURES, fixed_format=False)
    assert rows_as_dicts(free) == rows


@pytest.mark.parametrize("text,msg", [
    ("NAME\nROWS\n N C\nCOLUMNS\n    X C 1\n", "ENDATA"),
    ("NAME\nROWS\n N C\nCOLUMNS\n    X Q 1\nENDATA\n", "unknown row"),
    ("NAME\nROWS\n N C\n L R\nCOLUMNS\n    X R 1\nBOUNDS\n LO B X 5\n UP B X 1\nENDATA\n", "crossed bounds"),
    ("NAME\nROWS\n N C\nCOLUMNS\n    X C 1\nBOUNDS\n SC B X 5\nENDATA\n", "unsupported bound type"),
    ("NAME\nROWS\n N C\nQUADOBJ\nENDATA\n", "unsupported MPS section"),
])
def test_mps_rejects(text, msg):
    with pytest.raises(ValueError, match=msg):
        IO.load_mps_string(text)


def test_json_round_trip_and_quadratic():
    m = IO.load_json(path("portfolio_qp_600.json"))
    assert m.problem_type == "QP" and m.n_vars == 600
    again = IO.load_json_string(json.dumps(IO.model_to_json(m)))
    assert again.var_names == m.var_names and again.quadratic == m.quadratic
    assert rows_as_dicts(again) == rows_as_dicts(m)
    Q = m.hessian()
    np.testing.assert_allclose(Q, Q.T)
    # 1/2 x'Qx equals the stored-term sum of the verifier
    x = np.linspace(-1, 1, 600)
    direct = sum(0.5 * q * x[m.index()[a]] * x[m.index()[b]] for a, row in m.quadratic.items() for b, q in row.items())
    assert 0.5 * x @ Q @ x == pytest.approx(direct, rel=1e-12)


def test_load_model_from_file_dispatch(tmp_path):
    with pytest.raises(ValueError, match="no extension"):
        IO.load_model_from_file(str(tmp_path / "model"))
    with pytest.raises(ValueError, match="compressed"):
        IO.load_model_from_file(str(tmp_path / "model.mps.gz"))
    with pytest.raises(ValueError, match="unsupported model extension"):
        IO.load_model_from_file(str(tmp_path / "model.lp"))
    assert IO.load_model_from_file(path("afiro.mps")).n_vars == 32


@pytest.mark.parametrize("name,expected", [
    ("afiro.mps", -464.75314286),       # netlib AFIRO optimum -4.6475314286E+02
    ("afiro.json", -464.75314286),
    ("kuhn_degeneracy.json", -1.25),    # benchmarks/datasets/catalog.json
    ("illconditioned.json", 3.0e7),
])
def test_lp_fixtures_solved_by_oracle(name, expected):
    m = IO.load_model_from_file(path(name))
    r = O.lp_solve(m.to_lp_model())
    assert r["status"] == 0
    x = r["primal"]
    assert m.objective_of(x) == pytest.approx(expected, rel=1e-8)
    res = IO.Result(status="OPTIMAL", objective_value=m.objective_of(x),
                    primal={n: float(v) for n, v in zip(m.var_names, x)}, duality_gap=r["duality_gap"])
    v = IO.verify(m, res)
    assert v["is_valid"], v["issues"]


def test_verifier_catches_planted_errors():
    m = IO.load_json(path("knapsack6.json"))
    x = {n: 0.0 for n in m.var_names}
    good = IO.Result(status="OPTIMAL", objective_value=m.obj_constant, primal=dict(x))
    assert IO.verify(m, good)["is_valid"]
    frac = dict(x)
    frac[m.var_names[0]] = 0.5
    r = IO.verify(m, IO.Result(status="FEASIBLE", objective_value=m.objective_of([frac[n] for n in m.var_names]),
                                primal=frac))
    assert "Integrality violation exceeds tolerance." in r["issues"]
    over = {n: 1.0 for n in m.var_names}
    r = IO.verify(m, IO.Result(status="FEASIBLE", objective_value=m.objective_of([1.0] * m.n_vars), primal=over))
    assert "Constraint violation exceeds tolerance." in r["issues"]
    r = IO.verify(m, IO.Result(status="OPTIMAL", objective_value=123.0, primal=dict(x)))
    assert any("differs from recomputed" in s for s in r["issues"])
    r = IO.verify(m, IO.Result(status="OPTIMAL", objective_value=m.obj_constant, primal=dict(x), duality_gap=1e-3))
    assert any("contradicted" in s for s in r["issues"])
    assert not IO.verify(m, IO.Result(status="INFEASIBLE"))["is_valid"]


def test_result_json_fields():
    r = IO.Result(status="OPTIMAL", objective_value=1.0, primal={"a": 1.0})
    full = IO.result_to_json(r)
    short = IO.result_to_json(r, include_primal=False)
    for k in ("status", "objective_value", "primal", "dual", "slacks", "lp_diagnostics", "mip_diagnostics",
              "optimality_gap", "duality_gap", "primal_residual", "dual_residual", "presolve", "optimality_proven",
              "iterations", "nodes", "runtime_seconds", "message", "warnings"):
        assert k in full
    assert full["primal"] == {"a": 1.0} and short["primal"] is None and short["primal_variables"] == 1
    assert full["optimality_proven"] is True
    assert IO.result_to_json(IO.Result(status="FEASIBLE"))["optimality_proven"] is False
