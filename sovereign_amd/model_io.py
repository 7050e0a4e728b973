"""Model and result I/O in the reference's formats, so that callers of the
reference CLI (and the worker that parses its output) can use this build
unchanged.

  Model                 OptimizationModel (solver/core/include/sovereign/types.hpp:32-59)
  load_json / load_json_string
                        load_model_from_json_{file,string} (solver/core/src/json_io.cpp:14-73,
                        130-152): variables, objective (linear, constant, quadratic), constraints
  load_mps / load_mps_string
                        load_model_from_mps_{file,string} (solver/core/src/mps_io.cpp:71-249):
                        fixed or free records, OBJSENSE / OBJNAME, N/L/E/G rows, INTORG markers,
                        RHS (the objective row's entry is minus the constant), RANGES, and
                        UP/LO/FX/FR/MI/PL/BV/LI/UI bounds; the first RHS / RANGES / BOUNDS set
                        wins; integers without bounds get [0, 1]
  load_model_from_file  extension dispatch (mps_io.cpp:251-279)
  model_to_json         model_to_json_string (json_io.cpp:76-107)
  result_to_json        result_to_json_string (json_io.cpp:190-240)
  verify                SolutionVerifier::verify (solver/core/src/verifier.cpp:10-150)

The quadratic objective is 1/2 sum over stored terms q_ij x_i x_j
(verifier.cpp:71-79): a diagonal term enters Q_ii, an off-diagonal stored
term (i, j) adds q_ij / 2 to Q_ij and to Q_ji (qp_interior_point.cpp:65-100).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import problems as P

INF = 1e30
STATUS = {0: "OPTIMAL", 1: "FEASIBLE", 2: "INFEASIBLE", 3: "UNBOUNDED", 4: "TIME_LIMIT", 5: "ITERATION_LIMIT",
          6: "NUMERICAL_ERROR", 7: "ERROR", 8: "NOT_IMPLEMENTED"}
CONCLUSIVE = ("OPTIMAL", "INFEASIBLE", "UNBOUNDED")


@dataclass
class Model:
    """OptimizationModel: named variables and constraints, in file order."""
    problem_type: str = "LP"  # LP | QP | MILP
    sense: str = "minimize"
    var_names: list = field(default_factory=list)
    var_types: list = field(default_factory=list)  # continuous | integer | binary
    lb: list = field(default_factory=list)
    ub: list = field(default_factory=list)
    obj_linear: dict = field(default_factory=dict)
    obj_constant: float = 0.0
    quadratic: dict = field(default_factory=dict)  # name -> {name -> coefficient}
    con_names: list = field(default_factory=list)
    con_senses: list = field(default_factory=list)  # "<=" | ">=" | "="
    con_rhs: list = field(default_factory=list)
    con_linear: list = field(default_factory=list)  # [{name -> coefficient}]

    @property
    def n_vars(self):
        return len(self.var_names)

    def index(self):
        return {n: j for j, n in enumerate(self.var_names)}

    def is_int(self):
        return np.array([t in ("integer", "binary") for t in self.var_types], dtype=np.int8)

    def to_lp_model(self) -> P.LpModel:
        """The array image the C ABI takes (sov_lp_model)."""
        idx = self.index()
        nv, nr = self.n_vars, len(self.con_names)
        cols_r = [[] for _ in range(nv)]
        cols_v = [[] for _ in range(nv)]
        for r, lin in enumerate(self.con_linear):
            for name, val in lin.items():
                if name not in idx:
                    raise ValueError(f"constraint {self.con_names[r]!r} names unknown variable {name!r}")
                cols_r[idx[name]].append(r)
                cols_v[idx[name]].append(float(val))
        ptr, rows, vals = P.csc_from_columns(cols_r, cols_v, nr)
        obj = np.zeros(nv)
        for name, val in self.obj_linear.items():
            if name not in idx:
                raise ValueError(f"objective names unknown variable {name!r}")
            obj[idx[name]] = float(val)
        senses = "".join({"<=": "L", ">=": "G", "=": "E"}[s] for s in self.con_senses)
        m = P.LpModel(n_vars=nv, n_rows=nr, col_ptr=ptr, row_idx=rows, values=vals, obj=obj,
                      lb=np.array(self.lb, dtype=float), ub=np.array(self.ub, dtype=float), row_sense=senses,
                      rhs=np.array(self.con_rhs, dtype=float), obj_constant=float(self.obj_constant),
                      maximize=self.sense == "maximize")
        m.is_int = self.is_int()
        return m

    def hessian(self) -> np.ndarray:
        """Dense Q of 1/2 x'Qx (model sense)."""
        idx = self.index()
        Q = np.zeros((self.n_vars, self.n_vars))
        for a, row in self.quadratic.items():
            for b, v in row.items():
                i, j = idx[a], idx[b]
                if i == j:
                    Q[i, i] += v
                else:
                    Q[i, j] += 0.5 * v
                    Q[j, i] += 0.5 * v
        return Q

    def objective_of(self, x) -> float:
        idx = self.index()
        val = self.obj_constant + sum(v * x[idx[n]] for n, v in self.obj_linear.items())
        for a, row in self.quadratic.items():
            for b, q in row.items():
                val += 0.5 * q * x[idx[a]] * x[idx[b]]
        return float(val)


# --------------------------------------------------------------------- JSON --

_PTYPE = {"LP": "LP", "lp": "LP", "QP": "QP", "qp": "QP", "MILP": "MILP", "milp": "MILP"}
_SENSE = {"minimize": "minimize", "min": "minimize", "maximize": "maximize", "max": "maximize"}
_CSENSE = {"<=": "<=", "le": "<=", "LE": "<=", ">=": ">=", "ge": ">=", "GE": ">=", "=": "=", "==": "=", "eq": "=",
           "EQ": "="}
_VTYPE = ("continuous", "integer", "binary")


def _lookup(table, key, what):
    if key not in table:
        raise ValueError(f"Unknown {what}: {key}")
    return table[key]


def load_json_string(text: str) -> Model:
    return _parse_json(json.loads(text))


def load_json(path: str) -> Model:
    with open(path, "rb") as fh:
        return _parse_json(json.loads(fh.read()))


def _parse_json(j) -> Model:
    m = Model()
    m.problem_type = _lookup(_PTYPE, j["problem_type"], "problem_type")
    m.sense = _lookup(_SENSE, j.get("sense", "minimize"), "sense")
    for v in j["variables"]:
        t = v.get("type", "continuous")
        if t not in _VTYPE:
            raise ValueError(f"Unknown variable type: {t}")
        m.var_names.append(v["name"])
        m.var_types.append(t)
        m.lb.append(float(v.get("lower_bound", 0.0)))
        m.ub.append(float(v.get("upper_bound", 1.0 if t == "binary" else INF)))
    obj = j.get("objective", {})
    m.obj_constant = float(obj.get("constant", 0.0))
    m.obj_linear = {k: float(v) for k, v in obj.get("linear", {}).items()}
    m.quadratic = {a: {b: float(q) for b, q in row.items()} for a, row in obj.get("quadratic", {}).items()}
    for c in j.get("constraints", []):
        m.con_names.append(c.get("name", ""))
        m.con_senses.append(_lookup(_CSENSE, c["sense"], "constraint sense"))
        m.con_rhs.append(float(c["rhs"]))
        m.con_linear.append({k: float(v) for k, v in c.get("linear", {}).items()})
    return m


def model_to_json(m: Model) -> dict:
    out = {"problem_type": m.problem_type, "sense": m.sense,
           "variables": [{"name": n, "type": t, "lower_bound": l, "upper_bound": u}
                         for n, t, l, u in zip(m.var_names, m.var_types, m.lb, m.ub)],
           "objective": {"linear": dict(m.obj_linear), "constant": m.obj_constant}}
    if m.quadratic:
        out["objective"]["quadratic"] = m.quadratic
    out["constraints"] = [{"name": n, "linear": dict(lin), "sense": s, "rhs": r}
                          for n, lin, s, r in zip(m.con_names, m.con_linear, m.con_senses, m.con_rhs)]
    return out


# ---------------------------------------------------------------------- MPS --

def _number(text: str) -> float:
    s = text.replace("d", "E").replace("D", "E")
    for i in range(1, len(s) - 1):  # "1.5-3" -> "1.5E-3"
        if s[i] in "+-" and s[i - 1].isdigit():
            s = s[:i] + "E" + s[i:]
            break
    try:
        v = float(s)
    except ValueError:
        raise ValueError(f"invalid number: {text}") from None
    if not math.isfinite(v):
        raise ValueError(f"invalid number: {text}")
    return v


def _slice(line, a, n):
    return line[a:a + n].strip() if a < len(line) else ""


def _fields(line: str, section: str, fixed: bool):
    """Fixed-column fields when the card's separator columns are blank (names
    may then hold spaces), else whitespace-separated tokens."""
    if fixed and len(line) >= 5 and not _slice(line, 3, 1) and not _slice(line, 12, 2):
        if section == "ROWS" and not _slice(line, 0, 1) and len(_slice(line, 1, 2)) == 1:
            return [_slice(line, 1, 2), _slice(line, 4, 8)]
        if len(line) >= 22 and not _slice(line, 22, 2) and not _slice(line, 36, 3) and not _slice(line, 47, 2):
            if section == "BOUNDS":
                f = [_slice(line, 1, 2), _slice(line, 4, 8), _slice(line, 14, 8)]
                if _slice(line, 24, 12):
                    f.append(_slice(line, 24, 12))
                kind = f[0].upper()
                needs = kind not in ("FR", "MI", "PL", "BV")
                if f[2] and len(f) == (4 if needs else 3) and not line[36:].strip():
                    return f
                return line.split()
            if section in ("COLUMNS", "RHS", "RANGES"):
                if _slice(line, 14, 8) == "'MARKER'":
                    return [_slice(line, 4, 8), "'MARKER'", _slice(line, 39, 8)]
                f = [_slice(line, 4, 8), _slice(line, 14, 8), _slice(line, 24, 12)]
                if _slice(line, 39, 8):
                    f += [_slice(line, 39, 8), _slice(line, 49, 12)]
                try:
                    _number(f[2])
                    if len(f) == 5:
                        _number(f[4])
                    return f
                except ValueError:
                    pass
    return line.split()


def load_mps_string(text: str, fixed_format: bool = True, max_variables=5_000_000, max_constraints=5_000_000,
                    max_nonzeros=200_000_000) -> Model:
    m = Model()
    rows = []  # [name, type, rhs, range, ranged]
    row_id, col_id = {}, {}
    entries, bounded = [], []
    section = objective = requested = last_col = ""
    sets = {"RHS": None, "RANGES": None, "BOUNDS": None}
    int_block = saw_rows = saw_cols = ended = False
    nnz = 0

    def add_column(name):
        if not name:
            raise ValueError("missing column name")
        if name in col_id:
            return col_id[name]
        if len(m.var_names) >= max_variables:
            raise ValueError("variable limit exceeded")
        col_id[name] = len(m.var_names)
        m.var_names.append(name)
        m.var_types.append("continuous")
        m.lb.append(0.0)
        m.ub.append(INF)
        entries.append({})
        bounded.append(False)
        return col_id[name]

    def selected(kind, name):
        if sets[kind] is None:
            sets[kind] = name
        return name == sets[kind]

    for lineno, raw in enumerate(text.splitlines(), start=1):
        try:
            if lineno == 1 and raw.startswith("﻿"):
                raw = raw[1:]
            if not raw.strip() or raw[0] == "*" or raw.strip()[0] == "$":
                continue
            cut = raw.find(" $")
            if cut >= 0:
                raw = raw[:cut]
            if not raw[0].isspace():
                f = raw.split()
                head = f[0].upper()
                if head == "ENDATA":
                    ended = True
                    break
                if head not in ("NAME", "ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS", "OBJSENSE", "OBJNAME"):
                    raise ValueError(f"unsupported MPS section: {head}")
                section = head
                saw_rows = saw_rows or head == "ROWS"
                saw_cols = saw_cols or head == "COLUMNS"
                if head in ("OBJSENSE", "OBJNAME") and len(f) > 1:
                    if head == "OBJNAME":
                        requested = f[1]
                    else:
                        m.sense = _objsense(f[1])
                    section = "NAME"
                continue
            f = _fields(raw, section, fixed_format)
            if not f:
                continue
            if section in ("OBJSENSE", "OBJNAME"):
                if section == "OBJNAME":
                    requested = raw.strip()
                else:
                    m.sense = _objsense(f[0])
                section = "NAME"
            elif section == "ROWS":
                if len(f) != 2 or len(f[0]) != 1 or not f[1]:
                    raise ValueError("malformed ROWS record")
                kind = f[0].upper()
                if kind not in "NLEG":
                    raise ValueError("invalid row type")
                if f[1] in row_id:
                    raise ValueError(f"duplicate row: {f[1]}")
                if len(rows) >= max_constraints:
                    raise ValueError("row limit exceeded")
                row_id[f[1]] = len(rows)
                rows.append([f[1], kind, 0.0, 0.0, False])
                if kind == "N" and not objective:
                    objective = f[1]
            elif section == "COLUMNS":
                if "'MARKER'" in f:
                    k = f.index("'MARKER'")
                    if k + 1 >= len(f):
                        raise ValueError("missing integer marker kind")
                    kind = f[k + 1]
                    if kind == "'INTORG'" and not int_block:
                        int_block = True
                    elif kind == "'INTEND'" and int_block:
                        int_block = False
                    else:
                        raise ValueError("unmatched or nested integer marker")
                    last_col = ""
                    continue
                if len(f) < 3 or len(f) % 2 != 1:
                    raise ValueError("malformed COLUMNS record")
                if not f[0]:
                    f[0] = last_col
                if f[0] in col_id and last_col != f[0]:
                    raise ValueError(f"noncontiguous column: {f[0]}")
                c = add_column(f[0])
                last_col = f[0]
                if int_block:
                    m.var_types[c] = "integer"
                for k in range(1, len(f), 2):
                    if f[k] not in row_id:
                        raise ValueError(f"unknown row: {f[k]}")
                    nnz += 1
                    if nnz > max_nonzeros:
                        raise ValueError("nonzero limit exceeded")
                    entries[c][f[k]] = entries[c].get(f[k], 0.0) + _number(f[k + 1])
                    if not math.isfinite(entries[c][f[k]]):
                        raise ValueError("coefficient overflow")
            elif section in ("RHS", "RANGES"):
                start = len(f) % 2
                setname = f[0] if start else ""
                if len(f) - start < 2:
                    raise ValueError("malformed row/value pairs")
                if not selected(section, setname):
                    continue
                for k in range(start, len(f), 2):
                    if f[k] not in row_id:
                        raise ValueError(f"unknown row: {f[k]}")
                    r = rows[row_id[f[k]]]
                    val = _number(f[k + 1])
                    if section == "RHS":
                        r[2] = val
                    else:
                        r[3], r[4] = val, True
            elif section == "BOUNDS":
                kind = f[0].upper()
                valued = kind not in ("FR", "MI", "PL", "BV")
                expected = 4 if valued else 3
                if len(f) == expected - 1:
                    f.insert(1, "")
                if len(f) != expected:
                    raise ValueError("malformed BOUNDS record")
                if not selected("BOUNDS", f[1]):
                    continue
                c = add_column(f[2])
                val = _number(f[3]) if valued else 0.0
                if kind in ("UP", "UI"):
                    m.ub[c] = val
                    if not bounded[c] and val < 0:
                        m.lb[c] = -INF
                elif kind in ("LO", "LI"):
                    m.lb[c] = val
                elif kind == "FX":
                    m.lb[c] = m.ub[c] = val
                elif kind == "FR":
                    m.lb[c], m.ub[c] = -INF, INF
                elif kind == "MI":
                    m.lb[c] = -INF
                elif kind == "PL":
                    m.ub[c] = INF
                elif kind == "BV":
                    m.var_types[c], m.lb[c], m.ub[c] = "binary", 0.0, 1.0
                else:
                    raise ValueError(f"unsupported bound type: {kind} (semi-continuous bounds require an explicit "
                                     "reformulation)")
                if kind in ("LI", "UI"):
                    m.var_types[c] = "integer"
                bounded[c] = True
            else:
                raise ValueError("unexpected data record")
        except ValueError as e:
            raise ValueError(f"MPS line {lineno}: {e}") from None
    if not (saw_rows and saw_cols and ended):
        raise ValueError("MPS requires ROWS, COLUMNS and ENDATA")
    if int_block:
        raise ValueError("unterminated INTORG block")
    if requested:
        objective = requested
    if objective and (objective not in row_id or rows[row_id[objective]][1] != "N"):
        raise ValueError("OBJNAME must select an N row")
    for c, name in enumerate(m.var_names):
        if m.var_types[c] == "integer" and not bounded[c]:
            m.ub[c] = 1.0
        if m.lb[c] > m.ub[c]:
            raise ValueError(f"crossed bounds: {name}")
        if m.var_types[c] != "continuous":
            m.problem_type = "MILP"
        if objective in entries[c]:
            m.obj_linear[name] = entries[c][objective]
    if objective:
        m.obj_constant = -rows[row_id[objective]][2]
    row_entries = [dict() for _ in rows]
    for c, ent in enumerate(entries):
        for rname, val in ent.items():
            if val != 0:
                row_entries[row_id[rname]][m.var_names[c]] = val
    for r, (name, kind, rhs, rng, ranged) in enumerate(rows):
        if kind == "N":
            continue
        sense = {"L": "<=", "G": ">=", "E": "="}[kind]
        if ranged:
            lo = hi = rhs
            if kind == "G" or (kind == "E" and rng >= 0):
                hi += abs(rng)
            else:
                lo -= abs(rng)
            m.con_names.append(name)
            m.con_senses.append("<=")
            m.con_rhs.append(hi)
            m.con_linear.append(dict(row_entries[r]))
            name2 = name + "_lo"
            while name2 in row_id:
                name2 += "_lo"
            name, sense, rhs = name2, ">=", lo
        m.con_names.append(name)
        m.con_senses.append(sense)
        m.con_rhs.append(rhs)
        m.con_linear.append(dict(row_entries[r]))
        if len(m.con_names) > max_constraints:
            raise ValueError("expanded row limit exceeded")
    if sum(len(c) for c in m.con_linear) > max_nonzeros:
        raise ValueError("expanded nonzero limit exceeded")
    return m


def _objsense(word: str) -> str:
    w = word.upper()
    if w not in ("MAX", "MIN", "MAXIMIZE", "MINIMIZE"):
        raise ValueError("invalid OBJSENSE")
    return "maximize" if w.startswith("MAX") else "minimize"


def load_mps(path: str, fixed_format: bool = True) -> Model:
    with open(path, "rb") as fh:
        return load_mps_string(fh.read().decode("utf-8", errors="replace"), fixed_format=fixed_format)


def load_model_from_file(path: str) -> Model:
    ext = os.path.splitext(path)[1].lower().lstrip(".")
    if ext == "mps":
        return load_mps(path)
    if ext == "json":
        return load_json(path)
    if ext in ("gz", "bz2", "zip"):
        raise ValueError(f"compressed model files ({ext}) are not supported yet; decompress to .mps or .json first")
    if not ext:
        raise ValueError(f"cannot infer the model format: '{path}' has no extension. Use .mps or .json.")
    raise ValueError(f"unsupported model extension '.{ext}'. Supported: .mps, .json")


# ------------------------------------------------------------------ results --

@dataclass
class Result:
    """SolverResult (types.hpp:117-155) for the parts this build fills."""
    status: str = "NOT_IMPLEMENTED"
    objective_value: float | None = None
    primal: dict = field(default_factory=dict)
    dual: dict = field(default_factory=dict)
    slacks: dict = field(default_factory=dict)
    optimality_gap: float = 0.0
    duality_gap: float = 0.0
    primal_residual: float = 0.0
    dual_residual: float = 0.0
    iterations: int = 0
    nodes: int = 0
    runtime_seconds: float = 0.0
    message: str = ""
    warnings: list = field(default_factory=list)
    mip: dict = field(default_factory=dict)
    dual_certificate_space: str = ""


def result_to_json(r: Result, include_primal: bool = True) -> dict:
    mip = {"has_best_bound": False, "best_bound": None, "time_to_first_incumbent": None, "cuts_by_family": {},
           "presolve_fixed_variables": 0, "presolve_substituted_variables": 0, "presolve_removed_constraints": 0,
           "presolve_tightened_bounds": 0, "presolve_passes": 0, "node_lp_failures": 0, "dropped_subtrees": 0,
           "numerical_error_nodes": 0, "iteration_limit_nodes": 0, "unbounded_nodes": 0,
           "cut_validity_rejections": 0, "debug_solution_enabled": False, "debug_solution_violation": False,
           "debug_solution_checks": 0, "debug_solution_path": None, "debug_solution_first_violation": None}
    mip.update(r.mip)
    out = {"status": r.status, "objective_value": r.objective_value}
    if include_primal:
        out["primal"] = r.primal
    else:
        out["primal"] = None
        out["primal_variables"] = len(r.primal)
    out.update({
        "dual": r.dual, "slacks": r.slacks,
        "dual_certificate_space": r.dual_certificate_space or None,
        "lp_diagnostics": {}, "mip_diagnostics": mip, "optimality_gap": r.optimality_gap,
        "duality_gap": r.duality_gap, "primal_residual": r.primal_residual, "dual_residual": r.dual_residual,
        "presolve": {"fixed_variables": 0, "substituted_variables": 0, "removed_constraints": 0,
                     "tightened_bounds": 0, "passes": 0},
        "optimality_proven": r.status in CONCLUSIVE, "iterations": r.iterations, "nodes": r.nodes,
        "runtime_seconds": r.runtime_seconds, "message": r.message, "warnings": r.warnings,
    })
    return out


def verify(m: Model, r: Result, tol: float = 1e-6) -> dict:
    """SolutionVerifier::verify: bound, row and integrality violations of the
    reported point, the recomputed objective, and the status claims."""
    out = {"is_valid": False, "max_constraint_violation": 0.0, "max_bound_violation": 0.0,
           "max_integrality_violation": 0.0, "recomputed_objective": 0.0, "issues": [], "message": "Verifier executed."}
    if r.status in ("NOT_IMPLEMENTED", "ERROR", "NUMERICAL_ERROR", "TIME_LIMIT", "ITERATION_LIMIT"):
        out["issues"].append(f"Solver status does not provide a feasible primal solution: {r.status}")
        out["message"] = "Verification skipped for non-feasible status."
        return out
    if r.status in ("INFEASIBLE", "UNBOUNDED"):
        out["issues"].append(f"Status {r.status} has no independently checkable certificate.")
        out["message"] = "Status claim is unverified."
        return out
    x = r.primal
    for name, t, lo, hi in zip(m.var_names, m.var_types, m.lb, m.ub):
        if name not in x:
            out["issues"].append(f"Missing primal value for variable {name}")
            continue
        v = x[name]
        if not math.isfinite(v):
            out["issues"].append(f"Non-finite primal for {name}")
            continue
        out["max_bound_violation"] = max(out["max_bound_violation"], lo - v, v - hi)
        if t in ("integer", "binary"):
            out["max_integrality_violation"] = max(out["max_integrality_violation"], abs(v - round(v)))
    obj = m.obj_constant + sum(val * x[n] for n, val in m.obj_linear.items() if n in x)
    for a, row in m.quadratic.items():
        for b, q in row.items():
            if a in x and b in x:
                obj += 0.5 * q * x[a] * x[b]
    out["recomputed_objective"] = obj
    for name, lin, s, rhs in zip(m.con_names, m.con_linear, m.con_senses, m.con_rhs):
        lhs = sum(val * x[n] for n, val in lin.items() if n in x)
        if not math.isfinite(lhs):
            out["issues"].append(f"Non-finite constraint activity: {name}")
            continue
        viol = lhs - rhs if s == "<=" else rhs - lhs if s == ">=" else abs(lhs - rhs)
        out["max_constraint_violation"] = max(out["max_constraint_violation"], viol)
    if out["max_bound_violation"] > tol:
        out["issues"].append("Bound violation exceeds tolerance.")
    if out["max_constraint_violation"] > tol:
        out["issues"].append("Constraint violation exceeds tolerance.")
    if out["max_integrality_violation"] > tol:
        out["issues"].append("Integrality violation exceeds tolerance.")
    if r.objective_value is not None and abs(r.objective_value - obj) > tol * (1.0 + abs(obj)):
        out["issues"].append(f"Reported objective differs from recomputed value by {abs(r.objective_value - obj)}")
    if r.status == "OPTIMAL" and r.optimality_gap > 1e-2:
        out["issues"].append("OPTIMAL status with large optimality_gap.")
    if r.status == "OPTIMAL" and r.duality_gap > tol:
        out["issues"].append(f"OPTIMAL status contradicted by reported relative duality gap {r.duality_gap} "
                             f"(tolerance {tol}). The point may be feasible but optimality is not established.")
    if r.status not in CONCLUSIVE and r.status != "FEASIBLE" and r.objective_value is not None \
            and r.optimality_gap > 1e-2:
        out["issues"].append(f"Status {r.status} is not a proof, but an objective value is attached with gap "
                             f"{r.optimality_gap}.")
    out["is_valid"] = not out["issues"]
    out["message"] = "Solution verified." if out["is_valid"] else "Solution failed verification."
    return out
