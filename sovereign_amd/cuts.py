"""Cut separation of the tree search over the C ABI (include/sovereign_cuts.h),
mirroring the reference's generators and validity gate
(solver/cuts/include/sovereign/cuts.hpp:17-55):

    generate_cover_cuts(milp, x)  -> separate(model, kind, x, COVER)
    generate_mir_cuts(milp, x)    -> separate(model, kind, x, GOMORY)
    generate_cmir_cuts(milp, x)   -> separate(model, kind, x, CMIR)
    check_cut_validity(milp, cut, points) -> validity(model, cut, points)

Host code in libsovereign_hip.so; no device is needed. Variable kinds are
the reference's VariableType: 0 continuous, 1 integer, 2 binary.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi as K
from .problems import LpModel

COVER, GOMORY, CMIR = 1, 2, 4
FAMILY_NAMES = {COVER: "cover", GOMORY: "gomory", CMIR: "cmir"}
CONTINUOUS, INTEGER, BINARY = 0, 1, 2
ACCEPTED, REMOVES_POINT, BOX_INFEASIBLE = 0, 1, 2


@dataclass
class Cut:
    """sum coefs[j] x_j <= rhs (every generated cut is a '<=' row)."""
    coefs: dict
    rhs: float
    family: int
    efficacy: float = 0.0


def kinds_from(model: LpModel, is_int) -> np.ndarray:
    """Binary for integer columns within [0, 1], integer for the other integer
    columns, continuous otherwise (the JSON reader's 'binary' type)."""
    is_int = np.asarray(is_int, dtype=bool)
    binary = is_int & (model.lb >= 0.0) & (model.ub <= 1.0)
    return np.where(binary, BINARY, np.where(is_int, INTEGER, CONTINUOUS)).astype(np.int8)


def separate(model: LpModel, kind, x, family: int, int_tol: float = 1e-6, limit: int = 8,
             base_rows: int = -1) -> list[Cut]:
    lib = K.load()
    st = model.fill_struct(K.LpModelC)
    kind = np.ascontiguousarray(kind, dtype=np.int8)
    x = K.f64(x)
    cap_cuts, cap_nnz = max(1, limit), max(1, limit) * max(1, model.n_vars)
    for _ in range(2):
        row_ptr = np.zeros(cap_cuts + 1, dtype=np.int32)
        col = np.zeros(cap_nnz, dtype=np.int32)
        val = np.zeros(cap_nnz)
        sense = C.create_string_buffer(cap_cuts)
        rhs = np.zeros(cap_cuts)
        fam = np.zeros(cap_cuts, dtype=np.int32)
        eff = np.zeros(cap_cuts)
        out = K.CutList(max_cuts=cap_cuts, max_nnz=cap_nnz, row_ptr=K.pi(row_ptr), col=K.pi(col), val=K.pf(val),
                        sense=C.cast(sense, C.c_char_p), rhs=K.pf(rhs), family=K.pi(fam), efficacy=K.pf(eff))
        rc = lib.sov_separate_cuts(C.byref(st), kind.ctypes.data_as(C.POINTER(C.c_int8)), K.pf(x), family,
                                   int_tol, limit, base_rows, C.byref(out))
        if rc == K.SOV_OK:
            cuts = []
            for k in range(out.n_cuts):
                a, b = row_ptr[k], row_ptr[k + 1]
                cuts.append(Cut({int(j): float(v) for j, v in zip(col[a:b], val[a:b])}, float(rhs[k]), int(fam[k]),
                                float(eff[k])))
            return cuts
        if out.n_cuts <= cap_cuts and out.nnz <= cap_nnz:
            K.check(rc, "sov_separate_cuts")
        cap_cuts, cap_nnz = max(cap_cuts, out.n_cuts), max(cap_nnz, int(out.nnz))
    K.check(rc, "sov_separate_cuts")
    return []


def validity(model: LpModel, cut: Cut, points=(), sense: str = "L", tol: float = 1e-6) -> int:
    """check_cut_validity: ACCEPTED, REMOVES_POINT or BOX_INFEASIBLE."""
    lib = K.load()
    st = model.fill_struct(K.LpModelC)
    cols = np.array(list(cut.coefs.keys()), dtype=np.int32)
    vals = np.array(list(cut.coefs.values()), dtype=np.float64)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, model.n_vars) if len(points) else \
        np.zeros((0, model.n_vars))
    verdict = C.c_int32(-1)
    K.check(lib.sov_cut_validity(C.byref(st), len(cols), K.pi(cols), K.pf(vals), sense.encode(), float(cut.rhs),
                                 len(pts), K.pf(pts), float(tol), C.byref(verdict)), "sov_cut_validity")
    return verdict.value
