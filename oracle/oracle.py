"""ctypes wrapper of the CPU oracle (liboracle.so) and, where it was built,
of the reference's own CPU code (_ref/libsovref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsovref.so")

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class OrcLpModel(C.Structure):
    _fields_ = [
        ("n_vars", C.c_int32), ("n_rows", C.c_int32), ("col_ptr", _i32p), ("row_idx", _i32p),
        ("values", _f64p), ("obj", _f64p), ("obj_constant", C.c_double), ("maximize", C.c_int32),
        ("lb", _f64p), ("ub", _f64p), ("row_sense", C.c_char_p), ("rhs", _f64p),
    ]


class OrcOptions(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32), ("feasibility_tol", C.c_double), ("optimality_tol", C.c_double),
        ("fraction_to_boundary", C.c_double), ("enable_scaling", C.c_int32), ("normal_mode", C.c_int32),
    ]


class OrcResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iterations", C.c_int32), ("objective", C.c_double),
        ("duality_gap", C.c_double), ("primal_residual", C.c_double), ("dual_residual", C.c_double),
        ("factorizations", C.c_int32), ("used_sparse", C.c_int32), ("factor_seconds", C.c_double),
        ("total_seconds", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = {
    "orc_default_options": (None, [C.POINTER(OrcOptions)]),
    "orc_lp_solve": (C.c_int, [C.POINTER(OrcLpModel), C.POINTER(OrcOptions), C.POINTER(OrcResult), _f64p, _f64p, _f64p]),
    "orc_form_build": (C.c_void_p, [C.POINTER(OrcLpModel), C.c_int]),
    "orc_form_dims": (None, [C.c_void_p, _i32p, _i32p, _i32p, _i32p]),
    "orc_form_get": (None, [C.c_void_p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
    "orc_form_free": (None, [C.c_void_p]),
    "orc_solve_ipm_form": (C.c_int, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, C.POINTER(OrcOptions),
                                     C.POINTER(OrcResult), _f64p, _f64p, _f64p]),
    "orc_spmv_csc": (None, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]),
    "orc_dense_lu_factor": (C.c_int, [C.c_int32, _f64p, _i32p]),
    "orc_dense_lu_solve": (None, [C.c_int32, _f64p, _i32p, _f64p]),
    "orc_dense_lu_solve_transpose": (None, [C.c_int32, _f64p, _i32p, _f64p]),
    "orc_normal_eq_dense": (None, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p]),
    "orc_normal_eq_pattern_nnz": (C.c_int64, [C.c_int32, C.c_int32, _i32p, _i32p]),
    "orc_normal_eq_pattern": (None, [C.c_int32, C.c_int32, _i32p, _i32p, _i32p, _i32p]),
    "orc_ne_create": (C.c_void_p, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_int32]),
    "orc_ne_factor": (C.c_int, [C.c_void_p, _f64p]),
    "orc_ne_solve": (C.c_int, [C.c_void_p, _f64p]),
    "orc_ne_is_sparse": (C.c_int, [C.c_void_p]),
    "orc_ne_factor_nnz": (C.c_int64, [C.c_void_p]),
    "orc_ne_replaced_pivots": (C.c_int32, [C.c_void_p]),
    "orc_ne_free": (None, [C.c_void_p]),
    "orc_amd_order": (C.c_int, [C.c_int32, _i32p, _i32p, C.c_int64, _i32p, _i32p]),
    "orc_ldlt_factor_solve": (C.c_int, [C.c_int32, _i32p, _i32p, _f64p, C.c_double, C.c_double, _f64p, _i32p,
                                        _i64p, _i32p]),
    "orc_liswet_ipm": (C.c_int, [C.c_int32, _f64p, C.c_int32, C.c_int32, _f64p]),
    "orc_liswet_certify": (C.c_int, [C.c_int32, _f64p, _f64p, _f64p, _f64p]),
    "orc_liswet_polish": (C.c_int, [C.c_int32, _f64p, _f64p, _f64p, _f64p]),
}

_REF_SIGS = {
    "ref_separate_cuts": (C.c_int, [C.POINTER(OrcLpModel), C.POINTER(C.c_int8), _f64p, C.c_int32, C.c_double, C.c_int32,
                                    C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_int32), _f64p, _f64p]),
    "ref_cut_validity": (C.c_int, [C.POINTER(OrcLpModel), C.c_int32, C.POINTER(C.c_int32), _f64p, C.c_char, C.c_double,
                                   C.c_int32, _f64p, C.c_double]),
    "ref_qp_solve": (C.c_int, [C.c_int32, C.c_int32, _f64p, C.c_char_p, _f64p, _f64p, _f64p, _f64p, _f64p,
                               C.c_int32, C.c_double, _f64p, _f64p]),
    "ref_lp_solve": _SIGS["orc_lp_solve"],
    "ref_dense_lu_solve": (C.c_int, [C.c_int32, _f64p, _f64p, C.c_int]),
    "ref_spmv_csc": _SIGS["orc_spmv_csc"],
    "ref_normal_eq_pattern": (C.c_int64, [C.c_int32, C.c_int32, _i32p, _i32p, _i32p, _i32p]),
    "ref_sparse_normal_solve": (C.c_int, [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _i64p]),
    "ref_amd_order": (C.c_int, [C.c_int32, _i32p, _i32p, C.c_int64, _i32p, _i32p]),
}


def _bind(path, sigs):
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_oracle = None
_ref = None


def lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        _oracle = _bind(ORACLE_LIB, _SIGS)
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        _ref = _bind(REF_LIB, _REF_SIGS)
    return _ref


def pf(a):
    return a.ctypes.data_as(_f64p) if a is not None else None


def pi(a):
    return a.ctypes.data_as(_i32p) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def options(normal_mode: str = "auto", **kw) -> OrcOptions:
    o = OrcOptions()
    lib().orc_default_options(C.byref(o))
    o.normal_mode = {"auto": 0, "dense": 1, "sparse": 2}[normal_mode]
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _lp_solve(fn, model, opt):
    st = model.fill_struct(OrcLpModel)
    res = OrcResult()
    primal = np.zeros(model.n_vars)
    row_dual = np.zeros(model.n_rows)
    var_dual = np.zeros(model.n_vars)
    fn(C.byref(st), C.byref(opt), C.byref(res), pf(primal), pf(row_dual), pf(var_dual))
    out = res.as_dict()
    out.update(primal=primal, row_dual=row_dual, var_dual=var_dual)
    return out


def lp_solve(model, normal_mode="auto", **kw):
    """InteriorPointSolver::solve restated (oracle)."""
    return _lp_solve(lib().orc_lp_solve, model, options(normal_mode, **kw))


def ref_lp_solve(model, normal_mode="auto", **kw):
    """The reference's own InteriorPointSolver::solve (needs _ref)."""
    return _lp_solve(ref().ref_lp_solve, model, options(normal_mode, **kw))


def solve_form(lp, normal_mode="auto", **kw):
    o = options(normal_mode, **kw)
    res = OrcResult()
    x, y, s = np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n)
    lib().orc_solve_ipm_form(lp.m, lp.n, pi(i32(lp.col_ptr)), pi(i32(lp.row_idx)), pf(f64(lp.values)), pf(f64(lp.b)),
                             pf(f64(lp.c)), C.byref(o), C.byref(res), pf(x), pf(y), pf(s))
    out = res.as_dict()
    out.update(x=x, y=y, s=s)
    return out


def ipm_form(model, scaling=True):
    """build_ipm_form restated: dict of the IPM-form arrays."""
    st = model.fill_struct(OrcLpModel)
    h = lib().orc_form_build(C.byref(st), 1 if scaling else 0)
    if not h:
        raise ValueError("malformed model")
    m, n, nnz, ns = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    lib().orc_form_dims(h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(ns))
    m, n, nnz, ns = m.value, n.value, nnz.value, ns.value
    cp, ri, va = np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
    b, c, cs, rs, sh = np.zeros(m), np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(ns)
    lib().orc_form_get(h, pi(cp), pi(ri), pf(va), pf(b), pf(c), pf(cs), pf(rs), pf(sh))
    lib().orc_form_free(h)
    return dict(m=m, n=n, col_ptr=cp, row_idx=ri, values=va, b=b, c=c, col_scale=cs, row_scale=rs, shift=sh,
                n_structural=ns)


def spmv(nrows, ncols, col_ptr, row_idx, values, x, transpose=False, use_ref=False):
    y = np.zeros(ncols if transpose else nrows)
    fn = ref().ref_spmv_csc if use_ref else lib().orc_spmv_csc
    fn(nrows, ncols, pi(i32(col_ptr)), pi(i32(row_idx)), pf(f64(values)), pf(f64(x)), pf(y), 1 if transpose else 0)
    return y


def dense_lu(a):
    """DenseLU::factorize restated. a: numpy (n, n). Returns (status, lu, piv)."""
    n = a.shape[0]
    flat = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    piv = np.zeros(n, np.int32)
    st = lib().orc_dense_lu_factor(n, pf(flat), pi(piv))
    return st, flat.reshape((n, n), order="F"), piv


def dense_lu_solve(lu, piv, b, transpose=False):
    n = lu.shape[0]
    flat = np.asfortranarray(lu).reshape(-1, order="F").copy()
    x = f64(b).copy()
    fn = lib().orc_dense_lu_solve_transpose if transpose else lib().orc_dense_lu_solve
    fn(n, pf(flat), pi(i32(piv)), pf(x))
    return x


def ref_dense_lu_solve(a, b, transpose=False):
    n = a.shape[0]
    flat = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    x = f64(b).copy()
    st = ref().ref_dense_lu_solve(n, pf(flat), pf(x), 1 if transpose else 0)
    return st, x


def normal_eq_dense(m, n, col_ptr, row_idx, values, d):
    M = np.zeros(m * m)
    lib().orc_normal_eq_dense(m, n, pi(i32(col_ptr)), pi(i32(row_idx)), pf(f64(values)), pf(f64(d)), pf(M))
    return M.reshape((m, m), order="F")


def normal_eq_pattern(m, n, col_ptr, row_idx, use_ref=False):
    cp, ri = i32(col_ptr), i32(row_idx)
    if use_ref:
        nnz = ref().ref_normal_eq_pattern(m, n, pi(cp), pi(ri), None, None)
        pcp, pri = np.zeros(m + 1, np.int32), np.zeros(nnz, np.int32)
        ref().ref_normal_eq_pattern(m, n, pi(cp), pi(ri), pi(pcp), pi(pri))
        return pcp, pri
    nnz = lib().orc_normal_eq_pattern_nnz(m, n, pi(cp), pi(ri))
    pcp, pri = np.zeros(m + 1, np.int32), np.zeros(nnz, np.int32)
    lib().orc_normal_eq_pattern(m, n, pi(cp), pi(ri), pi(pcp), pi(pri))
    return pcp, pri


def amd_order(n, pcp, pri, fill_limit=0, use_ref=False):
    perm, iperm = np.zeros(n, np.int32), np.zeros(n, np.int32)
    fn = ref().ref_amd_order if use_ref else lib().orc_amd_order
    ok = fn(n, pi(i32(pcp)), pi(i32(pri)), fill_limit, pi(perm), pi(iperm))
    return ok, perm, iperm


class NormalEquations:
    """NormalEquations restated (CPU device policy)."""

    def __init__(self, m, n, col_ptr, row_idx, values, normal_mode="auto"):
        self._arrays = (i32(col_ptr), i32(row_idx), f64(values))
        self.m, self.n = m, n
        mode = {"auto": 0, "dense": 1, "sparse": 2}[normal_mode]
        self.h = lib().orc_ne_create(m, n, pi(self._arrays[0]), pi(self._arrays[1]), pf(self._arrays[2]), mode)

    def factor(self, d) -> bool:
        return lib().orc_ne_factor(self.h, pf(f64(d))) == 1

    def solve(self, rhs):
        x = f64(rhs).copy()
        if lib().orc_ne_solve(self.h, pf(x)) != 1:
            raise RuntimeError("solve failed")
        return x

    @property
    def is_sparse(self):
        return bool(lib().orc_ne_is_sparse(self.h))

    @property
    def replaced_pivots(self):
        return int(lib().orc_ne_replaced_pivots(self.h))

    def __del__(self):
        try:
            lib().orc_ne_free(self.h)
        except Exception:
            pass


def ldlt_factor_solve(pcp, pri, values, rhs=None, reg=0.0, tiny=0.0):
    """SparseLDLT restated on a lower-triangle pattern. Returns
    (status, x, replaced, factor_nnz, perm)."""
    n = len(pcp) - 1
    x = None if rhs is None else f64(rhs).copy()
    rep = C.c_int32(0)
    nnz = C.c_int64(0)
    perm = np.zeros(n, np.int32)
    st = lib().orc_ldlt_factor_solve(n, pi(i32(pcp)), pi(i32(pri)), pf(f64(values)), reg, tiny, pf(x), C.byref(rep),
                                     C.byref(nnz), pi(perm))
    return st, x, rep.value, nnz.value, perm


def ref_sparse_normal_solve(m, n, col_ptr, row_idx, values, d, rhs):
    x = f64(rhs).copy()
    lnnz = C.c_int64()
    st = ref().ref_sparse_normal_solve(m, n, pi(i32(col_ptr)), pi(i32(row_idx)), pf(f64(values)), pf(f64(d)), pf(x),
                                       C.byref(lnnz))
    return st, x, lnnz.value


LISWET_MODES = {"tolerance": 0, "reference_default": 1, "polish_last": 2}


def liswet_ipm(target, max_iterations=200, mode="polish_last"):
    """The reference's structured QP IPM restated (liswet_oracle.cpp).

    mode "tolerance": iterate to the structured tolerance (1e-10, 1e-12 for
    m <= 50) and return the last iterate; "reference_default": the
    reference's default (knot-space polish from complementarity 1e-6,
    :2473-2477); "polish_last": iterate to the tolerance or the iteration cap,
    then polish from the last iterate. Returns (iterations or a negative
    failure code: -2 method failure, -3 polish failure, x)."""
    t = f64(target)
    x = np.zeros(len(t))
    it = lib().orc_liswet_ipm(len(t), pf(t), max_iterations, LISWET_MODES[mode], pf(x))
    return it, x


def liswet_polish(target, s, z):
    """polish_liswet_knots restated: the certified x from the interior point
    (s, z), or None when the polish fails."""
    t = f64(target)
    x = np.zeros(len(t))
    rc = lib().orc_liswet_polish(len(t), pf(t), pf(f64(s)), pf(f64(z)), pf(x))
    return x if rc > 0 else None


def liswet_certify(target, x):
    """certify_liswet restated: dict of objective, primal, stationarity,
    complementarity, min_lambda, consistency, ok; and lambda."""
    t, xv = f64(target), f64(x)
    out = np.zeros(7)
    lam = np.zeros(len(t) - 2)
    lib().orc_liswet_certify(len(t), pf(t), pf(xv), pf(lam), pf(out))
    keys = ("objective", "primal", "stationarity", "complementarity", "min_lambda", "consistency", "ok")
    d = dict(zip(keys, out.tolist()))
    d["ok"] = bool(d["ok"])
    return d, lam


def ref_qp_solve(A, senses, rhs, lb, ub, c, Q, maximize=False, constant=0.0):
    """The reference's own QpInteriorPointSolver (needs _ref) on a dense QP:
    1/2 x'Qx + c'x + constant, rows A x (sense) rhs."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    m, n = A.shape
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    out = np.zeros(6)
    x = np.zeros(n)
    ref().ref_qp_solve(n, m, pf(A), "".join(senses).encode(), pf(f64(rhs)), pf(f64(lb)), pf(f64(ub)), pf(f64(c)),
                       pf(Q), 1 if maximize else 0, float(constant), pf(out), pf(x))
    return {"status": int(out[0]), "iterations": int(out[1]), "objective": out[2], "duality_gap": out[3],
            "primal_residual": out[4], "dual_residual": out[5], "x": x}


def ref_separate_cuts(model, kind, x, family, int_tol=1e-6, limit=8, base_rows=-1):
    """The reference's generate_cover_cuts (family 1), generate_mir_cuts (2) or
    generate_cmir_cuts (4) on `model` at x (needs _ref). kind: 0 continuous,
    1 integer, 2 binary. Returns [(dict var -> coef, rhs)], all '<=' rows."""
    st = model.fill_struct(OrcLpModel)
    kind = np.ascontiguousarray(kind, dtype=np.int8)
    x = f64(x)
    cap_cuts, cap_nnz = 256, 1 << 20
    n_cuts, nnz = C.c_int32(0), C.c_int64(0)
    row_ptr = np.zeros(cap_cuts + 1, dtype=np.int32)
    col = np.zeros(cap_nnz, dtype=np.int32)
    val = np.zeros(cap_nnz)
    rhs = np.zeros(cap_cuts)
    rc = ref().ref_separate_cuts(C.byref(st), kind.ctypes.data_as(C.POINTER(C.c_int8)), pf(x), family, int_tol, limit,
                                 base_rows, cap_cuts, cap_nnz, C.byref(n_cuts), C.byref(nnz), pi(row_ptr), pi(col),
                                 pf(val), pf(rhs))
    if rc != 0:
        raise RuntimeError("ref_separate_cuts: output capacity exceeded")
    out = []
    for k in range(n_cuts.value):
        a, b = row_ptr[k], row_ptr[k + 1]
        out.append(({int(j): float(v) for j, v in zip(col[a:b], val[a:b])}, float(rhs[k])))
    return out


def ref_cut_validity(model, coefs, sense, rhs, points, tol=1e-6):
    """check_cut_validity (needs _ref): 0 accepted, 1 a point is cut off, 2 the
    bound box cannot satisfy the cut."""
    st = model.fill_struct(OrcLpModel)
    cols = np.array(list(coefs.keys()), dtype=np.int32)
    vals = np.array(list(coefs.values()), dtype=np.float64)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, model.n_vars) if len(points) else np.zeros((0, model.n_vars))
    return int(ref().ref_cut_validity(C.byref(st), len(cols), pi(cols), pf(vals), sense.encode(), float(rhs), len(pts),
                                      pf(pts), float(tol)))
