"""CPU restatement of the reference's generic convex-QP interior point, for
tests only (the oracle): nothing in the product imports it.

    build_qp_form   build_qp_form, solver/qp/src/qp_interior_point.cpp:118-247
                    (lower-bound shift, finite upper bounds as extra rows with a
                    slack, Le slack / Ge surplus columns, free columns for
                    lb <= -1e29, the linear term c + Q shift, 1e-14 on the
                    structural diagonal of Q, :65-100)
    solve_qp_ipm    solve_qp_ipm, :2848-3116: the default starting point
                    (:2883-2947 without mehrotra_start) or Mehrotra's
                    (:2886-2932), the residuals, the stopping test
                    p_res, d_res < feasibility_tol and gap < optimality_tol
                    (:2972-2989), predictor-corrector with sigma = (mu_aff/mu)^3,
                    tau = 0.999 and a common step when Q != 0 (:3030-3104)
    kkt             QpKkt's dense path (:538-556): DenseLU of
                    [Q + X^-1 S, -A^T; A, -1e-12 I], singular below 1e-14
    qp_solve        QpInteriorPointSolver::solve's generic branch (:3166-3180):
                    the default start, then Mehrotra's when it fails

Inputs are dense arrays: the objective is 1/2 x'Qx + c'x + constant with Q
symmetric (model sense), rows A x (sense) rhs with senses 'L', 'G', 'E'.
Status codes follow SolverStatus: 0 Optimal, 7 Error.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

INF_BOUND = 1e29


def build_qp_form(A, senses, rhs, lb, ub, c, Q, maximize=False, constant=0.0):
    A = np.asarray(A, dtype=float)
    m0, ns = A.shape
    sign = -1.0 if maximize else 1.0
    lb = np.asarray(lb, dtype=float)
    ub = np.asarray(ub, dtype=float)
    free_s = ~(lb > -INF_BOUND)
    shift = np.where(free_s, 0.0, lb)
    ub_rows = [j for j in range(ns) if np.isfinite(ub[j]) and ub[j] < INF_BOUND]
    for j in ub_rows:
        if not free_s[j] and ub[j] - shift[j] < -1e-12:
            raise ValueError(f"Inconsistent bounds for variable {j}")
    m = m0 + len(ub_rows)
    rows = np.zeros((m, ns))
    rows[:m0] = A
    b = np.zeros(m)
    b[:m0] = np.asarray(rhs, dtype=float) - A @ shift
    kinds = list(senses) + ["L"] * len(ub_rows)
    for k, j in enumerate(ub_rows):
        rows[m0 + k, j] = 1.0
        b[m0 + k] = ub[j] - shift[j]
    nslack = sum(1 for kd in kinds if kd != "E")
    n = ns + nslack
    Af = np.zeros((m, n))
    Af[:, :ns] = rows
    col = ns
    for r, kd in enumerate(kinds):
        if kd == "L":
            Af[r, col] = 1.0
            col += 1
        elif kd == "G":
            Af[r, col] = -1.0
            col += 1
    Qf = np.zeros((n, n))
    Qf[:ns, :ns] = sign * np.asarray(Q, dtype=float)
    Qf[np.arange(ns), np.arange(ns)] += 1e-14
    shift_full = np.zeros(n)
    shift_full[:ns] = shift
    cf = np.zeros(n)
    cf[:ns] = sign * np.asarray(c, dtype=float) + (Qf @ shift_full)[:ns]
    free = np.zeros(n, dtype=bool)
    free[:ns] = free_s
    offset = cf @ shift_full - 0.5 * shift_full @ Qf @ shift_full + sign * constant
    return {"A": Af, "b": b, "c": cf, "Q": Qf, "free": free, "shift": shift, "ns": ns, "n": n, "m": m,
            "sign": sign, "offset": offset, "constant": constant, "c_model": np.asarray(c, dtype=float),
            "Q_model": np.asarray(Q, dtype=float)}


class Kkt:
    """[Q + H, -A^T; A, -1e-12 I] by dense LU; H = X^-1 S (0 on free
    columns), or the identity with Q left out for the starting point."""

    def __init__(self, f):
        self.f = f

    def factor(self, x, s, with_hessian=True):
        f = self.f
        n, m = f["n"], f["m"]
        h = np.ones(n)
        if with_hessian:
            h = np.where(f["free"], 0.0, np.maximum(s, 1e-16) / np.maximum(x, 1e-16))
        K = np.zeros((n + m, n + m))
        if with_hessian:
            K[:n, :n] = f["Q"]
        K[np.arange(n), np.arange(n)] += h
        K[n:, :n] = f["A"]
        K[:n, n:] = -f["A"].T
        K[np.arange(n, n + m), np.arange(n, n + m)] -= 1e-12
        self.lu = sla.lu_factor(K, check_finite=False)
        return bool(np.min(np.abs(np.diag(self.lu[0]))) >= 1e-14)

    def solve(self, r1, r2):
        z = sla.lu_solve(self.lu, np.concatenate([r1, r2]), check_finite=False)
        n = self.f["n"]
        return z[:n], z[n:]


def _step(x, dx, free):
    mask = (~free) & (dx < 0)
    return min(1.0, float(np.min(-x[mask] / dx[mask]))) if mask.any() else 1.0


def solve_qp_ipm(f, mehrotra_start=False, max_iterations=100, feasibility_tol=1e-8, optimality_tol=1e-8,
                 tau=0.999, common_step=True):
    n, m, A, b, c, Q, free = f["n"], f["m"], f["A"], f["b"], f["c"], f["Q"], f["free"]
    x = np.ones(n)
    s = np.ones(n)
    y = np.zeros(m)
    kkt = Kkt(f)
    if m > 0 and kkt.factor(np.ones(n), np.ones(n), False):
        r2 = b - A @ np.ones(n)
        _, dy = kkt.solve(np.zeros(n), r2)
        if mehrotra_start:
            xt = 1.0 + A.T @ dy
            u, yls = kkt.solve(-(c + Q @ xt), np.zeros(m))
            bnd = ~free
            min_x = np.min(xt[bnd]) if bnd.any() else np.inf
            min_s = np.min(-u[bnd]) if bnd.any() else np.inf
            sx, ss = max(-1.5 * min_x, 0.0), max(-1.5 * min_s, 0.0)
            xh, sh = xt[bnd] + sx, -u[bnd] + ss
            xs, sum_x, sum_s = float(xh @ sh), float(xh.sum()), float(sh.sum())
            bx = 0.5 * xs / sum_s if sum_s > 0 else 1.0
            bs = 0.5 * xs / sum_x if sum_x > 0 else 1.0
            x = np.where(free, xt, np.maximum(1e-8, xt + sx + bx))
            s = np.where(free, s, np.maximum(1e-8, -u + ss + bs))
            y = yls
        else:
            atdy = A.T @ dy
            qones = Q @ np.ones(n)
            x = np.where(free, 1.0 + atdy, np.maximum(1.0, np.abs(atdy)))
            s = np.where(free, s, np.maximum(1.0, np.abs(qones + c - atdy)))
            y = dy
    s = np.where(free, 0.0, s)
    count = max(1, int(np.sum(~free)))
    has_q = bool(np.any(np.abs(Q) > 1e-12))
    bnorm = max(1.0, float(np.max(np.abs(b))) if m else 0.0)
    cnorm = max(1.0, float(np.max(np.abs(c))))
    it = 0
    for it in range(max_iterations):
        rp = b - A @ x
        qx = Q @ x
        rd = qx + c - A.T @ y - s
        mu = float(x @ s) / count
        p_res = (float(np.max(np.abs(rp))) if m else 0.0) / bnorm
        d_res = float(np.max(np.abs(rd))) / max(1.0, cnorm + float(np.max(np.abs(qx))))
        gap = mu * count / (1.0 + abs(float(c @ x) + 0.5 * float(x @ qx) + f["offset"]))
        if p_res < feasibility_tol and d_res < feasibility_tol and gap < optimality_tol:
            return _result(f, 0, it + 1, x, y, s, gap, p_res, d_res)
        if not kkt.factor(x, s):
            return _result(f, 7, it, x, y, s, gap, p_res, d_res, "QP-IPM Newton factorization failed.")

        def newton(rxs):
            r1 = np.where(free, -rd, -rd + rxs / np.maximum(x, 1e-16))
            dx, dy = kkt.solve(r1, rp)
            ds = np.where(free, 0.0, (rxs - s * dx) / np.maximum(x, 1e-16))
            return dx, dy, ds

        dxa, dya, dsa = newton(-x * s)
        ap, ad = _step(x, dxa, free), _step(s, dsa, free)
        bnd = ~free
        mua = float(np.sum((x[bnd] + ap * dxa[bnd]) * (s[bnd] + ad * dsa[bnd]))) / count
        sigma = min(1.0, (mua / mu) ** 3) if mu > 0 else 0.0
        rxs = np.where(free, 0.0, -x * s - dxa * dsa + sigma * mu)
        dx, dy, ds = newton(rxs)
        ap = min(1.0, tau * _step(x, dx, free))
        ad = min(1.0, tau * _step(s, ds, free))
        if has_q and common_step:
            ap = ad = min(ap, ad)
        x = np.where(free, x + ap * dx, np.maximum(1e-14, x + ap * dx))
        s = np.where(free, s, np.maximum(1e-14, s + ad * ds))
        y = y + ad * dy
    return _result(f, 7, max_iterations, x, y, s, np.nan, np.nan, np.nan,
                   "QP-IPM iteration limit reached without convergence.")


def _result(f, status, iterations, x, y, s, gap, p_res, d_res, message=""):
    ns = f["ns"]
    xs = f["shift"] + x[:ns]
    obj = f["constant"] + float(f["c_model"] @ xs) + 0.5 * float(xs @ f["Q_model"] @ xs)
    return {"status": status, "iterations": iterations, "x": xs, "objective": obj, "duality_gap": gap,
            "primal_residual": p_res, "dual_residual": d_res, "y": y, "s": s[:ns], "message": message}


def qp_solve(A, senses, rhs, lb, ub, c, Q, maximize=False, constant=0.0, **kw):
    """The default start, then Mehrotra's when the first attempt fails."""
    f = build_qp_form(A, senses, rhs, lb, ub, c, Q, maximize, constant)
    first = solve_qp_ipm(f, False, **kw)
    if first["status"] == 0:
        return first
    second = solve_qp_ipm(f, True, **kw)
    if second["status"] == 0:
        second["iterations"] += first["iterations"]
        return second
    return first
