// Host side of the structured QP: knot-space polish and certificate, in long
// double (the reference's arithmetic for this path).
//
//   min 0.5 |x|^2 - t'x   s.t.  (A x)_r = x_r - 2 x_{r+1} + x_{r+2} >= 0
//
// With a set W of active rows fixed, the minimizer is the least-squares
// piecewise-linear fit of t whose knots are 0, n-1 and r + 1 for every row r
// outside W (solve_liswet_knot_projection, solver/qp/src/
// qp_interior_point.cpp:1509-1580). The polish is a primal active-set method
// on W, warm-started from the interior point's guess W = {r : s_r < z_r} (the
// guess of polish_liswet_knots, :2269-2377):
//   1. grow W by the rows the fit violates until the fit is feasible;
//   2. release the row of W with the most negative multiplier, move from x
//      toward the fit of the smaller W and stop at the first row whose
//      second difference would turn negative (it joins W); repeat until every
//      multiplier of W is >= 0.
// Each step lowers the objective, so the method ends, and a few wrong rows in
// the guess cost a few steps (the reference's passes flip every blamed row at
// once and can oscillate when the guess is off). The result is then checked
// by certify_liswet's certificate (:1410-1480).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

namespace sovqp {

typedef long double ld;

inline void apply_a(const std::vector<ld>& x, std::vector<ld>& out) {
  const std::size_t m = x.size() - 2;
  out.assign(m, 0.0L);
  for (std::size_t r = 0; r < m; ++r) out[r] = x[r] - 2.0L * x[r + 1] + x[r + 2];
}

inline void apply_at(const std::vector<ld>& z, std::size_t n, std::vector<ld>& out) {
  out.assign(n, 0.0L);
  for (std::size_t r = 0; r < z.size(); ++r) {
    out[r] += z[r];
    out[r + 1] -= 2.0L * z[r];
    out[r + 2] += z[r];
  }
}

inline ld scale_of(const std::vector<ld>& a, ld s = 1.0L) {
  for (ld v : a) s = std::max(s, std::abs(v));
  return s;
}

// Multipliers of x: the solution of A' lambda = x - t on its first n - 2
// rows, lambda_i = sum_{k <= i} (i - k + 1) res_k, evaluated as
// (i + 1) S0_i - S1_i with compensated (Neumaier) prefix sums S0 = sum res_k,
// S1 = sum k res_k. Same value as the recurrence lambda_i = res_i +
// 2 lambda_{i-1} - lambda_{i-2} (:995-1040), without its error growth.
inline void multipliers(const std::vector<ld>& t, const std::vector<ld>& x, std::vector<ld>& lam) {
  const std::size_t m = t.size() - 2;
  lam.assign(m, 0.0L);
  ld s0 = 0, c0 = 0, s1 = 0, c1 = 0;
  auto add = [](ld& s, ld& c, ld v) {
    const ld u = s + v;
    c += std::abs(s) >= std::abs(v) ? (s - u) + v : (v - u) + s;
    s = u;
  };
  for (std::size_t i = 0; i < m; ++i) {
    const ld r = x[i] - t[i];
    add(s0, c0, r);
    add(s1, c1, static_cast<ld>(i) * r);
    const ld k = static_cast<ld>(i + 1);
    lam[i] = (k * s0 - s1) + (k * c0 - c1);
  }
}

struct Certificate {
  ld primal = 0, stationarity = 0, complementarity = 0, min_lambda = 0, consistency = 0;
  bool ok = false;
};

// certify_liswet (:1410-1480): multipliers from the second-difference
// recurrence of x - t, then primal feasibility, dual sign, stationarity,
// consistency of the last two rows and complementarity, each within
// 1e-8 x max(1, |t|, |A' lambda|, |x|). The recurrence runs in __float128
// (the reference's high-precision variant, :1001-1029): it has a double root
// at 1, so a rounding error at row k grows like (i - k) by row i.
inline Certificate certify(const std::vector<ld>& t, const std::vector<ld>& x, std::vector<ld>& lam,
                           std::vector<ld>& slack) {
  const std::size_t N = t.size(), m = N - 2;
  std::vector<ld> res(N), atr;
  lam.assign(m, 0.0L);
  for (std::size_t j = 0; j < N; ++j) res[j] = x[j] - t[j];
  {
    __float128 a = static_cast<__float128>(res[0]), b = 0;
    lam[0] = res[0];
    if (m > 1) {
      b = static_cast<__float128>(res[1]) + 2 * a;
      lam[1] = static_cast<ld>(b);
    }
    for (std::size_t i = 2; i < m; ++i) {
      const __float128 c = static_cast<__float128>(res[i]) + 2 * b - a;
      lam[i] = static_cast<ld>(c);
      a = b;
      b = c;
    }
  }
  apply_a(x, slack);
  apply_at(lam, N, atr);
  Certificate c;
  c.min_lambda = std::numeric_limits<ld>::infinity();
  for (std::size_t j = 0; j < N; ++j) c.stationarity = std::max(c.stationarity, std::abs(res[j] - atr[j]));
  for (std::size_t r = 0; r < m; ++r) {
    c.primal = std::max(c.primal, std::max(0.0L, -slack[r]));
    c.complementarity = std::max(c.complementarity, std::abs(lam[r] * slack[r]));
    c.min_lambda = std::min(c.min_lambda, lam[r]);
  }
  if (m > 1)
    c.consistency =
        std::max(std::abs(res[N - 2] - (lam[m - 2] - 2.0L * lam[m - 1])), std::abs(res[N - 1] - lam[m - 1]));
  const ld scale = scale_of(x, scale_of(atr, scale_of(t)));
  const ld tol = 1e-8L * scale;
  c.ok = std::isfinite(c.stationarity) && std::isfinite(c.complementarity) && c.primal <= tol &&
         c.min_lambda >= -tol && c.stationarity <= tol && c.consistency <= tol && c.complementarity <= tol;
  return c;
}

// Least-squares piecewise-linear fit of t with knots 0, n - 1 and r + 1 for
// every row r outside W: hat-basis normal equations (tridiagonal), Cholesky,
// interpolation. False when the system is not positive definite.
inline bool knot_fit(const std::vector<char>& active, const std::vector<ld>& t, std::vector<ld>& x) {
  const int n = static_cast<int>(t.size()), m = n - 2;
  std::vector<int> kn{0};
  for (int r = 0; r < m; ++r)
    if (!active[r]) kn.push_back(r + 1);
  kn.push_back(n - 1);
  kn.erase(std::unique(kn.begin(), kn.end()), kn.end());
  const std::size_t k = kn.size();
  std::vector<ld> dg(k, 0.0L), off(k - 1, 0.0L), b(k, 0.0L);
  for (std::size_t g = 0; g + 1 < k; ++g) {
    const int a = kn[g], e = kn[g + 1], len = e - a;
    const int last = (g + 2 == k) ? e : e - 1;
    for (int j = a; j <= last; ++j) {
      const ld u = static_cast<ld>(j - a) / static_cast<ld>(len);
      dg[g] += (1.0L - u) * (1.0L - u);
      b[g] += (1.0L - u) * t[j];
      if (u != 0.0L) {
        dg[g + 1] += u * u;
        b[g + 1] += u * t[j];
        off[g] += (1.0L - u) * u;
      }
    }
  }
  std::vector<ld> L(k), sub(k - 1);
  for (std::size_t i = 0; i < k; ++i) {
    ld p = dg[i];
    if (i > 0) {
      sub[i - 1] = off[i - 1] / L[i - 1];
      p -= sub[i - 1] * sub[i - 1];
    }
    if (!(p > 0.0L) || !std::isfinite(p)) return false;
    L[i] = std::sqrt(p);
  }
  for (std::size_t i = 0; i < k; ++i) {
    if (i > 0) b[i] -= sub[i - 1] * b[i - 1];
    b[i] /= L[i];
  }
  for (std::size_t i = k; i-- > 0;) {
    if (i + 1 < k) b[i] -= sub[i] * b[i + 1];
    b[i] /= L[i];
  }
  x.assign(n, 0.0L);
  for (std::size_t g = 0; g + 1 < k; ++g) {
    const int a = kn[g], e = kn[g + 1], len = e - a;
    const int last = (g + 2 == k) ? e : e - 1;
    for (int j = a; j <= last; ++j) {
      const ld u = static_cast<ld>(j - a) / static_cast<ld>(len);
      x[j] = (1.0L - u) * b[g] + u * b[g + 1];
    }
  }
  return true;
}

struct PolishStats {
  int fits = 0;      // knot fits computed
  int released = 0;  // rows released from W
  int added = 0;     // rows added to W
};

// The polish described at the top, from the active-set guess `active`.
// Returns true with x certified (cert.ok); false when the method breaks
// down or exceeds its step budget.
inline bool active_set_polish(const std::vector<ld>& t, std::vector<char> active, std::vector<ld>& x,
                              Certificate& cert, PolishStats& st) {
  const int n = static_cast<int>(t.size()), m = n - 2;
  const ld tscale = scale_of(t);
  std::vector<ld> xs, slack, sk, lam;
  // 1. feasibility: rows the fit violates join W (W only grows, and the
  //    all-active fit, a straight line, is feasible)
  for (int guard = 0;; ++guard) {
    if (guard > m + 1 || !knot_fit(active, t, x)) return false;
    ++st.fits;
    apply_a(x, slack);
    const ld tolf = 1e-14L * scale_of(x, tscale);
    int grown = 0;
    for (int r = 0; r < m; ++r)
      if (!active[r] && slack[r] < -tolf) {
        active[r] = 1;
        ++grown;
      }
    st.added += grown;
    if (grown == 0) break;
  }
  // 2. release / block steps
  const int budget = 4 * m + 64;
  for (int step = 0; step < budget; ++step) {
    multipliers(t, x, lam);
    const ld tol = 1e-10L * scale_of(x, tscale);
    int rel = -1;
    ld worst = -tol;
    for (int r = 0; r < m; ++r)
      if (active[r] && lam[r] < worst) {
        worst = lam[r];
        rel = r;
      }
    if (rel < 0) break;
    active[rel] = 0;
    ++st.released;
    for (;;) {  // move toward the fit of the current W, adding blocking rows
      if (!knot_fit(active, t, xs)) return false;
      ++st.fits;
      apply_a(xs, sk);
      apply_a(x, slack);
      const ld tolf = 1e-14L * scale_of(xs, tscale);
      ld alpha = 1.0L;
      int block = -1;
      for (int r = 0; r < m; ++r) {
        if (active[r] || !(sk[r] < -tolf)) continue;
        const ld s0 = std::max(0.0L, slack[r]);
        const ld a = s0 / (s0 - sk[r]);
        if (a < alpha) {
          alpha = a;
          block = r;
        }
      }
      if (block < 0) {
        x.swap(xs);
        break;
      }
      for (int j = 0; j < n; ++j) x[j] += alpha * (xs[j] - x[j]);
      active[block] = 1;
      ++st.added;
      if (++step >= budget) return false;
    }
  }
  std::vector<ld> cl, cs;
  cert = certify(t, x, cl, cs);
  return cert.ok;
}

}  // namespace sovqp
