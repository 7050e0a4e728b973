// Test infrastructure, not product: a CPU restatement of the reference's
// structured interior-point method for the separable convex QP
//
//     min 0.5 ||x||^2 - target' x   s.t.  x_r - 2 x_{r+1} + x_{r+2} >= 0,
//                                          r = 0 .. n-3, x free,
//
// used only by tests/ and the CPU baseline of bench.py to check the GPU QP
// solver. It follows solve_liswet_structured_ipm
// (solver/qp/src/qp_interior_point.cpp:2378-2826): start x = target,
// s = max(1, A x), z = 1 / s (:2400-2406); residuals and stopping test
// (:2437-2497); the x-space normal matrix I + A' diag(z/s) A as three bands
// (:2499-2526); LiswetPentadiagonalCholesky (:1042-1149: Jacobi scaling by
// sqrt(max(1, diag)), pivot test v > 1e-18 max(1, |row|), one retry with
// regularization 1e-12 max diag capped at 1e-6); the direction with two
// refinement steps (:2590-2651); Mehrotra sigma = clamp((mu_aff/mu)^3, 0.01,
// 0.99) and a common step tau = 0.995 (:2666-2716). It runs to the
// structured tolerance 1e-10 (1e-12 for m <= 50) (mode 0); or, in the
// reference's default mode, stops at complementarity 1e-6 (:2473) and
// polishes in knot space (polish_liswet_knots :2269-2377,
// solve_liswet_knot_projection :1509-1580) (mode 1); or runs to the
// tolerance or the iteration cap and then polishes from the last iterate
// (mode 2, the method of the GPU solver). The result
// is certified as certify_liswet does (:1410-1480): lambda from the
// second-difference recurrence of x - target (:995-1040, evaluated in
// __float128 as the reference's high-precision option), then primal
// feasibility, dual sign, stationarity, consistency and complementarity
// within 1e-8 x scale. Arithmetic is long double, as in the reference.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

typedef long double ld;

void apply_a(const std::vector<ld>& x, std::vector<ld>& out) {  // rows r: x_r - 2 x_{r+1} + x_{r+2}
  const std::size_t m = x.size() - 2;
  out.assign(m, 0.0L);
  for (std::size_t r = 0; r < m; ++r) out[r] = x[r] - 2.0L * x[r + 1] + x[r + 2];
}

void apply_at(const std::vector<ld>& z, std::size_t n, std::vector<ld>& out) {
  out.assign(n, 0.0L);
  for (std::size_t r = 0; r < z.size(); ++r) {
    out[r] += z[r];
    out[r + 1] -= 2.0L * z[r];
    out[r + 2] += z[r];
  }
}

ld max_abs(const std::vector<ld>& v) {
  ld a = 0.0L;
  for (ld x : v) a = std::max(a, std::abs(x));
  return a;
}

ld positive_step(const std::vector<ld>& v, const std::vector<ld>& d) {
  ld a = 1.0L;
  for (std::size_t i = 0; i < v.size(); ++i)
    if (d[i] < 0.0L) a = std::min(a, -v[i] / d[i]);
  return std::isfinite(a) ? std::max(0.0L, a) : 0.0L;
}

struct Penta {
  std::vector<ld> scale, d, l1, l2;
  ld reg = 0.0L;
  bool factor(const std::vector<ld>& diag, const std::vector<ld>& s1, const std::vector<ld>& s2) {
    const std::size_t n = diag.size();
    const ld maxd = std::max(1.0L, *std::max_element(diag.begin(), diag.end()));
    scale.assign(n, 1.0L);
    for (std::size_t i = 0; i < n; ++i) scale[i] = std::sqrt(std::max(1.0L, diag[i]));
    auto attempt = [&](ld r) {
      d.assign(n, 0.0L);
      l1.assign(n, 0.0L);
      l2.assign(n, 0.0L);
      for (std::size_t i = 0; i < n; ++i) {
        const ld a = (diag[i] + r) / (scale[i] * scale[i]);
        const ld b1 = i > 0 ? s1[i] / (scale[i] * scale[i - 1]) : 0.0L;
        const ld b2 = i > 1 ? s2[i] / (scale[i] * scale[i - 2]) : 0.0L;
        const ld row = std::max(1.0L, std::abs(a) + std::abs(b1) + std::abs(b2));
        if (i > 1) {
          l2[i] = b2 / d[i - 2];
          if (!std::isfinite(l2[i])) return false;
        }
        if (i > 0) {
          ld w = b1;
          if (i > 1) w -= l2[i] * l1[i - 1];
          l1[i] = w / d[i - 1];
          if (!std::isfinite(l1[i])) return false;
        }
        ld v = a;
        if (i > 0) v -= l1[i] * l1[i];
        if (i > 1) v -= l2[i] * l2[i];
        if (!(v > 1e-18L * row) || !std::isfinite(v)) return false;
        d[i] = std::sqrt(v);
      }
      reg = r;
      return true;
    };
    if (attempt(0.0L)) return true;
    const ld r = 1e-12L * maxd;
    if (r >= 1e-6L) return false;
    return attempt(r);
  }
  bool solve(const std::vector<ld>& rhs, std::vector<ld>& x) const {
    const std::size_t n = d.size();
    std::vector<ld> y(rhs);
    for (std::size_t i = 0; i < n; ++i) y[i] /= scale[i];
    for (std::size_t i = 0; i < n; ++i) {
      if (i > 0) y[i] -= l1[i] * y[i - 1];
      if (i > 1) y[i] -= l2[i] * y[i - 2];
      y[i] /= d[i];
      if (!std::isfinite(y[i])) return false;
    }
    x.assign(n, 0.0L);
    for (std::size_t k = n; k-- > 0;) {
      ld v = y[k];
      if (k + 1 < n) v -= l1[k + 1] * x[k + 1];
      if (k + 2 < n) v -= l2[k + 2] * x[k + 2];
      x[k] = v / d[k];
      if (!std::isfinite(x[k])) return false;
    }
    for (std::size_t i = 0; i < n; ++i) x[i] /= scale[i];
    return true;
  }
};


// Piecewise-linear least-squares fit through the knots of the inactive rows
// (solve_liswet_knot_projection, :1509-1580): hat-basis normal equations,
// tridiagonal Cholesky, then interpolation.
bool knot_projection(const std::vector<char>& active, const std::vector<ld>& t, std::vector<ld>& x) {
  const int n = static_cast<int>(t.size()), m = n - 2;
  std::vector<int> knots{0};
  for (int r = 0; r < m; ++r)
    if (!active[r]) knots.push_back(r + 1);
  knots.push_back(n - 1);
  knots.erase(std::unique(knots.begin(), knots.end()), knots.end());
  if (knots.size() < 2) return false;
  const std::size_t k = knots.size();
  std::vector<ld> dg(k, 0.0L), off(k - 1, 0.0L), rhs(k, 0.0L);
  for (std::size_t g = 0; g + 1 < k; ++g) {
    const int left = knots[g], right = knots[g + 1], len = right - left;
    if (len <= 0) return false;
    const int last = (g + 2 == k) ? right : right - 1;
    for (int j = left; j <= last; ++j) {
      const ld u = static_cast<ld>(j - left) / static_cast<ld>(len);
      const ld hl = 1.0L - u, hr = u;
      dg[g] += hl * hl;
      rhs[g] += hl * t[j];
      if (hr != 0.0L) {
        dg[g + 1] += hr * hr;
        rhs[g + 1] += hr * t[j];
        off[g] += hl * hr;
      }
    }
  }
  std::vector<ld> ldg(k), lsub(k - 1);
  for (std::size_t i = 0; i < k; ++i) {
    ld p = dg[i];
    if (i > 0) {
      lsub[i - 1] = off[i - 1] / ldg[i - 1];
      p -= lsub[i - 1] * lsub[i - 1];
    }
    if (!(p > 0.0L) || !std::isfinite(p)) return false;
    ldg[i] = std::sqrt(p);
  }
  for (std::size_t i = 0; i < k; ++i) {
    if (i > 0) rhs[i] -= lsub[i - 1] * rhs[i - 1];
    rhs[i] /= ldg[i];
  }
  for (std::size_t i = k; i-- > 0;) {
    if (i + 1 < k) rhs[i] -= lsub[i] * rhs[i + 1];
    rhs[i] /= ldg[i];
  }
  x.assign(n, 0.0L);
  for (std::size_t g = 0; g + 1 < k; ++g) {
    const int left = knots[g], right = knots[g + 1], len = right - left;
    const int last = (g + 2 == k) ? right : right - 1;
    for (int j = left; j <= last; ++j) {
      const ld u = static_cast<ld>(j - left) / static_cast<ld>(len);
      x[j] = (1.0L - u) * rhs[g] + u * rhs[g + 1];
    }
  }
  return true;
}

struct Cert {
  ld prim = 0, stat = 0, comp = 0, minl = 0, cons = 0;
  bool ok = false;
};

// Multipliers from the second-difference recurrence of res = x - target
// (liswet_recover_dual_recurrence, :995-1040) in __float128, the reference's
// high-precision variant (:1001-1029): the recurrence has a double root at 1,
// so a rounding error made at row k grows like (i - k) by row i, and long
// double is not enough once the multipliers reach ~1e6 (n = 20002 with a
// target of order one).
void recover_lambda(const std::vector<ld>& res, std::vector<ld>& lam) {
  const std::size_t m = res.size() - 2;
  lam.assign(m, 0.0L);
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

// certify_liswet (:1410-1480) in long double; lam and slack filled.
Cert certify(const std::vector<ld>& t, const std::vector<ld>& x, std::vector<ld>& lam, std::vector<ld>& slack) {
  const std::size_t N = t.size(), m = N - 2;
  std::vector<ld> res(N), atr;
  lam.assign(m, 0.0L);
  for (std::size_t j = 0; j < N; ++j) res[j] = x[j] - t[j];
  recover_lambda(res, lam);
  apply_a(x, slack);
  apply_at(lam, N, atr);
  Cert c;
  c.minl = std::numeric_limits<ld>::infinity();
  for (std::size_t j = 0; j < N; ++j) c.stat = std::max(c.stat, std::abs(res[j] - atr[j]));
  for (std::size_t r = 0; r < m; ++r) {
    c.prim = std::max(c.prim, std::max(0.0L, -slack[r]));
    c.comp = std::max(c.comp, std::abs(lam[r] * slack[r]));
    c.minl = std::min(c.minl, lam[r]);
  }
  if (m > 1)
    c.cons = std::max(std::abs(res[N - 2] - (lam[m - 2] - 2.0L * lam[m - 1])), std::abs(res[N - 1] - lam[m - 1]));
  ld scale = 1.0L;
  for (std::size_t j = 0; j < N; ++j) scale = std::max({scale, std::abs(t[j]), std::abs(atr[j]), std::abs(x[j])});
  const ld tol = 1e-8L * scale;
  c.ok = std::isfinite(c.stat) && std::isfinite(c.comp) && c.prim <= tol && c.minl >= -tol && c.stat <= tol &&
         c.cons <= tol && c.comp <= tol;
  return c;
}

// polish_liswet_knots (:2269-2377): active rows start as s < z; each pass
// projects on the knots, certifies, and flips the rows the certificate
// blames (negative multipliers leave, violated or complementarity-breaking
// rows enter), at most min(1024, max(8, m / 10)) of them by evidence.
bool polish(const std::vector<ld>& t, const std::vector<ld>& s, const std::vector<ld>& z, std::vector<ld>& x) {
  const int m = static_cast<int>(s.size());
  std::vector<char> active(m);
  for (int r = 0; r < m; ++r) active[r] = s[r] < z[r];
  for (int pass = 0; pass < 50; ++pass) {
    std::vector<ld> cand, lam, cs;
    if (!knot_projection(active, t, cand)) return false;
    const Cert c = certify(t, cand, lam, cs);
    ld sc = 1.0L;
    for (ld v : t) sc = std::max(sc, std::abs(v));
    for (ld v : cand) sc = std::max(sc, std::abs(v));
    const ld tol = 1e-8L * sc;
    if (c.ok) {
      x.swap(cand);
      return true;
    }
    std::vector<char> next = active;
    std::vector<std::pair<ld, int>> ev;
    int changes = 0;
    for (int r = 0; r < m; ++r) {
      if (active[r] && lam[r] < -tol) {
        next[r] = 0;
        ev.emplace_back(std::abs(lam[r]), r);
      } else if (!active[r] && (cs[r] < -tol || (lam[r] > tol && lam[r] * cs[r] > tol))) {
        next[r] = 1;
        ev.emplace_back(std::max(-cs[r], lam[r] * cs[r]), r);
      }
      if (next[r] != active[r]) ++changes;
    }
    const int cap = std::min(1024, std::max(8, m / 10));
    if (changes > cap) {
      std::sort(ev.begin(), ev.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
      next = active;
      changes = 0;
      for (int q = 0; q < cap && q < static_cast<int>(ev.size()); ++q) {
        next[ev[q].second] = active[ev[q].second] ? 0 : 1;
        ++changes;
      }
    }
    if (changes == 0) return false;
    active.swap(next);
  }
  return false;
}

}  // namespace

extern "C" {

// Certificate of a point x (certify_liswet, :1410-1480). out[0..6]:
// objective 0.5|x|^2 - target'x, primal violation, stationarity,
// complementarity, min lambda, consistency, ok (1/0). lambda (n-2) may be NULL.
int orc_liswet_certify(int32_t n, const double* target, const double* x_in, double* lambda_out, double* out) {
  if (n < 3 || !target || !x_in || !out) return -1;
  const std::size_t N = static_cast<std::size_t>(n), m = N - 2;
  std::vector<ld> t(target, target + N), x(x_in, x_in + N), lam, slack;
  const Cert c = certify(t, x, lam, slack);
  ld obj = 0.0L;
  for (std::size_t j = 0; j < N; ++j) obj += 0.5L * x[j] * x[j] - t[j] * x[j];
  const double v[7] = {static_cast<double>(obj),    static_cast<double>(c.prim), static_cast<double>(c.stat),
                       static_cast<double>(c.comp), static_cast<double>(c.minl), static_cast<double>(c.cons),
                       c.ok ? 1.0 : 0.0};
  for (int k = 0; k < 7; ++k) out[k] = v[k];
  if (lambda_out)
    for (std::size_t r = 0; r < m; ++r) lambda_out[r] = static_cast<double>(lam[r]);
  return 0;
}

// Knot-space polish from an interior point (s, z): passes used (> 0) or -1;
// x_out receives the certified point.
int orc_liswet_polish(int32_t n, const double* target, const double* s_in, const double* z_in, double* x_out) {
  if (n < 3 || !target || !s_in || !z_in || !x_out) return -2;
  const std::size_t N = static_cast<std::size_t>(n), m = N - 2;
  std::vector<ld> t(target, target + N), s(s_in, s_in + m), z(z_in, z_in + m), x;
  if (!polish(t, s, z, x)) return -1;
  for (std::size_t j = 0; j < N; ++j) x_out[j] = static_cast<double>(x[j]);
  return 1;
}

// The structured IPM in mode 0, 1 or 2 (above). Returns iterations (>= 0),
// -1 on bad input, -2 when the method failed (factorization, non-finite
// step), -3 when the polish failed; x_out gets the last iterate either way.
int orc_liswet_ipm(int32_t n, const double* target, int32_t max_iterations, int32_t mode,
                   double* x_out) {
  if (n < 3 || !target || !x_out) return -1;
  const std::size_t N = static_cast<std::size_t>(n), m = N - 2;
  const ld tol = m <= 50 ? 1e-12L : 1e-10L;
  const ld tau = 0.995L;
  std::vector<ld> t(target, target + N), x(t), s, z(m);
  apply_a(x, s);
  for (ld& v : s) v = std::max(1.0L, v);
  for (std::size_t r = 0; r < m; ++r) z[r] = 1.0L / s[r];
  int status = -2;
  int it = 0;
  for (; it < max_iterations; ++it) {
    std::vector<ld> ax, atz, rp(m), rd(N);
    apply_a(x, ax);
    apply_at(z, N, atz);
    for (std::size_t r = 0; r < m; ++r) rp[r] = ax[r] - s[r];
    for (std::size_t j = 0; j < N; ++j) rd[j] = x[j] - t[j] - atz[j];
    ld mu = 0.0L, maxc = 0.0L;
    for (std::size_t r = 0; r < m; ++r) {
      mu += s[r] * z[r];
      maxc = std::max(maxc, std::abs(s[r] * z[r]));
    }
    mu /= static_cast<ld>(m);
    const ld pres = max_abs(rp) / std::max(1.0L, max_abs(ax));
    const ld dres = max_abs(rd) / std::max(1.0L, max_abs(t) + max_abs(atz));
    const ld cres = maxc / std::max(1.0L, max_abs(t));
    if (!std::isfinite(mu) || !std::isfinite(pres) || !std::isfinite(dres) || !std::isfinite(cres)) break;
    if (mode == 1 && cres <= 1e-6L) {  // the default mode: knot-space polish from here (:2473-2477)
      status = polish(t, s, z, x) ? it : -3;
      break;
    }
    if (pres <= tol && dres <= tol && cres <= tol) {
      status = it;
      break;
    }
    std::vector<ld> diag(N, 1.0L), s1(N, 0.0L), s2(N, 0.0L), d(m);
    for (std::size_t r = 0; r < m; ++r) {
      d[r] = z[r] / s[r];
      const ld a[3] = {1.0L, -2.0L, 1.0L};
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
          const std::size_t i = r + p, j = r + q;
          const ld v = d[r] * a[p] * a[q];
          if (i == j) diag[i] += v;
          else if (i == j + 1) s1[i] += v;
          else if (i == j + 2) s2[i] += v;
        }
    }
    Penta f;
    if (!f.factor(diag, s1, s2)) break;
    auto direction = [&](const std::vector<ld>& rc, std::vector<ld>& dx, std::vector<ld>& ds, std::vector<ld>& dz) {
      std::vector<ld> w(m), atw, rhs(N);
      for (std::size_t r = 0; r < m; ++r) w[r] = (rc[r] + z[r] * rp[r]) / s[r];
      apply_at(w, N, atw);
      for (std::size_t j = 0; j < N; ++j) rhs[j] = -rd[j] - atw[j];
      if (!f.solve(rhs, dx)) return false;
      for (int k = 0; k < 2; ++k) {
        std::vector<ld> adx, dadx, res(N), corr;
        apply_a(dx, adx);
        for (std::size_t r = 0; r < m; ++r) adx[r] *= d[r];
        apply_at(adx, N, dadx);
        ld rinf = 0.0L;
        for (std::size_t j = 0; j < N; ++j) {
          res[j] = rhs[j] - dx[j] - dadx[j] - f.reg * dx[j];
          rinf = std::max(rinf, std::abs(res[j]));
        }
        if (rinf <= 1e-24L) break;
        if (!f.solve(res, corr)) return false;
        for (std::size_t j = 0; j < N; ++j) dx[j] += corr[j];
      }
      std::vector<ld> adx;
      apply_a(dx, adx);
      ds.assign(m, 0.0L);
      dz.assign(m, 0.0L);
      for (std::size_t r = 0; r < m; ++r) {
        ds[r] = adx[r] + rp[r];
        dz[r] = (-rc[r] - z[r] * ds[r]) / s[r];
      }
      return true;
    };
    std::vector<ld> rca(m), dxa, dsa, dza;
    for (std::size_t r = 0; r < m; ++r) rca[r] = s[r] * z[r];
    if (!direction(rca, dxa, dsa, dza)) break;
    const ld as = std::min(1.0L, positive_step(s, dsa)), az = std::min(1.0L, positive_step(z, dza));
    ld mua = 0.0L;
    for (std::size_t r = 0; r < m; ++r) mua += (s[r] + as * dsa[r]) * (z[r] + az * dza[r]);
    mua /= static_cast<ld>(m);
    const ld ratio = mu > 0.0L ? std::max(0.0L, mua / mu) : 0.0L;
    const ld sigma = std::max(0.01L, std::min(0.99L, ratio * ratio * ratio));
    std::vector<ld> rc(m), dx, ds, dz;
    for (std::size_t r = 0; r < m; ++r) rc[r] = s[r] * z[r] - sigma * mu + dsa[r] * dza[r];
    if (!direction(rc, dx, ds, dz)) break;
    const ld alpha = std::min(std::min(1.0L, tau * positive_step(s, ds)), std::min(1.0L, tau * positive_step(z, dz)));
    if (!(alpha > 0.0L) || !std::isfinite(alpha)) break;
    for (std::size_t j = 0; j < N; ++j) x[j] += alpha * dx[j];
    for (std::size_t r = 0; r < m; ++r) {
      s[r] += alpha * ds[r];
      z[r] += alpha * dz[r];
    }
  }
  if (mode == 2) {  // polish from the last iterate
    std::vector<ld> xp = x;
    if (polish(t, s, z, xp)) {
      x.swap(xp);
      if (status < 0) status = it;
    } else {
      status = -3;
    }
  }
  for (std::size_t j = 0; j < N; ++j) x_out[j] = static_cast<double>(x[j]);
  return status;
}

}  // extern "C"
