// Mehrotra predictor-corrector interior-point method with every iterate
// resident in HBM.
//
// Follows solve_ipm (solver/lp/src/interior_point.cpp:454-740) step for step:
// the starting point (:488-506), residuals and the relative gap (:572-602),
// best-iterate tracking (:603-615), termination (:633), the affine predictor
// (:643-673), the corrector (:675-699) and the update with its 1e-14 floors
// (:709-717); solve_newton (:418-452) is split into a row kernel
// (rhs = rp + A tmp), the normal-equation solve and a column kernel
// (dx, ds and the ratio test). Scalars (residual norms, step lengths, sigma)
// are reduced on the device; the host reads them once per iteration to take
// the termination and best-iterate decisions the reference takes.

#include "normal_eq.hpp"
#include "spmv_device.hpp"

#include <chrono>
#include <cmath>
#include <limits>
#include <vector>
#include <cstdio>
#include <cstdlib>

namespace sov {
namespace {

constexpr int kThreads = 256;
// CG iterations (predictor + corrector) of one FP32-preconditioned IPM
// iteration past which the remaining factorizations are FP64: the next
// iteration's count, about twice this one, would pass the break-even below.
constexpr int kPcgLockIterations = 8;

struct IpmScalars {
  double pres, dres, gap, mu, xs, cx;
  double ap_aff, ad_aff, mu_aff, sigma, ap, ad;
};

// ---- residuals --------------------------------------------------------------

template <int G>
__global__ __launch_bounds__(kThreads) void resid_rows_kernel(int m, const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ idx,
                                                              const double* __restrict__ val,
                                                              const double* __restrict__ x,
                                                              const double* __restrict__ b, double* __restrict__ rp,
                                                              double* __restrict__ partial) {
  __shared__ double sh[kThreads / 64];
  const int sub = threadIdx.x & (G - 1);
  const int i = (blockIdx.x * kThreads + threadIdx.x) / G;
  const double ax = csr_row_dot<G>(ptr, idx, val, x, i, m, sub);
  double mx = 0.0;
  if (i < m && sub == 0) {
    const double r = b[i] - ax;
    rp[i] = r;
    mx = fabs(r);
  }
  mx = block_reduce<1>(mx, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = mx;
}

template <int G>
__global__ __launch_bounds__(kThreads) void resid_cols_kernel(int n, const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ idx,
                                                              const double* __restrict__ val,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ c,
                                                              const double* __restrict__ x,
                                                              const double* __restrict__ s, double* __restrict__ rd,
                                                              double* __restrict__ d, double* __restrict__ rxs,
                                                              double* __restrict__ tmp, double* __restrict__ partial) {
  __shared__ double sh[kThreads / 64];
  const int sub = threadIdx.x & (G - 1);
  const int j = (blockIdx.x * kThreads + threadIdx.x) / G;
  const double aty = csr_row_dot<G>(ptr, idx, val, y, j, n, sub);
  double mx = 0.0, xs = 0.0, cx = 0.0;
  if (j < n && sub == 0) {
    const double xj = x[j], sj = s[j];
    const double r = c[j] - (aty + sj);
    rd[j] = r;
    const double sf = fmax(sj, 1e-16);
    const double dj = fmax(xj, 1e-16) / sf;
    d[j] = dj;
    const double q = -xj * sj;
    rxs[j] = q;
    tmp[j] = dj * r - q / sf;
    mx = fabs(r);
    xs = xj * sj;
    cx = c[j] * xj;
  }
  mx = block_reduce<1>(mx, sh);
  xs = block_reduce<0>(xs, sh);
  cx = block_reduce<0>(cx, sh);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x + 0] = mx;
    partial[3 * blockIdx.x + 1] = xs;
    partial[3 * blockIdx.x + 2] = cx;
  }
}

__global__ __launch_bounds__(kThreads) void resid_finish_kernel(int nrow_parts, const double* __restrict__ prow,
                                                                int ncol_parts, const double* __restrict__ pcol, int n,
                                                                double bnorm, double cnorm, IpmScalars* sc) {
  __shared__ double sh[kThreads / 64];
  double rmax = 0.0, dmax = 0.0, xs = 0.0, cx = 0.0;
  for (int k = threadIdx.x; k < nrow_parts; k += kThreads) rmax = fmax(rmax, prow[k]);
  for (int k = threadIdx.x; k < ncol_parts; k += kThreads) {
    dmax = fmax(dmax, pcol[3 * k]);
    xs += pcol[3 * k + 1];
    cx += pcol[3 * k + 2];
  }
  rmax = block_reduce<1>(rmax, sh);
  dmax = block_reduce<1>(dmax, sh);
  xs = block_reduce<0>(xs, sh);
  cx = block_reduce<0>(cx, sh);
  if (threadIdx.x == 0) {
    sc->pres = rmax / bnorm;
    sc->dres = dmax / cnorm;
    sc->xs = xs;
    sc->cx = cx;
    sc->mu = xs / static_cast<double>(n);
    sc->gap = fabs(xs) / (1.0 + fabs(cx));
  }
}

// ---- Newton system ------------------------------------------------------------

// out_i = base_i + (A v)_i
template <int G>
__global__ __launch_bounds__(kThreads) void rows_axpy_kernel(int m, const int32_t* __restrict__ ptr,
                                                             const int32_t* __restrict__ idx,
                                                             const double* __restrict__ val,
                                                             const double* __restrict__ v,
                                                             const double* __restrict__ base, double* __restrict__ out) {
  const int sub = threadIdx.x & (G - 1);
  const int i = (blockIdx.x * kThreads + threadIdx.x) / G;
  const double av = csr_row_dot<G>(ptr, idx, val, v, i, m, sub);
  if (i < m && sub == 0) out[i] = base[i] + av;
}

// dx = d (A^T dy - rd) + rxs / max(s, 1e-16); ds = rd - A^T dy; partial
// ratio tests min(-x/dx | dx < 0) and min(-s/ds | ds < 0), starting at 1.
template <int G>
__global__ __launch_bounds__(kThreads) void direction_cols_kernel(
    int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx, const double* __restrict__ val,
    const double* __restrict__ dy, const double* __restrict__ d, const double* __restrict__ rd,
    const double* __restrict__ rxs, const double* __restrict__ x, const double* __restrict__ s,
    double* __restrict__ dx, double* __restrict__ ds, double* __restrict__ partial) {
  __shared__ double sh[kThreads / 64];
  const int sub = threadIdx.x & (G - 1);
  const int j = (blockIdx.x * kThreads + threadIdx.x) / G;
  const double atdy = csr_row_dot<G>(ptr, idx, val, dy, j, n, sub);
  double ap = 1.0, ad = 1.0;
  if (j < n && sub == 0) {
    const double sj = s[j];
    const double dxj = d[j] * (atdy - rd[j]) + rxs[j] / fmax(sj, 1e-16);
    const double dsj = rd[j] - atdy;
    dx[j] = dxj;
    ds[j] = dsj;
    if (dxj < 0.0) ap = fmin(ap, -x[j] / dxj);
    if (dsj < 0.0) ad = fmin(ad, -sj / dsj);
  }
  ap = block_reduce<2>(ap, sh);
  ad = block_reduce<2>(ad, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = ap;
    partial[2 * blockIdx.x + 1] = ad;
  }
}

// corrector: 0 -> affine steps; 1 -> tau-scaled, capped at 1.
__global__ __launch_bounds__(kThreads) void alpha_finish_kernel(int nparts, const double* __restrict__ partial,
                                                                double tau, int corrector, IpmScalars* sc) {
  __shared__ double sh[kThreads / 64];
  double ap = 1.0, ad = 1.0;
  for (int k = threadIdx.x; k < nparts; k += kThreads) {
    ap = fmin(ap, partial[2 * k]);
    ad = fmin(ad, partial[2 * k + 1]);
  }
  ap = block_reduce<2>(ap, sh);
  ad = block_reduce<2>(ad, sh);
  if (threadIdx.x == 0) {
    if (corrector) {
      sc->ap = fmin(1.0, tau * ap);
      sc->ad = fmin(1.0, tau * ad);
    } else {
      sc->ap_aff = ap;
      sc->ad_aff = ad;
    }
  }
}

__global__ __launch_bounds__(kThreads) void mu_aff_kernel(int n, const double* __restrict__ x,
                                                          const double* __restrict__ s,
                                                          const double* __restrict__ dxa,
                                                          const double* __restrict__ dsa, const IpmScalars* sc,
                                                          double* __restrict__ partial) {
  __shared__ double sh[kThreads / 64];
  const double ap = sc->ap_aff, ad = sc->ad_aff;
  double acc = 0.0;
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads)
    acc += (x[j] + ap * dxa[j]) * (s[j] + ad * dsa[j]);
  acc = block_reduce<0>(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kThreads) void sigma_finish_kernel(int nparts, const double* __restrict__ partial, int n,
                                                                IpmScalars* sc) {
  __shared__ double sh[kThreads / 64];
  double acc = 0.0;
  for (int k = threadIdx.x; k < nparts; k += kThreads) acc += partial[k];
  acc = block_reduce<0>(acc, sh);
  if (threadIdx.x == 0) {
    const double mu_aff = acc / static_cast<double>(n);
    sc->mu_aff = mu_aff;
    const double mu = sc->mu;
    sc->sigma = mu > 0.0 ? fmin(1.0, pow(mu_aff / mu, 3.0)) : 0.0;
  }
}

// rxs = -x s - dx_aff ds_aff + sigma mu; tmp = d rd - rxs / max(s, 1e-16).
__global__ __launch_bounds__(kThreads) void corrector_cols_kernel(int n, const double* __restrict__ x,
                                                                  const double* __restrict__ s,
                                                                  const double* __restrict__ dxa,
                                                                  const double* __restrict__ dsa,
                                                                  const double* __restrict__ d,
                                                                  const double* __restrict__ rd, const IpmScalars* sc,
                                                                  double* __restrict__ rxs, double* __restrict__ tmp) {
  const double sm = sc->sigma * sc->mu;
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const double q = -x[j] * s[j] - dxa[j] * dsa[j] + sm;
    rxs[j] = q;
    tmp[j] = d[j] * rd[j] - q / fmax(s[j], 1e-16);
  }
}

__global__ __launch_bounds__(kThreads) void update_kernel(int n, int m, const IpmScalars* sc, double* __restrict__ x,
                                                          double* __restrict__ s, double* __restrict__ y,
                                                          const double* __restrict__ dx,
                                                          const double* __restrict__ ds,
                                                          const double* __restrict__ dy) {
  const double ap = sc->ap, ad = sc->ad;
  const int total = n + m;
  for (int k = blockIdx.x * kThreads + threadIdx.x; k < total; k += gridDim.x * kThreads) {
    if (k < n) {
      x[k] = fmax(x[k] + ap * dx[k], 1e-14);
      s[k] = fmax(s[k] + ad * ds[k], 1e-14);
    } else {
      y[k - n] += ad * dy[k - n];
    }
  }
}

// ---- starting point (:488-506) -------------------------------------------------

__global__ void fill_kernel(int n, double v, double* __restrict__ out) {
  for (int k = blockIdx.x * kThreads + threadIdx.x; k < n; k += gridDim.x * kThreads) out[k] = v;
}

// x = max(1, |A^T dy|), s = max(1, |c - A^T dy|)
template <int G>
__global__ __launch_bounds__(kThreads) void start_cols_kernel(int n, const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ idx,
                                                              const double* __restrict__ val,
                                                              const double* __restrict__ dy,
                                                              const double* __restrict__ c, double* __restrict__ x,
                                                              double* __restrict__ s) {
  const int sub = threadIdx.x & (G - 1);
  const int j = (blockIdx.x * kThreads + threadIdx.x) / G;
  const double v = csr_row_dot<G>(ptr, idx, val, dy, j, n, sub);
  if (j < n && sub == 0) {
    x[j] = fmax(1.0, fabs(v));
    s[j] = fmax(1.0, fabs(c[j] - v));
  }
}

// b - A ones
template <int G>
__global__ __launch_bounds__(kThreads) void start_rows_kernel(int m, const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ idx,
                                                              const double* __restrict__ val,
                                                              const double* __restrict__ ones,
                                                              const double* __restrict__ b, double* __restrict__ out) {
  const int sub = threadIdx.x & (G - 1);
  const int i = (blockIdx.x * kThreads + threadIdx.x) / G;
  const double v = csr_row_dot<G>(ptr, idx, val, ones, i, m, sub);
  if (i < m && sub == 0) out[i] = b[i] - v;
}

using Clock = std::chrono::steady_clock;

struct DevBuf {
  double* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(std::size_t n) {
    return hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(double)) == hipSuccess ? SOV_OK : SOV_ERR_NOMEM;
  }
};

struct Events {
  hipEvent_t e[6] = {};
  Events() {
    for (auto& v : e) (void)hipEventCreate(&v);
  }
  ~Events() {
    for (auto& v : e)
      if (v) (void)hipEventDestroy(v);
  }
  static double ms(hipEvent_t a, hipEvent_t b) {
    float t = 0.f;
    return hipEventElapsedTime(&t, a, b) == hipSuccess ? t : 0.0;
  }
};

}  // namespace

}  // namespace sov

// Resident solver: the matrix and its normal-equation plan.
struct sov_ipm {
  sov_mat* A = nullptr;
  sov::NormalSolver ne;
};

namespace sov {

// Runs the Mehrotra loop on an LP in IPM form. Outputs on the host.
int ipm_run(sov_ipm* h, const double* b_h, const double* c_h, const sov_ipm_options& opt, sov_ipm_result* res,
            double* x_out, double* y_out, double* s_out) {
  const auto t_start = Clock::now();
  *res = sov_ipm_result{};
  res->status = 7;
  sov_mat* A = h->A;
  NormalSolver& ne = h->ne;
  const int m = A->nrows, n = A->ncols;
  hipStream_t st = stream();
  if (n <= 0 || m <= 0) {
    set_error("IPM: empty model");
    return SOV_ERR_ARG;
  }
  int rc = SOV_OK;

  DevBuf b, c, x, s, y, rp, rd, d, rxs, tmp, dy, dx, ds, dxa, dsa, bx, by, bs, rhs_save, prow, pcol;
  for (auto* v : {&c, &x, &s, &rd, &d, &rxs, &tmp, &dx, &ds, &dxa, &dsa, &bx, &bs})
    if ((rc = v->alloc(n)) != SOV_OK) return rc;
  for (auto* v : {&b, &y, &rp, &dy, &by, &rhs_save})
    if ((rc = v->alloc(m)) != SOV_OK) return rc;
  const int gr = A->group_rows, gc = A->group_cols;
  const unsigned row_blocks = blocks_for(static_cast<long long>(m) * gr, kThreads);
  const unsigned col_blocks = blocks_for(static_cast<long long>(n) * gc, kThreads);
  const unsigned flat_blocks = std::max(1u, std::min(1024u, blocks_for(std::max(n, m), kThreads)));
  if ((rc = prow.alloc(row_blocks)) != SOV_OK) return rc;
  if ((rc = pcol.alloc(3 * static_cast<std::size_t>(std::max(col_blocks, flat_blocks)))) != SOV_OK) return rc;
  IpmScalars* sc = nullptr;
  IpmScalars* sc_h = nullptr;
  SOV_HIP_TRY(hipMalloc(&sc, sizeof(IpmScalars)));
  SOV_HIP_TRY(hipHostMalloc(&sc_h, sizeof(IpmScalars)));
  struct Cleanup {
    IpmScalars* a;
    IpmScalars* b;
    ~Cleanup() {
      if (a) (void)hipFree(a);
      if (b) (void)hipHostFree(b);
    }
  } cleanup{sc, sc_h};

  SOV_HIP_TRY(hipMemcpyAsync(b.p, b_h, m * sizeof(double), hipMemcpyHostToDevice, st));
  SOV_HIP_TRY(hipMemcpyAsync(c.p, c_h, n * sizeof(double), hipMemcpyHostToDevice, st));
  double bnorm = 1.0, cnorm = 1.0;
  for (int i = 0; i < m; ++i) bnorm = std::max(bnorm, std::abs(b_h[i]));
  for (int j = 0; j < n; ++j) cnorm = std::max(cnorm, std::abs(c_h[j]));
  const double tau = opt.fraction_to_boundary;
  // The sparse factorization is FP64 with direct solves: no mixed scheme.
  const bool mixed = opt.precision == SOV_PREC_FP32 && !ne.sparse_mode();
  res->sparse_factor = ne.sparse_mode() ? 1 : 0;
  bool fp64_locked = !mixed;
  const double solve_tol = opt.refine_tol > 0 ? opt.refine_tol : 1e-12;
  const int solve_max = opt.refine_max > 0 ? opt.refine_max : 30;
  Events ev;
  const char* trace_env = std::getenv("SOVEREIGN_IPM_TRACE");
  const bool trace = trace_env && trace_env[0] == '1';
  double factor_ms = 0.0, solve_ms = 0.0;

  auto launch_rows_axpy = [&](const double* v, const double* base, double* out) {
    SOV_DISPATCH_GROUP(gr, {
      rows_axpy_kernel<G><<<row_blocks, kThreads, 0, st>>>(m, A->r_ptr, A->r_col, A->r_val, v, base, out);
    });
  };
  auto launch_direction = [&](const double* dyv, double* dxv, double* dsv) {
    SOV_DISPATCH_GROUP(gc, {
      direction_cols_kernel<G><<<col_blocks, kThreads, 0, st>>>(n, A->c_ptr, A->c_row, A->c_val, dyv, d.p, rd.p,
                                                               rxs.p, x.p, s.p, dxv, dsv, pcol.p);
    });
  };
  // Normal-equation solve with the precision of the current factor; a
  // mixed-precision solve that cannot reach its tolerance refactors in FP64.
  // The preconditioned CG of an FP32 factor stops at the relative residual
  // max(refine_tol, min(1e-6, 1e-3 mu)): an inexact Newton step while mu is
  // large, exact to refine_tol as mu falls (the factorizations turn FP64 at
  // mixed_switch_mu, where the solves are direct).
  int cur_prec = SOV_PREC_FP64;
  double pcg_tol = solve_tol;
  auto newton_solve = [&](double* v) -> int {
    if (cur_prec == SOV_PREC_FP32) {
      SOV_HIP_TRY(hipMemcpyAsync(rhs_save.p, v, m * sizeof(double), hipMemcpyDeviceToDevice, st));
      int r = ne.solve(v, st, pcg_tol, solve_max);
      if (r < 0) return r;
      res->refine_iterations += ne.last_pcg_iterations();
      if (r == SOV_OK) {
        res->max_solve_residual = std::max(res->max_solve_residual, ne.last_residual());
        return SOV_OK;
      }
      // PCG stalled: this and every later factorization in FP64.
      fp64_locked = true;
      cur_prec = SOV_PREC_FP64;
      if ((r = ne.factor(d.p, SOV_PREC_FP64, st)) != SOV_OK) return r;
      ++res->factorizations;
      if (ne.breakdown(st) != SOV_OK) return SOV_SINGULAR;
      SOV_HIP_TRY(hipMemcpyAsync(v, rhs_save.p, m * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    return ne.solve(v, st, solve_tol, solve_max);
  };

  // Starting point.
  fill_kernel<<<flat_blocks, kThreads, 0, st>>>(n, 1.0, x.p);
  fill_kernel<<<flat_blocks, kThreads, 0, st>>>(n, 1.0, s.p);
  fill_kernel<<<flat_blocks, kThreads, 0, st>>>(m, 0.0, y.p);
  fill_kernel<<<flat_blocks, kThreads, 0, st>>>(n, 1.0, d.p);
  SOV_LAUNCH_CHECK();
  cur_prec = mixed ? SOV_PREC_FP32 : SOV_PREC_FP64;
  if ((rc = ne.factor(d.p, cur_prec, st)) != SOV_OK) return rc;
  ++res->factorizations;
  if (cur_prec == SOV_PREC_FP32) ++res->fp32_factorizations;
  if (ne.breakdown(st) == SOV_OK) {
    SOV_DISPATCH_GROUP(gr, {
      start_rows_kernel<G><<<row_blocks, kThreads, 0, st>>>(m, A->r_ptr, A->r_col, A->r_val, d.p, b.p, dy.p);
    });
    SOV_LAUNCH_CHECK();
    rc = newton_solve(dy.p);
    if (rc < 0) return rc;
    if (rc == SOV_OK) {
      SOV_DISPATCH_GROUP(gc, {
        start_cols_kernel<G><<<col_blocks, kThreads, 0, st>>>(n, A->c_ptr, A->c_row, A->c_val, dy.p, c.p, x.p, s.p);
      });
      SOV_LAUNCH_CHECK();
      SOV_HIP_TRY(hipMemcpyAsync(y.p, dy.p, m * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
  }

  double best_metric = std::numeric_limits<double>::infinity();
  double best_gap = best_metric, best_p = best_metric, best_d = best_metric;
  int best_it = 0;
  SOV_HIP_TRY(hipMemcpyAsync(bx.p, x.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  SOV_HIP_TRY(hipMemcpyAsync(by.p, y.p, m * sizeof(double), hipMemcpyDeviceToDevice, st));
  SOV_HIP_TRY(hipMemcpyAsync(bs.p, s.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));

  bool factor_pending = false;  // a factorization whose breakdown flag is unread
  bool finished = false;
  bool pending_timing = false;
  auto publish_best = [&]() {
    res->status = best_p <= opt.feasibility_tol ? 1 : 6;
    res->iterations = best_it;
    res->duality_gap = best_gap;
    res->primal_residual = best_p;
    res->dual_residual = best_d;
    return SOV_OK;
  };
  const double* out_x = bx.p;
  const double* out_y = by.p;
  const double* out_s = bs.p;

  for (int it = 0; it < opt.max_iterations && !finished; ++it) {
    SOV_DISPATCH_GROUP(gr, {
      resid_rows_kernel<G><<<row_blocks, kThreads, 0, st>>>(m, A->r_ptr, A->r_col, A->r_val, x.p, b.p, rp.p, prow.p);
    });
    SOV_DISPATCH_GROUP(gc, {
      resid_cols_kernel<G><<<col_blocks, kThreads, 0, st>>>(n, A->c_ptr, A->c_row, A->c_val, y.p, c.p, x.p, s.p,
                                                           rd.p, d.p, rxs.p, tmp.p, pcol.p);
    });
    resid_finish_kernel<<<1, kThreads, 0, st>>>(row_blocks, prow.p, col_blocks, pcol.p, n, bnorm, cnorm, sc);
    SOV_LAUNCH_CHECK();
    SOV_HIP_TRY(hipMemcpyAsync(sc_h, sc, sizeof(IpmScalars), hipMemcpyDeviceToHost, st));
    SOV_HIP_TRY(hipStreamSynchronize(st));
    if (pending_timing) {
      factor_ms += Events::ms(ev.e[0], ev.e[1]);
      solve_ms += Events::ms(ev.e[2], ev.e[3]) + Events::ms(ev.e[4], ev.e[5]);
      pending_timing = false;
    }
    if (factor_pending) {
      factor_pending = false;
      if (ne.breakdown(st) != SOV_OK) {  // the previous iteration's factorization failed
        publish_best();
        finished = true;
        break;
      }
    }
    const IpmScalars S = *sc_h;
    if (trace)
      std::fprintf(stderr, "[ipm] it=%d prec=%s pres=%.3e dres=%.3e gap=%.3e mu=%.3e ap=%.3f ad=%.3f pcg=%d res=%.2e\n",
                   it, cur_prec == SOV_PREC_FP32 ? "fp32" : "fp64", S.pres, S.dres, S.gap, S.mu, S.ap, S.ad,
                   res->refine_iterations, res->max_solve_residual);
    const double metric = std::max(S.pres, std::max(S.dres, S.gap));
    if (metric < best_metric) {
      best_metric = metric;
      SOV_HIP_TRY(hipMemcpyAsync(bx.p, x.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      SOV_HIP_TRY(hipMemcpyAsync(by.p, y.p, m * sizeof(double), hipMemcpyDeviceToDevice, st));
      SOV_HIP_TRY(hipMemcpyAsync(bs.p, s.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      best_gap = S.gap;
      best_p = S.pres;
      best_d = S.dres;
      best_it = it + 1;
    }
    if (S.gap < best_gap) best_gap = S.gap;
    if (S.pres < opt.feasibility_tol && S.dres < opt.feasibility_tol && S.gap < opt.optimality_tol) {
      res->status = 0;
      res->iterations = it + 1;
      res->duality_gap = S.gap;
      res->primal_residual = S.pres;
      res->dual_residual = S.dres;
      out_x = x.p;
      out_y = y.p;
      out_s = s.p;
      finished = true;
      break;
    }

    // Factorization of A D A^T, D = x / s (d was written by resid_cols).
    cur_prec = (mixed && !fp64_locked && S.mu >= opt.mixed_switch_mu) ? SOV_PREC_FP32 : SOV_PREC_FP64;
    pcg_tol = std::max(solve_tol, std::min(1e-6, 1e-3 * S.mu));
    (void)hipEventRecord(ev.e[0], st);
    if ((rc = ne.factor(d.p, cur_prec, st)) != SOV_OK) return rc;
    (void)hipEventRecord(ev.e[1], st);
    ++res->factorizations;
    if (cur_prec == SOV_PREC_FP32) {
      ++res->fp32_factorizations;
      const int broke = ne.breakdown(st);
      if (trace)
        std::fprintf(stderr, "[ipm] it=%d fp32 factor: replaced pivots %d%s\n", it, ne.replaced_pivots(),
                     broke != SOV_OK ? ", breakdown (redone in fp64)" : "");
      if (broke != SOV_OK) {  // FP32 breakdown: redo this one in FP64
        fp64_locked = true;
        cur_prec = SOV_PREC_FP64;
        if ((rc = ne.factor(d.p, cur_prec, st)) != SOV_OK) return rc;
        (void)hipEventRecord(ev.e[1], st);
        ++res->factorizations;
        factor_pending = true;
      }
    } else {
      factor_pending = true;
    }

    // Affine predictor.
    const int pcg_before = res->refine_iterations;
    launch_rows_axpy(tmp.p, rp.p, dy.p);
    SOV_LAUNCH_CHECK();
    (void)hipEventRecord(ev.e[2], st);
    rc = newton_solve(dy.p);
    (void)hipEventRecord(ev.e[3], st);
    if (rc < 0) return rc;
    if (rc == SOV_SINGULAR) {
      publish_best();
      finished = true;
      break;
    }
    launch_direction(dy.p, dxa.p, dsa.p);
    alpha_finish_kernel<<<1, kThreads, 0, st>>>(col_blocks, pcol.p, tau, 0, sc);
    mu_aff_kernel<<<flat_blocks, kThreads, 0, st>>>(n, x.p, s.p, dxa.p, dsa.p, sc, pcol.p);
    sigma_finish_kernel<<<1, kThreads, 0, st>>>(flat_blocks, pcol.p, n, sc);
    corrector_cols_kernel<<<flat_blocks, kThreads, 0, st>>>(n, x.p, s.p, dxa.p, dsa.p, d.p, rd.p, sc, rxs.p, tmp.p);
    SOV_LAUNCH_CHECK();

    // Corrector.
    launch_rows_axpy(tmp.p, rp.p, dy.p);
    SOV_LAUNCH_CHECK();
    (void)hipEventRecord(ev.e[4], st);
    rc = newton_solve(dy.p);
    (void)hipEventRecord(ev.e[5], st);
    if (rc < 0) return rc;
    if (rc == SOV_SINGULAR) {
      publish_best();
      finished = true;
      break;
    }
    // The preconditioned CG's iteration count roughly doubles per
    // iteration once mu is small. An FP32 factorization plus the CG costs
    // more than an FP64 factorization with direct solves past about 11 CG
    // iterations (m = 8192: 5.2 ms + 0.28 ms per CG iteration against
    // 8.3 ms), so once the next iteration is expected past that, the
    // remaining factorizations are FP64, before a CG stall forces one to be
    // redone.
    if (cur_prec == SOV_PREC_FP32 && res->refine_iterations - pcg_before > kPcgLockIterations) fp64_locked = true;
    pending_timing = true;
    launch_direction(dy.p, dx.p, ds.p);
    alpha_finish_kernel<<<1, kThreads, 0, st>>>(col_blocks, pcol.p, tau, 1, sc);
    update_kernel<<<flat_blocks, kThreads, 0, st>>>(n, m, sc, x.p, s.p, y.p, dx.p, ds.p, dy.p);
    SOV_LAUNCH_CHECK();
    res->iterations = it + 1;
  }
  if (!finished) publish_best();

  SOV_HIP_TRY(hipStreamSynchronize(st));
  if (x_out) SOV_HIP_TRY(hipMemcpy(x_out, out_x, n * sizeof(double), hipMemcpyDeviceToHost));
  if (y_out) SOV_HIP_TRY(hipMemcpy(y_out, out_y, m * sizeof(double), hipMemcpyDeviceToHost));
  if (s_out) SOV_HIP_TRY(hipMemcpy(s_out, out_s, n * sizeof(double), hipMemcpyDeviceToHost));
  {
    std::vector<double> xv(static_cast<std::size_t>(n));
    SOV_HIP_TRY(hipMemcpy(xv.data(), out_x, n * sizeof(double), hipMemcpyDeviceToHost));
    double obj = 0.0;
    for (int j = 0; j < n; ++j) obj += c_h[j] * xv[j];
    res->objective = obj;
  }
  res->factor_seconds = factor_ms * 1e-3;
  res->solve_seconds = solve_ms * 1e-3;
  res->seconds = std::chrono::duration<double>(Clock::now() - t_start).count();
  count_operation();
  return SOV_OK;
}

}  // namespace sov

extern "C" {

void sov_ipm_default_options(sov_ipm_options* o) {
  o->max_iterations = 250;
  o->feasibility_tol = 1e-9;
  o->optimality_tol = 1e-9;
  o->fraction_to_boundary = 0.999;
  o->enable_scaling = 1;
  o->precision = SOV_PREC_FP64;
  o->mixed_switch_mu = 1e-6;
  o->refine_max = 30;
  o->refine_tol = 1e-12;
  o->normal_path = SOV_NE_AUTO;
}

int sov_ipm_create(sov_mat* a, sov_ipm** out) { return sov_ipm_create_path(a, SOV_NE_AUTO, out); }

int sov_ipm_create_path(sov_mat* a, int32_t path, sov_ipm** out) {
  if (!a || !out || path < SOV_NE_AUTO || path > SOV_NE_SPARSE) return SOV_ERR_ARG;
  *out = nullptr;
  if (!sov::device_ready()) {
    sov::set_error("no usable GPU");
    return SOV_ERR_NODEV;
  }
  auto* h = new sov_ipm;
  h->A = a;
  const int rc = h->ne.init(a, path);
  if (rc != SOV_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return SOV_OK;
}

int sov_ipm_solve(sov_ipm* h, const double* b, const double* c, const sov_ipm_options* opt, sov_ipm_result* res,
                  double* x, double* y, double* s) {
  if (!h || !b || !c || !opt || !res) return SOV_ERR_ARG;
  return sov::ipm_run(h, b, c, *opt, res, x, y, s);
}

int sov_ipm_free(sov_ipm* h) {
  delete h;
  return SOV_OK;
}

int sov_ipm_solve_form(int32_t m, int32_t n, const int32_t* col_ptr, const int32_t* row_idx, const double* values,
                       const double* b, const double* c, const sov_ipm_options* opt, sov_ipm_result* res, double* x,
                       double* y, double* s) {
  if (!col_ptr || !b || !c || !opt || !res) return SOV_ERR_ARG;
  *res = sov_ipm_result{};
  res->status = 7;
  if (!sov::device_ready()) {
    sov::set_error("no usable GPU");
    return SOV_ERR_NODEV;
  }
  sov_mat* A = nullptr;
  int rc = sov::mat_upload(m, n, col_ptr, row_idx, values, &A);
  if (rc != SOV_OK) return rc;
  sov_ipm* h = nullptr;
  rc = sov_ipm_create_path(A, opt->normal_path, &h);
  if (rc == SOV_OK) rc = sov::ipm_run(h, b, c, *opt, res, x, y, s);
  sov_ipm_free(h);
  sov::mat_release(A);
  return rc;
}

}  // extern "C"
