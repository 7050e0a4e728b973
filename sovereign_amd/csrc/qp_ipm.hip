// Generic convex QP interior point on the GPU: the generic branch of
// QpInteriorPointSolver::solve (solver/qp/src/qp_interior_point.cpp:
// 3166-3180), i.e. solve_qp_ipm (:2848-3116) on the form of build_qp_form
// (:118-247), with the Newton system of QpKkt (:398-656) solved in
// Schur-complement form on the device:
//
//   [Q + H  -A^T] [dx]   [r1]      K = Q + H + delta I  (blocked Cholesky, MFMA)
//   [A       0  ] [dy] = [r2]      V' = A L^-T          (MFMA, through the factor's
//                                                        diagonal-block inverses)
//   S = V'V + delta I (MFMA SYRK), factored by the same Cholesky;
//   dy = S^-1 (r2 - A K^-1 r1),  dx = K^-1 (r1 + A^T dy),
//
// refined against the unregularised system (at most 3 steps, each kept only
// when it halves the residual). Residuals, the stopping test, the step rule
// and the updates follow solve_qp_ipm; vectors live in HBM and the host reads
// back one small record of block partial sums per reduction to take the
// reference's decisions. Everything is FP64.

#include "chol.hpp"
#include "internal.hpp"
#include "mfma_gemm.hpp"
#include "sovereign_qp.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

namespace {

using sov::DenseCholesky;
using sov::FactorFlags;
using sov::kSuper;

constexpr double kRegInit = 1e-9;  // qp_interior_point.cpp:386-388
constexpr double kRegMax = 1e-5;
constexpr int kRefineSteps = 3;
constexpr int kRedBlocks = 64, kRedThreads = 256, kRedSlots = 8;

// ------------------------------------------------------------- the form --
// build_qp_form: lower bounds shifted out (free when lb <= -1e29), finite
// upper bounds as extra rows with a slack, a slack per Le row and a surplus
// per Ge row, Q of 1/2 x'Qx over the form's columns with 1e-14 on the
// structural diagonal (build_hessian, :65-100), c = sense c + Q shift.
struct QpForm {
  int ns = 0, n = 0, m = 0;
  std::vector<int32_t> acp, ari;  // A, m x n CSC
  std::vector<double> aval;
  std::vector<int32_t> qcp, qri;  // Q, n x n CSC (both triangles)
  std::vector<double> qval;
  std::vector<double> b, c, shift, freev;
  double offset = 0.0, sign = 1.0;
  bool has_q = false;
};

int build_form(const sov_qp_model& md, QpForm& F) {
  F.ns = md.n_vars;
  F.sign = md.maximize ? -1.0 : 1.0;
  const int m0 = md.n_rows, ns = md.n_vars;
  F.shift.assign(static_cast<std::size_t>(ns), 0.0);
  std::vector<char> free_s(static_cast<std::size_t>(ns), 0);
  std::vector<int> ub_var;
  std::vector<double> ub_val;
  for (int j = 0; j < ns; ++j) {
    const bool no_lower = !(md.lb[j] > -1e29);
    free_s[j] = no_lower ? 1 : 0;
    F.shift[j] = no_lower ? 0.0 : md.lb[j];
    if (std::isfinite(md.ub[j]) && md.ub[j] < 1e29) {
      const double u = md.ub[j] - F.shift[j];
      if (!no_lower && u < -1e-12) {
        sov::set_error("sov_qp_ipm_solve: inconsistent bounds for variable " + std::to_string(j));
        return SOV_ERR_ARG;
      }
      ub_var.push_back(j);
      ub_val.push_back(u);
    }
  }
  F.m = m0 + static_cast<int>(ub_var.size());
  F.b.assign(static_cast<std::size_t>(F.m), 0.0);
  std::vector<char> kind(static_cast<std::size_t>(F.m), 'L');
  for (int r = 0; r < m0; ++r) {
    const char s = md.row_sense[r];
    kind[r] = (s == 'G' || s == 'g') ? 'G' : (s == 'E' || s == 'e') ? 'E' : 'L';
    F.b[r] = md.rhs[r];
  }
  std::vector<std::vector<std::pair<int, double>>> cols(static_cast<std::size_t>(ns));
  for (int j = 0; j < ns; ++j)
    for (int p = md.col_ptr[j]; p < md.col_ptr[j + 1]; ++p) {
      const int r = md.row_idx[p];
      if (r < 0 || r >= m0) {
        sov::set_error("sov_qp_ipm_solve: row index out of range");
        return SOV_ERR_ARG;
      }
      if (md.values[p] == 0.0) continue;
      cols[j].push_back({r, md.values[p]});
      F.b[r] -= md.values[p] * F.shift[j];
    }
  for (std::size_t k = 0; k < ub_var.size(); ++k) {
    cols[ub_var[k]].push_back({m0 + static_cast<int>(k), 1.0});
    F.b[m0 + k] = ub_val[k];
  }
  int nslack = 0;
  for (char k : kind) nslack += k != 'E';
  F.n = ns + nslack;
  cols.resize(static_cast<std::size_t>(F.n));
  int sc = ns;
  for (int r = 0; r < F.m; ++r) {
    if (kind[r] == 'L') cols[sc++].push_back({r, 1.0});
    else if (kind[r] == 'G') cols[sc++].push_back({r, -1.0});
  }
  F.acp.assign(static_cast<std::size_t>(F.n) + 1, 0);
  for (int j = 0; j < F.n; ++j) {
    auto& e = cols[j];
    std::sort(e.begin(), e.end());
    std::vector<std::pair<int, double>> merged;
    for (const auto& t : e) {
      if (!merged.empty() && merged.back().first == t.first) merged.back().second += t.second;
      else merged.push_back(t);
    }
    for (const auto& t : merged)
      if (t.second != 0.0) {
        F.ari.push_back(t.first);
        F.aval.push_back(t.second);
      }
    F.acp[j + 1] = static_cast<int32_t>(F.ari.size());
  }
  // Q over the form's columns
  std::vector<std::vector<std::pair<int, double>>> qc(static_cast<std::size_t>(F.n));
  for (int64_t t = 0; t < md.q_nnz; ++t) {
    const int i = md.q_row[t], j = md.q_col[t];
    if (i < 0 || i >= ns || j < 0 || j >= ns) {
      sov::set_error("sov_qp_ipm_solve: quadratic term index out of range");
      return SOV_ERR_ARG;
    }
    const double v = F.sign * md.q_val[t];
    if (i == j) qc[i].push_back({i, v});
    else {
      qc[j].push_back({i, 0.5 * v});
      qc[i].push_back({j, 0.5 * v});
    }
  }
  for (int i = 0; i < ns; ++i) qc[i].push_back({i, 1e-14});
  F.qcp.assign(static_cast<std::size_t>(F.n) + 1, 0);
  for (int j = 0; j < F.n; ++j) {
    auto& e = qc[j];
    std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (std::size_t k = 0; k < e.size();) {
      const int r = e[k].first;
      double v = 0.0;
      for (; k < e.size() && e[k].first == r; ++k) v += e[k].second;
      if (v != 0.0) {
        F.qri.push_back(r);
        F.qval.push_back(v);
        if (std::abs(v) > 1e-12) F.has_q = true;
      }
    }
    F.qcp[j + 1] = static_cast<int32_t>(F.qri.size());
  }
  // c = sense c + Q shift; the objective offset of the gap test (:2956-2969)
  std::vector<double> sf(static_cast<std::size_t>(F.n), 0.0), qs(static_cast<std::size_t>(F.n), 0.0);
  for (int j = 0; j < ns; ++j) sf[j] = F.shift[j];
  for (int j = 0; j < F.n; ++j)
    for (int p = F.qcp[j]; p < F.qcp[j + 1]; ++p) qs[F.qri[p]] += F.qval[p] * sf[j];
  F.c.assign(static_cast<std::size_t>(F.n), 0.0);
  for (int j = 0; j < ns; ++j) F.c[j] = F.sign * md.obj[j] + qs[j];
  double cs = 0.0, sqs = 0.0;
  for (int j = 0; j < F.n; ++j) {
    cs += F.c[j] * sf[j];
    sqs += sf[j] * qs[j];
  }
  F.offset = cs - 0.5 * sqs + F.sign * md.obj_constant;
  F.freev.assign(static_cast<std::size_t>(F.n), 0.0);
  for (int j = 0; j < ns; ++j) F.freev[j] = free_s[j] ? 1.0 : 0.0;
  return SOV_OK;
}

// -------------------------------------------------------------- kernels --
__device__ __forceinline__ void block_reduce(double* vals, int nv, const bool* is_max, double* out) {
  __shared__ double red[kRedSlots][kRedThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < nv; ++k) {
    double v = is_max[k] ? wave_max(vals[k]) : wave_sum(vals[k]);
    if (lane == 0) red[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    const int k = threadIdx.x;
    double v = red[k][0];
    for (int q = 1; q < kRedThreads / 64; ++q) v = is_max[k] ? fmax(v, red[k][q]) : v + red[k][q];
    out[blockIdx.x * kRedSlots + k] = v;
  }
}

// rp = b - Ax, rd = Qx + c - A'y - s; partials: max|rp|, max|rd|, max|Qx|, x's, c'x, x'Qx
__global__ __launch_bounds__(kRedThreads) void resid_kernel(int n, int m, const double* __restrict__ b,
                                                            const double* __restrict__ ax, const double* __restrict__ qx,
                                                            const double* __restrict__ c, const double* __restrict__ aty,
                                                            const double* __restrict__ s, const double* __restrict__ x,
                                                            double* __restrict__ rp, double* __restrict__ rd,
                                                            double* __restrict__ out) {
  double v[6] = {0, 0, 0, 0, 0, 0};
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const double r = b[i] - ax[i];
    rp[i] = r;
    v[0] = fmax(v[0], fabs(r));
  }
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const double r = qx[j] + c[j] - aty[j] - s[j];
    rd[j] = r;
    v[1] = fmax(v[1], fabs(r));
    v[2] = fmax(v[2], fabs(qx[j]));
    v[3] += x[j] * s[j];
    v[4] += c[j] * x[j];
    v[5] += x[j] * qx[j];
  }
  const bool mx[6] = {true, true, true, false, false, false};
  block_reduce(v, 6, mx, out);
}

__global__ void fill_kernel(double* __restrict__ p, double v, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// K = [Q] + diag(h + reg) (column-major n x n, both triangles of Q); diag0 = diag(K)
__global__ void assemble_k_kernel(int n, const int32_t* __restrict__ qcp, const int32_t* __restrict__ qri,
                                  const double* __restrict__ qval, int with_q, const double* __restrict__ h, double reg,
                                  double* __restrict__ K, double* __restrict__ diag0) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double* col = K + static_cast<long long>(j) * n;
  if (with_q)
    for (int p = qcp[j]; p < qcp[j + 1]; ++p) col[qri[p]] += qval[p];
  col[j] += h[j] + reg;
  diag0[j] = col[j];
}

// h = X^-1 S (0 on free columns), or 1 for the starting point
__global__ void hdiag_kernel(int n, const double* __restrict__ x, const double* __restrict__ s,
                             const double* __restrict__ fr, int with_hessian, double* __restrict__ h) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  h[j] = with_hessian ? (fr[j] != 0.0 ? 0.0 : fmax(s[j], 1e-16) / fmax(x[j], 1e-16)) : 1.0;
}

__global__ void diag_add_kernel(int m, double* __restrict__ S, double reg, double* __restrict__ diag0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  S[i + static_cast<long long>(i) * m] += reg;
  diag0[i] = S[i + static_cast<long long>(i) * m];
}

// z = a + alpha * w
__global__ void axpy_kernel(int n, const double* __restrict__ a, double alpha, const double* __restrict__ w,
                            double* __restrict__ z) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) z[j] = a[j] + alpha * w[j];
}

// e1 = r1 - ([Q] dx + h dx - A'dy), e2 = r2 - A dx; partials max|e1|, max|e2|
__global__ __launch_bounds__(kRedThreads) void kkt_resid_kernel(int n, int m, const double* __restrict__ r1,
                                                                const double* __restrict__ qdx, int with_q,
                                                                const double* __restrict__ h,
                                                                const double* __restrict__ dx,
                                                                const double* __restrict__ atdy,
                                                                const double* __restrict__ r2,
                                                                const double* __restrict__ adx, double* __restrict__ e1,
                                                                double* __restrict__ e2, double* __restrict__ out) {
  double v[2] = {0, 0};
  const int stride = gridDim.x * blockDim.x;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const double e = r1[j] - ((with_q ? qdx[j] : 0.0) + h[j] * dx[j] - atdy[j]);
    e1[j] = e;
    v[0] = fmax(v[0], fabs(e));
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const double e = r2[i] - adx[i];
    e2[i] = e;
    v[1] = fmax(v[1], fabs(e));
  }
  const bool mx[2] = {true, true};
  block_reduce(v, 2, mx, out);
}

// r1 = -rd + rxs / x (free: -rd)
__global__ void r1_kernel(int n, const double* __restrict__ fr, const double* __restrict__ rd,
                          const double* __restrict__ rxs, const double* __restrict__ x, double* __restrict__ r1) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) r1[j] = fr[j] != 0.0 ? -rd[j] : -rd[j] + rxs[j] / fmax(x[j], 1e-16);
}

// ds = (rxs - s dx) / x (free: 0)
__global__ void ds_kernel(int n, const double* __restrict__ fr, const double* __restrict__ rxs,
                          const double* __restrict__ s, const double* __restrict__ dx, const double* __restrict__ x,
                          double* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) ds[j] = fr[j] != 0.0 ? 0.0 : (rxs[j] - s[j] * dx[j]) / fmax(x[j], 1e-16);
}

// partials: min over bounded columns of -x/dx (dx < 0) and -s/ds (ds < 0)
__global__ __launch_bounds__(kRedThreads) void step_kernel(int n, const double* __restrict__ fr,
                                                           const double* __restrict__ x, const double* __restrict__ dx,
                                                           const double* __restrict__ s, const double* __restrict__ ds,
                                                           double* __restrict__ out) {
  double v[2] = {-1.0, -1.0};  // maxima of -(step): the host negates
  const int stride = gridDim.x * blockDim.x;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    if (fr[j] != 0.0) continue;
    if (dx[j] < 0.0) v[0] = fmax(v[0], x[j] / dx[j]);
    if (ds[j] < 0.0) v[1] = fmax(v[1], s[j] / ds[j]);
  }
  const bool mx[2] = {true, true};
  block_reduce(v, 2, mx, out);
}

// partial: sum over bounded columns of (x + ap dx)(s + ad ds)
__global__ __launch_bounds__(kRedThreads) void muaff_kernel(int n, const double* __restrict__ fr,
                                                            const double* __restrict__ x, const double* __restrict__ dx,
                                                            double ap, const double* __restrict__ s,
                                                            const double* __restrict__ ds, double ad,
                                                            double* __restrict__ out) {
  double v[1] = {0.0};
  const int stride = gridDim.x * blockDim.x;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    if (fr[j] == 0.0) v[0] += (x[j] + ap * dx[j]) * (s[j] + ad * ds[j]);
  const bool mx[1] = {false};
  block_reduce(v, 1, mx, out);
}

// rxs = -x s (predictor) or -x s - dxa dsa + sigma mu (corrector); 0 on free columns
__global__ void rxs_kernel(int n, const double* __restrict__ fr, const double* __restrict__ x,
                           const double* __restrict__ s, const double* __restrict__ dxa, const double* __restrict__ dsa,
                           double sigma_mu, int corrector, double* __restrict__ rxs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (!corrector) rxs[j] = -x[j] * s[j];
  else rxs[j] = fr[j] != 0.0 ? 0.0 : -x[j] * s[j] - dxa[j] * dsa[j] + sigma_mu;
}

__global__ void update_kernel(int n, const double* __restrict__ fr, double* __restrict__ x,
                              const double* __restrict__ dx, double ap, double* __restrict__ s,
                              const double* __restrict__ ds, double ad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (fr[j] != 0.0) {
    x[j] += ap * dx[j];
    return;
  }
  x[j] = fmax(1e-14, x[j] + ap * dx[j]);
  s[j] = fmax(1e-14, s[j] + ad * ds[j]);
}

// The default starting point (:2933-2947): x = max(1, |A'dy|) (free: 1 + A'dy),
// s = max(1, |Q 1 + c - A'dy|)
__global__ void start_default_kernel(int n, const double* __restrict__ fr, const double* __restrict__ atdy,
                                     const double* __restrict__ q1, const double* __restrict__ c,
                                     double* __restrict__ x, double* __restrict__ s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (fr[j] != 0.0) {
    x[j] = 1.0 + atdy[j];
    return;
  }
  x[j] = fmax(1.0, fabs(atdy[j]));
  s[j] = fmax(1.0, fabs(q1[j] + c[j] - atdy[j]));
}

__global__ void zero_free_kernel(int n, const double* __restrict__ fr, double* __restrict__ s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n && fr[j] != 0.0) s[j] = 0.0;
}

// Mehrotra's start (:2886-2932), second half: with xt = 1 + A'dy and u the
// least-squares dual direction, x = max(1e-8, xt + shift_x + bal_x),
// s = max(1e-8, -u + shift_s + bal_s) (free: x = xt)
__global__ void start_mehrotra_kernel(int n, const double* __restrict__ fr, const double* __restrict__ xt,
                                      const double* __restrict__ u, double sx, double bx, double ss, double bs,
                                      double* __restrict__ x, double* __restrict__ s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (fr[j] != 0.0) {
    x[j] = xt[j];
    return;
  }
  x[j] = fmax(1e-8, xt[j] + sx + bx);
  s[j] = fmax(1e-8, -u[j] + ss + bs);
}

// partials over bounded columns: max(-xt), max(u) (= -min(-u)), then
// (with the shifts) sum xh sh, sum xh, sum sh
__global__ __launch_bounds__(kRedThreads) void mehrotra_stats_kernel(int n, const double* __restrict__ fr,
                                                                     const double* __restrict__ xt,
                                                                     const double* __restrict__ u, double sx, double ss,
                                                                     double* __restrict__ out) {
  double v[5] = {-1e300, -1e300, 0.0, 0.0, 0.0};
  const int stride = gridDim.x * blockDim.x;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    if (fr[j] != 0.0) continue;
    v[0] = fmax(v[0], -xt[j]);
    v[1] = fmax(v[1], u[j]);
    const double xh = xt[j] + sx, sh = -u[j] + ss;
    v[2] += xh * sh;
    v[3] += xh;
    v[4] += sh;
  }
  const bool mx[5] = {true, true, false, false, false};
  block_reduce(v, 5, mx, out);
}

unsigned blocks(long long n, int per = 256) { return static_cast<unsigned>(std::max(1LL, (n + per - 1) / per)); }

// ---------------------------------------------------------------- state --
struct Dev {
  int n = 0, m = 0;
  sov_mat* A = nullptr;
  sov_mat* Q = nullptr;
  std::vector<void*> owned;
  double *Ad = nullptr, *K = nullptr, *dK = nullptr, *Vt = nullptr, *Vw = nullptr, *S = nullptr, *dS = nullptr;
  double *b = nullptr, *c = nullptr, *fr = nullptr, *x = nullptr, *s = nullptr, *y = nullptr, *h = nullptr;
  double *ax = nullptr, *qx = nullptr, *aty = nullptr, *rp = nullptr, *rd = nullptr, *rxs = nullptr, *r1 = nullptr;
  double *t = nullptr, *u = nullptr, *w = nullptr, *dx = nullptr, *dy = nullptr, *ds = nullptr;
  double *dxa = nullptr, *dya = nullptr, *dsa = nullptr, *e1 = nullptr, *e2 = nullptr, *cx = nullptr, *cy = nullptr;
  double *fx = nullptr, *fy = nullptr, *ones = nullptr, *zn = nullptr, *zm = nullptr, *red = nullptr;
  int32_t *qcp = nullptr, *qri = nullptr;
  double* qval = nullptr;
  DenseCholesky<double> chK, chS;
  double reg = 0.0;
  bool with_q = true;
  int factorizations = 0, refinements = 0;
  ~Dev() {
    for (void* p : owned) (void)hipFree(p);
    if (A) sov::mat_release(A);
    if (Q) sov::mat_release(Q);
  }
  template <typename T>
  int alloc(T** p, std::size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<std::size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      sov::set_error("sov_qp_ipm_solve: out of device memory");
      return SOV_ERR_NOMEM;
    }
    owned.push_back(q);
    *p = static_cast<T*>(q);
    return SOV_OK;
  }
};

// Sums (or maxima) over the block partials of the last reduction.
int read_partials(Dev& D, int nv, const bool* is_max, double* out, hipStream_t s) {
  std::vector<double> part(static_cast<std::size_t>(kRedBlocks) * kRedSlots);
  SOV_HIP_TRY(hipMemcpyAsync(part.data(), D.red, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  for (int k = 0; k < nv; ++k) {
    double v = part[k];
    for (int bk = 1; bk < kRedBlocks; ++bk) v = is_max[k] ? std::max(v, part[bk * kRedSlots + k]) : v + part[bk * kRedSlots + k];
    out[k] = v;
  }
  return SOV_OK;
}

int factor_status(const DenseCholesky<double>& ch, hipStream_t s, int* status) {
  FactorFlags f{};
  SOV_HIP_TRY(hipMemcpyAsync(&f, ch.flags(), sizeof(FactorFlags), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  *status = f.status;
  return SOV_OK;
}

// K = [Q] + H + reg I, its Cholesky factor, V' = A L^-T and the factored
// S = V'V + reg I, raising reg 100-fold on a breakdown of either factor.
int factor_kkt(Dev& D, bool with_hessian, hipStream_t s) {
  D.with_q = with_hessian;
  const int n = D.n, m = D.m;
  hdiag_kernel<<<blocks(n), 256, 0, s>>>(n, D.x, D.s, D.fr, with_hessian ? 1 : 0, D.h);
  SOV_LAUNCH_CHECK();
  for (double reg = std::max(D.reg, kRegInit); reg <= kRegMax * 1.0000001; reg *= 100.0) {
    fill_kernel<<<1024, 256, 0, s>>>(D.K, 0.0, static_cast<long long>(n) * n);
    assemble_k_kernel<<<blocks(n), 256, 0, s>>>(n, D.qcp, D.qri, D.qval, with_hessian ? 1 : 0, D.h, reg, D.K, D.dK);
    SOV_LAUNCH_CHECK();
    int rc = D.chK.factor(D.K, n, D.dK, s);
    if (rc != SOV_OK) return rc;
    ++D.factorizations;
    sov::count_factorization();
    int st = 0;
    if ((rc = factor_status(D.chK, s, &st)) != SOV_OK) return rc;
    if (st != 0) continue;
    if (m > 0) {
      // V' (m x n) = A L^-T, block column by block column: V'_P = (A_P - V'_<P L_P,<P^T) inv(L_PP)^T
      SOV_HIP_TRY(hipMemcpyAsync(D.Vt, D.Ad, sizeof(double) * m * static_cast<std::size_t>(n), hipMemcpyDeviceToDevice, s));
      const double* inv = D.chK.super_inverses();
      for (int p0 = 0, P = 0; p0 < n; p0 += kSuper, ++P) {
        const int bp = std::min(kSuper, n - p0);
        double* vp = D.Vt + static_cast<long long>(p0) * m;
        if (p0 > 0)
          SOV_HIP_TRY((sov::gemm_update<double, false, false>(m, bp, p0, D.Vt, m, D.K + p0, n, vp, m, -1.0, s)));
        SOV_HIP_TRY((sov::gemm_update<double, false, false, true>(m, bp, bp, vp, m,
                                                                  inv + static_cast<long long>(P) * kSuper * kSuper,
                                                                  kSuper, D.Vw, m, 1.0, s)));
        SOV_HIP_TRY(hipMemcpyAsync(vp, D.Vw, sizeof(double) * m * static_cast<std::size_t>(bp), hipMemcpyDeviceToDevice, s));
      }
      // S = V'V (lower) + reg I
      SOV_HIP_TRY((sov::gemm_update<double, false, true, true>(m, m, n, D.Vt, m, D.Vt, m, D.S, m, 1.0, s)));
      diag_add_kernel<<<blocks(m), 256, 0, s>>>(m, D.S, reg, D.dS);
      SOV_LAUNCH_CHECK();
      if ((rc = D.chS.factor(D.S, m, D.dS, s)) != SOV_OK) return rc;
      ++D.factorizations;
      sov::count_factorization();
      if ((rc = factor_status(D.chS, s, &st)) != SOV_OK) return rc;
      if (st != 0) continue;
    }
    D.reg = reg;
    return SOV_OK;
  }
  return SOV_SINGULAR;
}

// One regularised solve: dy = S^-1 (r2 - A K^-1 r1), dx = K^-1 (r1 + A'dy).
int schur_solve(Dev& D, const double* r1, const double* r2, double* dx, double* dy, hipStream_t s) {
  const int n = D.n, m = D.m;
  SOV_HIP_TRY(hipMemcpyAsync(D.t, r1, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  int rc = D.chK.solve(D.K, n, D.t, nullptr, s);
  if (rc != SOV_OK) return rc;
  if (m > 0) {
    if ((rc = sov::spmv_launch(D.A, false, D.t, D.u, s)) != SOV_OK) return rc;
    axpy_kernel<<<blocks(m), 256, 0, s>>>(m, r2, -1.0, D.u, dy);
    SOV_LAUNCH_CHECK();
    if ((rc = D.chS.solve(D.S, m, dy, nullptr, s)) != SOV_OK) return rc;
    if ((rc = sov::spmv_launch(D.A, true, dy, D.w, s)) != SOV_OK) return rc;
    axpy_kernel<<<blocks(n), 256, 0, s>>>(n, r1, 1.0, D.w, dx);
    SOV_LAUNCH_CHECK();
    return D.chK.solve(D.K, n, dx, nullptr, s);
  }
  SOV_HIP_TRY(hipMemcpyAsync(dx, D.t, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  return SOV_OK;
}

// Residual of the unregularised system for (dx, dy) into e1, e2; returns its max-norm.
int kkt_residual(Dev& D, const double* r1, const double* r2, const double* dx, const double* dy, double* err,
                 hipStream_t s) {
  const int n = D.n, m = D.m;
  int rc;
  if (D.with_q && (rc = sov::spmv_launch(D.Q, false, dx, D.qx, s)) != SOV_OK) return rc;
  if (m > 0) {
    if ((rc = sov::spmv_launch(D.A, true, dy, D.w, s)) != SOV_OK) return rc;
    if ((rc = sov::spmv_launch(D.A, false, dx, D.u, s)) != SOV_OK) return rc;
  } else {
    fill_kernel<<<blocks(n), 256, 0, s>>>(D.w, 0.0, n);
  }
  kkt_resid_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, m, r1, D.qx, D.with_q ? 1 : 0, D.h, dx, D.w, r2, D.u, D.e1,
                                                      D.e2, D.red);
  SOV_LAUNCH_CHECK();
  const bool mx[2] = {true, true};
  double v[2];
  if ((rc = read_partials(D, 2, mx, v, s)) != SOV_OK) return rc;
  *err = std::max(v[0], v[1]);
  return SOV_OK;
}

// QpKkt::solve: the regularised solve, then refinement steps while each
// halves the residual.
int kkt_solve(Dev& D, const double* r1, const double* r2, double* dx, double* dy, bool refine, hipStream_t s) {
  int rc = schur_solve(D, r1, r2, dx, dy, s);
  if (rc != SOV_OK || !refine) return rc;
  double err = 0.0;
  if ((rc = kkt_residual(D, r1, r2, dx, dy, &err, s)) != SOV_OK) return rc;
  const int n = D.n, m = D.m;
  for (int step = 0; step < kRefineSteps && err > 0.0; ++step) {
    // correction from (e1, e2) (copied: the residual below overwrites them)
    SOV_HIP_TRY(hipMemcpyAsync(D.fx, D.e1, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (m > 0) SOV_HIP_TRY(hipMemcpyAsync(D.fy, D.e2, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    if ((rc = schur_solve(D, D.fx, D.fy, D.cx, D.cy, s)) != SOV_OK) return rc;
    axpy_kernel<<<blocks(n), 256, 0, s>>>(n, dx, 1.0, D.cx, D.cx);
    if (m > 0) axpy_kernel<<<blocks(m), 256, 0, s>>>(m, dy, 1.0, D.cy, D.cy);
    SOV_LAUNCH_CHECK();
    double next = 0.0;
    if ((rc = kkt_residual(D, r1, r2, D.cx, D.cy, &next, s)) != SOV_OK) return rc;
    if (!(next < 0.5 * err)) break;
    SOV_HIP_TRY(hipMemcpyAsync(dx, D.cx, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (m > 0) SOV_HIP_TRY(hipMemcpyAsync(dy, D.cy, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    ++D.refinements;
    err = next;
  }
  return SOV_OK;
}

int setup(Dev& D, const QpForm& F, hipStream_t s) {
  const int n = F.n, m = F.m;
  D.n = n;
  D.m = m;
  int rc;
  if ((rc = sov::mat_upload(m, n, F.acp.data(), F.ari.data(), F.aval.data(), &D.A)) != SOV_OK) return rc;
  if ((rc = sov::mat_upload(n, n, F.qcp.data(), F.qri.data(), F.qval.data(), &D.Q)) != SOV_OK) return rc;
  double** vn[] = {&D.c, &D.fr, &D.x, &D.s, &D.h, &D.qx, &D.aty, &D.rd, &D.rxs, &D.r1, &D.t, &D.w, &D.dx,
                   &D.ds, &D.dxa, &D.dsa, &D.e1, &D.cx, &D.fx, &D.ones, &D.zn, &D.dK};
  for (double** p : vn)
    if ((rc = D.alloc(p, static_cast<std::size_t>(n))) != SOV_OK) return rc;
  double** vm[] = {&D.b, &D.y, &D.ax, &D.rp, &D.u, &D.dy, &D.dya, &D.e2, &D.cy, &D.fy, &D.zm, &D.dS};
  for (double** p : vm)
    if ((rc = D.alloc(p, static_cast<std::size_t>(std::max(m, 1)))) != SOV_OK) return rc;
  if ((rc = D.alloc(&D.K, static_cast<std::size_t>(n) * n)) != SOV_OK) return rc;
  if ((rc = D.alloc(&D.red, static_cast<std::size_t>(kRedBlocks) * kRedSlots)) != SOV_OK) return rc;
  if (m > 0) {
    if ((rc = D.alloc(&D.Ad, static_cast<std::size_t>(m) * n)) != SOV_OK) return rc;
    if ((rc = D.alloc(&D.Vt, static_cast<std::size_t>(m) * n)) != SOV_OK) return rc;
    if ((rc = D.alloc(&D.Vw, static_cast<std::size_t>(m) * std::min(n, kSuper))) != SOV_OK) return rc;
    if ((rc = D.alloc(&D.S, static_cast<std::size_t>(m) * m)) != SOV_OK) return rc;
    std::vector<double> ad(static_cast<std::size_t>(m) * n, 0.0);
    for (int j = 0; j < n; ++j)
      for (int p = F.acp[j]; p < F.acp[j + 1]; ++p) ad[F.ari[p] + static_cast<std::size_t>(j) * m] = F.aval[p];
    SOV_HIP_TRY(hipMemcpyAsync(D.Ad, ad.data(), ad.size() * sizeof(double), hipMemcpyHostToDevice, s));
    SOV_HIP_TRY(hipMemcpyAsync(D.b, F.b.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    SOV_HIP_TRY(hipStreamSynchronize(s));
  }
  if ((rc = D.alloc(&D.qcp, F.qcp.size())) != SOV_OK) return rc;
  if ((rc = D.alloc(&D.qri, std::max<std::size_t>(F.qri.size(), 1))) != SOV_OK) return rc;
  if ((rc = D.alloc(&D.qval, std::max<std::size_t>(F.qval.size(), 1))) != SOV_OK) return rc;
  SOV_HIP_TRY(hipMemcpy(D.qcp, F.qcp.data(), F.qcp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (!F.qri.empty()) {
    SOV_HIP_TRY(hipMemcpy(D.qri, F.qri.data(), F.qri.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    SOV_HIP_TRY(hipMemcpy(D.qval, F.qval.data(), F.qval.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  SOV_HIP_TRY(hipMemcpy(D.c, F.c.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  SOV_HIP_TRY(hipMemcpy(D.fr, F.freev.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  fill_kernel<<<blocks(n), 256, 0, s>>>(D.ones, 1.0, n);
  fill_kernel<<<blocks(n), 256, 0, s>>>(D.zn, 0.0, n);
  fill_kernel<<<blocks(std::max(m, 1)), 256, 0, s>>>(D.zm, 0.0, std::max(m, 1));
  SOV_LAUNCH_CHECK();
  return SOV_OK;
}

struct Attempt {
  int status = 7, iterations = 0;
  double gap = 0, pres = 0, dres = 0;
};

// solve_qp_ipm (:2848-3116) from the default or Mehrotra's starting point.
int run_ipm(Dev& D, const QpForm& F, const sov_qp_ipm_options& o, bool mehrotra, Attempt& A, hipStream_t s) {
  const int n = D.n, m = D.m;
  const bool refine = o.iterative_refinement != 0;
  int rc;
  fill_kernel<<<blocks(n), 256, 0, s>>>(D.x, 1.0, n);
  fill_kernel<<<blocks(n), 256, 0, s>>>(D.s, 1.0, n);
  fill_kernel<<<blocks(std::max(m, 1)), 256, 0, s>>>(D.y, 0.0, std::max(m, 1));
  SOV_LAUNCH_CHECK();
  D.reg = 0.0;
  if (m > 0) {
    // H = I with Q left out: dy from [I -A'; A 0] [.; dy] = [0; b - A 1]
    rc = factor_kkt(D, false, s);
    if (rc < 0) return rc;
    if (rc == SOV_OK) {
      if ((rc = sov::spmv_launch(D.A, false, D.ones, D.ax, s)) != SOV_OK) return rc;
      axpy_kernel<<<blocks(m), 256, 0, s>>>(m, D.b, -1.0, D.ax, D.rp);
      SOV_LAUNCH_CHECK();
      if ((rc = kkt_solve(D, D.zn, D.rp, D.dxa, D.dya, refine, s)) != SOV_OK) return rc;
      if ((rc = sov::spmv_launch(D.A, true, D.dya, D.aty, s)) != SOV_OK) return rc;  // A'dy
      if (!mehrotra) {
        if ((rc = sov::spmv_launch(D.Q, false, D.ones, D.qx, s)) != SOV_OK) return rc;
        start_default_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.aty, D.qx, D.c, D.x, D.s);
        SOV_LAUNCH_CHECK();
        SOV_HIP_TRY(hipMemcpyAsync(D.y, D.dya, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
      } else {
        // xt = 1 + A'dy; (u, yls) solve [I -A'; A 0] = [-(c + Q xt); 0]
        double* xt = D.dsa;
        axpy_kernel<<<blocks(n), 256, 0, s>>>(n, D.ones, 1.0, D.aty, xt);
        SOV_LAUNCH_CHECK();
        if ((rc = sov::spmv_launch(D.Q, false, xt, D.qx, s)) != SOV_OK) return rc;
        axpy_kernel<<<blocks(n), 256, 0, s>>>(n, D.c, 1.0, D.qx, D.r1);
        axpy_kernel<<<blocks(n), 256, 0, s>>>(n, D.zn, -1.0, D.r1, D.r1);
        SOV_LAUNCH_CHECK();
        if ((rc = kkt_solve(D, D.r1, D.zm, D.dx, D.dy, refine, s)) != SOV_OK) return rc;
        const bool mx[5] = {true, true, false, false, false};
        double v[5];
        mehrotra_stats_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, D.fr, xt, D.dx, 0.0, 0.0, D.red);
        SOV_LAUNCH_CHECK();
        if ((rc = read_partials(D, 5, mx, v, s)) != SOV_OK) return rc;
        const double min_x = v[0] > -1e299 ? -v[0] : INFINITY, min_s = v[1] > -1e299 ? -v[1] : INFINITY;
        const double sx = std::max(-1.5 * min_x, 0.0), ss = std::max(-1.5 * min_s, 0.0);
        mehrotra_stats_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, D.fr, xt, D.dx, std::isfinite(sx) ? sx : 0.0,
                                                                 std::isfinite(ss) ? ss : 0.0, D.red);
        SOV_LAUNCH_CHECK();
        if ((rc = read_partials(D, 5, mx, v, s)) != SOV_OK) return rc;
        const double xs = v[2], sum_x = v[3], sum_s = v[4];
        const double bx = sum_s > 0.0 ? 0.5 * xs / sum_s : 1.0, bs = sum_x > 0.0 ? 0.5 * xs / sum_x : 1.0;
        start_mehrotra_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, xt, D.dx, std::isfinite(sx) ? sx : 0.0, bx,
                                                        std::isfinite(ss) ? ss : 0.0, bs, D.x, D.s);
        SOV_LAUNCH_CHECK();
        SOV_HIP_TRY(hipMemcpyAsync(D.y, D.dy, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
      }
    }
  }
  zero_free_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.s);
  SOV_LAUNCH_CHECK();
  int n_bounded = 0;
  for (double f : F.freev) n_bounded += f == 0.0;
  const double count = std::max(1, n_bounded);
  double bnorm = 1.0, cnorm = 1.0;
  for (double v : F.b) bnorm = std::max(bnorm, std::abs(v));
  for (double v : F.c) cnorm = std::max(cnorm, std::abs(v));
  const double tau = o.fraction_to_boundary;
  for (int it = 0; it < o.max_iterations; ++it) {
    // residuals and the stopping test (:2950-2989)
    if (m > 0) {
      if ((rc = sov::spmv_launch(D.A, false, D.x, D.ax, s)) != SOV_OK) return rc;
      if ((rc = sov::spmv_launch(D.A, true, D.y, D.aty, s)) != SOV_OK) return rc;
    } else {
      fill_kernel<<<blocks(n), 256, 0, s>>>(D.aty, 0.0, n);
    }
    if ((rc = sov::spmv_launch(D.Q, false, D.x, D.qx, s)) != SOV_OK) return rc;
    resid_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, m, D.b, D.ax, D.qx, D.c, D.aty, D.s, D.x, D.rp, D.rd, D.red);
    SOV_LAUNCH_CHECK();
    const bool mx6[6] = {true, true, true, false, false, false};
    double v[6];
    if ((rc = read_partials(D, 6, mx6, v, s)) != SOV_OK) return rc;
    const double mu = v[3] / count;
    const double p_res = v[0] / bnorm;
    const double d_res = v[1] / std::max(1.0, cnorm + v[2]);
    const double gap = mu * count / (1.0 + std::abs(v[4] + 0.5 * v[5] + F.offset));
    A.iterations = it + 1;
    A.gap = gap;
    A.pres = p_res;
    A.dres = d_res;
    if (p_res < o.feasibility_tol && d_res < o.feasibility_tol && gap < o.optimality_tol) {
      A.status = 0;
      return SOV_OK;
    }
    rc = factor_kkt(D, true, s);
    if (rc < 0) return rc;
    if (rc != SOV_OK) {  // "QP-IPM Newton factorization failed."
      A.status = 7;
      A.iterations = it;
      return SOV_OK;
    }
    // affine predictor
    rxs_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.x, D.s, nullptr, nullptr, 0.0, 0, D.rxs);
    r1_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.rd, D.rxs, D.x, D.r1);
    SOV_LAUNCH_CHECK();
    if ((rc = kkt_solve(D, D.r1, D.rp, D.dxa, D.dya, refine, s)) != SOV_OK) return rc;
    ds_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.rxs, D.s, D.dxa, D.x, D.dsa);
    step_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, D.fr, D.x, D.dxa, D.s, D.dsa, D.red);
    SOV_LAUNCH_CHECK();
    const bool mx2[2] = {true, true};
    double st[2];
    if ((rc = read_partials(D, 2, mx2, st, s)) != SOV_OK) return rc;
    const double ap_aff = st[0] < 0.0 ? std::min(1.0, -st[0]) : 1.0;
    const double ad_aff = st[1] < 0.0 ? std::min(1.0, -st[1]) : 1.0;
    muaff_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, D.fr, D.x, D.dxa, ap_aff, D.s, D.dsa, ad_aff, D.red);
    SOV_LAUNCH_CHECK();
    const bool mx1[1] = {false};
    double mua = 0.0;
    if ((rc = read_partials(D, 1, mx1, &mua, s)) != SOV_OK) return rc;
    mua /= count;
    const double sigma = mu > 0.0 ? std::min(1.0, std::pow(mua / mu, 3.0)) : 0.0;
    // corrector
    rxs_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.x, D.s, D.dxa, D.dsa, sigma * mu, 1, D.rxs);
    r1_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.rd, D.rxs, D.x, D.r1);
    SOV_LAUNCH_CHECK();
    if ((rc = kkt_solve(D, D.r1, D.rp, D.dx, D.dy, refine, s)) != SOV_OK) return rc;
    ds_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.rxs, D.s, D.dx, D.x, D.ds);
    step_kernel<<<kRedBlocks, kRedThreads, 0, s>>>(n, D.fr, D.x, D.dx, D.s, D.ds, D.red);
    SOV_LAUNCH_CHECK();
    if ((rc = read_partials(D, 2, mx2, st, s)) != SOV_OK) return rc;
    double ap = std::min(1.0, tau * (st[0] < 0.0 ? std::min(1.0, -st[0]) : 1.0));
    double ad = std::min(1.0, tau * (st[1] < 0.0 ? std::min(1.0, -st[1]) : 1.0));
    if (F.has_q && o.common_step) ap = ad = std::min(ap, ad);
    update_kernel<<<blocks(n), 256, 0, s>>>(n, D.fr, D.x, D.dx, ap, D.s, D.ds, ad);
    if (m > 0) axpy_kernel<<<blocks(m), 256, 0, s>>>(m, D.y, ad, D.dy, D.y);
    SOV_LAUNCH_CHECK();
  }
  A.status = 7;  // "QP-IPM iteration limit reached without convergence."
  return SOV_OK;
}

}  // namespace

extern "C" {

void sov_qp_ipm_default_options(sov_qp_ipm_options* o) {
  if (!o) return;
  o->max_iterations = 100;
  o->iterative_refinement = 1;
  o->feasibility_tol = 1e-8;
  o->optimality_tol = 1e-8;
  o->fraction_to_boundary = 0.999;
  o->common_step = 1;
  o->retry_mehrotra_start = 1;
}

int sov_qp_ipm_solve(const sov_qp_model* md, const sov_qp_ipm_options* opt, sov_qp_ipm_result* res,
                     double* primal, double* row_dual, double* var_dual) {
  if (!md || !res || md->n_vars <= 0 || md->n_rows < 0 || (md->n_rows > 0 && (!md->row_sense || !md->rhs)) ||
      !md->col_ptr || !md->obj || !md->lb || !md->ub || (md->q_nnz > 0 && (!md->q_row || !md->q_col || !md->q_val)))
    return SOV_ERR_ARG;
  *res = sov_qp_ipm_result{};
  res->status = 7;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  const auto t0 = std::chrono::steady_clock::now();
  sov_qp_ipm_options o;
  sov_qp_ipm_default_options(&o);
  if (opt) o = *opt;
  QpForm F;
  int rc = build_form(*md, F);
  if (rc != SOV_OK) return rc;
  res->n_form = F.n;
  res->m_form = F.m;
  hipStream_t s = sov::stream();
  auto D = std::make_unique<Dev>();
  if ((rc = setup(*D, F, s)) != SOV_OK) return rc;
  Attempt first;
  if ((rc = run_ipm(*D, F, o, false, first, s)) != SOV_OK) return rc;
  Attempt used = first;
  if (first.status != 0 && o.retry_mehrotra_start) {
    Attempt second;
    if ((rc = run_ipm(*D, F, o, true, second, s)) != SOV_OK) return rc;
    if (second.status == 0) {
      second.iterations += first.iterations;
      used = second;
      res->mehrotra_start = 1;
    } else {
      // the reference reports the first attempt; its iterate is overwritten here, so run it again
      if ((rc = run_ipm(*D, F, o, false, first, s)) != SOV_OK) return rc;
      used = first;
    }
  }
  // the point in the model's variables, the objective on the model
  const int n = F.n, m = F.m, ns = F.ns;
  std::vector<double> x(static_cast<std::size_t>(n)), sv(static_cast<std::size_t>(n)), y(static_cast<std::size_t>(std::max(m, 1)));
  SOV_HIP_TRY(hipMemcpyAsync(x.data(), D->x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipMemcpyAsync(sv.data(), D->s, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (m > 0) SOV_HIP_TRY(hipMemcpyAsync(y.data(), D->y, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  std::vector<double> xm(static_cast<std::size_t>(ns));
  for (int j = 0; j < ns; ++j) xm[j] = F.shift[j] + x[j];
  double obj = md->obj_constant;
  for (int j = 0; j < ns; ++j) obj += md->obj[j] * xm[j];
  for (int64_t t = 0; t < md->q_nnz; ++t) obj += 0.5 * md->q_val[t] * xm[md->q_row[t]] * xm[md->q_col[t]];
  if (primal)
    for (int j = 0; j < ns; ++j) primal[j] = xm[j];
  if (var_dual)
    for (int j = 0; j < ns; ++j) var_dual[j] = sv[j];
  if (row_dual)
    for (int i = 0; i < m; ++i) row_dual[i] = y[i];
  res->status = used.status;
  res->iterations = used.iterations;
  res->objective = obj;
  res->duality_gap = used.gap;
  res->primal_residual = used.pres;
  res->dual_residual = used.dres;
  res->factorizations = D->factorizations;
  res->refinement_steps = D->refinements;
  res->regularization = D->reg;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  sov::count_operation();
  return SOV_OK;
}

}  // extern "C"
