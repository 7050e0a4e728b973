// Normal-equation assembly M = A diag(d) A^T + 1e-12 I, its factorization
// and solves, and the C ABI of the NormalEquations interface
// (solver/lp/src/interior_point.cpp:129-212).
//
// Assembly (reference: build_normal_eq, interior_point.cpp:74-101, and
// build_normal_eq_pattern/values, solver/numerical/src/sparse_symmetric.cpp:
// 39-169): for sparse A a plan built once lists, for every lower-triangle
// position of M, the (A entry, A entry, column) triples that feed it in the
// reference's column order; one thread per position sums them, so the
// assembly is deterministic and needs no atomics. When that plan would be
// larger than M itself (dense rows) the assembly is an MFMA SYRK of
// W = A diag(sqrt d) instead.

#include "mfma_gemm.hpp"
#include "normal_eq.hpp"
#include "spmv_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

namespace sov {
namespace {

constexpr int kRed = 256;     // threads of reduction kernels
constexpr int kRedBlocks = 256;  // partial sums per reduction

template <typename T>
__global__ __launch_bounds__(256) void fill_lower_kernel(T* __restrict__ M, int m) {
  const int j = blockIdx.x;
  T* col = M + static_cast<long long>(j) * m;
  for (int i = j + threadIdx.x; i < m; i += blockDim.x) col[i] = T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void assemble_plan_kernel(long long entries, const int32_t* __restrict__ eptr,
                                                            const int32_t* __restrict__ ta,
                                                            const int32_t* __restrict__ tb,
                                                            const int32_t* __restrict__ tj,
                                                            const long long* __restrict__ eoff,
                                                            const int32_t* __restrict__ ediag,
                                                            const double* __restrict__ val,
                                                            const double* __restrict__ d, T* __restrict__ M,
                                                            double* __restrict__ diag0, double reg) {
  const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  double acc = 0.0;
  for (int t = eptr[e]; t < eptr[e + 1]; ++t) acc = fma(val[ta[t]] * d[tj[t]], val[tb[t]], acc);
  const int drow = ediag[e];
  if (drow >= 0) {
    acc += reg;
    diag0[drow] = acc;
  }
  M[eoff[e]] = static_cast<T>(acc);
}

template <typename T>
__global__ __launch_bounds__(256) void scale_sqrt_kernel(const double* __restrict__ Ad, const double* __restrict__ d,
                                                         T* __restrict__ W, int m, int n) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(m) * n) return;
  const int j = static_cast<int>(idx / m);
  W[idx] = static_cast<T>(Ad[idx] * sqrt(fmax(d[j], 0.0)));
}

template <typename T>
__global__ void diag_fix_kernel(T* __restrict__ M, int m, double reg, double* __restrict__ diag0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const long long o = i + static_cast<long long>(i) * m;
  const double v = static_cast<double>(M[o]) + reg;
  M[o] = static_cast<T>(v);
  diag0[i] = v;
}

// t_j = d_j (A^T p)_j, one group of G lanes per column.
template <int G>
__global__ __launch_bounds__(256) void at_scaled_kernel(int n, const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ p, const double* __restrict__ d,
                                                        double* __restrict__ t) {
  const int sub = threadIdx.x & (G - 1);
  const int j = (blockIdx.x * 256 + threadIdx.x) / G;
  const double v = csr_row_dot<G>(ptr, idx, val, p, j, n, sub);
  if (sub == 0 && j < n) t[j] = d[j] * v;
}

// q_i = (A t)_i + reg p_i.
template <int G>
__global__ __launch_bounds__(256) void a_plus_reg_kernel(int m, const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ idx,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ t, const double* __restrict__ p,
                                                         double reg, double* __restrict__ q) {
  const int sub = threadIdx.x & (G - 1);
  const int i = (blockIdx.x * 256 + threadIdx.x) / G;
  const double v = csr_row_dot<G>(ptr, idx, val, t, i, m, sub);
  if (sub == 0 && i < m) q[i] = v + reg * p[i];
}

// r = b - q, partial max |r| and max |b|.
__global__ __launch_bounds__(kRed) void residual_kernel(int m, const double* __restrict__ b,
                                                        const double* __restrict__ q, double* __restrict__ r,
                                                        double* __restrict__ partial) {
  __shared__ double sh[kRed / 64];
  double rmax = 0.0, bmax = 0.0;
  for (int i = blockIdx.x * kRed + threadIdx.x; i < m; i += gridDim.x * kRed) {
    const double v = b[i] - q[i];
    r[i] = v;
    rmax = fmax(rmax, fabs(v));
    bmax = fmax(bmax, fabs(b[i]));
  }
  rmax = block_reduce<1>(rmax, sh);
  bmax = block_reduce<1>(bmax, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = rmax;
    partial[gridDim.x + blockIdx.x] = bmax;
  }
}

__global__ void residual_finish_kernel(int nparts, const double* __restrict__ partial, PcgScalars* sc) {
  __shared__ double sh[kRed / 64];
  double r = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    r = fmax(r, partial[k]);
    b = fmax(b, partial[nparts + k]);
  }
  r = block_reduce<1>(r, sh);
  b = block_reduce<1>(b, sh);
  if (threadIdx.x == 0) {
    sc->rnorm = r;
    sc->bnorm = b;
  }
}

__global__ __launch_bounds__(kRed) void dot_kernel(int m, const double* __restrict__ x, const double* __restrict__ y,
                                                   double* __restrict__ partial) {
  __shared__ double sh[kRed / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kRed + threadIdx.x; i < m; i += gridDim.x * kRed) s = fma(x[i], y[i], s);
  s = block_reduce<0>(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// which: 0 -> rz (initial), 1 -> pq and alpha = rz / pq, 2 -> new rz and beta.
__global__ void dot_finish_kernel(int nparts, const double* __restrict__ partial, PcgScalars* sc, int which) {
  __shared__ double sh[kRed / 64];
  double s = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) s += partial[k];
  s = block_reduce<0>(s, sh);
  if (threadIdx.x == 0) {
    if (which == 0) {
      sc->rz = s;
    } else if (which == 1) {
      sc->pq = s;
      sc->alpha = s != 0.0 ? sc->rz / s : 0.0;
    } else {
      sc->beta = sc->rz != 0.0 ? s / sc->rz : 0.0;
      sc->rz = s;
    }
  }
}

// x += alpha p; r -= alpha q; partial max |r|.
__global__ __launch_bounds__(kRed) void pcg_step_kernel(int m, const PcgScalars* __restrict__ sc,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        double* __restrict__ x, double* __restrict__ r,
                                                        double* __restrict__ partial) {
  __shared__ double sh[kRed / 64];
  const double alpha = sc->alpha;
  double rmax = 0.0;
  for (int i = blockIdx.x * kRed + threadIdx.x; i < m; i += gridDim.x * kRed) {
    x[i] = fma(alpha, p[i], x[i]);
    const double v = fma(-alpha, q[i], r[i]);
    r[i] = v;
    rmax = fmax(rmax, fabs(v));
  }
  rmax = block_reduce<1>(rmax, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = rmax;
}

__global__ void max_finish_kernel(int nparts, const double* __restrict__ partial, PcgScalars* sc) {
  __shared__ double sh[kRed / 64];
  double r = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) r = fmax(r, partial[k]);
  r = block_reduce<1>(r, sh);
  if (threadIdx.x == 0) sc->rnorm = r;
}

// p = z + beta p; p = z on the first pass (p is not read then: it may hold
// stale bits, and 0 * NaN would not be 0).
__global__ __launch_bounds__(kRed) void direction_kernel(int m, const PcgScalars* __restrict__ sc, int first,
                                                         const double* __restrict__ z, double* __restrict__ p) {
  const double beta = first ? 0.0 : sc->beta;
  for (int i = blockIdx.x * kRed + threadIdx.x; i < m; i += gridDim.x * kRed) p[i] = first ? z[i] : fma(beta, p[i], z[i]);
}

unsigned red_blocks(int m) { return std::max(1u, std::min<unsigned>(kRedBlocks, blocks_for(m, kRed))); }

template <typename T>
void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

NormalSolver::~NormalSolver() {
  free_dev(eptr_);
  free_dev(ta_);
  free_dev(tb_);
  free_dev(tj_);
  free_dev(eoff_);
  free_dev(ediag_);
  free_dev(slot_);
  free_dev(ediag_perm_);
  free_dev(dense_a_);
  if (wbuf_) (void)hipFree(wbuf_);
  free_dev(M64_);
  free_dev(M32_);
  free_dev(diag0_);
  free_dev(r_);
  free_dev(z_);
  free_dev(p_);
  free_dev(q_);
  free_dev(t_);
  free_dev(b_);
  free_dev(partial_);
  free_dev(sc_);
  if (sc_host_) (void)hipHostFree(sc_host_);
}

int NormalSolver::init(sov_mat* a, int path) {
  pcg_graphs_.clear();
  A_ = a;
  m_ = a->nrows;
  n_ = a->ncols;
  if (path == SOV_NE_AUTO) {
    const char* e = std::getenv("SOVEREIGN_NE_PATH");
    if (e && std::strcmp(e, "dense") == 0) path = SOV_NE_DENSE;
    if (e && std::strcmp(e, "sparse") == 0) path = SOV_NE_SPARSE;
  }
  const std::vector<int32_t>& cp = a->h_cptr;
  const std::vector<int32_t>& ri = a->h_crow;
  // Plan size: pairs (p, q) of a column with row(p) >= row(q).
  long long triples = 0;
  for (int j = 0; j < n_; ++j) {
    const long long k = cp[j + 1] - cp[j];
    triples += k * (k + 1) / 2;
  }
  const double dense_entries = static_cast<double>(m_) * m_;
  if (static_cast<double>(triples) > 2.0 * dense_entries && m_ > 0 && path != SOV_NE_SPARSE) {
    std::vector<double> Ad(static_cast<std::size_t>(m_) * n_, 0.0);
    for (int j = 0; j < n_; ++j)
      for (int p = cp[j]; p < cp[j + 1]; ++p) Ad[ri[p] + static_cast<std::size_t>(j) * m_] += a->h_cval[p];
    SOV_HIP_TRY(hipMalloc(&dense_a_, Ad.size() * sizeof(double)));
    SOV_HIP_TRY(hipMemcpy(dense_a_, Ad.data(), Ad.size() * sizeof(double), hipMemcpyHostToDevice));
    SOV_HIP_TRY(hipMalloc(&wbuf_, Ad.size() * sizeof(double)));
  } else {
    struct Triple {
      long long key;
      int32_t p, q, j;
    };
    std::vector<Triple> tri;
    tri.reserve(static_cast<std::size_t>(triples) + static_cast<std::size_t>(m_));
    for (int j = 0; j < n_; ++j) {
      for (int p = cp[j]; p < cp[j + 1]; ++p) {
        for (int q = cp[j]; q < cp[j + 1]; ++q) {
          if (ri[p] < ri[q]) continue;  // equal rows: both orders, as the reference sums them
          tri.push_back({static_cast<long long>(ri[q]) * m_ + ri[p], p, q, j});
        }
      }
    }
    // Every diagonal position takes the regularization, even with no entries.
    std::vector<char> has_diag(static_cast<std::size_t>(m_), 0);
    for (const Triple& t : tri)
      if (t.key % m_ == t.key / m_) has_diag[t.key / m_] = 1;
    for (int i = 0; i < m_; ++i)
      if (!has_diag[i]) tri.push_back({static_cast<long long>(i) * m_ + i, -1, -1, -1});
    std::stable_sort(tri.begin(), tri.end(), [](const Triple& x, const Triple& y) { return x.key < y.key; });
    std::vector<int32_t> eptr(1, 0), ta, tb, tj, ediag;
    std::vector<long long> eoff;
    for (std::size_t t = 0; t < tri.size(); ++t) {
      if (t == 0 || tri[t].key != tri[t - 1].key) {
        if (t > 0) eptr.push_back(static_cast<int32_t>(ta.size()));
        const long long key = tri[t].key;
        const int col = static_cast<int>(key / m_), row = static_cast<int>(key % m_);
        eoff.push_back(row + static_cast<long long>(col) * m_);
        ediag.push_back(row == col ? row : -1);
      }
      if (tri[t].p >= 0) {
        ta.push_back(tri[t].p);
        tb.push_back(tri[t].q);
        tj.push_back(tri[t].j);
      }
    }
    eptr.push_back(static_cast<int32_t>(ta.size()));
    entries_ = static_cast<long long>(eoff.size());
    triples_ = static_cast<long long>(ta.size());
    auto up = [](auto** dst, const auto& v) -> bool {
      using E = typename std::decay<decltype(v)>::type::value_type;
      if (hipMalloc(reinterpret_cast<void**>(dst), std::max<std::size_t>(v.size(), 1) * sizeof(E)) != hipSuccess)
        return false;
      return v.empty() || hipMemcpy(*dst, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(&eptr_, eptr) || !up(&ta_, ta) || !up(&tb_, tb) || !up(&tj_, tj) || !up(&eoff_, eoff) ||
        !up(&ediag_, ediag)) {
      set_error("normal equations: plan upload failed");
      return SOV_ERR_NOMEM;
    }
    if (path != SOV_NE_DENSE && m_ > 0) {
      const int rc = init_sparse(eoff, path);
      if (rc < 0) return rc;
    }
  }
  if (!sp_ && path == SOV_NE_SPARSE && m_ > 0) {
    set_error("normal equations: the sparse factorization was asked for and could not be set up");
    return SOV_ERR_ARG;
  }
  if (sp_) {  // FP64 only: no dense matrix, no dense factor workspace
    const std::size_t vec = static_cast<std::size_t>(std::max(std::max(m_, n_), 1));
    for (double** v : {&r_, &z_, &p_, &q_, &t_, &b_}) SOV_HIP_TRY(hipMalloc(v, vec * sizeof(double)));
    SOV_HIP_TRY(hipMalloc(&partial_, 2 * kRedBlocks * sizeof(double)));
    SOV_HIP_TRY(hipMalloc(&sc_, sizeof(PcgScalars)));
    SOV_HIP_TRY(hipHostMalloc(&sc_host_, sizeof(PcgScalars)));
    return SOV_OK;
  }
  const std::size_t mm = static_cast<std::size_t>(m_) * m_;
  SOV_HIP_TRY(hipMalloc(&M64_, std::max<std::size_t>(mm, 1) * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&diag0_, std::max(m_, 1) * sizeof(double)));
  const std::size_t vec = static_cast<std::size_t>(std::max(std::max(m_, n_), 1));
  for (double** v : {&r_, &z_, &p_, &q_, &t_, &b_}) SOV_HIP_TRY(hipMalloc(v, vec * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&partial_, 2 * kRedBlocks * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&sc_, sizeof(PcgScalars)));
  SOV_HIP_TRY(hipHostMalloc(&sc_host_, sizeof(PcgScalars)));
  int rc = ch64.reserve(std::max(m_, 1));
  if (rc != SOV_OK) return rc;
  return SOV_OK;
}

// Sparse factorization set-up from the plan's entries (their dense offsets
// row + col m, row >= col, give the lower pattern). AUTO keeps the dense
// path when the pattern is dense-like (mean degree above 64), or when the
// ordering predicts a factor above a tenth of the dense triangle and the
// dense matrix fits the device; returns SOV_OK either way (sp_ set or not).
int NormalSolver::init_sparse(const std::vector<long long>& eoff, int path) {
  const long long entries = static_cast<long long>(eoff.size());
  const double mean_degree = 2.0 * static_cast<double>(entries - m_) / m_;
  std::size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  const double dense_bytes = 1.6 * 8.0 * static_cast<double>(m_) * m_;  // FP64 + FP32 matrices, factor workspace
  const bool dense_fits = free_b == 0 || dense_bytes < 0.8 * static_cast<double>(free_b);
  const double tri = 0.5 * static_cast<double>(m_) * (m_ + 1);
  int64_t limit = 0;
  if (path == SOV_NE_AUTO && dense_fits) {
    if (mean_degree > 64.0 || m_ < 200) return SOV_OK;
    limit = static_cast<int64_t>(0.1 * tri);
  }
  std::vector<int32_t> lcp(static_cast<std::size_t>(m_) + 1, 0), lri(static_cast<std::size_t>(entries));
  for (long long e = 0; e < entries; ++e) ++lcp[static_cast<int>(eoff[e] / m_) + 1];  // column = offset / m
  for (int j = 0; j < m_; ++j) lcp[j + 1] += lcp[j];
  {
    std::vector<int32_t> next(lcp.begin(), lcp.end() - 1);
    for (long long e = 0; e < entries; ++e) lri[next[static_cast<int>(eoff[e] / m_)]++] = static_cast<int32_t>(eoff[e] % m_);
  }
  auto sp = std::make_unique<SparseCholesky>();
  const int rc = sp->sym.analyse(m_, lcp, lri, limit);
  if (rc == SOV_SINGULAR) return SOV_OK;  // the fill limit: dense
  if (rc != SOV_OK) {
    set_error("normal equations: sparse symbolic analysis failed");
    return rc;
  }
  if (path == SOV_NE_AUTO && dense_fits && static_cast<double>(sp->sym.nnz_l) > 0.1 * tri) return SOV_OK;
  std::vector<long long> slot(static_cast<std::size_t>(entries));
  std::vector<int32_t> edp(static_cast<std::size_t>(entries));
  for (long long e = 0; e < entries; ++e) {
    const int col = static_cast<int>(eoff[e] / m_), row = static_cast<int>(eoff[e] % m_);
    slot[e] = sp->sym.slot(row, col);
    if (slot[e] < 0) {
      set_error("normal equations: an entry outside the sparse factor's structure");
      return SOV_ERR_ARG;
    }
    edp[e] = row == col ? sp->sym.iperm[row] : -1;
  }
  int r2 = sp->upload();
  if (r2 != SOV_OK) return r2;
  SOV_HIP_TRY(hipMalloc(&slot_, std::max<std::size_t>(slot.size(), 1) * sizeof(long long)));
  SOV_HIP_TRY(hipMalloc(&ediag_perm_, std::max<std::size_t>(edp.size(), 1) * sizeof(int32_t)));
  if (entries > 0) {
    SOV_HIP_TRY(hipMemcpy(slot_, slot.data(), slot.size() * sizeof(long long), hipMemcpyHostToDevice));
    SOV_HIP_TRY(hipMemcpy(ediag_perm_, edp.data(), edp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  sp_ = std::move(sp);
  return SOV_OK;
}

template <typename T>
int NormalSolver::assemble(const double* d_dev, T* M, hipStream_t s) {
  if (m_ == 0) return SOV_OK;
  TimedPair* tp = timed_begin(s, kTimeAssemble);
  fill_lower_kernel<T><<<m_, 256, 0, s>>>(M, m_);
  SOV_LAUNCH_CHECK();
  if (dense_a_) {
    T* W = static_cast<T*>(wbuf_);
    const long long total = static_cast<long long>(m_) * n_;
    scale_sqrt_kernel<T><<<blocks_for(total, 256), 256, 0, s>>>(dense_a_, d_dev, W, m_, n_);
    SOV_LAUNCH_CHECK();
    SOV_HIP_TRY((gemm_update<T, false, true>(m_, m_, n_, W, m_, W, m_, M, m_, T(1), s)));
    diag_fix_kernel<T><<<blocks_for(m_, 256), 256, 0, s>>>(M, m_, kNormalReg, diag0_);
    SOV_LAUNCH_CHECK();
  } else {
    assemble_plan_kernel<T><<<blocks_for(entries_, 256), 256, 0, s>>>(entries_, eptr_, ta_, tb_, tj_, eoff_, ediag_,
                                                                      A_->c_val, d_dev, M, diag0_, kNormalReg);
    SOV_LAUNCH_CHECK();
  }
  timed_end(tp, s, dense_a_ ? 4 : 2, 0.0);
  return SOV_OK;
}

int NormalSolver::factor(const double* d_dev, int precision, hipStream_t s) {
  d_ = d_dev;
  prec_ = precision;
  if (m_ == 0) return SOV_OK;
  if (sp_) {
    prec_ = SOV_PREC_FP64;
    TimedPair* tp = timed_begin(s, kTimeAssemble);
    int rc = sp_->zero_fronts(s);
    if (rc != SOV_OK) return rc;
    assemble_plan_kernel<double><<<blocks_for(entries_, 256), 256, 0, s>>>(
        entries_, eptr_, ta_, tb_, tj_, slot_, ediag_perm_, A_->c_val, d_dev, sp_->fronts(), sp_->diag0(), kNormalReg);
    SOV_LAUNCH_CHECK();
    timed_end(tp, s, 2, 0.0);
    return sp_->factor(s);
  }
  if (precision == SOV_PREC_FP32) {
    if (!M32_) SOV_HIP_TRY(hipMalloc(&M32_, static_cast<std::size_t>(m_) * m_ * sizeof(float)));
    int rc = ch32.reserve(m_);
    if (rc != SOV_OK) return rc;
    rc = assemble<float>(d_dev, M32_, s);
    if (rc != SOV_OK) return rc;
    return ch32.factor(M32_, m_, diag0_, s);
  }
  int rc = assemble<double>(d_dev, M64_, s);
  if (rc != SOV_OK) return rc;
  return ch64.factor(M64_, m_, diag0_, s);
}

int NormalSolver::breakdown(hipStream_t s) {
  if (m_ == 0) return SOV_OK;
  FactorFlags f{};
  FactorFlags* src = sp_ ? sp_->flags() : prec_ == SOV_PREC_FP32 ? ch32.flags() : ch64.flags();
  SOV_HIP_TRY(hipMemcpyAsync(&f, src, sizeof(f), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  replaced_ = f.replaced;
  return f.status != 0 ? SOV_SINGULAR : SOV_OK;
}

int NormalSolver::matvec(const double* p, double* q, hipStream_t s) {
  const int gc = A_->group_cols, gr = A_->group_rows;
  if (n_ > 0) {
    SOV_DISPATCH_GROUP(gc, {
      at_scaled_kernel<G><<<blocks_for(static_cast<long long>(n_) * G, 256), 256, 0, s>>>(
          n_, A_->c_ptr, A_->c_row, A_->c_val, p, d_, t_);
    });
    SOV_LAUNCH_CHECK();
  }
  SOV_DISPATCH_GROUP(gr, {
    a_plus_reg_kernel<G><<<blocks_for(static_cast<long long>(m_) * G, 256), 256, 0, s>>>(
        m_, A_->r_ptr, A_->r_col, A_->r_val, t_, p, kNormalReg, q);
  });
  SOV_LAUNCH_CHECK();
  return SOV_OK;
}

int NormalSolver::solve(double* x, hipStream_t s, double tol, int max_iter) {
  last_iters_ = 0;
  last_res_ = 0.0;
  if (m_ == 0) return SOV_OK;
  if (sp_) return sp_->solve(x, s);
  if (prec_ != SOV_PREC_FP32) return ch64.solve(M64_, m_, x, nullptr, s);
  // Preconditioned CG on M x = b, preconditioner = the FP32 Cholesky factor.
  const unsigned rb = red_blocks(m_);
  const std::size_t bytes = static_cast<std::size_t>(m_) * sizeof(double);
  SOV_HIP_TRY(hipMemcpyAsync(b_, x, bytes, hipMemcpyDeviceToDevice, s));
  int rc = ch32.solve(M32_, m_, x, nullptr, s);  // x0 = P b
  if (rc != SOV_OK) return rc;
  auto check = [&](double* out_ratio) -> int {
    SOV_HIP_TRY(hipMemcpyAsync(sc_host_, sc_, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    SOV_HIP_TRY(hipStreamSynchronize(s));
    const double bn = sc_host_->bnorm;
    *out_ratio = bn > 0.0 ? sc_host_->rnorm / bn : sc_host_->rnorm;
    return SOV_OK;
  };
  if ((rc = matvec(x, q_, s)) != SOV_OK) return rc;
  residual_kernel<<<rb, kRed, 0, s>>>(m_, b_, q_, r_, partial_);
  residual_finish_kernel<<<1, kRed, 0, s>>>(rb, partial_, sc_);
  SOV_LAUNCH_CHECK();
  double ratio = 0.0;
  if ((rc = check(&ratio)) != SOV_OK) return rc;
  last_res_ = ratio;
  if (!std::isfinite(ratio)) return SOV_SINGULAR;
  if (ratio <= tol) return SOV_OK;
  SOV_HIP_TRY(hipMemcpyAsync(z_, r_, bytes, hipMemcpyDeviceToDevice, s));
  if ((rc = ch32.solve(M32_, m_, z_, nullptr, s)) != SOV_OK) return rc;
  direction_kernel<<<rb, kRed, 0, s>>>(m_, sc_, 1, z_, p_);
  dot_kernel<<<rb, kRed, 0, s>>>(m_, r_, z_, partial_);
  dot_finish_kernel<<<1, kRed, 0, s>>>(rb, partial_, sc_, 0);
  SOV_LAUNCH_CHECK();
  // The step of an iteration: q = M p, alpha, x += alpha p, r -= alpha q,
  // and the residual norm the host reads.
  auto step = [&]() -> int {
    int r2 = matvec(p_, q_, s);
    if (r2 != SOV_OK) return r2;
    dot_kernel<<<rb, kRed, 0, s>>>(m_, p_, q_, partial_);
    dot_finish_kernel<<<1, kRed, 0, s>>>(rb, partial_, sc_, 1);
    pcg_step_kernel<<<rb, kRed, 0, s>>>(m_, sc_, p_, q_, x, r_, partial_);
    max_finish_kernel<<<1, kRed, 0, s>>>(rb, partial_, sc_);
    SOV_LAUNCH_CHECK();
    return SOV_OK;
  };
  if ((rc = step()) != SOV_OK) return rc;
  for (int it = 1; it <= max_iter; ++it) {
    if ((rc = check(&ratio)) != SOV_OK) return rc;
    last_iters_ = it;
    last_res_ = ratio;
    if (!std::isfinite(ratio)) break;
    if (ratio <= tol) return SOV_OK;
    if (it == max_iter) break;
    // Between two checks: z = P r, beta, p = z + beta p, then the next
    // step; one graph launch per iteration (the solve is captured inline).
    rc = pcg_graphs_.run(x, M32_, d_, m_, s, [&]() -> int {  // d_: the product M p reads it
      SOV_HIP_TRY(hipMemcpyAsync(z_, r_, bytes, hipMemcpyDeviceToDevice, s));
      int r2 = ch32.solve(M32_, m_, z_, nullptr, s);
      if (r2 != SOV_OK) return r2;
      dot_kernel<<<rb, kRed, 0, s>>>(m_, r_, z_, partial_);
      dot_finish_kernel<<<1, kRed, 0, s>>>(rb, partial_, sc_, 2);
      direction_kernel<<<rb, kRed, 0, s>>>(m_, sc_, 0, z_, p_);
      SOV_LAUNCH_CHECK();
      return step();
    });
    if (rc != SOV_OK) return rc;
  }
  return SOV_SINGULAR;
}

}  // namespace sov

// ---------------------------------------------------------------------------
// C ABI: NormalEquations{factor(d); solve(rhs)} on host vectors.
// ---------------------------------------------------------------------------
struct sov_ne {
  sov_mat* A = nullptr;
  sov::NormalSolver solver;
  double* d_dev = nullptr;
  double* x_dev = nullptr;
  bool factored = false;
  double last_residual = 0.0;
  ~sov_ne() {
    if (d_dev) (void)hipFree(d_dev);
    if (x_dev) (void)hipFree(x_dev);
  }
};

extern "C" {

int sov_ne_create(sov_mat* a, sov_ne** out) { return sov_ne_create_path(a, SOV_NE_AUTO, out); }

int sov_ne_create_path(sov_mat* a, int32_t path, sov_ne** out) {
  if (!a || !out || path < SOV_NE_AUTO || path > SOV_NE_SPARSE) return SOV_ERR_ARG;
  *out = nullptr;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  auto* h = new sov_ne;
  h->A = a;
  int rc = h->solver.init(a, path);
  if (rc == SOV_OK && hipMalloc(&h->d_dev, std::max(a->ncols, 1) * sizeof(double)) != hipSuccess) rc = SOV_ERR_NOMEM;
  if (rc == SOV_OK && hipMalloc(&h->x_dev, std::max(a->nrows, 1) * sizeof(double)) != hipSuccess) rc = SOV_ERR_NOMEM;
  if (rc != SOV_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return SOV_OK;
}

int sov_ne_factor(sov_ne* h, const double* d_host, int32_t precision) {
  if (!h || !d_host) return SOV_ERR_ARG;
  if (precision != SOV_PREC_FP64 && precision != SOV_PREC_FP32) return SOV_ERR_ARG;
  hipStream_t s = sov::stream();
  h->factored = false;
  SOV_HIP_TRY(hipMemcpyAsync(h->d_dev, d_host, h->A->ncols * sizeof(double), hipMemcpyHostToDevice, s));
  int rc = h->solver.factor(h->d_dev, precision, s);
  if (rc != SOV_OK) return rc;
  rc = h->solver.breakdown(s);
  if (rc == SOV_OK) {
    h->factored = true;
    sov::count_operation();
    sov::count_factorization();
  }
  return rc;
}

int sov_ne_solve(sov_ne* h, double* rhs_host) {
  if (!h || !rhs_host) return SOV_ERR_ARG;
  if (!h->factored) return SOV_SINGULAR;
  hipStream_t s = sov::stream();
  const std::size_t bytes = h->A->nrows * sizeof(double);
  SOV_HIP_TRY(hipMemcpyAsync(h->x_dev, rhs_host, bytes, hipMemcpyHostToDevice, s));
  const int rc = h->solver.solve(h->x_dev, s, 1e-12, 50);
  if (rc < 0) return rc;
  SOV_HIP_TRY(hipMemcpyAsync(rhs_host, h->x_dev, bytes, hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  h->last_residual = h->solver.last_residual();
  return SOV_OK;
}

double sov_ne_last_residual(const sov_ne* h) { return h ? h->last_residual : -1.0; }

int sov_ne_get_info(const sov_ne* h, sov_ne_info* info) {
  if (!h || !info) return SOV_ERR_ARG;
  *info = sov_ne_info{};
  const sov::SparseSymbolic* y = h->solver.sparse_symbolic();
  if (!y) return SOV_OK;
  info->sparse = 1;
  info->supernodes = static_cast<int32_t>(y->first.size());
  info->levels = static_cast<int32_t>(y->level_ptr.size()) - 1;
  info->max_front = y->max_front;
  info->nnz_l = y->nnz_l;
  info->factor_flops = y->flops;
  info->front_doubles = y->front_total;
  return SOV_OK;
}

int sov_min_degree_order(int32_t n, const int32_t* col_ptr, const int32_t* row_idx, int64_t fill_limit,
                         int32_t* perm, int32_t* iperm, int64_t* predicted_nnz) {
  if (n < 0 || !col_ptr || (n > 0 && (!row_idx || !perm || !iperm))) return SOV_ERR_ARG;
  for (int j = 0; j < n; ++j)
    if (col_ptr[j + 1] < col_ptr[j]) return SOV_ERR_ARG;
  for (int p = 0; p < col_ptr[n]; ++p)
    if (row_idx[p] < 0 || row_idx[p] >= n) return SOV_ERR_ARG;
  std::vector<int> pv, ipv;
  int64_t pred = 0;
  const bool ok = sov::min_degree_order(n, col_ptr, row_idx, fill_limit, pv, ipv, &pred);
  if (predicted_nnz) *predicted_nnz = pred;
  for (int k = 0; k < n; ++k) {
    perm[k] = pv[k];
    iperm[k] = ipv[k];
  }
  return ok ? SOV_OK : SOV_SINGULAR;
}
int32_t sov_ne_replaced_pivots(const sov_ne* h) { return h ? h->solver.replaced_pivots() : 0; }

int sov_ne_free(sov_ne* h) {
  delete h;
  return SOV_OK;
}

}  // extern "C"
