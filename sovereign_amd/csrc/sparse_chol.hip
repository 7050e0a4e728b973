// Device side of the sparse Cholesky (see sparse_chol.hpp): multifrontal
// factorization and solves, one launch per level of the supernodal tree,
// one workgroup per front. A front of order at most kLdsFront is factored in
// LDS (copied in, children's updates added, factored, written back); larger
// fronts are factored in place in HBM.
#include "sparse_chol.hpp"

#include <algorithm>
#include <cmath>

namespace sov {

namespace {

constexpr int kFrontThreads = 256;
constexpr int kFrontWaves = kFrontThreads / 64;
constexpr int kLdsFront = 128;  // 128 x 128 doubles = 128 KiB of LDS

__global__ __launch_bounds__(256) void zero_kernel(double* __restrict__ p, long long count) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0;
}

__global__ void reset_sparse_flags_kernel(FactorFlags* f) {
  f->status = 0;
  f->replaced = 0;
}

struct FrontMeta {
  const int32_t* first;
  const int32_t* ncol;
  const int32_t* fsize;
  const int64_t* fbase;
  const int64_t* vbase;
  const int32_t* child_ptr;
  const int32_t* child_idx;
  const int64_t* map_base;
  const int32_t* maps;
  const int32_t* rows;
};

// One front: add the children's update matrices, factor the leading ncol
// columns (pivot rule of the reference's LDL^T in Cholesky form), then
// F22 -= L21 L21^T, the update matrix the parent adds.
__global__ __launch_bounds__(kFrontThreads) void front_factor_kernel(const int32_t* __restrict__ level_fronts,
                                                                     FrontMeta fm, double* __restrict__ fronts,
                                                                     const double* __restrict__ diag0,
                                                                     FactorFlags* __restrict__ flags, int in_lds) {
  extern __shared__ double lds_front[];
  __shared__ double s_rp;
  const int s = level_fronts[blockIdx.x];
  const int f = fm.fsize[s], nc = fm.ncol[s], c0 = fm.first[s];
  const long long lf = f;
  double* G = fronts + fm.fbase[s];
  double* F = in_lds ? lds_front : G;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (in_lds) {
    for (long long q = tid; q < lf * lf; q += kFrontThreads) F[q] = G[q];
    __syncthreads();
  }
  for (int q = fm.child_ptr[s]; q < fm.child_ptr[s + 1]; ++q) {
    const int c = fm.child_idx[q];
    const int fc = fm.fsize[c], ncc = fm.ncol[c];
    const long long u = fc - ncc;
    const double* U = fronts + fm.fbase[c] + static_cast<long long>(ncc) * fc + ncc;  // child's F22, ld fc
    const int32_t* mp = fm.maps + fm.map_base[c];
    // lower triangle of the u x u update, column by column (distinct rows of
    // one child land in distinct rows of the parent, so no two threads add
    // to the same entry)
    for (long long q2 = tid; q2 < u * u; q2 += kFrontThreads) {
      const long long j = q2 / u, i = q2 - j * u;
      if (i >= j) F[mp[i] + static_cast<long long>(mp[j]) * lf] += U[i + j * fc];
    }
    __syncthreads();
  }
  // Leading columns, right-looking within the nc x nc block and L21.
  for (int k = 0; k < nc; ++k) {
    if (tid == 0) {
      const double dk = F[k + k * lf];
      const double thr = 1e-13 * fabs(diag0[c0 + k]);
      double piv;
      if (isfinite(dk) && dk <= thr) {
        piv = 1e64;  // the reference's replaced pivot d = 1e128
        atomicAdd(&flags->replaced, 1);
      } else if (!isfinite(dk) || dk < 1e-30) {
        piv = 1.0;
        atomicExch(&flags->status, 1);
      } else {
        piv = sqrt(dk);
      }
      F[k + k * lf] = piv;
      s_rp = 1.0 / piv;
    }
    __syncthreads();
    const double rp = s_rp;
    for (int i = k + 1 + tid; i < f; i += kFrontThreads) F[i + k * lf] *= rp;
    __syncthreads();
    for (int j = k + 1 + wv; j < nc; j += kFrontWaves) {  // columns k+1 .. nc-1 of the block
      const double ljk = F[j + k * lf];
      for (int i = j + lane; i < f; i += 64) F[i + j * lf] -= F[i + k * lf] * ljk;
    }
    __syncthreads();
  }
  // Update matrix: F22(i, j) -= sum_k L(i, k) L(j, k), nc <= j <= i < f.
  for (int j = nc + wv; j < f; j += kFrontWaves) {
    for (int i = j + lane; i < f; i += 64) {
      double acc = 0.0;
      for (int k = 0; k < nc; ++k) acc = fma(F[i + k * lf], F[j + k * lf], acc);
      F[i + j * lf] -= acc;
    }
  }
  __syncthreads();
  if (in_lds) {
    for (long long q = tid; q < lf * lf; q += kFrontThreads) G[q] = F[q];
  }
}

// Forward solve of one front: its vector (own columns from xp, zero below)
// plus the children's update vectors, then L11 y = w(0:nc) with
// w(nc:f) -= L21 y left as the front's update vector; y into xp.
__global__ __launch_bounds__(kFrontThreads) void front_forward_kernel(const int32_t* __restrict__ level_fronts,
                                                                      FrontMeta fm,
                                                                      const double* __restrict__ fronts,
                                                                      double* __restrict__ vf,
                                                                      double* __restrict__ xp) {
  const int s = level_fronts[blockIdx.x];
  const int f = fm.fsize[s], nc = fm.ncol[s], c0 = fm.first[s];
  const long long lf = f;
  const double* L = fronts + fm.fbase[s];
  double* w = vf + fm.vbase[s];
  const int tid = threadIdx.x;
  for (int p = tid; p < f; p += kFrontThreads) w[p] = p < nc ? xp[c0 + p] : 0.0;
  __syncthreads();
  for (int q = fm.child_ptr[s]; q < fm.child_ptr[s + 1]; ++q) {
    const int c = fm.child_idx[q];
    const int ncc = fm.ncol[c], u = fm.fsize[c] - ncc;
    const double* wc = vf + fm.vbase[c] + ncc;
    const int32_t* mp = fm.maps + fm.map_base[c];
    for (int t = tid; t < u; t += kFrontThreads) w[mp[t]] += wc[t];
    __syncthreads();
  }
  for (int k = 0; k < nc; ++k) {
    const double yk = w[k] / L[k + k * lf];
    for (int p = k + 1 + tid; p < f; p += kFrontThreads) w[p] -= L[p + k * lf] * yk;
    if (tid == 0) xp[c0 + k] = yk;
    __syncthreads();
  }
}

// Backward solve of one front: t_k = y_k - sum_{p >= nc} L(p, k) x(rows p)
// (ancestors' entries, final), then L11^T x = t, x into xp.
__global__ __launch_bounds__(kFrontThreads) void front_backward_kernel(const int32_t* __restrict__ level_fronts,
                                                                       FrontMeta fm,
                                                                       const double* __restrict__ fronts,
                                                                       double* __restrict__ vf,
                                                                       double* __restrict__ xp) {
  const int s = level_fronts[blockIdx.x];
  const int f = fm.fsize[s], nc = fm.ncol[s], c0 = fm.first[s];
  const long long lf = f;
  const double* L = fronts + fm.fbase[s];
  const int32_t* rows = fm.rows + fm.vbase[s];
  double* w = vf + fm.vbase[s];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int k = wv; k < nc; k += kFrontWaves) {
    double acc = 0.0;
    for (int p = nc + lane; p < f; p += 64) acc = fma(L[p + k * lf], xp[rows[p]], acc);
    acc = wave_sum(acc);
    if (lane == 0) w[k] = xp[c0 + k] - acc;
  }
  __syncthreads();
  for (int k = nc - 1; k >= 0; --k) {
    const double xk = w[k] / L[k + k * lf];
    for (int q = tid; q < k; q += kFrontThreads) w[q] -= L[k + q * lf] * xk;
    if (tid == 0) xp[c0 + k] = xk;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void permute_in_kernel(int m, const int32_t* __restrict__ perm,
                                                         const double* __restrict__ x, double* __restrict__ xp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) xp[k] = x[perm[k]];
}

__global__ __launch_bounds__(256) void permute_out_kernel(int m, const int32_t* __restrict__ perm,
                                                          const double* __restrict__ xp, double* __restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) x[perm[k]] = xp[k];
}

template <typename T>
int upload_vec(T** dst, const std::vector<T>& v) {
  if (hipMalloc(reinterpret_cast<void**>(dst), std::max<std::size_t>(v.size(), 1) * sizeof(T)) != hipSuccess) {
    set_error("sparse Cholesky: device allocation failed");
    return SOV_ERR_NOMEM;
  }
  if (!v.empty() && hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("sparse Cholesky: upload failed");
    return SOV_ERR_DEVICE;
  }
  return SOV_OK;
}

}  // namespace

SparseCholesky::~SparseCholesky() {
  for (void* p : {static_cast<void*>(fronts_), static_cast<void*>(diag0_), static_cast<void*>(vf_),
                  static_cast<void*>(xp_), static_cast<void*>(flags_),
                  static_cast<void*>(d_perm_), static_cast<void*>(d_first_), static_cast<void*>(d_ncol_),
                  static_cast<void*>(d_front_), static_cast<void*>(d_rows_), static_cast<void*>(d_child_ptr_),
                  static_cast<void*>(d_child_idx_), static_cast<void*>(d_maps_), static_cast<void*>(d_level_idx_),
                  static_cast<void*>(d_front_base_), static_cast<void*>(d_vec_base_),
                  static_cast<void*>(d_map_base_)})
    if (p) (void)hipFree(p);
}

int SparseCholesky::upload() {
  const SparseSymbolic& y = sym;
  int rc;
  std::vector<int32_t> perm32(y.perm.begin(), y.perm.end());
  if ((rc = upload_vec(&d_perm_, perm32)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_first_, y.first)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_ncol_, y.ncol)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_front_, y.front)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_rows_, y.rows)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_child_ptr_, y.child_ptr)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_child_idx_, y.child_idx)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_maps_, y.maps)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_level_idx_, y.level_idx)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_front_base_, y.front_base)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_vec_base_, y.vec_base)) != SOV_OK) return rc;
  if ((rc = upload_vec(&d_map_base_, y.map_base)) != SOV_OK) return rc;
  SOV_HIP_TRY(hipMalloc(&fronts_, std::max<int64_t>(y.front_total, 1) * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&diag0_, std::max(y.m, 1) * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&vf_, std::max<int64_t>(y.vec_total, 1) * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&xp_, std::max(y.m, 1) * sizeof(double)));
  SOV_HIP_TRY(hipMalloc(&flags_, sizeof(FactorFlags)));
  SOV_HIP_TRY(hipMemset(flags_, 0, sizeof(FactorFlags)));
  static bool lds_ok = hipFuncSetAttribute(reinterpret_cast<const void*>(front_factor_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           kLdsFront * kLdsFront * static_cast<int>(sizeof(double))) == hipSuccess;
  if (!lds_ok) {
    set_error("sparse Cholesky: could not raise the LDS limit of the front kernel");
    return SOV_ERR_DEVICE;
  }
  return SOV_OK;
}

int SparseCholesky::zero_fronts(hipStream_t s) {
  if (sym.front_total > 0) {
    zero_kernel<<<blocks_for(sym.front_total, 256), 256, 0, s>>>(fronts_, sym.front_total);
    SOV_LAUNCH_CHECK();
  }
  return SOV_OK;
}

int SparseCholesky::factor(hipStream_t s) {
  reset_sparse_flags_kernel<<<1, 1, 0, s>>>(flags_);
  SOV_LAUNCH_CHECK();
  const FrontMeta fm{d_first_, d_ncol_, d_front_, d_front_base_, d_vec_base_, d_child_ptr_,
                     d_child_idx_, d_map_base_, d_maps_, d_rows_};
  const int nlev = static_cast<int>(sym.level_ptr.size()) - 1;
  for (int l = 0; l < nlev; ++l) {
    const int b = sym.level_ptr[l], cnt = sym.level_ptr[l + 1] - b;
    int fmax = 0;
    for (int q = b; q < b + cnt; ++q) fmax = std::max(fmax, sym.front[sym.level_idx[q]]);
    const bool in_lds = fmax <= kLdsFront;
    const std::size_t lds = in_lds ? static_cast<std::size_t>(fmax) * fmax * sizeof(double) : 0;
    front_factor_kernel<<<cnt, kFrontThreads, lds, s>>>(d_level_idx_ + b, fm, fronts_, diag0_, flags_,
                                                        in_lds ? 1 : 0);
    SOV_LAUNCH_CHECK();
  }
  return SOV_OK;
}

int SparseCholesky::solve(double* x, hipStream_t s) {
  const int m = sym.m;
  if (m == 0) return SOV_OK;
  const FrontMeta fm{d_first_, d_ncol_, d_front_, d_front_base_, d_vec_base_, d_child_ptr_,
                     d_child_idx_, d_map_base_, d_maps_, d_rows_};
  permute_in_kernel<<<blocks_for(m, 256), 256, 0, s>>>(m, d_perm_, x, xp_);
  SOV_LAUNCH_CHECK();
  const int nlev = static_cast<int>(sym.level_ptr.size()) - 1;
  for (int l = 0; l < nlev; ++l) {
    const int b = sym.level_ptr[l], cnt = sym.level_ptr[l + 1] - b;
    front_forward_kernel<<<cnt, kFrontThreads, 0, s>>>(d_level_idx_ + b, fm, fronts_, vf_, xp_);
  }
  SOV_LAUNCH_CHECK();
  for (int l = nlev - 1; l >= 0; --l) {
    const int b = sym.level_ptr[l], cnt = sym.level_ptr[l + 1] - b;
    front_backward_kernel<<<cnt, kFrontThreads, 0, s>>>(d_level_idx_ + b, fm, fronts_, vf_, xp_);
  }
  SOV_LAUNCH_CHECK();
  permute_out_kernel<<<blocks_for(m, 256), 256, 0, s>>>(m, d_perm_, xp_, x);
  SOV_LAUNCH_CHECK();
  return SOV_OK;
}

}  // namespace sov
