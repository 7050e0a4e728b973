// Blocked right-looking Cholesky on MFMA and persistent triangular solves.
//
// Replaces, for the interior-point normal equations, the reference's dense
// path (build_normal_eq + DenseLU, solver/lp/src/interior_point.cpp:74-101,
// 176-179, whose GPU form is 4 launches per column in
// solver/gpu/src/cuda_driver.cpp:461-502) and its sparse LDL^T
// (solver/numerical/src/sparse_ldlt.cpp:105-223), keeping the LDL^T's pivot
// rule: a pivot <= 1e-13 x |its diagonal before elimination| drops that
// direction, a non-finite or < 1e-30 pivot is a breakdown.
//
// Per panel of kPanel columns:
//   potrf_panel  one workgroup factors the diagonal block in LDS;
//   trsm_panel   L21 = A21 L11^-T, 64 rows per workgroup, L11 in LDS;
//   gemm_update  A22 -= L21 L21^T on MFMA (lower tiles only).
// After the last panel the 64x64 diagonal blocks are inverted in parallel
// so the triangular solves become products.

#include "chol.hpp"
#include "mfma_gemm.hpp"

#include <algorithm>
#include <atomic>

namespace sov {
namespace {

std::atomic<bool> g_timing{false};

constexpr int kPanelLd = kPanel + 1;  // LDS leading dimension of a panel block

// ---------------------------------------------------------------------------
// Diagonal block: right-looking unblocked Cholesky in LDS, 1024 threads.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void potrf_panel_kernel(T* __restrict__ a, long long lda, int k0, int nb,
                                                           const double* __restrict__ diag0,
                                                           FactorFlags* __restrict__ flags) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* L = reinterpret_cast<T*>(smem_raw);
  double* d0 = reinterpret_cast<double*>(smem_raw + sizeof(T) * kPanel * kPanelLd);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < nb * nb; idx += blockDim.x) {
    const int i = idx % nb, j = idx / nb;
    if (i >= j) L[j * kPanelLd + i] = a[(k0 + i) + static_cast<long long>(k0 + j) * lda];
  }
  if (tid < nb) d0[tid] = diag0[k0 + tid];
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    // Every thread takes the same pivot decision from the same LDS value.
    const double dj = static_cast<double>(L[j * kPanelLd + j]);
    const bool finite = isfinite(dj);
    const bool drop = finite && dj <= 1e-13 * fabs(d0[j]);
    const bool bad = !drop && (!finite || dj < 1e-30);
    const T piv = drop ? DropPivot<T>::value : bad ? T(1) : static_cast<T>(sqrt(dj));
    if (tid > j && tid < nb) L[j * kPanelLd + tid] /= piv;
    __syncthreads();
    if (tid == 0) {
      L[j * kPanelLd + j] = piv;
      if (drop) atomicAdd(&flags->replaced, 1);
      if (bad) atomicOr(&flags->status, 1);
    }
    const int w = nb - j - 1;
    for (int idx = tid; idx < w * w; idx += blockDim.x) {
      const int i = j + 1 + idx % w, t = j + 1 + idx / w;
      if (i >= t) L[t * kPanelLd + i] -= L[j * kPanelLd + i] * L[j * kPanelLd + t];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nb * nb; idx += blockDim.x) {
    const int i = idx % nb, j = idx / nb;
    if (i >= j) a[(k0 + i) + static_cast<long long>(k0 + j) * lda] = L[j * kPanelLd + i];
  }
}

// ---------------------------------------------------------------------------
// Panel rows below the diagonal block: L21 = A21 L11^-T by forward
// substitution along the row. 1024 threads = 16 wavefronts; lane = row
// (64 rows per workgroup), wavefront w owns columns w, w+16, ...; the value
// of column j is passed through LDS (double-buffered, one barrier per column).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void trsm_panel_kernel(T* __restrict__ a, long long lda, int k0, int nb,
                                                          int m) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* L = reinterpret_cast<T*>(smem_raw);
  T* X = L + kPanel * kPanelLd;  // [2][64]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = k0 + nb + blockIdx.x * 64 + lane;
  const bool live = row < m;
  for (int idx = tid; idx < nb * nb; idx += blockDim.x) {
    const int i = idx % nb, j = idx / nb;
    if (i >= j) L[j * kPanelLd + i] = a[(k0 + i) + static_cast<long long>(k0 + j) * lda];
  }
  constexpr int Q = kPanel / 16;
  T v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int c = wave + 16 * q;
    v[q] = (live && c < nb) ? a[row + static_cast<long long>(k0 + c) * lda] : T(0);
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const int owner = j & 15, qj = j >> 4;
    T* xb = X + (j & 1) * 64;
    if (wave == owner) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (q == qj) {
          v[q] = v[q] / L[j * kPanelLd + j];
          xb[lane] = v[q];
        }
      }
    }
    __syncthreads();
    const T xj = xb[lane];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int c = wave + 16 * q;
      if (c > j && c < nb) v[q] -= xj * L[j * kPanelLd + c];
    }
  }
  if (live) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int c = wave + 16 * q;
      if (c < nb) a[row + static_cast<long long>(k0 + c) * lda] = v[q];
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse of each kSolveBlock diagonal block of L (in-place lower-triangular
// inversion in LDS, LAPACK trti2 order), written to dinv (block p at
// p*64*64, column-major, zero above the diagonal).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void diag_inverse_kernel(const T* __restrict__ a, long long lda, int m,
                                                           T* __restrict__ dinv) {
  constexpr int B = kSolveBlock, LD = B + 1;
  __shared__ double W[B * LD];
  const int p = blockIdx.x, base = p * B;
  const int nbp = min(B, m - base);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < B * B; idx += blockDim.x) {
    const int i = idx % B, j = idx / B;
    double v = 0.0;
    if (i < nbp && j < nbp && i >= j) v = static_cast<double>(a[(base + i) + static_cast<long long>(base + j) * lda]);
    if (i == j && i >= nbp) v = 1.0;
    W[j * LD + i] = v;
  }
  __syncthreads();
  for (int j = B - 1; j >= 0; --j) {
    const double inv = 1.0 / W[j * LD + j];
    // new column below the diagonal: -inv * Winv[j+1:, j+1:] * W[j+1:, j]
    double out = 0.0;
    const int i = tid;
    if (i > j && i < B) {
      for (int k = j + 1; k <= i; ++k) out += W[k * LD + i] * W[j * LD + k];
      out *= -inv;
    }
    __syncthreads();
    if (i > j && i < B) W[j * LD + i] = out;
    if (tid == 0) W[j * LD + j] = inv;
    __syncthreads();
  }
  T* dst = dinv + static_cast<long long>(p) * B * B;
  for (int idx = tid; idx < B * B; idx += blockDim.x) {
    const int i = idx % B, j = idx / B;
    dst[idx] = i >= j ? static_cast<T>(W[j * LD + i]) : T(0);
  }
}

__global__ void reset_flags_kernel(FactorFlags* f) {
  f->status = 0;
  f->replaced = 0;
}

// Bounded wait for a row block published in generation `gen`. A wait that
// exceeds the budget marks the solve failed (status 2) instead of hanging.
__device__ __forceinline__ void wait_block(const unsigned* flag, unsigned gen, FactorFlags* fl) {
  long long spins = 0;
  while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > (1LL << 26)) {
      atomicOr(&fl->status, 2);
      return;
    }
  }
}

__device__ __forceinline__ double load_fresh(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Forward solve L y = b in place on x, persistent: one workgroup per row
// block, block order taken from a ticket so a workgroup only ever waits on
// blocks whose workgroups started earlier. 256 threads: lane = row of the
// block, wavefront h = a quarter of the 64 columns of each source block.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void trsv_forward_kernel(const T* __restrict__ a, long long lda, int m,
                                                           const T* __restrict__ dinv, double* x,
                                                           unsigned* blk_flags, FactorFlags* fl, unsigned gen,
                                                           unsigned ticket_base) {
  constexpr int B = kSolveBlock;
  __shared__ int s_blk;
  __shared__ double yj[B];
  __shared__ double part[4][B];
  const int tid = threadIdx.x, lane = tid & 63, h = tid >> 6;
  if (tid == 0) s_blk = static_cast<int>(atomicAdd(&fl->ticket, 1u) - ticket_base);
  __syncthreads();
  const int i = s_blk;
  const int row = i * B + lane;
  const bool live = row < m;
  double acc = 0.0;
  for (int j = 0; j < i; ++j) {
    if (tid == 0) wait_block(&blk_flags[j], gen, fl);
    __syncthreads();
    if (tid < B) yj[tid] = load_fresh(&x[j * B + tid]);
    __syncthreads();
    if (live) {
#pragma unroll 4
      for (int q = 0; q < 16; ++q) {
        const int c = h * 16 + q;
        acc += static_cast<double>(a[row + static_cast<long long>(j * B + c) * lda]) * yj[c];
      }
    }
    __syncthreads();
  }
  part[h][lane] = acc;
  __syncthreads();
  if (tid < B) {
    const double bval = live ? x[row] : 0.0;
    yj[tid] = bval - (part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]);
  }
  __syncthreads();
  // y_i = inv(L_ii) c: row `lane`, columns of quarter h.
  const T* inv = dinv + static_cast<long long>(i) * B * B;
  double y = 0.0;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int c = h * 16 + q;
    y += static_cast<double>(inv[lane + c * B]) * yj[c];
  }
  part[h][lane] = y;
  __syncthreads();
  if (tid < B && live) x[row] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
  __threadfence();
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&blk_flags[i], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Backward solve L^T x = y in place, persistent, blocks from the last one
// up. Block i needs column block i of L below the diagonal: wavefront w owns
// 16 of its columns and its lanes run down the rows, so every load is a
// contiguous 512-byte column segment; lane partial sums are reduced once at
// the end.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void trsv_backward_kernel(const T* __restrict__ a, long long lda, int m,
                                                            const T* __restrict__ dinv, double* x,
                                                            unsigned* blk_flags, FactorFlags* fl, unsigned gen,
                                                            unsigned ticket_base) {
  constexpr int B = kSolveBlock;
  __shared__ int s_blk;
  __shared__ double xj[B];
  __shared__ double c[B];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nblk = (m + B - 1) / B;
  if (tid == 0) s_blk = nblk - 1 - static_cast<int>(atomicAdd(&fl->ticket, 1u) - ticket_base);
  __syncthreads();
  const int i = s_blk;
  double p[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) p[q] = 0.0;
  for (int j = nblk - 1; j > i; --j) {
    if (tid == 0) wait_block(&blk_flags[j], gen, fl);
    __syncthreads();
    const int r = j * B + lane;
    if (tid < B) xj[tid] = (j * B + tid < m) ? load_fresh(&x[j * B + tid]) : 0.0;
    __syncthreads();
    if (r < m) {
      const double xr = xj[lane];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int col = i * B + w * 16 + q;
        p[q] += static_cast<double>(a[r + static_cast<long long>(col) * lda]) * xr;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const double s = wave_sum(p[q]);
    if (lane == 0) c[w * 16 + q] = s;
  }
  __syncthreads();
  if (tid < B) {
    const int r = i * B + tid;
    c[tid] = (r < m ? x[r] : 0.0) - c[tid];
  }
  __syncthreads();
  // x_i = inv(L_ii)^T c: output k = w*16+q sums inv[r + k*B] c[r] over r = lane.
  const T* inv = dinv + static_cast<long long>(i) * B * B;
  double outv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int k = w * 16 + q;
    outv[q] = wave_sum(static_cast<double>(inv[lane + k * B]) * c[lane]);
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int r = i * B + w * 16 + q;
      if (r < m) x[r] = outv[q];
    }
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&blk_flags[i], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
constexpr std::size_t panel_lds_bytes() {
  return sizeof(T) * kPanel * kPanelLd + sizeof(double) * kPanel;
}

template <typename T>
constexpr std::size_t trsm_lds_bytes() {
  return sizeof(T) * kPanel * kPanelLd + sizeof(T) * 128;
}

template <typename T>
bool configure_lds() {
  static const bool ok = [] {
    bool r = hipFuncSetAttribute(reinterpret_cast<const void*>(potrf_panel_kernel<T>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(panel_lds_bytes<T>())) == hipSuccess;
    r = r && hipFuncSetAttribute(reinterpret_cast<const void*>(trsm_panel_kernel<T>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(trsm_lds_bytes<T>())) == hipSuccess;
    return r;
  }();
  return ok;
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  float ms() {
    float t = 0.f;
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&t, a, b);
    return t;
  }
};

}  // namespace

void set_kernel_timing(bool on) { g_timing.store(on); }
bool kernel_timing() { return g_timing.load(); }
KernelTimes& kernel_times() {
  static KernelTimes t;
  return t;
}

template <typename T>
DenseCholesky<T>::~DenseCholesky() {
  if (dinv_) (void)hipFree(dinv_);
  if (solve_flags_) (void)hipFree(solve_flags_);
  if (flags_) (void)hipFree(flags_);
}

template <typename T>
int DenseCholesky<T>::reserve(int m) {
  if (m <= m_ && flags_) return SOV_OK;
  if (dinv_) (void)hipFree(dinv_);
  if (solve_flags_) (void)hipFree(solve_flags_);
  if (flags_) (void)hipFree(flags_);
  dinv_ = nullptr;
  solve_flags_ = nullptr;
  flags_ = nullptr;
  const int nblk = (m + kSolveBlock - 1) / kSolveBlock;
  SOV_HIP_TRY(hipMalloc(&dinv_, sizeof(T) * static_cast<std::size_t>(nblk) * kSolveBlock * kSolveBlock));
  SOV_HIP_TRY(hipMalloc(&solve_flags_, sizeof(unsigned) * static_cast<std::size_t>(nblk)));
  SOV_HIP_TRY(hipMalloc(&flags_, sizeof(FactorFlags)));
  SOV_HIP_TRY(hipMemset(solve_flags_, 0, sizeof(unsigned) * static_cast<std::size_t>(nblk)));
  SOV_HIP_TRY(hipMemset(flags_, 0, sizeof(FactorFlags)));
  generation_ = 0;
  tickets_used_ = 0;
  m_ = m;
  if (!configure_lds<T>()) {
    set_error("could not raise the LDS limit of the panel kernels");
    return SOV_ERR_DEVICE;
  }
  return SOV_OK;
}

template <typename T>
int DenseCholesky<T>::factor(T* a, int m, const double* diag0, hipStream_t s) {
  int rc = reserve(m);
  if (rc != SOV_OK) return rc;
  const bool timing = kernel_timing();
  EventPair ev;
  reset_flags_kernel<<<1, 1, 0, s>>>(flags_);
  SOV_LAUNCH_CHECK();
  const long long lda = m;
  for (int k0 = 0; k0 < m; k0 += kPanel) {
    const int nb = std::min(kPanel, m - k0);
    if (timing) (void)hipEventRecord(ev.a, s);
    potrf_panel_kernel<T><<<1, 1024, panel_lds_bytes<T>(), s>>>(a, lda, k0, nb, diag0, flags_);
    SOV_LAUNCH_CHECK();
    const int below = m - k0 - nb;
    if (below > 0) {
      trsm_panel_kernel<T><<<blocks_for(below, 64), 1024, trsm_lds_bytes<T>(), s>>>(a, lda, k0, nb, m);
      SOV_LAUNCH_CHECK();
    }
    if (timing) {
      (void)hipEventRecord(ev.b, s);
      kernel_times().panel_ms += ev.ms();
      kernel_times().panel_launches += below > 0 ? 2 : 1;
    }
    if (below > 0) {
      const T* l21 = a + (k0 + nb) + static_cast<long long>(k0) * lda;
      T* a22 = a + (k0 + nb) + static_cast<long long>(k0 + nb) * lda;
      if (timing) (void)hipEventRecord(ev.a, s);
      SOV_HIP_TRY((gemm_update<T, false, true>(below, below, nb, l21, lda, l21, lda, a22, lda, T(-1), s)));
      if (timing) {
        (void)hipEventRecord(ev.b, s);
        KernelTimes& kt = kernel_times();
        kt.syrk_ms += ev.ms();
        kt.syrk_launches += 1;
        kt.syrk_flops += static_cast<double>(below) * (below + 1) * nb;
      }
    }
  }
  diag_inverse_kernel<T><<<blocks_for(m, kSolveBlock), 256, 0, s>>>(a, lda, m, dinv_);
  SOV_LAUNCH_CHECK();
  return SOV_OK;
}

template <typename T>
int DenseCholesky<T>::solve(const T* a, int m, double* x, double* /*work*/, hipStream_t s) {
  if (m > m_) return SOV_ERR_ARG;
  const unsigned nblk = static_cast<unsigned>((m + kSolveBlock - 1) / kSolveBlock);
  const bool timing = kernel_timing();
  EventPair ev;
  if (timing) (void)hipEventRecord(ev.a, s);
  unsigned gen = ++generation_;
  trsv_forward_kernel<T><<<nblk, 256, 0, s>>>(a, m, m, dinv_, x, solve_flags_, flags_, gen, tickets_used_);
  SOV_LAUNCH_CHECK();
  tickets_used_ += nblk;
  gen = ++generation_;
  trsv_backward_kernel<T><<<nblk, 256, 0, s>>>(a, m, m, dinv_, x, solve_flags_, flags_, gen, tickets_used_);
  SOV_LAUNCH_CHECK();
  tickets_used_ += nblk;
  if (timing) {
    (void)hipEventRecord(ev.b, s);
    kernel_times().solve_ms += ev.ms();
  }
  return SOV_OK;
}

template class DenseCholesky<double>;
template class DenseCholesky<float>;

}  // namespace sov

extern "C" {

void sov_set_kernel_timing(int32_t on) { sov::set_kernel_timing(on != 0); }

int sov_kernel_times(double* out) {
  if (!out) return SOV_ERR_ARG;
  sov::KernelTimes& t = sov::kernel_times();
  const double v[7] = {t.syrk_ms, t.syrk_launches, t.syrk_flops, t.panel_ms, t.solve_ms, t.assemble_ms,
                       t.panel_launches};
  for (int i = 0; i < 7; ++i) out[i] = v[i];
  t = sov::KernelTimes{};
  return SOV_OK;
}

}  // extern "C"
