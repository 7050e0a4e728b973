// Blocked right-looking Cholesky on MFMA and blocked triangular solves.
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
//   gemm (store) L21 = A21 inv(L11)^T on MFMA, copied back into the factor;
//   gemm_update  A22 -= L21 L21^T on MFMA (lower tiles only).
// The panel kernel also inverts its diagonal block; the inverses serve the
// panel TRSM (as a GEMM) and the blocked triangular solves.

#include "chol.hpp"
#include "mfma_gemm.hpp"

#include <algorithm>
#include <atomic>
#include <vector>

namespace sov {
namespace {

std::atomic<bool> g_timing{false};

constexpr int kPanelLd = kPanel + 1;  // LDS leading dimension of a panel block

// ---------------------------------------------------------------------------
// Diagonal block: Cholesky of the nb x nb block (nb <= kPanel = 128) and the
// inverse of its factor, register-tiled. The 1024 threads form a 32 x 32 grid
// of 4 x 4 tiles held in registers; per column only the finished column of L
// goes through LDS (Lst, which ends up holding all of L), so LDS traffic is a
// broadcast column per step rather than the whole trailing block. The inverse
// X = L^-1 is the forward substitution L X = I, right-looking on the same
// tiles. L11 goes back to `a`, X to inv_out (leading dimension kPanel, zero
// above the diagonal), which turns the panel TRSM into a GEMM.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void potrf_panel_kernel(T* __restrict__ a, long long lda, int k0, int nb,
                                                           const double* __restrict__ diag0,
                                                           FactorFlags* __restrict__ flags, T* __restrict__ inv_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Lst = reinterpret_cast<T*>(smem_raw);           // [kPanel][kPanelLd], column j at j*kPanelLd
  T* rowbuf = Lst + kPanel * kPanelLd;               // [2][kPanel]
  double* d0 = reinterpret_cast<double*>(rowbuf + 2 * kPanel);
  double* dv = d0 + kPanel;  // running diagonal (see below)
  const int tid = threadIdx.x;
  const int tr = tid & 31, tc = tid >> 5;  // tile row / tile column
  const int r0 = 4 * tr, c0 = 4 * tc;
  const bool lower_tile = tr >= tc;
  for (int idx = tid; idx < kPanel * kPanelLd; idx += blockDim.x) Lst[idx] = T(0);
  if (tid < nb) d0[tid] = diag0[k0 + tid];
  T v[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = r0 + r, k = c0 + c;
      v[r][c] = (lower_tile && i >= k && i < nb && k < nb) ? a[(k0 + i) + static_cast<long long>(k0 + k) * lda]
                                                          : T(0);
    }
  __syncthreads();
  // dv[i] tracks the diagonal entry i as the elimination proceeds (column
  // owners update it while publishing their column), so every thread can
  // take the pivot decision itself and a column step needs one barrier.
  for (int i = tid; i < kPanel; i += blockDim.x)
    dv[i] = i < nb ? static_cast<double>(a[(k0 + i) + static_cast<long long>(k0 + i) * lda]) : 1.0;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const int jt = j >> 2, jr = j & 3;
    const double dj = dv[j];
    const bool finite = isfinite(dj);
    const bool drop = finite && dj <= 1e-13 * fabs(d0[j]);
    const bool bad = !drop && (!finite || dj < 1e-30);
    const T piv = drop ? DropPivot<T>::value : bad ? T(1) : static_cast<T>(sqrt(dj));
    const T rpiv = T(1) / piv;
    if (tid == 0) {
      if (drop) atomicAdd(&flags->replaced, 1);
      if (bad) atomicOr(&flags->status, 1);
    }
    if (tc == jt && lower_tile) {  // column owners scale and publish column j
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c != jr) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = r0 + r;
          if (i == j) {
            v[r][c] = piv;
            Lst[j * kPanelLd + j] = piv;
          } else if (i > j && i < nb) {
            const T l = v[r][c] * rpiv;
            v[r][c] = l;
            Lst[j * kPanelLd + i] = l;
            dv[i] -= static_cast<double>(l) * static_cast<double>(l);
          }
        }
      }
    }
    __syncthreads();
    if (lower_tile && c0 + 3 > j) {  // rank-1 update of the trailing triangle
      T li[4], lk[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) li[r] = Lst[j * kPanelLd + r0 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) lk[c] = Lst[j * kPanelLd + c0 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c0 + c > j && r0 + r >= c0 + c) v[r][c] -= li[r] * lk[c];
    }
  }
  // L11 back to global memory.
  if (lower_tile) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = r0 + r, k = c0 + c;
        if (i >= k && i < nb && k < nb) a[(k0 + i) + static_cast<long long>(k0 + k) * lda] = v[r][c];
      }
  }
  // X = L^-1 by right-looking substitution on the identity (X stays lower).
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) v[r][c] = (r0 + r == c0 + c) ? T(1) : T(0);
  for (int k = 0; k < nb; ++k) {
    T* rb = rowbuf + (k & 1) * kPanel;
    const T lkk = Lst[k * kPanelLd + k];
    if (tr == (k >> 2)) {  // owners of row k scale it and publish it
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r != (k & 3)) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          v[r][c] = v[r][c] / lkk;
          rb[c0 + c] = v[r][c];
        }
      }
    }
    __syncthreads();
    if (lower_tile && r0 + 3 > k) {
      T xr[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) xr[c] = rb[c0 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = r0 + r;
        if (i <= k) continue;
        const T lik = Lst[k * kPanelLd + i];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] -= lik * xr[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = r0 + r, k = c0 + c;
      inv_out[i + k * kPanel] = (lower_tile && i >= k && i < nb && k < nb) ? v[r][c] : T(0);
    }
}


// L21 = A21 inv(L11)^T, computed out of place into P (rows below the
// diagonal block, leading dimension lda), copied back into the factor here.
template <typename T>
__global__ __launch_bounds__(256) void copy_panel_kernel(const T* __restrict__ P, T* __restrict__ a, long long lda,
                                                         int r0, int k0, int rows, int nb) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(rows) * nb) return;
  const int r = static_cast<int>(idx % rows), c = static_cast<int>(idx / rows);
  a[(r0 + r) + static_cast<long long>(k0 + c) * lda] = P[r + static_cast<long long>(c) * lda];
}

__global__ void reset_flags_kernel(FactorFlags* f) {
  f->status = 0;
  f->replaced = 0;
}

#include "trsv.hpp"

template <typename T>
constexpr std::size_t panel_lds_bytes() {
  return sizeof(T) * (kPanel * kPanelLd + 2 * kPanel) + sizeof(double) * 2 * kPanel;
}

template <typename T>
bool configure_lds() {
  static const bool ok = [] {
    bool r = hipFuncSetAttribute(reinterpret_cast<const void*>(potrf_panel_kernel<T>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(panel_lds_bytes<T>())) == hipSuccess;
    return r;
  }();
  return ok;
}

// Event pairs recorded around kernels while timing is on. Recording does
// not wait; the elapsed times are collected (after one stream sync) by
// sov_kernel_times, so timing can stay on inside a timed region.
enum TimedKind { kTimeSyrk = 0, kTimePanel = 1, kTimeSolve = 2 };
struct TimedPair {
  hipEvent_t a = nullptr, b = nullptr;
  int kind = 0;
  double flops = 0.0;
  int launches = 0;
};
struct EventPool {
  std::vector<TimedPair> pairs;
  std::size_t used = 0;
  TimedPair* next() {
    if (used == pairs.size()) {
      TimedPair p;
      if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return nullptr;
      pairs.push_back(p);
    }
    return &pairs[used++];
  }
};
EventPool& pool() {
  static EventPool p;
  return p;
}

// Opens a timed pair on `s` when timing is on (nullptr otherwise).
TimedPair* timed_begin(hipStream_t s, int kind) {
  if (!g_timing.load()) return nullptr;
  TimedPair* p = pool().next();
  if (!p) return nullptr;
  p->kind = kind;
  p->flops = 0.0;
  p->launches = 0;
  (void)hipEventRecord(p->a, s);
  return p;
}
void timed_end(TimedPair* p, hipStream_t s, int launches, double flops) {
  if (!p) return;
  p->launches = launches;
  p->flops = flops;
  (void)hipEventRecord(p->b, s);
}

}  // namespace

void set_kernel_timing(bool on) { g_timing.store(on); }
bool kernel_timing() { return g_timing.load(); }
KernelTimes& kernel_times() {
  static KernelTimes t;
  return t;
}

// Folds the recorded pairs into the totals (waits for the stream once).
void collect_kernel_times(hipStream_t s) {
  EventPool& p = pool();
  if (p.used == 0) return;
  (void)hipStreamSynchronize(s);
  KernelTimes& t = kernel_times();
  for (std::size_t k = 0; k < p.used; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.pairs[k].a, p.pairs[k].b) != hipSuccess) ms = 0.f;
    const TimedPair& e = p.pairs[k];
    if (e.kind == kTimeSyrk) {
      t.syrk_ms += ms;
      t.syrk_launches += e.launches;
      t.syrk_flops += e.flops;
    } else if (e.kind == kTimePanel) {
      t.panel_ms += ms;
      t.panel_launches += e.launches;
    } else {
      t.solve_ms += ms;
    }
  }
  p.used = 0;
}

template <typename T>
DenseCholesky<T>::~DenseCholesky() {
  if (side_) (void)hipStreamSynchronize(side_);
  if (ev_panel_) (void)hipEventDestroy(ev_panel_);
  if (ev_bulk_) (void)hipEventDestroy(ev_bulk_);
  if (side_) (void)hipStreamDestroy(side_);
  for (void* p : {static_cast<void*>(panel_), static_cast<void*>(panel_inv_), static_cast<void*>(work_),
                  static_cast<void*>(flags_)})
    if (p) (void)hipFree(p);
}

template <typename T>
int DenseCholesky<T>::reserve(int m) {
  if (m <= m_ && flags_) return SOV_OK;
  for (void* p : {static_cast<void*>(panel_), static_cast<void*>(panel_inv_), static_cast<void*>(work_),
                  static_cast<void*>(flags_)})
    if (p) (void)hipFree(p);
  panel_ = panel_inv_ = nullptr;
  work_ = nullptr;
  flags_ = nullptr;
  const int np = (m + kPanel - 1) / kPanel;
  SOV_HIP_TRY(hipMalloc(&panel_, sizeof(T) * static_cast<std::size_t>(m) * kPanel * 2));
  SOV_HIP_TRY(hipMalloc(&panel_inv_, sizeof(T) * static_cast<std::size_t>(np) * kPanel * kPanel));
  SOV_HIP_TRY(hipMalloc(&work_, sizeof(double) * static_cast<std::size_t>(m)));
  SOV_HIP_TRY(hipMalloc(&flags_, sizeof(FactorFlags)));
  SOV_HIP_TRY(hipMemset(flags_, 0, sizeof(FactorFlags)));
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
  if (!side_) {
    SOV_HIP_TRY(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    SOV_HIP_TRY(hipEventCreateWithFlags(&ev_panel_, hipEventDisableTiming));
    SOV_HIP_TRY(hipEventCreateWithFlags(&ev_bulk_, hipEventDisableTiming));
  }
  reset_flags_kernel<<<1, 1, 0, s>>>(flags_);
  SOV_LAUNCH_CHECK();
  // The side stream starts after everything already queued on s.
  SOV_HIP_TRY(hipEventRecord(ev_bulk_, s));
  SOV_HIP_TRY(hipStreamWaitEvent(side_, ev_bulk_, 0));
  const long long lda = m;
  const std::size_t pstride = static_cast<std::size_t>(m) * kPanel;
  for (int k0 = 0, p = 0; k0 < m; k0 += kPanel, ++p) {
    const int nb = std::min(kPanel, m - k0);
    const int below = m - k0 - nb;
    T* pinv = panel_inv_ + static_cast<long long>(p) * kPanel * kPanel;
    T* L21 = panel_ + (p & 1) * pstride;
    TimedPair* tp = timed_begin(s, kTimePanel);
    potrf_panel_kernel<T><<<1, 1024, panel_lds_bytes<T>(), s>>>(a, lda, k0, nb, diag0, flags_, pinv);
    SOV_LAUNCH_CHECK();
    if (below > 0) {
      // L21 = A21 inv(L11)^T on MFMA, out of place, then back into the factor.
      const T* a21 = a + (k0 + nb) + static_cast<long long>(k0) * lda;
      SOV_HIP_TRY((gemm_update<T, false, false, true>(below, nb, nb, a21, lda, pinv, kPanel, L21, lda, T(1), s)));
      copy_panel_kernel<T><<<blocks_for(static_cast<long long>(below) * nb, 256), 256, 0, s>>>(L21, a, lda, k0 + nb,
                                                                                               k0, below, nb);
      SOV_LAUNCH_CHECK();
    }
    timed_end(tp, s, below > 0 ? 3 : 1, 0.0);
    if (below <= 0) continue;
    // Look-ahead: the next panel's column strip first (on s, the critical
    // path), the rest of the trailing matrix on the side stream, where it
    // overlaps the next panel's factorization. The strip update also writes
    // the strictly upper part of its diagonal block, which nothing reads.
    const int strip = std::min(kPanel, below);
    const int rest = below - strip;
    SOV_HIP_TRY(hipStreamWaitEvent(s, ev_bulk_, 0));  // the strip was part of the previous bulk update
    TimedPair* ts = timed_begin(s, kTimeSyrk);
    T* a_strip = a + (k0 + nb) + static_cast<long long>(k0 + nb) * lda;
    SOV_HIP_TRY((gemm_update<T, false, false>(below, strip, nb, L21, lda, L21, lda, a_strip, lda, T(-1), s)));
    timed_end(ts, s, 1, static_cast<double>(below) * strip * nb * 2.0 - static_cast<double>(strip) * (strip - 1) * nb);
    if (rest > 0) {
      SOV_HIP_TRY(hipEventRecord(ev_panel_, s));
      SOV_HIP_TRY(hipStreamWaitEvent(side_, ev_panel_, 0));
      T* a22 = a + (k0 + nb + strip) + static_cast<long long>(k0 + nb + strip) * lda;
      const T* l21r = L21 + strip;
      TimedPair* tb = timed_begin(side_, kTimeSyrk);
      SOV_HIP_TRY((gemm_update<T, false, true>(rest, rest, nb, l21r, lda, l21r, lda, a22, lda, T(-1), side_)));
      timed_end(tb, side_, 1, static_cast<double>(rest) * (rest + 1) * nb);
      SOV_HIP_TRY(hipEventRecord(ev_bulk_, side_));
    }
  }
  SOV_HIP_TRY(hipEventRecord(ev_bulk_, side_));
  SOV_HIP_TRY(hipStreamWaitEvent(s, ev_bulk_, 0));
  return SOV_OK;
}

template <typename T>
int DenseCholesky<T>::solve(const T* a, int m, double* x, double* /*work*/, hipStream_t s) {
  if (m > m_) return SOV_ERR_ARG;
  const int np = (m + kPanel - 1) / kPanel;
  TimedPair* tp = timed_begin(s, kTimeSolve);
  for (int p = 0; p < np; ++p) {  // L y = x: y into work_, x updated in place
    const int below = std::max(0, m - (p + 1) * kPanel);
    trsv_fwd_step_kernel<T><<<std::max(1u, blocks_for(below, kSolveRows)), 256, 0, s>>>(a, m, m, p, panel_inv_, x,
                                                                                         work_);
  }
  SOV_LAUNCH_CHECK();
  for (int p = np - 1; p >= 0; --p) {  // L^T x = y: x back into x, work_ updated in place
    const int above = p * kPanel;
    trsv_bwd_step_kernel<T><<<std::max(1u, blocks_for(above, kSolveRows)), 256, 0, s>>>(a, m, m, p, panel_inv_, work_,
                                                                                         x);
  }
  SOV_LAUNCH_CHECK();
  timed_end(tp, s, 2 * np, 0.0);
  return SOV_OK;
}

template class DenseCholesky<double>;
template class DenseCholesky<float>;

}  // namespace sov

extern "C" {

void sov_set_kernel_timing(int32_t on) { sov::set_kernel_timing(on != 0); }

int sov_kernel_times(double* out) {
  if (!out) return SOV_ERR_ARG;
  if (sov::device_ready()) sov::collect_kernel_times(sov::stream());
  sov::KernelTimes& t = sov::kernel_times();
  const double v[7] = {t.syrk_ms, t.syrk_launches, t.syrk_flops, t.panel_ms, t.solve_ms, t.assemble_ms,
                       t.panel_launches};
  for (int i = 0; i < 7; ++i) out[i] = v[i];
  t = sov::KernelTimes{};
  return SOV_OK;
}

}  // extern "C"
