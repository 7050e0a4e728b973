// Dense LU with partial pivoting, DenseLU's layout and pivot rule.
//
// Replaces run_dense_lu (solver/gpu/src/cuda_driver.cpp:461-502), which
// issues four launches per column (pivot, swap, scale, rank-1 update, the PTX
// kernels of :137-365), by a blocked right-looking LU:
//   lu_panel   one workgroup factors 64 columns (pivot search over the whole
//              column with ties to the lowest row, row swap inside the panel,
//              scale, rank-1 updates of the panel);
//   lu_swap    applies the panel's swaps to the other columns;
//   lu_trsm    U12 = L11^-1 A12 (unit lower, L11 in LDS);
//   gemm       A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64).
// The pivot rule and the singular test (largest remaining |a| < 1e-14) are
// those of DenseLU::factorize (solver/numerical/src/dense_lu.cpp:22-34).

#include "internal.hpp"
#include "mfma_gemm.hpp"

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

namespace {

constexpr int kNb = 64;

__global__ __launch_bounds__(1024) void lu_panel_kernel(double* __restrict__ a, long long n, int k0, int nb,
                                                        int32_t* __restrict__ ipiv, int* __restrict__ status) {
  __shared__ double s_val[16];
  __shared__ int s_idx[16];
  __shared__ int s_piv;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int jj = 0; jj < nb; ++jj) {
    const int col = k0 + jj;
    const double* ccol = a + static_cast<long long>(col) * n;
    // First row holding the largest |a(i, col)|, i >= col.
    double best = -1.0;
    int bidx = static_cast<int>(n);
    for (long long i = col + tid; i < n; i += blockDim.x) {
      const double v = fabs(ccol[i]);
      if (v > best) {
        best = v;
        bidx = static_cast<int>(i);
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bidx, off, 64);
      if (ov > best || (ov == best && oi < bidx)) {
        best = ov;
        bidx = oi;
      }
    }
    if (lane == 0) {
      s_val[wid] = best;
      s_idx[wid] = bidx;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = s_val[0];
      int bi = s_idx[0];
      for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
        if (s_val[w] > bv || (s_val[w] == bv && s_idx[w] < bi)) {
          bv = s_val[w];
          bi = s_idx[w];
        }
      }
      if (!(bv >= 1e-14) || bi >= n) {
        *status = 1;
        bi = col;
      }
      s_piv = bi;
      ipiv[col] = bi;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != col && tid < nb) {
      double* x = a + col + static_cast<long long>(k0 + tid) * n;
      double* y = a + p + static_cast<long long>(k0 + tid) * n;
      const double t = *x;
      *x = *y;
      *y = t;
    }
    __syncthreads();
    const double akk = ccol[col];
    for (long long i = col + 1 + tid; i < n; i += blockDim.x) {
      const double l = a[i + static_cast<long long>(col) * n] / akk;
      a[i + static_cast<long long>(col) * n] = l;
      for (int t = jj + 1; t < nb; ++t) {
        const long long o = static_cast<long long>(k0 + t) * n;
        a[i + o] -= l * a[col + o];
      }
    }
    __syncthreads();
  }
}

// Apply the swaps ipiv[k0 .. k0+nb) in order to every column outside the panel.
__global__ __launch_bounds__(256) void lu_swap_kernel(double* __restrict__ a, long long n, int k0, int nb,
                                                      const int32_t* __restrict__ ipiv) {
  const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n || (c >= k0 && c < k0 + nb)) return;
  double* col = a + c * n;
  for (int k = k0; k < k0 + nb; ++k) {
    const int p = ipiv[k];
    if (p != k) {
      const double t = col[k];
      col[k] = col[p];
      col[p] = t;
    }
  }
}

// U12 = L11^-1 A12 with L11 unit lower (nb x nb), one column per thread.
__global__ __launch_bounds__(256) void lu_trsm_kernel(double* __restrict__ a, long long n, int k0, int nb) {
  __shared__ double L[kNb * (kNb + 1)];
  for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
    const int i = idx % nb, j = idx / nb;
    L[j * (kNb + 1) + i] = a[(k0 + i) + static_cast<long long>(k0 + j) * n];
  }
  __syncthreads();
  const long long c = k0 + nb + static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double* col = a + c * n + k0;
  double v[kNb];
#pragma unroll
  for (int i = 0; i < kNb; ++i) v[i] = i < nb ? col[i] : 0.0;
#pragma unroll
  for (int j = 0; j < kNb; ++j) {
#pragma unroll
    for (int i = j + 1; i < kNb; ++i) v[i] -= L[j * (kNb + 1) + i] * v[j];
  }
#pragma unroll
  for (int i = 0; i < kNb; ++i)
    if (i < nb) col[i] = v[i];
}

// Clears the singular flag: a kernel rather than hipMemsetAsync, so that the
// sequence stays free of memset nodes.
__global__ void lu_clear_kernel(int* __restrict__ status) { *status = 0; }

// LAPACK-style swap list -> DenseLU's permutation (row i of P*A is row
// piv[i] of A; cuda_driver.cpp:496-500): a sequential walk, one thread.
__global__ void lu_perm_kernel(const int32_t* __restrict__ rows, int32_t* __restrict__ piv, long long n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (long long i = 0; i < n; ++i) piv[i] = static_cast<int32_t>(i);
  for (long long k = 0; k < n; ++k) {
    const int32_t r = rows[k];
    const int32_t t = piv[k];
    piv[k] = piv[r];
    piv[r] = t;
  }
}

// Scratch of the allocation-free entry point: the swap list (kNoAllocMaxN
// entries) and, after it, the singular flag. Each GPU holds its own copy;
// host code reaches it only through the address hipGetSymbolAddress returns.
constexpr int64_t kNoAllocMaxN = 32768;
__device__ __align__(256) int32_t scratch[kNoAllocMaxN + 64];

}  // namespace

namespace sov {

// The factorization's launches on device memory: a (n x n, column-major) in
// place, the swap list in ipiv (n), the singular flag in *status (cleared
// here). Nothing is allocated.
int dense_lu_launch(int64_t n, double* a, int32_t* ipiv, int* status, hipStream_t s) {
  lu_clear_kernel<<<1, 1, 0, s>>>(status);
  SOV_LAUNCH_CHECK();
  for (int64_t k0 = 0; k0 < n; k0 += kNb) {
    const int nb = static_cast<int>(std::min<int64_t>(kNb, n - k0));
    lu_panel_kernel<<<1, 1024, 0, s>>>(a, n, static_cast<int>(k0), nb, ipiv, status);
    SOV_LAUNCH_CHECK();
    lu_swap_kernel<<<blocks_for(n, 256), 256, 0, s>>>(a, n, static_cast<int>(k0), nb, ipiv);
    SOV_LAUNCH_CHECK();
    const int64_t rest = n - k0 - nb;
    if (rest > 0) {
      lu_trsm_kernel<<<blocks_for(rest, 256), 256, 0, s>>>(a, n, static_cast<int>(k0), nb);
      SOV_LAUNCH_CHECK();
      const double* l21 = a + (k0 + nb) + k0 * n;
      const double* u12 = a + k0 + (k0 + nb) * n;
      double* a22 = a + (k0 + nb) + (k0 + nb) * n;
      SOV_HIP_TRY((gemm_update<double, true, false>(static_cast<int>(rest), static_cast<int>(rest), nb, l21, n, u12,
                                                    n, a22, n, -1.0, s)));
    }
  }
  return SOV_OK;
}

int dense_lu_run(int64_t n, double* a_host, int32_t* piv_host) {
  if (n <= 0 || !a_host || !piv_host) return SOV_ERR_ARG;
  hipStream_t s = stream();
  const std::size_t bytes = static_cast<std::size_t>(n) * static_cast<std::size_t>(n) * sizeof(double);
  double* a = nullptr;
  int32_t* ipiv = nullptr;
  int* status = nullptr;
  struct Free {
    double** a;
    int32_t** p;
    int** s;
    ~Free() {
      if (*a) (void)hipFree(*a);
      if (*p) (void)hipFree(*p);
      if (*s) (void)hipFree(*s);
    }
  } guard{&a, &ipiv, &status};
  SOV_HIP_TRY(hipMalloc(&a, bytes));
  SOV_HIP_TRY(hipMalloc(&ipiv, static_cast<std::size_t>(n) * sizeof(int32_t)));
  SOV_HIP_TRY(hipMalloc(&status, sizeof(int)));
  SOV_HIP_TRY(hipMemcpyAsync(a, a_host, bytes, hipMemcpyHostToDevice, s));
  if (const int rc = dense_lu_launch(n, a, ipiv, status, s); rc != SOV_OK) return rc;
  int st = 0;
  SOV_HIP_TRY(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  if (st != 0) return SOV_SINGULAR;
  std::vector<int32_t> rows(static_cast<std::size_t>(n));
  SOV_HIP_TRY(hipMemcpy(a_host, a, bytes, hipMemcpyDeviceToHost));
  SOV_HIP_TRY(hipMemcpy(rows.data(), ipiv, static_cast<std::size_t>(n) * sizeof(int32_t), hipMemcpyDeviceToHost));
  // LAPACK-style swap list -> DenseLU's permutation (cuda_driver.cpp:496-500).
  for (int64_t i = 0; i < n; ++i) piv_host[i] = static_cast<int32_t>(i);
  for (int64_t k = 0; k < n; ++k) std::swap(piv_host[k], piv_host[rows[k]]);
  return SOV_OK;
}

}  // namespace sov

extern "C" int sov_dense_lu(int64_t n, double* a_host, int32_t* piv_host) {
  if (!sov::device_ready()) {
    sov::set_error("no usable GPU");
    return SOV_ERR_NODEV;
  }
  const int rc = sov::dense_lu_run(n, a_host, piv_host);
  if (rc == SOV_OK) {
    sov::count_operation();
    sov::count_factorization();
  }
  return rc;
}

namespace {

// Per-device address of `scratch`, looked up once per device with that
// device current.
std::mutex g_scratch_mu;
int32_t* g_scratch_dev[64] = {};

int scratch_for_current_device(int32_t** out) {
  int dev = 0;
  SOV_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    sov::set_error("sov_dense_lu_dev_noalloc: device index out of range");
    return SOV_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      sov::set_error(std::string("sov_dense_lu_dev_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return SOV_ERR_DEVICE;
    }
    g_scratch_dev[dev] = static_cast<int32_t*>(p);
  }
  *out = g_scratch_dev[dev];
  return SOV_OK;
}

}  // namespace

extern "C" int sov_dense_lu_dev_noalloc(int64_t n, double* a_dev, int32_t* piv_dev) {
  if (n <= 0 || !a_dev || !piv_dev) return SOV_ERR_ARG;
  if (n > kNoAllocMaxN) {
    sov::set_error("sov_dense_lu_dev_noalloc: n exceeds the scratch limit of 32768");
    return SOV_ERR_ARG;
  }
  if (!sov::device_ready()) {
    sov::set_error("no usable GPU");
    return SOV_ERR_NODEV;
  }
  int32_t* scr = nullptr;
  if (const int rc = scratch_for_current_device(&scr); rc != SOV_OK) return rc;
  int32_t* ipiv = scr;
  int* status = reinterpret_cast<int*>(scr + kNoAllocMaxN);
  hipStream_t s = sov::stream();
  if (const int rc = sov::dense_lu_launch(n, a_dev, ipiv, status, s); rc != SOV_OK) return rc;
  lu_perm_kernel<<<1, 64, 0, s>>>(ipiv, piv_dev, n);
  SOV_LAUNCH_CHECK();
  int st = 0;
  SOV_HIP_TRY(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  sov::count_operation();
  sov::count_factorization();
  return st != 0 ? SOV_SINGULAR : SOV_OK;
}
