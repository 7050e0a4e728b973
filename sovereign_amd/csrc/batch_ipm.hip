// Batched interior-point solves of many small dense LPs in one persistent
// launch - the strong-branching children of a branch-and-bound node.
//
// The reference solves the 2k child LPs of a node one after another with a
// warm-started dual simplex (pick_strong_branch / pick_reliability_branch,
// solver/milp/src/branch_and_bound.cpp:726-899; solve_two_lps_parallel
// :554-638 is serial on Linux). Here every child LP
//
//     min c'x   s.t.  A x = b,  0 <= x <= u   (u_j >= 1e29: no upper bound,
//                                            u_j == 0: x_j fixed at 0)
//
// is solved by a bounded-variable Mehrotra predictor-corrector method run
// entirely by one 1024-thread workgroup; workgroups take LPs from a queue, so
// LPs of uneven size share one launch. Per iteration:
//   M = A D A^T          FP32 MFMA (v_mfma_f32_16x16x4f32), FP32 accumulation,
//                        written to LDS as a packed lower triangle (m <= 272);
//   M = L L^T            FP32, in LDS, with the LDL^T pivot rule of the
//                        reference (drop a pivot <= 1e-13 |diagonal|);
//   M dy = rhs           FP64 preconditioned CG with the FP32 factor as the
//                        preconditioner and the FP64 product A (D (A^T p)).
// Residuals, step lengths and the termination test are FP64.

#include "internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>

typedef float flt4_t __attribute__((ext_vector_type(4)));
typedef double dbl4_t __attribute__((ext_vector_type(4)));

namespace {

// 16 wavefronts. At this size the kernel spills about 850 B per thread (128
// registers a wave); 512 threads remove the spills but take longer per LP
// (85 against 77 ms for the 128 knapsack children): the LDS factorization
// and inversion steps have half the threads.
constexpr int kThreads = 1024;
static_assert(kThreads % 256 == 0 && kThreads <= 1024, "gemv_n slices the columns in 256-row groups");
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRows = 272;
constexpr int kStage = 1024;  // doubles of LDS staging (vectors, partial sums)
constexpr int kTraceRows = 200, kTraceCols = 16;  // SOVEREIGN_BATCH_TRACE record

struct BatchMatrix {
  int m, n;
  const double* a64;  // column-major m x n
  const float* a32;
};

struct BatchDesc {
  int mat;
  int pad;
  long long b_off, x_off, ws_off;
};

struct BatchOpts {
  double tol, tau, pcg_tol;
  int max_iter, pcg_max;
};

}  // namespace

typedef sov_batch_result sov_batch_result_dev;

namespace {

__device__ __forceinline__ long long packed_index(int i, int j, int m) {
  // column-major packed lower triangle, i >= j
  return static_cast<long long>(j) * (2 * m - j + 1) / 2 + (i - j);
}

// Block-wide reduction of up to 4 values; every thread gets the results.
template <int OP>
__device__ __forceinline__ double bred(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = OP == 0 ? wave_sum(v) : OP == 1 ? wave_max(v) : wave_min(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
  for (int k = 1; k < kWaves; ++k) r = OP == 0 ? r + red[k] : OP == 1 ? fmax(r, red[k]) : fmin(r, red[k]);
  return r;
}

struct Lp {
  int m, n;
  const double* A;
  const float* Af;
  const double *b, *c, *u;
  // workspace vectors (global, this LP only)
  double *x, *w, *z, *v, *rc, *ru, *D, *dx, *dw, *dz, *dv, *dxa, *dwa, *dza, *dva, *t;
  double *y, *rb, *dy, *r, *p, *q, *zz, *rhs, *sc;  // sc: Jacobi scale of A D A^T
  double* M64;  // packed lower FP64 normal matrix (global), for LPs whose FP32 solves stall
};

// Storage of the packed normal matrix and its factor: FP32 in LDS, or FP64
// in global memory for an LP that has switched to FP64 factorizations.
template <typename T>
struct Prec;
template <>
struct Prec<float> {
  static constexpr float drop_pivot = 1e18f;
  __device__ static __forceinline__ flt4_t mfma(float a, float b, flt4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  typedef flt4_t acc_t;
  __device__ static __forceinline__ int row(int lk, int g) { return 4 * lk + g; }  // output row of register g
  __device__ static __forceinline__ float a(const Lp& L, long long idx) { return L.Af[idx]; }
};
template <>
struct Prec<double> {
  static constexpr double drop_pivot = 1e64;
  __device__ static __forceinline__ dbl4_t mfma(double a, double b, dbl4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  typedef dbl4_t acc_t;
  __device__ static __forceinline__ int row(int lk, int g) { return lk + 4 * g; }
  __device__ static __forceinline__ double a(const Lp& L, long long idx) { return L.A[idx]; }
};

// out_j = (A^T vec)_j for j < n; vec (m) is staged in LDS first.
__device__ void gemv_t(const Lp& L, const double* vec, double* out, double* stage) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __syncthreads();
  for (int i = tid; i < L.m; i += kThreads) stage[i] = vec[i];
  __syncthreads();
  for (int j0 = wv * 4; j0 < L.n; j0 += kWaves * 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = j0 + q;
      if (j < L.n) {
        const double* col = L.A + static_cast<long long>(j) * L.m;
        for (int i = lane; i < L.m; i += 64) s[q] = fma(col[i], stage[i], s[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double tot = wave_sum(s[q]);
      if (lane == 0 && j0 + q < L.n) out[j0 + q] = tot;
    }
  }
  __syncthreads();
}

// out_i = base_i + (A vec)_i for i < m (base may be null). vec (n) global.
__device__ void gemv_n(const Lp& L, const double* vec, const double* base, double* out, double* stage) {
  constexpr int kSlices = kThreads / 256;  // slices of the columns, 256 rows each
  const int tid = threadIdx.x;
  const int row = tid & 255, h = tid >> 8;
  for (int r0 = 0; r0 < L.m; r0 += 256) {
    const int i = r0 + row;
    double acc = 0.0;
    if (i < L.m)
      for (int j = h; j < L.n; j += kSlices) acc = fma(L.A[i + static_cast<long long>(j) * L.m], vec[j], acc);
    __syncthreads();
    stage[h * 256 + row] = acc;
    __syncthreads();
    if (h == 0 && i < L.m) {
      double s = stage[row];
#pragma unroll
      for (int q = 1; q < kSlices; ++q) s += stage[q * 256 + row];
      out[i] = (base ? base[i] : 0.0) + s;
    }
  }
  __syncthreads();
}

// Jacobi-scaled normal matrix in LDS (packed lower, FP32):
//   Mt = S A D A^T S,  S = diag(sc),  sc_i = (A D A^T)_ii^(-1/2),
// assembled on MFMA from the operands A(i, k) sqrt(D_k) sc_i, so Mt has a
// unit diagonal whatever the spread of D; the FP32 factorization of Mt is
// then accurate relative to its own scale. L.t holds sqrt(D) meanwhile.
template <typename T>
__device__ void assemble(const Lp& L, T* M, double reg) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  for (int j = tid; j < L.n; j += kThreads) L.t[j] = sqrt(fmax(L.D[j], 0.0));
  __syncthreads();
  {  // diagonal in FP64: four threads per row
    const int row = tid >> 2, part = tid & 3;
    for (int i0 = 0; i0 < L.m; i0 += kThreads / 4) {
      const int i = i0 + row;
      double d = 0.0;
      if (i < L.m)
        for (int k = part; k < L.n; k += 4) {
          const double a = L.A[i + static_cast<long long>(k) * L.m];
          d = fma(a * a, L.D[k], d);
        }
      d += __shfl_xor(d, 1);
      d += __shfl_xor(d, 2);
      if (part == 0 && i < L.m) L.sc[i] = d + reg > 0.0 ? 1.0 / sqrt(d + reg) : 1.0;
    }
  }
  __syncthreads();
  const int nt = (L.m + 15) / 16;
  const int tiles = nt * (nt + 1) / 2;
  for (int t = wv; t < tiles; t += kWaves) {
    int ti = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while (ti * (ti + 1) / 2 > t) --ti;
    const int tj = t - ti * (ti + 1) / 2;
    const int r0 = ti * 16, c0 = tj * 16;
    const int ra = r0 + lrow, cb = c0 + lrow;
    const double sa = ra < L.m ? L.sc[ra] : 0.0, sbv = cb < L.m ? L.sc[cb] : 0.0;
    typename Prec<T>::acc_t acc = {0, 0, 0, 0};
    for (int k0 = 0; k0 < L.n; k0 += 16) {
      T av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u + lk;
        const bool kin = k < L.n;
        const double sd = kin ? L.t[k] : 0.0;
        av[u] = (kin && ra < L.m) ? Prec<T>::a(L, ra + static_cast<long long>(k) * L.m) * static_cast<T>(sd * sa) : T(0);
        bv[u] = (kin && cb < L.m) ? Prec<T>::a(L, cb + static_cast<long long>(k) * L.m) * static_cast<T>(sd * sbv) : T(0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = Prec<T>::mfma(av[u], bv[u], acc);
    }
    // lane holds column c0 + lrow, rows r0 + Prec<T>::row(lk, g)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int i = r0 + Prec<T>::row(lk, g), j = c0 + lrow;
      if (i < L.m && j < L.m && i >= j)
        M[packed_index(i, j, L.m)] = acc[g] + (i == j ? static_cast<T>(reg * L.sc[i] * L.sc[i]) : T(0));
    }
  }
  __syncthreads();
}

// In-place FP32 Cholesky of the packed lower triangle in LDS (unit diagonal
// before elimination, so the reference's drop rule reads pivot <= 1e-13).
// Returns the number of dropped pivots.
template <typename T>
__device__ int cholesky(const Lp& L, T* M) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = L.m;
  int dropped = 0;
  for (int k = 0; k < m; ++k) {
    const long long kk = packed_index(k, k, m);
    const double dk = static_cast<double>(M[kk]);
    const bool drop = !(dk > 1e-13) && isfinite(dk);
    const T piv = drop ? Prec<T>::drop_pivot : (isfinite(dk) ? static_cast<T>(sqrt(dk)) : T(1));
    dropped += drop ? 1 : 0;
    for (int i = k + 1 + tid; i < m; i += kThreads) M[kk + (i - k)] /= piv;
    __syncthreads();
    if (tid == 0) M[kk] = piv;
    for (int j = k + 1 + wv; j < m; j += kWaves) {
      const T ljk = M[kk + (j - k)];
      const long long cj = packed_index(j, j, m);
      for (int i = j + lane; i < m; i += 64) M[cj + (i - j)] -= M[kk + (i - k)] * ljk;
    }
    __syncthreads();
  }
  return dropped;
}

// In place: the packed factor L becomes X = L^-1 (lower). Column j, from
// the last: X(j, j) = 1 / L(j, j), X(i, j) = -X(j, j) sum_{k=j+1..i} X(i, k)
// L(k, j), with the columns right of j already inverted. Four threads per
// row share the sum; tmp (LDS, m doubles) holds the column until all reads
// of L(., j) are done.
template <typename T>
__device__ void invert_lower(const Lp& L, T* M, double* tmp) {
  const int tid = threadIdx.x, m = L.m;
  const int row = tid >> 2, part = tid & 3;
  for (int j = m - 1; j >= 0; --j) {
    const long long jj = packed_index(j, j, m);
    const T xjj = T(1) / M[jj];  // read before anyone rewrites it below
    for (int i0 = j + 1; i0 < m; i0 += kThreads / 4) {
      const int i = i0 + row;
      T acc = T(0);
      if (i < m) {
        // X(i, k) at packed_index(i, k): the stride to k + 1 is m - k - 1
        long long idx = packed_index(i, j + 1 + part, m);
        for (int k = j + 1 + part; k <= i; k += 4) {
          acc += M[idx] * M[jj + (k - j)];
          idx += (m - k - 1) + (m - k - 2) + (m - k - 3) + (m - k - 4);
        }
      }
      acc += __shfl_xor(acc, 1);
      acc += __shfl_xor(acc, 2);
      if (part == 0 && i < m) tmp[i] = acc;
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < m; i += kThreads) M[jj + (i - j)] = -xjj * static_cast<T>(tmp[i]);
    if (tid == 0) M[jj] = xjj;
    __syncthreads();
  }
}

// vec <- M^-1 vec = S X^T X S vec (vec in LDS, FP64), with X = inverse of
// the FP32 factor of the scaled matrix; tmp (LDS, m doubles) is scratch.
template <typename T>
__device__ void precondition(const Lp& L, const T* M, double* vec, double* tmp) {
  const int tid = threadIdx.x, m = L.m;
  const int row = tid >> 2, part = tid & 3;
  __syncthreads();
  for (int i = tid; i < m; i += kThreads) vec[i] *= L.sc[i];
  __syncthreads();
  for (int i0 = 0; i0 < m; i0 += kThreads / 4) {  // tmp = X vec: row i of X
    const int i = i0 + row;
    double acc = 0.0;
    if (i < m) {
      long long idx = packed_index(i, part, m);
      for (int k = part; k <= i; k += 4) {
        acc = fma(static_cast<double>(M[idx]), vec[k], acc);
        idx += (m - k - 1) + (m - k - 2) + (m - k - 3) + (m - k - 4);
      }
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (part == 0 && i < m) tmp[i] = acc;
  }
  __syncthreads();
  for (int k0 = 0; k0 < m; k0 += kThreads / 4) {  // vec = X^T tmp: column k of X
    const int k = k0 + row;
    double acc = 0.0;
    if (k < m) {
      const long long base = packed_index(k, k, m);
      for (int i = k + part; i < m; i += 4) acc = fma(static_cast<double>(M[base + (i - k)]), tmp[i], acc);
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (part == 0 && k < m) vec[k] = acc * L.sc[k];
  }
  __syncthreads();
}

// q = A D A^T p (FP64, matrix-free); p, q global (m).
__device__ void normal_matvec(const Lp& L, const double* p, double* q, double* stage) {
  gemv_t(L, p, L.t, stage);
  for (int j = threadIdx.x; j < L.n; j += kThreads) L.t[j] *= L.D[j];
  __syncthreads();
  gemv_n(L, L.t, nullptr, q, stage);
}

// dy = M^-1 rhs by PCG. Returns iterations.
template <typename T>
__device__ int solve_normal(const Lp& L, const T* M, double* stage, double* red, double tol, int max_it) {
  const int tid = threadIdx.x, m = L.m;
  double* pv = stage;       // [0, m)  preconditioner vector (LDS)
  double* tv = stage + 512;  // [512, 512+m) preconditioner temp (LDS)
  // x0 = P rhs
  for (int i = tid; i < m; i += kThreads) pv[i] = L.rhs[i];
  precondition(L, M, pv, tv);
  for (int i = tid; i < m; i += kThreads) L.dy[i] = pv[i];
  __syncthreads();
  normal_matvec(L, L.dy, L.q, stage);
  double rn = 0.0, bn = 0.0;
  for (int i = tid; i < m; i += kThreads) {
    const double r = L.rhs[i] - L.q[i];
    L.r[i] = r;
    rn = fmax(rn, fabs(r));
    bn = fmax(bn, fabs(L.rhs[i]));
  }
  rn = bred<1>(rn, red);
  bn = bred<1>(bn, red);
  if (!(rn > tol * fmax(bn, 1e-300))) return 0;
  for (int i = tid; i < m; i += kThreads) pv[i] = L.r[i];
  precondition(L, M, pv, tv);
  double rz = 0.0;
  for (int i = tid; i < m; i += kThreads) {
    L.zz[i] = pv[i];
    L.p[i] = pv[i];
    rz += L.r[i] * pv[i];
  }
  rz = bred<0>(rz, red);
  int it = 0;
  for (it = 1; it <= max_it; ++it) {
    __syncthreads();
    normal_matvec(L, L.p, L.q, stage);
    double pq = 0.0;
    for (int i = tid; i < m; i += kThreads) pq += L.p[i] * L.q[i];
    pq = bred<0>(pq, red);
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    rn = 0.0;
    for (int i = tid; i < m; i += kThreads) {
      L.dy[i] += alpha * L.p[i];
      const double r = L.r[i] - alpha * L.q[i];
      L.r[i] = r;
      rn = fmax(rn, fabs(r));
    }
    rn = bred<1>(rn, red);
    if (!(rn > tol * bn) || !isfinite(rn)) break;
    for (int i = tid; i < m; i += kThreads) pv[i] = L.r[i];
    precondition(L, M, pv, tv);
    double rz2 = 0.0;
    for (int i = tid; i < m; i += kThreads) rz2 += L.r[i] * pv[i];
    rz2 = bred<0>(rz2, red);
    const double beta = rz != 0.0 ? rz2 / rz : 0.0;
    rz = rz2;
    for (int i = tid; i < m; i += kThreads) L.p[i] = pv[i] + beta * L.p[i];
  }
  __syncthreads();
  return it;
}

__device__ __forceinline__ bool active(const Lp& L, int j) { return L.u[j] > 0.0; }
__device__ __forceinline__ bool bounded(const Lp& L, int j) { return L.u[j] < 1e29; }

// Newton direction for complementarity right-hand sides r1 (x z) and r2 (w v),
// given the current residuals; writes (dx, dw, dz, dv) and dy.
template <typename T>
__device__ int newton(const Lp& L, const T* M, double* stage, double* red, const double* r1, const double* r2,
                      double* dx, double* dw, double* dz, double* dv, double pcg_tol, const BatchOpts& o) {
  const int tid = threadIdx.x;
  for (int j = tid; j < L.n; j += kThreads) {
    double tj = 0.0;
    if (active(L, j)) {
      double g = L.rc[j] - r1[j] / L.x[j];
      if (bounded(L, j)) g += (r2[j] - L.v[j] * L.ru[j]) / L.w[j];
      tj = L.D[j] * g;
    }
    L.t[j] = tj;
  }
  __syncthreads();
  gemv_n(L, L.t, L.rb, L.rhs, stage);
  // keep t: gemv_t / normal_matvec overwrite L.t, so stash it in dx first
  for (int j = tid; j < L.n; j += kThreads) dx[j] = L.t[j];
  __syncthreads();
  const int its = solve_normal(L, M, stage, red, pcg_tol, o.pcg_max);
  gemv_t(L, L.dy, L.t, stage);  // t = A^T dy
  for (int j = tid; j < L.n; j += kThreads) {
    if (!active(L, j)) {
      dx[j] = dw[j] = dz[j] = dv[j] = 0.0;
      continue;
    }
    const double dxj = L.D[j] * L.t[j] - dx[j];
    dx[j] = dxj;
    dz[j] = (r1[j] - L.z[j] * dxj) / L.x[j];
    if (bounded(L, j)) {
      const double dwj = L.ru[j] - dxj;
      dw[j] = dwj;
      dv[j] = (r2[j] - L.v[j] * dwj) / L.w[j];
    } else {
      dw[j] = dv[j] = 0.0;
    }
  }
  __syncthreads();
  return its;
}

// Largest step in [0, 1] keeping (x, w) and (z, v) nonnegative.
__device__ void max_steps(const Lp& L, const double* dx, const double* dw, const double* dz, const double* dv,
                          double* red, double* ap, double* ad) {
  double a = 1.0, b = 1.0;
  for (int j = threadIdx.x; j < L.n; j += kThreads) {
    if (!active(L, j)) continue;
    if (dx[j] < 0.0) a = fmin(a, -L.x[j] / dx[j]);
    if (dz[j] < 0.0) b = fmin(b, -L.z[j] / dz[j]);
    if (bounded(L, j)) {
      if (dw[j] < 0.0) a = fmin(a, -L.w[j] / dw[j]);
      if (dv[j] < 0.0) b = fmin(b, -L.v[j] / dv[j]);
    }
  }
  *ap = bred<2>(a, red);
  *ad = bred<2>(b, red);
}

__global__ __launch_bounds__(kThreads) void batch_ipm_kernel(const BatchMatrix* __restrict__ mats,
                                                             const BatchDesc* __restrict__ descs, int count,
                                                             unsigned* queue, const double* __restrict__ b_all,
                                                             const double* __restrict__ c_all,
                                                             const double* __restrict__ u_all, double* ws,
                                                             double* x_all, sov_batch_result_dev* out, BatchOpts o,
                                                             double* trace, int trace_lp) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* stage = reinterpret_cast<double*>(smem);           // kStage doubles
  double* red = stage + kStage;                              // 64 doubles
  float* M = reinterpret_cast<float*>(red + 64);             // packed lower triangle
  __shared__ int s_lp;
  __shared__ Lp s_L;
  const int tid = threadIdx.x;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_lp = static_cast<int>(atomicAdd(queue, 1u));
    __syncthreads();
    const int lp = s_lp;
    if (lp >= count) break;
    const BatchDesc dsc = descs[lp];
    const BatchMatrix mat = mats[dsc.mat];
    // The LP's sizes and pointers live in LDS: every thread reads the same
    // values, and keeping them out of registers leaves those to the loops.
    __syncthreads();
    if (tid == 0) {
      Lp& W = s_L;
      W.m = mat.m;
      W.n = mat.n;
      W.A = mat.a64;
      W.Af = mat.a32;
      W.b = b_all + dsc.b_off;
      W.c = c_all + dsc.x_off;
      W.u = u_all + dsc.x_off;
      double* wsp = ws + dsc.ws_off;
      double** nv[] = {&W.x, &W.w, &W.z, &W.v, &W.rc, &W.ru, &W.D, &W.dx, &W.dw, &W.dz, &W.dv,
                       &W.dxa, &W.dwa, &W.dza, &W.dva, &W.t};
      for (double** pp : nv) {
        *pp = wsp;
        wsp += W.n;
      }
      double** mv[] = {&W.y, &W.rb, &W.dy, &W.r, &W.p, &W.q, &W.zz, &W.rhs, &W.sc};
      for (double** pp : mv) {
        *pp = wsp;
        wsp += W.m;
      }
      W.M64 = wsp;  // m (m + 1) / 2
    }
    __syncthreads();
    const Lp& L = s_L;
    const int m = L.m, n = L.n;
    // Starting point: x mid-box (or 1), w = u - x, z = v = 1, y = 0.
    double nb = 0.0;
    for (int j = tid; j < n; j += kThreads) {
      const double uj = L.u[j];
      double xj = 0.0, wj = 0.0, vj = 0.0, zj = 0.0;
      if (uj > 0.0) {
        xj = uj < 1e29 ? 0.5 * uj : 1.0;
        wj = uj < 1e29 ? uj - xj : 0.0;
        zj = 1.0;
        vj = uj < 1e29 ? 1.0 : 0.0;
        nb += uj < 1e29 ? 1.0 : 0.0;
      }
      L.x[j] = xj;
      L.w[j] = wj;
      L.z[j] = zj;
      L.v[j] = vj;
    }
    for (int i = tid; i < m; i += kThreads) L.y[i] = 0.0;
    double nact = 0.0;
    for (int j = tid; j < n; j += kThreads) nact += active(L, j) ? 1.0 : 0.0;
    nact = bred<0>(nact, red);
    nb = bred<0>(nb, red);
    double bnorm = 0.0, cnorm = 0.0;
    for (int i = tid; i < m; i += kThreads) bnorm = fmax(bnorm, fabs(L.b[i]));
    for (int j = tid; j < n; j += kThreads) cnorm = fmax(cnorm, fabs(L.c[j]));
    bnorm = bred<1>(bnorm, red);
    cnorm = bred<1>(cnorm, red);

    int status = 5, iters = 0, pcg_total = 0;
    double pobj = 0.0, gap = 0.0, pres = 0.0, dres = 0.0;
    // optional per-iteration record of one LP (kTraceCols doubles a row)
    double* tr = (trace && lp == trace_lp) ? trace : nullptr;
    auto note = [&](int it, int col, double v) {
      if (tr && tid == 0 && it < kTraceRows) tr[it * kTraceCols + col] = v;
    };
    int stalled = 0;  // consecutive FP64 iterations whose two PCG solves both hit the cap
    bool fp64 = false;  // FP64 factorizations (global memory) after the FP32 solves stalled
    int fp64_from = -1;
    double best = 1e300;
    int best_it = 0;
    for (int it = 0; it < o.max_iter; ++it) {
      __syncthreads();
      long long t0 = clock64();
      // residuals
      gemv_n(L, L.x, nullptr, L.rb, stage);  // rb = A x (then b - A x)
      gemv_t(L, L.y, L.t, stage);            // t = A^T y
      double rbmax = 0.0, rcmax = 0.0, rumax = 0.0, cx = 0.0, by = 0.0, uv = 0.0, xz = 0.0;
      for (int i = tid; i < m; i += kThreads) {
        const double r = L.b[i] - L.rb[i];
        L.rb[i] = r;
        rbmax = fmax(rbmax, fabs(r));
        by += L.b[i] * L.y[i];
      }
      for (int j = tid; j < n; j += kThreads) {
        if (!active(L, j)) {
          L.rc[j] = 0.0;
          L.ru[j] = 0.0;
          continue;
        }
        const double rcj = L.c[j] - L.t[j] - L.z[j] + L.v[j];
        L.rc[j] = rcj;
        rcmax = fmax(rcmax, fabs(rcj));
        cx += L.c[j] * L.x[j];
        xz += L.x[j] * L.z[j];
        if (bounded(L, j)) {
          const double ruj = L.u[j] - L.x[j] - L.w[j];
          L.ru[j] = ruj;
          rumax = fmax(rumax, fabs(ruj));
          uv += L.u[j] * L.v[j];
          xz += L.w[j] * L.v[j];
        } else {
          L.ru[j] = 0.0;
        }
      }
      rbmax = bred<1>(rbmax, red);
      rcmax = bred<1>(rcmax, red);
      rumax = bred<1>(rumax, red);
      cx = bred<0>(cx, red);
      by = bred<0>(by, red);
      uv = bred<0>(uv, red);
      xz = bred<0>(xz, red);
      pobj = cx;
      const double dobj = by - uv;
      pres = fmax(rbmax / (1.0 + bnorm), rumax);
      dres = rcmax / (1.0 + cnorm);
      gap = fabs(pobj - dobj) / (1.0 + fabs(pobj));
      iters = it;
      note(it, 0, it);
      note(it, 1, pres);
      note(it, 2, dres);
      note(it, 3, gap);
      if (pres < o.tol && dres < o.tol && gap < o.tol) {
        status = 0;
        break;
      }
      if (!isfinite(pobj) || !isfinite(dobj)) {
        status = 6;
        break;
      }
      // When an FP32-preconditioned solve hits its cap (below, where the
      // iteration is then redone) or 30 iterations pass without progress the
      // LP switches to FP64 factorizations; stalling in FP64 (both solves at
      // the cap in three iterations) ends it with status 6, which the caller
      // finishes on the per-LP path.
      const double metric = fmax(pres, fmax(dres, gap));
      if (metric < best) {
        best = metric;
        best_it = it;
      }
      if ((fp64 ? stalled >= 3 : stalled >= 1) || it - best_it > 30) {
        if (fp64) {
          status = 6;
          break;
        }
        // The FP32 factor no longer preconditions well enough: from here on
        // this LP's normal matrix is factored in FP64 (global memory).
        fp64 = true;
        fp64_from = it;
        stalled = 0;
        best_it = it;
        note(it, 15, -1.0);
      }
      const double mu = xz / fmax(nact + nb, 1.0);
      note(it, 4, mu);
      // scaling D = (z/x + v/w)^-1
      for (int j = tid; j < n; j += kThreads) {
        double dj = 0.0;
        if (active(L, j)) {
          double g = L.z[j] / L.x[j];
          if (bounded(L, j)) g += L.v[j] / L.w[j];
          dj = 1.0 / g;
        }
        L.D[j] = dj;
      }
      __syncthreads();
      long long t1 = clock64();
      note(it, 9, static_cast<double>(t1 - t0));
      double ap, ad;
      // Factor and solve; an FP32-preconditioned solve that hits the CG cap
      // is not used: the LP switches to FP64 factorizations and the
      // iteration is redone (a step along the unconverged direction can undo
      // the primal feasibility of a nearly solved LP).
      for (;;) {
        if (fp64) assemble(L, L.M64, 0.0);
        else assemble(L, M, 0.0);
        t0 = clock64();
        note(it, 10, static_cast<double>(t0 - t1));
        note(it, 15, fp64 ? cholesky(L, L.M64) : cholesky(L, M));
        t1 = clock64();
        note(it, 11, static_cast<double>(t1 - t0));
        if (fp64) invert_lower(L, L.M64, stage);
        else invert_lower(L, M, stage);
        // inexact Newton while mu is large: CG relative residual max(pcg_tol, min(1e-6, 1e-3 mu))
        const double ptol = fp64 ? o.pcg_tol : fmax(o.pcg_tol, fmin(1e-6, 1e-3 * mu));
        t0 = clock64();
        note(it, 12, static_cast<double>(t0 - t1));
        // predictor: r1 = -x z, r2 = -w v (in dxa/dwa scratch: use dza/dva as rhs storage)
        for (int j = tid; j < n; j += kThreads) {
          L.dza[j] = -L.x[j] * L.z[j];
          L.dva[j] = -L.w[j] * L.v[j];
        }
        __syncthreads();
        // r1 -> p-buffer copies: newton reads r1/r2 after writing dz/dv, so pass copies.
        for (int j = tid; j < n; j += kThreads) {
          L.dxa[j] = L.dza[j];
          L.dwa[j] = L.dva[j];
        }
        __syncthreads();
        const int k1 = fp64 ? newton(L, L.M64, stage, red, L.dxa, L.dwa, L.dx, L.dw, L.dz, L.dv, ptol, o)
                            : newton(L, M, stage, red, L.dxa, L.dwa, L.dx, L.dw, L.dz, L.dv, ptol, o);
        pcg_total += k1;
        note(it, 7, k1);
        t1 = clock64();
        note(it, 13, static_cast<double>(t1 - t0));
        if (!fp64 && k1 >= o.pcg_max) {
          fp64 = true;
          fp64_from = it;
          stalled = 0;
          best_it = it;
          note(it, 15, -1.0);
          continue;
        }
        max_steps(L, L.dx, L.dw, L.dz, L.dv, red, &ap, &ad);
        double mua = 0.0;
        for (int j = tid; j < n; j += kThreads) {
          if (!active(L, j)) continue;
          mua += (L.x[j] + ap * L.dx[j]) * (L.z[j] + ad * L.dz[j]);
          if (bounded(L, j)) mua += (L.w[j] + ap * L.dw[j]) * (L.v[j] + ad * L.dv[j]);
        }
        mua = bred<0>(mua, red) / fmax(nact + nb, 1.0);
        const double sigma = mu > 0.0 ? fmin(1.0, pow(mua / mu, 3.0)) : 0.0;
        // corrector: r1 = sigma mu - x z - dx_a dz_a, r2 = sigma mu - w v - dw_a dv_a
        for (int j = tid; j < n; j += kThreads) {
          L.dxa[j] = sigma * mu - L.x[j] * L.z[j] - L.dx[j] * L.dz[j];
          L.dwa[j] = bounded(L, j) ? sigma * mu - L.w[j] * L.v[j] - L.dw[j] * L.dv[j] : 0.0;
        }
        __syncthreads();
        const int k2 = fp64 ? newton(L, L.M64, stage, red, L.dxa, L.dwa, L.dx, L.dw, L.dz, L.dv, ptol, o)
                            : newton(L, M, stage, red, L.dxa, L.dwa, L.dx, L.dw, L.dz, L.dv, ptol, o);
        pcg_total += k2;
        note(it, 8, k2);
        t0 = clock64();
        note(it, 14, static_cast<double>(t0 - t1));
        if (!fp64 && k2 >= o.pcg_max) {
          fp64 = true;
          fp64_from = it;
          stalled = 0;
          best_it = it;
          note(it, 15, -1.0);
          continue;
        }
        stalled = (k2 >= o.pcg_max && k1 >= o.pcg_max) ? stalled + 1 : 0;
        break;
      }
      max_steps(L, L.dx, L.dw, L.dz, L.dv, red, &ap, &ad);
      ap = fmin(1.0, o.tau * ap);
      ad = fmin(1.0, o.tau * ad);
      note(it, 5, ap);
      note(it, 6, ad);
      for (int j = tid; j < n; j += kThreads) {
        if (!active(L, j)) continue;
        L.x[j] = fmax(L.x[j] + ap * L.dx[j], 1e-14);
        L.z[j] = fmax(L.z[j] + ad * L.dz[j], 1e-14);
        if (bounded(L, j)) {
          L.w[j] = fmax(L.w[j] + ap * L.dw[j], 1e-14);
          L.v[j] = fmax(L.v[j] + ad * L.dv[j], 1e-14);
        }
      }
      for (int i = tid; i < m; i += kThreads) L.y[i] += ad * L.dy[i];
      iters = it + 1;
    }
    __syncthreads();
    if (x_all)
      for (int j = tid; j < n; j += kThreads) x_all[dsc.x_off + j] = L.x[j];
    if (tid == 0) {
      sov_batch_result_dev r;
      r.status = status;
      r.iterations = iters;
      r.pcg_iterations = pcg_total;
      r.fp64_from = fp64_from;
      r.objective = pobj;
      r.duality_gap = gap;
      r.primal_residual = pres;
      r.dual_residual = dres;
      out[lp] = r;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side: matrices are uploaded once (FP64 and an FP32 copy for MFMA);
// each solve ships the per-LP vectors and runs one launch.
// ---------------------------------------------------------------------------
struct sov_batch {
  std::vector<BatchMatrix> mats;      // device pointers
  std::vector<void*> owned;           // allocations to free
  float last_ms = 0.f;
  ~sov_batch() {
    for (void* p : owned) (void)hipFree(p);
  }
};

namespace {

std::size_t lds_bytes(int mmax) {
  return sizeof(double) * (kStage + 64) + sizeof(float) * static_cast<std::size_t>(mmax) * (mmax + 1) / 2;
}

template <typename T>
T* dev_alloc(sov_batch* h, std::size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<std::size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
  h->owned.push_back(p);
  return static_cast<T*>(p);
}

// The pair of timing events of one call, destroyed on every return path.
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// Device buffers owned by one call.
struct TmpBuffers {
  std::vector<void*> p;
  ~TmpBuffers() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  T* get(std::size_t cnt) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<std::size_t>(cnt, 1) * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(q);
    return static_cast<T*>(q);
  }
};

}  // namespace

extern "C" {

int sov_batch_create(sov_batch** out) {
  if (!out) return SOV_ERR_ARG;
  *out = nullptr;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  static const bool lds_ok =
      hipFuncSetAttribute(reinterpret_cast<const void*>(batch_ipm_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          static_cast<int>(lds_bytes(kMaxRows))) == hipSuccess;
  if (!lds_ok) {
    sov::set_error("sov_batch_create: could not raise the LDS limit");
    return SOV_ERR_DEVICE;
  }
  *out = new sov_batch;
  return SOV_OK;
}

int sov_batch_add_matrix(sov_batch* h, int32_t m, int32_t n, const double* a_colmajor, int32_t* id) {
  if (!h || !a_colmajor || !id || m <= 0 || n <= 0) return SOV_ERR_ARG;
  if (m > kMaxRows) {
    sov::set_error("sov_batch_add_matrix: at most 272 rows per batched LP");
    return SOV_ERR_ARG;
  }
  const std::size_t cnt = static_cast<std::size_t>(m) * n;
  std::vector<float> af(cnt);
  for (std::size_t k = 0; k < cnt; ++k) af[k] = static_cast<float>(a_colmajor[k]);
  double* a64 = dev_alloc<double>(h, cnt);
  float* a32 = dev_alloc<float>(h, cnt);
  if (!a64 || !a32) return SOV_ERR_NOMEM;
  SOV_HIP_TRY(hipMemcpy(a64, a_colmajor, cnt * sizeof(double), hipMemcpyHostToDevice));
  SOV_HIP_TRY(hipMemcpy(a32, af.data(), cnt * sizeof(float), hipMemcpyHostToDevice));
  h->mats.push_back({m, n, a64, a32});
  *id = static_cast<int32_t>(h->mats.size() - 1);
  return SOV_OK;
}

int sov_batch_solve(sov_batch* h, int32_t count, const int32_t* matrix_id, const double* b_all,
                    const int64_t* b_off, const double* c_all, const double* u_all, const int64_t* x_off,
                    const sov_ipm_options* opt, sov_batch_result* res, double* x_all) {
  if (!h || count < 0 || !matrix_id || !b_all || !b_off || !c_all || !u_all || !x_off || !opt || !res)
    return SOV_ERR_ARG;
  if (count == 0) return SOV_OK;
  hipStream_t s = sov::stream();
  std::vector<BatchDesc> descs(static_cast<std::size_t>(count));
  long long b_total = 0, x_total = 0, ws_total = 0;
  int mmax = 0;
  for (int k = 0; k < count; ++k) {
    const int id = matrix_id[k];
    if (id < 0 || id >= static_cast<int>(h->mats.size())) return SOV_ERR_ARG;
    if (b_off[k] < 0 || x_off[k] < 0) {
      sov::set_error("sov_batch_solve: negative b_off / x_off");
      return SOV_ERR_ARG;
    }
    const BatchMatrix& mt = h->mats[id];
    mmax = std::max(mmax, mt.m);
    descs[k] = {id, 0, b_off[k], x_off[k], ws_total};
    b_total = std::max<long long>(b_total, b_off[k] + mt.m);
    x_total = std::max<long long>(x_total, x_off[k] + mt.n);
    ws_total += 16LL * mt.n + 9LL * mt.m + static_cast<long long>(mt.m) * (mt.m + 1) / 2;
  }
  TmpBuffers tmp;  // per-solve device buffers, freed at the end of the call
  auto* d_mats = tmp.get<BatchMatrix>(h->mats.size());
  auto* d_desc = tmp.get<BatchDesc>(descs.size());
  auto* d_b = tmp.get<double>(static_cast<std::size_t>(b_total));
  auto* d_c = tmp.get<double>(static_cast<std::size_t>(x_total));
  auto* d_u = tmp.get<double>(static_cast<std::size_t>(x_total));
  auto* d_ws = tmp.get<double>(static_cast<std::size_t>(ws_total));
  auto* d_x = x_all ? tmp.get<double>(static_cast<std::size_t>(x_total)) : nullptr;
  auto* d_out = tmp.get<sov_batch_result_dev>(static_cast<std::size_t>(count));
  auto* d_queue = tmp.get<unsigned>(1);
  if (!d_mats || !d_desc || !d_b || !d_c || !d_u || !d_ws || (x_all && !d_x) || !d_out || !d_queue)
    return SOV_ERR_NOMEM;
  SOV_HIP_TRY(hipMemcpyAsync(d_mats, h->mats.data(), h->mats.size() * sizeof(BatchMatrix), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemcpyAsync(d_desc, descs.data(), descs.size() * sizeof(BatchDesc), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemcpyAsync(d_b, b_all, b_total * sizeof(double), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemcpyAsync(d_c, c_all, x_total * sizeof(double), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemcpyAsync(d_u, u_all, x_total * sizeof(double), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemsetAsync(d_queue, 0, sizeof(unsigned), s));
  BatchOpts o;
  o.tol = std::min(opt->feasibility_tol, opt->optimality_tol);
  o.tau = opt->fraction_to_boundary;
  o.max_iter = opt->max_iterations > 0 ? std::min(opt->max_iterations, 200) : 100;
  o.pcg_tol = opt->refine_tol > 0 ? opt->refine_tol : 1e-12;
  o.pcg_max = opt->refine_max > 0 ? opt->refine_max : 30;
  int dev = 0, cus = 0;
  SOV_HIP_TRY(hipGetDevice(&dev));
  SOV_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = static_cast<unsigned>(std::min(count, std::max(cus, 1)));
  EventPair ev;
  SOV_HIP_TRY(hipEventCreate(&ev.e0));
  SOV_HIP_TRY(hipEventCreate(&ev.e1));
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  (void)hipEventRecord(e0, s);
  // SOVEREIGN_BATCH_TRACE=k records the iterations of LP k (stderr).
  const char* tenv = std::getenv("SOVEREIGN_BATCH_TRACE");
  const int trace_lp = tenv ? std::atoi(tenv) : -1;
  double* d_trace = nullptr;
  if (trace_lp >= 0 && trace_lp < count) {
    d_trace = tmp.get<double>(kTraceRows * kTraceCols);
    if (!d_trace) return SOV_ERR_NOMEM;
    SOV_HIP_TRY(hipMemsetAsync(d_trace, 0, sizeof(double) * kTraceRows * kTraceCols, s));
  }
  batch_ipm_kernel<<<grid, kThreads, lds_bytes(mmax), s>>>(d_mats, d_desc, count, d_queue, d_b, d_c, d_u, d_ws, d_x,
                                                          d_out, o, d_trace, trace_lp);
  const hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, s);
  if (le != hipSuccess) {
    sov::set_error(std::string("batch_ipm_kernel launch: ") + hipGetErrorString(le));
    return SOV_ERR_DEVICE;
  }
  std::vector<sov_batch_result_dev> host(static_cast<std::size_t>(count));
  SOV_HIP_TRY(hipMemcpyAsync(host.data(), d_out, count * sizeof(sov_batch_result_dev), hipMemcpyDeviceToHost, s));
  if (x_all) SOV_HIP_TRY(hipMemcpyAsync(x_all, d_x, x_total * sizeof(double), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  (void)hipEventElapsedTime(&h->last_ms, e0, e1);
  if (d_trace) {
    std::vector<double> t(kTraceRows * kTraceCols);
    SOV_HIP_TRY(hipMemcpy(t.data(), d_trace, sizeof(double) * t.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < kTraceRows; ++r) {
      const double* q = t.data() + r * kTraceCols;
      if (r > 0 && q[0] == 0.0) break;
      std::fprintf(stderr,
                   "[batch %d] it=%d pres=%.2e dres=%.2e gap=%.2e mu=%.2e ap=%.3f ad=%.3f pcg=%d/%d drop=%d "
                   "cycles res=%.0f asm=%.0f chol=%.0f inv=%.0f newton=%.0f/%.0f\n",
                   trace_lp, static_cast<int>(q[0]), q[1], q[2], q[3], q[4], q[5], q[6], static_cast<int>(q[7]),
                   static_cast<int>(q[8]), static_cast<int>(q[15]), q[9], q[10], q[11], q[12], q[13], q[14]);
    }
  }
  for (int k = 0; k < count; ++k) res[k] = host[k];
  sov::count_operation();
  return SOV_OK;
}

double sov_batch_last_kernel_ms(const sov_batch* h) { return h ? h->last_ms : -1.0; }

int sov_batch_free(sov_batch* h) {
  delete h;
  return SOV_OK;
}

}  // extern "C"
