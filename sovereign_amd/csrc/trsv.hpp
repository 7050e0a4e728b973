// Blocked triangular solves with the Cholesky factor (included by chol.hip
// inside its anonymous namespace).
//
// After a factorization the inverse of every kSuper x kSuper diagonal block
// of L is formed explicitly (recursive doubling on MFMA, from the panel
// inverses). A solve is then a sequence of matrix-vector products, two per
// super-block and direction:
//   forward  y_P = X_P c_P;            c_below -= L(below, P) y_P
//   backward x_P = X_P^T c_P;          c_above -= L(P, above)^T x_P
// The products are HBM-bound: each reads its block of L (or X_P) once, in
// the factor's precision, and accumulates in FP64.
#pragma once

// Expects to be included inside namespace sov (anonymous namespace) of chol.hip.

constexpr int kGemvRows = 64;    // rows of y per workgroup (one per lane)
constexpr int kGemvWaves = 16;   // wavefronts per workgroup, splitting the columns

// y[r] = s (SUB: y[r] -= s), s = sum_{k < K} A(r, k) x[k], r < M; A column-
// major. TRI: A is lower triangular (columns k > r are skipped). K <= kSuper.
// A workgroup owns 64 rows; its 16 wavefronts take every 16th column (many
// loads in flight per CU: the products are latency-bound at these sizes)
// and their partial sums are added in a fixed order (deterministic).
template <typename T, bool SUB, bool TRI>
__global__ __launch_bounds__(64 * kGemvWaves) void gemv_n_kernel(int M, int K, const T* __restrict__ A,
                                                                 long long lda, const double* __restrict__ x,
                                                                 double* __restrict__ y) {
  __shared__ double xs[kSuper];
  __shared__ double part[kGemvWaves][kGemvRows];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = blockIdx.x * kGemvRows;
  const int kend = TRI ? min(K, r0 + kGemvRows) : K;
  for (int k = threadIdx.x; k < kend; k += 64 * kGemvWaves) xs[k] = x[k];
  __syncthreads();
  const int r = r0 + lane;
  double s0 = 0.0, s1 = 0.0;
  if (r < M) {
    const T* col = A + r;
    constexpr int W = kGemvWaves;
    int k = w;
    for (; k + 3 * W < kend; k += 4 * W) {
      T v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = col[static_cast<long long>(k + W * q) * lda];
      s0 = fma(static_cast<double>(v[0]), xs[k], s0);
      s1 = fma(static_cast<double>(v[1]), xs[k + W], s1);
      s0 = fma(static_cast<double>(v[2]), xs[k + 2 * W], s0);
      s1 = fma(static_cast<double>(v[3]), xs[k + 3 * W], s1);
    }
    for (; k < kend; k += W) s0 = fma(static_cast<double>(col[static_cast<long long>(k) * lda]), xs[k], s0);
  }
  part[w][lane] = s0 + s1;
  __syncthreads();
  if (w == 0 && r < M) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < kGemvWaves; ++q) t += part[q][lane];
    y[r] = SUB ? y[r] - t : t;
  }
}

// y[c] = s (SUB: y[c] -= s), s = sum_{r < M} A(r, c) x[r], c < N; one
// wavefront per column, lanes down the rows. TRI: A is lower triangular
// (rows r < c are skipped). M <= kSuper.
template <typename T, bool SUB, bool TRI>
__global__ __launch_bounds__(256) void gemv_t_kernel(int M, int N, const T* __restrict__ A, long long lda,
                                                     const double* __restrict__ x, double* __restrict__ y) {
  __shared__ double xs[kSuper];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int r = threadIdx.x; r < M; r += 256) xs[r] = x[r];
  __syncthreads();
  const int c = blockIdx.x * 4 + w;
  if (c >= N) return;
  const T* col = A + static_cast<long long>(c) * lda;
  int r = TRI ? (c & ~63) + lane : lane;
  double s0 = 0.0, s1 = 0.0;
  for (; r + 192 < M; r += 256) {
    const T v0 = col[r], v1 = col[r + 64], v2 = col[r + 128], v3 = col[r + 192];
    s0 = fma(static_cast<double>(v0), xs[r], s0);
    s1 = fma(static_cast<double>(v1), xs[r + 64], s1);
    s0 = fma(static_cast<double>(v2), xs[r + 128], s0);
    s1 = fma(static_cast<double>(v3), xs[r + 192], s1);
  }
  for (; r < M; r += 64) s0 = fma(static_cast<double>(col[r]), xs[r], s0);
  const double t = wave_sum(s0 + s1);
  if (lane == 0) y[c] = SUB ? y[c] - t : t;
}
