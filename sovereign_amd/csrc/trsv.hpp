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

constexpr int kGemvRows = 16;     // rows of y per workgroup
constexpr int kGemvPhases = 64;   // threads per row, splitting the columns
constexpr int kGemvInFlight = 8;  // loads a thread issues before it uses them

// y[r] = s (SUB: y[r] -= s), s = sum_{k < K} A(r, k) x[k], r < M; A column-
// major. TRI: A is lower triangular (columns k > r are skipped). K <= kSuper.
// A workgroup owns 16 rows; lane & 15 is the row and the workgroup's 64
// column phases (four per wavefront) take every 64th column, kGemvInFlight
// loads in flight each, so that a thread's chain of dependent load rounds is
// short (the products are latency-bound at these sizes: 1024 columns are 16
// per thread, two rounds). The 64 partial sums of a row are added in a fixed
// order (deterministic).
template <typename T, bool SUB, bool TRI>
__global__ __launch_bounds__(kGemvRows * kGemvPhases) void gemv_n_kernel(int M, int K, const T* __restrict__ A,
                                                                          long long lda,
                                                                          const double* __restrict__ x,
                                                                          double* __restrict__ y) {
  __shared__ double xs[kSuper];
  __shared__ double part[kGemvPhases][kGemvRows];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int row = lane & 15, ph = (lane >> 4) + 4 * w;  // phase 0 .. kGemvPhases - 1
  const int r0 = blockIdx.x * kGemvRows;
  const int kend = TRI ? min(K, r0 + kGemvRows) : K;
  for (int k = t; k < kend; k += kGemvRows * kGemvPhases) xs[k] = x[k];
  __syncthreads();
  const int r = r0 + row;
  double s0 = 0.0, s1 = 0.0;
  if (r < M) {
    const T* col = A + r;
    constexpr int P = kGemvPhases, F = kGemvInFlight;
    int k = ph;
    for (; k + (F - 1) * P < kend; k += F * P) {
      T v[F];
#pragma unroll
      for (int q = 0; q < F; ++q) v[q] = col[static_cast<long long>(k + P * q) * lda];
#pragma unroll
      for (int q = 0; q < F; q += 2) {
        s0 = fma(static_cast<double>(v[q]), xs[k + P * q], s0);
        s1 = fma(static_cast<double>(v[q + 1]), xs[k + P * (q + 1)], s1);
      }
    }
    for (; k < kend; k += P) s0 = fma(static_cast<double>(col[static_cast<long long>(k) * lda]), xs[k], s0);
  }
  part[ph][row] = s0 + s1;
  __syncthreads();
  if (t < kGemvRows && r0 + t < M) {
    double s = 0.0;
#pragma unroll 8
    for (int q = 0; q < kGemvPhases; ++q) s += part[q][t];
    y[r0 + t] = SUB ? y[r0 + t] - s : s;
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
  for (; r + 448 < M; r += 512) {
    T v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = col[r + 64 * q];
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
      s0 = fma(static_cast<double>(v[q]), xs[r + 64 * q], s0);
      s1 = fma(static_cast<double>(v[q + 1]), xs[r + 64 * (q + 1)], s1);
    }
  }
  for (; r < M; r += 64) s0 = fma(static_cast<double>(col[r]), xs[r], s0);
  const double t = wave_sum(s0 + s1);
  if (lane == 0) y[c] = SUB ? y[c] - t : t;
}
