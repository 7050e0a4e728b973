// Blocked triangular solves with the Cholesky factor (included by chol.hip
// inside its anonymous namespace).
//
// The factor is split into kPanel-row blocks whose diagonal inverses the
// panel kernel already produced. Each block step is one launch:
//   forward  y_p = inv(L_pp) c_p;  c_q -= L_qp y_p for the rows below,
//   backward x_p = inv(L_pp)^T c_p; c_q -= L_pq^T x_p for the rows above.
// Every workgroup of a step recomputes the small 128 x 128 product for block
// p (the inverse is read from L2) and then updates its own 64 entries, so a
// step needs no inter-workgroup synchronisation; launches are ordered by the
// stream. The right-hand side is updated in place in `c`; the block results go
// to `out` (a different vector), so no workgroup reads what another writes.
// Each thread issues all of its loads before it accumulates, so one step
// costs a few memory latencies rather than one per loop trip.
#pragma once

// Expects to be included inside namespace sov (anonymous namespace) of chol.hip.

constexpr int kSolveRows = 64;  // entries of c updated by one workgroup per step

// Forward step p: out[P] = inv(L_pp) c[P]; c[below] -= L[below, P] out[P].
template <typename T>
__global__ __launch_bounds__(256) void trsv_fwd_step_kernel(const T* __restrict__ a, long long lda, int m, int p,
                                                            const T* __restrict__ pinv, double* __restrict__ c,
                                                            double* __restrict__ out) {
  __shared__ double cp[kPanel];
  __shared__ double part[2][kPanel];
  __shared__ double y[kPanel];
  __shared__ double red[4][kSolveRows];
  const int tid = threadIdx.x;
  const int base = p * kPanel;
  const int nbp = min(kPanel, m - base);
  const int row = tid & (kPanel - 1), half = tid >> 7;
  const T* inv = pinv + static_cast<long long>(p) * kPanel * kPanel;
  T iv[64];
#pragma unroll
  for (int k = 0; k < 64; ++k) iv[k] = inv[row + (half * 64 + k) * kPanel];
  // this workgroup's rows below the block: 64 rows, 4 slices of 32 columns
  const int r = base + nbp + blockIdx.x * kSolveRows + (tid & 63);
  const int slice = tid >> 6;
  T lv[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const int col = slice * 32 + k;
    lv[k] = (r < m && col < nbp) ? a[r + static_cast<long long>(base + col) * lda] : T(0);
  }
  if (tid < kPanel) cp[tid] = tid < nbp ? c[base + tid] : 0.0;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 64; ++k) s += static_cast<double>(iv[k]) * cp[half * 64 + k];
  part[half][row] = s;
  __syncthreads();
  if (tid < kPanel) {
    y[tid] = part[0][tid] + part[1][tid];
    if (blockIdx.x == 0 && tid < nbp) out[base + tid] = y[tid];
  }
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < 32; ++k) t += static_cast<double>(lv[k]) * y[slice * 32 + k];
  red[slice][tid & 63] = t;
  __syncthreads();
  if (tid < kSolveRows) {
    const int rr = base + nbp + blockIdx.x * kSolveRows + tid;
    if (rr < m) c[rr] -= red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
  }
}

// Backward step p: out[P] = inv(L_pp)^T c[P]; c[above] -= L[P, above]^T out[P].
// pinvt holds the transposed inverses, so the product reads them coalesced.
template <typename T>
__global__ __launch_bounds__(256) void trsv_bwd_step_kernel(const T* __restrict__ a, long long lda, int m, int p,
                                                            const T* __restrict__ pinvt, double* __restrict__ c,
                                                            double* __restrict__ out) {
  __shared__ double cp[kPanel];
  __shared__ double part[2][kPanel];
  __shared__ double xp[kPanel];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int base = p * kPanel;
  const int nbp = min(kPanel, m - base);
  // x_k = sum_r inv(r, k) cp_r = sum_r invt(k, r) cp_r; thread (k, half)
  const int k = tid & (kPanel - 1), half = tid >> 7;
  const T* invt = pinvt + static_cast<long long>(p) * kPanel * kPanel;
  T iv[64];
#pragma unroll
  for (int q = 0; q < 64; ++q) iv[q] = invt[k + (half * 64 + q) * kPanel];
  // columns above: 64 per workgroup, wavefront w takes 16 of them; lanes run
  // down the block's rows (two 64-row halves) of each column.
  const int col0 = blockIdx.x * kSolveRows + w * 16;
  T l0[16], l1[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int col = col0 + q;
    const T* colp = a + static_cast<long long>(col) * lda + base;
    const bool ok = col < base;
    l0[q] = (ok && lane < nbp) ? colp[lane] : T(0);
    l1[q] = (ok && 64 + lane < nbp) ? colp[64 + lane] : T(0);
  }
  if (tid < kPanel) cp[tid] = tid < nbp ? c[base + tid] : 0.0;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 64; ++q) s += static_cast<double>(iv[q]) * cp[half * 64 + q];
  part[half][k] = s;
  __syncthreads();
  if (tid < kPanel) {
    xp[tid] = part[0][tid] + part[1][tid];
    if (blockIdx.x == 0 && tid < nbp) out[base + tid] = xp[tid];
  }
  __syncthreads();
  const double x0 = xp[lane], x1 = xp[64 + lane];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const double v = wave_sum(static_cast<double>(l0[q]) * x0 + static_cast<double>(l1[q]) * x1);
    if (lane == 0 && col0 + q < base) c[col0 + q] -= v;
  }
}
