// MFMA tile kernel for the rank-K updates of the dense factorizations:
//
//   C(r, c) += alpha * sum_k A(r, k) * B(c, k)        r < M, c < N, k < K
//
// A(r, k) = A[r + k*lda] (a block of rows of a column-major matrix).
// B(c, k) = B[c + k*ldb] in NT form (Cholesky's L21 * L21^T), or
// B(c, k) = B[k + c*ldb] in NN form (LU's L21 * U12).
// C(r, c) = C[r + c*ldc]. LOWER restricts the update to r >= c (symmetric
// rank-K update of a lower triangle) and launches only tiles on or below the
// diagonal.
//
// Tiling for CDNA4: a 256-thread workgroup owns a 64x64 tile of C as four
// wavefronts of 32x32, each a 2x2 grid of 16x16 MFMA blocks
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32). K is staged through LDS in
// chunks of 32 with a register prefetch of the next chunk. LDS rows are padded
// to 80 elements so the four k-rows one MFMA operand read touches fall in
// different banks. Workgroups are renumbered so each XCD (round-robin
// dispatch over 8) works on a contiguous run of tiles, which share operand
// rows in that XCD's L2.
#pragma once

#include "common.hpp"

namespace sov {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef dbl4 acc_t;
  __device__ static __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Output row held by accumulator register g of a lane in lane group lk
  // (lane >> 4): D(i, j) sits on lane j + 16 (i mod 4), register i / 4.
  __device__ static __forceinline__ int row(int lk, int g) { return 4 * g + lk; }
};

template <>
struct Mfma<float> {
  typedef flt4 acc_t;
  __device__ static __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D(i, j) sits on lane j + 16 (i / 4), register i mod 4.
  __device__ static __forceinline__ int row(int lk, int g) { return 4 * lk + g; }
};

constexpr int kTile = 64;
constexpr int kChunk = 32;
constexpr int kLdsStride = 80;

// Block index with XCD-contiguous renumbering (bijective on the grid).
__device__ __forceinline__ int xcd_swizzle(int b, int grid) {
  const int full = grid & ~7;
  if (b >= full) return b;
  return (b & 7) * (full >> 3) + (b >> 3);
}

// Element offsets of the operands of batch entry (blockIdx.y, blockIdx.z).
struct GemmBatch {
  long long a_y = 0, a_z = 0, b_y = 0, b_z = 0, c_y = 0, c_z = 0;
};

// Tile coordinates of linear tile t: lower triangle (ti >= tj) or full.
__device__ __forceinline__ void tile_coords(int t, bool lower, int tiles_m, int* ti, int* tj) {
  if (lower) {
    int i = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    *ti = i;
    *tj = t - i * (i + 1) / 2;
  } else {
    *ti = t % tiles_m;
    *tj = t / tiles_m;
  }
}

template <typename T, bool NN, bool LOWER, bool STORE = false>
__global__ __launch_bounds__(256) void gemm_update_kernel(int M, int N, int K, const T* __restrict__ A,
                                                          long long lda, const T* __restrict__ B,
                                                          long long ldb, T* __restrict__ C,
                                                          long long ldc, T alpha, GemmBatch bt) {
  typedef typename Mfma<T>::acc_t acc_t;
  A += blockIdx.y * bt.a_y + blockIdx.z * bt.a_z;
  B += blockIdx.y * bt.b_y + blockIdx.z * bt.b_z;
  C += blockIdx.y * bt.c_y + blockIdx.z * bt.c_z;
  __shared__ T As[kChunk * kLdsStride];
  __shared__ T Bs[kChunk * kLdsStride];

  const int tiles_m = (M + kTile - 1) / kTile;
  int ti, tj;
  tile_coords(xcd_swizzle(blockIdx.x, gridDim.x), LOWER, tiles_m, &ti, &tj);
  const int r0 = ti * kTile, c0 = tj * kTile;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
  const int lrow = lane & 15, lk = lane >> 4;

  acc_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // Global -> register staging of one chunk: 8 elements of A and of B per thread.
  T ra[8], rb[8];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int r = tid & 63, k = (tid >> 6) + 4 * q;
      const int gr = r0 + r, gk = k0 + k;
      ra[q] = (gr < M && gk < K) ? A[gr + static_cast<long long>(gk) * lda] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (NN) {
        const int k = tid & 31, c = (tid >> 5) + 8 * q;
        const int gc = c0 + c, gk = k0 + k;
        rb[q] = (gc < N && gk < K) ? B[gk + static_cast<long long>(gc) * ldb] : T(0);
      } else {
        const int c = tid & 63, k = (tid >> 6) + 4 * q;
        const int gc = c0 + c, gk = k0 + k;
        rb[q] = (gc < N && gk < K) ? B[gc + static_cast<long long>(gk) * ldb] : T(0);
      }
    }
  };
  auto store_chunk = [&]() {
#pragma unroll
    for (int q = 0; q < 8; ++q) As[((tid >> 6) + 4 * q) * kLdsStride + (tid & 63)] = ra[q];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (NN) Bs[(tid & 31) * kLdsStride + (tid >> 5) + 8 * q] = rb[q];
      else Bs[((tid >> 6) + 4 * q) * kLdsStride + (tid & 63)] = rb[q];
    }
  };

  load_chunk(0);
  for (int k0 = 0; k0 < K; k0 += kChunk) {
    __syncthreads();
    store_chunk();
    __syncthreads();
    if (k0 + kChunk < K) load_chunk(k0 + kChunk);
#pragma unroll
    for (int kk = 0; kk < kChunk; kk += 4) {
      const int krow = (kk + lk) * kLdsStride;
      const T a0 = As[krow + wr + lrow], a1 = As[krow + wr + 16 + lrow];
      const T b0 = Bs[krow + wc + lrow], b1 = Bs[krow + wc + 16 + lrow];
      // The B tile goes in as the instruction's first operand, so each
      // 16 x 16 result comes out transposed: a lane holds one row of C.
      acc[0][0] = Mfma<T>::step(b0, a0, acc[0][0]);
      acc[0][1] = Mfma<T>::step(b1, a0, acc[0][1]);
      acc[1][0] = Mfma<T>::step(b0, a1, acc[1][0]);
      acc[1][1] = Mfma<T>::step(b1, a1, acc[1][1]);
    }
  }

  // Operands X(i, k) and Y(k, j) of an instruction sit on lanes i + 16 k and
  // j + 16 k; result (i, j) on lane j + 16 (...), register g with i =
  // Mfma<T>::row(lane >> 4, g). Here X = B-tile rows (C columns) and Y =
  // A-tile rows (C rows), so lane l holds C row (l & 15) of the block for
  // four C columns: one store instruction writes 16 consecutive rows of each
  // of four columns.
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gr = r0 + wr + 16 * i + lrow;
    if (gr >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int gc = c0 + wc + 16 * j + Mfma<T>::row(lk, g);
        if (gc >= N || (LOWER && gr < gc)) continue;
        if (STORE) C[gr + static_cast<long long>(gc) * ldc] = alpha * acc[i][j][g];
        else C[gr + static_cast<long long>(gc) * ldc] += alpha * acc[i][j][g];
      }
    }
  }
}

// Host launcher: one product, or ny x nz of them offset by `bt`. Returns a
// hipError_t from the launch.
template <typename T, bool NN, bool LOWER, bool STORE = false>
inline hipError_t gemm_update(int M, int N, int K, const T* A, long long lda, const T* B, long long ldb, T* C,
                              long long ldc, T alpha, hipStream_t s, int ny = 1, int nz = 1,
                              GemmBatch bt = GemmBatch{}) {
  if (M <= 0 || N <= 0 || K <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  const long long tm = (M + kTile - 1) / kTile, tn = (N + kTile - 1) / kTile;
  const long long tiles = LOWER ? tm * (tm + 1) / 2 : tm * tn;
  const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(ny), static_cast<unsigned>(nz));
  gemm_update_kernel<T, NN, LOWER, STORE><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, bt);
  return hipGetLastError();
}

}  // namespace sov
