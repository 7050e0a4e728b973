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
// Tiling for CDNA4: a 256-thread workgroup owns a BM x BN tile of C (32x32,
// 64x64, 128x64 or 128x128) as a 2x2 grid of wavefronts, each a grid of 16x16 MFMA
// blocks (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32). K is staged through
// LDS in chunks of 32 with a register prefetch of the next chunk; FP32
// operands are loaded 16 bytes per thread where their layout allows (bulk
// SYRK of the m = 8192 factorization: 51.4 -> 48.3 us per launch on average
// against element-wise loads, same GPU). LDS rows are padded by 16 elements so the four
// k-rows one MFMA operand read touches fall in different banks. Workgroups are renumbered so each XCD (round-robin
// dispatch over 8) works on a contiguous run of tiles, which share operand
// rows in that XCD's L2.
#pragma once

#include "common.hpp"

#include <hip/hip_ext.h>

#include <cstdint>

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

// Elements one thread stages per load from an operand read along its rows:
// 16 bytes for FP32. FP64 stays at one element per thread and load (its
// 16-byte form made the bulk SYRK of the factorization 8% slower on MI355X).
template <typename T>
constexpr int stage_width() {
  return sizeof(T) == 4 ? 4 : 1;
}
template <typename T, int VW>
struct StageVec {
  typedef T type __attribute__((ext_vector_type(VW)));
};
template <typename T>
struct StageVec<T, 1> {
  typedef T type;
};

// Stages one kChunk-column chunk of an operand stored along its rows
// (X(x, k) = X[x + k*ldx], rows x0 .. x0 + R*VW - 1 of the tile): thread tid
// takes VW consecutive rows of k-column k0 + tid / R + (256 / R) q. For
// VW > 1 with `full` (every row of the tile exists and the columns are
// 16-byte aligned) that is one vector load; otherwise VW element loads,
// where out-of-range elements read element 0 and are replaced by zero.
// k-columns at or past K are zero either way.
template <typename T, typename V, int VW, int R, int NV>
__device__ __forceinline__ void stage_rows(V* dst, const T* __restrict__ X, long long ldx, int x0, int rows, int k0,
                                           int K, bool full, int tid) {
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const int r = (tid % R) * VW, gk = k0 + tid / R + (256 / R) * q;
    if constexpr (VW == 1) {
      const bool in = x0 + r < rows && gk < K;
      const T t = X[in ? (x0 + r) + static_cast<long long>(gk) * ldx : 0];
      dst[q] = in ? t : T(0);
    } else if (full) {
      const V v = *reinterpret_cast<const V*>(X + (x0 + r) + static_cast<long long>(gk < K ? gk : 0) * ldx);
      dst[q] = gk < K ? v : V{};
    } else {
      V v;
#pragma unroll
      for (int e = 0; e < VW; ++e) {
        const bool in = x0 + r + e < rows && gk < K;
        const T t = X[in ? (x0 + r + e) + static_cast<long long>(gk) * ldx : 0];
        v[e] = in ? t : T(0);
      }
      dst[q] = v;
    }
  }
}

// Updates (C read-modify-write) with 64 x 64 tiles ask for at least two
// workgroups' worth of waves per SIMD (the register budget that buys): the
// C traffic and the operand staging overlap the MFMA work of other
// workgroups only with enough of them resident (FP64, M = 8064, K = 128:
// 275 -> 228 us per launch). Stores (TRSM, inverses) keep the default budget.
template <typename T, bool NN, bool LOWER, bool STORE = false, int BM = kTile, int BN = kTile>
__global__ __launch_bounds__(256, !STORE && BM * BN <= kTile * kTile ? 2 : 1) void gemm_update_kernel(int M, int N, int K, const T* __restrict__ A,
                                                          long long lda, const T* __restrict__ B,
                                                          long long ldb, T* __restrict__ C,
                                                          long long ldc, T alpha, GemmBatch bt, int vec) {
  static_assert(BM % 32 == 0 && BN % 32 == 0 && BM <= 128 && BN <= 128, "tile shape");
  static_assert(!LOWER || BM == BN, "triangular tiling needs square tiles");
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int IM = BM / 32, JN = BN / 32;          // 16 x 16 blocks per wavefront (2 x 2 wavefronts)
  constexpr int LA = BM + 16, LB = BN + 16;          // padded LDS rows
  constexpr int QA = BM * kChunk / 256, QB = BN * kChunk / 256;  // staged elements per thread
  A += blockIdx.y * bt.a_y + blockIdx.z * bt.a_z;
  B += blockIdx.y * bt.b_y + blockIdx.z * bt.b_z;
  C += blockIdx.y * bt.c_y + blockIdx.z * bt.c_z;
  __shared__ __align__(16) T As[kChunk * LA];
  __shared__ __align__(16) T Bs[kChunk * LB];

  const int tiles_m = (M + BM - 1) / BM;
  int ti, tj;
  tile_coords(xcd_swizzle(blockIdx.x, gridDim.x), LOWER, tiles_m, &ti, &tj);
  const int r0 = ti * BM, c0 = tj * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave & 1) * (BM / 2), wc = (wave >> 1) * (BN / 2);
  const int lrow = lane & 15, lk = lane >> 4;

  acc_t acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // Offset of C element (i, j, g) of this lane (see the epilogue), or -1
  // where the element is outside C or above the diagonal of a LOWER tile.
  auto c_offset = [&](int i, int j, int g) -> long long {
    const int gr = r0 + wr + 16 * i + lrow;
    const int gc = c0 + wc + 16 * j + Mfma<T>::row(lk, g);
    const bool in = gr < M && gc < N && (!LOWER || gr >= gc);
    return in ? gr + static_cast<long long>(gc) * ldc : -1;
  };
  // The old values of C for an update (!STORE), loaded after the K loop
  // with every load in flight at once, out-of-range lanes reading C[0]
  // instead of branching. (Loading them before the K loop holds their
  // registers through it, which costs more occupancy than it hides latency.)
  T cold[IM][JN][4];
  auto load_c = [&]() {
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const long long o = c_offset(i, j, g);
          cold[i][j][g] = C[o < 0 ? 0 : o];
        }
  };

  // Global -> register staging of one chunk of kChunk k-columns. A and, in
  // NT form, B are read along their rows, stage_width<T>() elements per
  // thread (stage_rows: one load when `vec` marks the operand aligned and the
  // tile is full). NN's B is read along k, element by element (out-of-range
  // elements read element 0 and are replaced by zero).
  constexpr int VW = stage_width<T>();
  typedef typename StageVec<T, VW>::type vec_t;
  constexpr int RA = BM / VW, RB = BN / VW;              // threads per staged k-column
  constexpr int VA = QA / VW, VB = QB / VW;
  static_assert(256 % RA == 0 && 256 % RB == 0 && VA >= 1 && VB >= 1, "staging map");
  const bool full_a = (vec & 1) && r0 + BM <= M;
  const bool full_b = (vec & 2) && c0 + BN <= N;
  vec_t ra[VA], rb[NN ? 1 : VB];
  T rn[NN ? QB : 1];
  auto load_chunk = [&](int k0) {
    stage_rows<T, vec_t, VW, RA, VA>(ra, A, lda, r0, M, k0, K, full_a, tid);
    if constexpr (NN) {
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const int k = tid % kChunk, c = tid / kChunk + (256 / kChunk) * q;
        const int gc = c0 + c, gk = k0 + k;
        const bool in = gc < N && gk < K;
        const T v = B[in ? gk + static_cast<long long>(gc) * ldb : 0];
        rn[q] = in ? v : T(0);
      }
    } else {
      stage_rows<T, vec_t, VW, RB, VB>(rb, B, ldb, c0, N, k0, K, full_b, tid);
    }
  };
  auto store_chunk = [&]() {
#pragma unroll
    for (int q = 0; q < VA; ++q)
      *reinterpret_cast<vec_t*>(&As[(tid / RA + (256 / RA) * q) * LA + (tid % RA) * VW]) = ra[q];
    if constexpr (NN) {
#pragma unroll
      for (int q = 0; q < QB; ++q) Bs[(tid % kChunk) * LB + tid / kChunk + (256 / kChunk) * q] = rn[q];
    } else {
#pragma unroll
      for (int q = 0; q < VB; ++q)
        *reinterpret_cast<vec_t*>(&Bs[(tid / RB + (256 / RB) * q) * LB + (tid % RB) * VW]) = rb[q];
    }
  };

  load_chunk(0);
  for (int k0 = 0; k0 < K; k0 += kChunk) {
    __syncthreads();
    store_chunk();
    __syncthreads();
    if (k0 + kChunk < K) load_chunk(k0 + kChunk);
    // Operands of kGroup k-steps are read from LDS together, so their
    // latencies overlap, then the group's MFMAs run.
    constexpr int kGroup = (IM + JN) * sizeof(T) <= 16 ? 8 : 4;
#pragma unroll
    for (int kg = 0; kg < kChunk; kg += 4 * kGroup) {
      T av[kGroup][IM], bv[kGroup][JN];
#pragma unroll
      for (int s = 0; s < kGroup; ++s) {
        const int ka = (kg + 4 * s + lk) * LA, kb = (kg + 4 * s + lk) * LB;
#pragma unroll
        for (int i = 0; i < IM; ++i) av[s][i] = As[ka + wr + 16 * i + lrow];
#pragma unroll
        for (int j = 0; j < JN; ++j) bv[s][j] = Bs[kb + wc + 16 * j + lrow];
      }
      // The B tile goes in as the instruction's first operand, so each
      // 16 x 16 result comes out transposed: a lane holds one row of C.
#pragma unroll
      for (int s = 0; s < kGroup; ++s)
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
          for (int j = 0; j < JN; ++j) acc[i][j] = Mfma<T>::step(bv[s][j], av[s][i], acc[i][j]);
    }
  }

  // Operands X(i, k) and Y(k, j) of an instruction sit on lanes i + 16 k and
  // j + 16 k; result (i, j) on lane j + 16 (...), register g with i =
  // Mfma<T>::row(lane >> 4, g). Here X = B-tile rows (C columns) and Y =
  // A-tile rows (C rows), so lane l holds C row (l & 15) of the block for
  // four C columns: one store instruction writes 16 consecutive rows of each
  // of four columns.
  if constexpr (!STORE) load_c();
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const long long o = c_offset(i, j, g);
        if (o < 0) continue;
        if constexpr (STORE) C[o] = alpha * acc[i][j][g];
        else C[o] = cold[i][j][g] + alpha * acc[i][j][g];
      }
}

// Host launcher: one product, or ny x nz of them offset by `bt`, with BM x BN
// tiles. `done` (and `start`), when given, are recorded by the launch itself
// as the kernel's completion (and start): hipExtLaunchKernel's stop and
// start events, which spare the stream separate markers around the kernel.
// Returns a hipError_t from the launch.
template <typename T, bool NN, bool LOWER, bool STORE = false, int BM = kTile, int BN = kTile>
inline hipError_t gemm_update(int M, int N, int K, const T* A, long long lda, const T* B, long long ldb, T* C,
                              long long ldc, T alpha, hipStream_t s, int ny = 1, int nz = 1,
                              GemmBatch bt = GemmBatch{}, hipEvent_t done = nullptr, hipEvent_t start = nullptr) {
  if (M <= 0 || N <= 0 || K <= 0 || ny <= 0 || nz <= 0) {
    if (start) (void)hipEventRecord(start, s);
    return done ? hipEventRecord(done, s) : hipSuccess;
  }
  const long long tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  const long long tiles = LOWER ? tm * (tm + 1) / 2 : tm * tn;
  const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(ny), static_cast<unsigned>(nz));
  // Operands whose every staged column starts on a boundary of
  // stage_width<T>() elements take one load per thread and chunk in full
  // tiles (bit 0: A, bit 1: B in NT form).
  constexpr long long vw = stage_width<T>();
  auto aligned = [](const T* p, long long ld, long long oy, long long oz) {
    return reinterpret_cast<std::uintptr_t>(p) % (vw * sizeof(T)) == 0 && ld % vw == 0 && oy % vw == 0 &&
           oz % vw == 0;
  };
  const int vec = (aligned(A, lda, bt.a_y, bt.a_z) ? 1 : 0) | (!NN && aligned(B, ldb, bt.b_y, bt.b_z) ? 2 : 0);
  if (done || start) {
    hipExtLaunchKernelGGL(gemm_update_kernel<T, NN, LOWER, STORE, BM, BN>, grid, dim3(256), 0, s, start, done, 0, M,
                          N, K, A, lda, B, ldb, C, ldc, alpha, bt, vec);
  } else {
    gemm_update_kernel<T, NN, LOWER, STORE, BM, BN>
        <<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, bt, vec);
  }
  return hipGetLastError();
}

// Tile shapes the autotuner chooses between for a given product class.
// 32x32 suits the latency-bound panel products (their time is set by one
// tile's MFMA chain and loads, not by their width).
enum GemmShape { kGemm64x64 = 0, kGemm128x64 = 1, kGemm128x128 = 2, kGemm32x32 = 3, kGemmShapes = 4 };

// gemm_update with a run-time tile shape (square shapes only when LOWER).
template <typename T, bool NN, bool LOWER, bool STORE = false>
inline hipError_t gemm_update_shape(int shape, int M, int N, int K, const T* A, long long lda, const T* B,
                                    long long ldb, T* C, long long ldc, T alpha, hipStream_t s,
                                    hipEvent_t done = nullptr, hipEvent_t start = nullptr) {
  const GemmBatch one{};
  switch (shape) {
    case kGemm128x128:
      return gemm_update<T, NN, LOWER, STORE, 128, 128>(M, N, K, A, lda, B, ldb, C, ldc, alpha, s, 1, 1, one, done,
                                                        start);
    case kGemm32x32:
      return gemm_update<T, NN, LOWER, STORE, 32, 32>(M, N, K, A, lda, B, ldb, C, ldc, alpha, s, 1, 1, one, done,
                                                      start);
    case kGemm128x64:
      if constexpr (!LOWER)
        return gemm_update<T, NN, LOWER, STORE, 128, 64>(M, N, K, A, lda, B, ldb, C, ldc, alpha, s, 1, 1, one, done,
                                                         start);
      else return gemm_update<T, NN, LOWER, STORE>(M, N, K, A, lda, B, ldb, C, ldc, alpha, s, 1, 1, one, done, start);
    default:
      return gemm_update<T, NN, LOWER, STORE>(M, N, K, A, lda, B, ldb, C, ldc, alpha, s, 1, 1, one, done, start);
  }
}

}  // namespace sov
