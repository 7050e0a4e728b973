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
// entirely by one 512-thread workgroup; workgroups take LPs from a queue, so
// LPs of uneven size share one launch. Per iteration:
//   M = A D A^T          FP32 MFMA (v_mfma_f32_16x16x4f32), FP32 accumulation,
//                        written to LDS as a packed lower triangle (m <= 272);
//   M = L L^T            FP32, in LDS, blocked 16 columns a step (diagonal
//                        block in registers, panel and trailing update on
//                        MFMA), with the LDL^T pivot rule of the reference
//                        (drop a pivot <= 1e-13 |diagonal|);
//   M dy = rhs           FP64 preconditioned CG with the FP32 factor as the
//                        preconditioner and the FP64 product A (D (A^T p)).
// Residuals, step lengths and the termination test are FP64. An LP whose
// FP32 factor drops a pivot, or whose CG solves stall, continues with FP64
// factorizations of its normal matrix (packed in global memory, FP64 MFMA).

#include "batch_layout.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>

typedef float flt4_t __attribute__((ext_vector_type(4)));
typedef double dbl4_t __attribute__((ext_vector_type(4)));

namespace {

namespace lay = sov_batch_layout;
using lay::kMaxRows;
using lay::kStage;
using lay::row_ld;

// 8 wavefronts: 256 registers a lane, about 300 B of scratch a lane. (At
// 1024 threads, 128 registers a lane, the kernel spilled over 1 KB a lane;
// that size was preferred while the LDS factorization went a column at a
// time and wanted every thread.)
constexpr int kThreads = 512;
static_assert(kThreads % 256 == 0 && kThreads <= 1024, "gemv_n slices the columns in 256-row groups");
constexpr int kWaves = kThreads / 64;
static_assert(kWaves <= lay::kRed, "block reductions keep one partial per wavefront");
static_assert(kThreads / 256 * 256 <= kStage && kMaxRows <= kStage, "staged vectors and partial sums fit kStage");
constexpr int kTraceRows = 200, kTraceCols = 16;  // SOVEREIGN_BATCH_TRACE record

struct BatchMatrix {
  int m, n;
  const double* a64;   // column-major m x n
  const float* a32;    // row-major m x ldr, zero beyond column n (FP32 operands of the assembly)
  int ldr;             // n rounded up to a multiple of 32
  const double* a64r;  // row-major m x ldr (FP64, for A^T v)
};

struct BatchDesc {
  int mat;
  int pad;
  long long b_off, x_off, ws_off;
  long long in_off, out_off;  // warm start / final iterate (x, w, z, v, y) offsets, -1: none
};

struct BatchOpts {
  double tol, tau, pcg_tol;
  int max_iter, pcg_max;
  double save_mu;  // the iterate written to state_out: the first with mu below this (<= 0: the final one)
  long long lds_matrix_bytes;  // size of the LDS matrix region (lay::lds_matrix_bytes of the largest LP)
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
  const float* Af;   // row-major, leading dimension ldr
  const double* Ar;  // row-major FP64, leading dimension ldr
  int ldr;
  const double *b, *c, *u;
  // workspace vectors (global, this LP only)
  double *x, *w, *z, *v, *rc, *ru, *D, *dx, *dw, *dz, *dv, *dxa, *dwa, *dza, *dva, *t;
  double *y, *rb, *dy, *r, *p, *q, *zz, *rhs, *sc;  // sc: Jacobi scale of A D A^T
  double* M64;  // packed lower FP64 normal matrix (global), for LPs in FP64 mode
  float* xs;    // inverses of the FP32 factor's 16 x 16 diagonal blocks (global, 256 floats each)
  float* sd;    // sqrt(D) in FP32, ldr entries (zero beyond n)
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
};
template <>
struct Prec<double> {
  static constexpr double drop_pivot = 1e64;
  __device__ static __forceinline__ dbl4_t mfma(double a, double b, dbl4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  typedef dbl4_t acc_t;
  __device__ static __forceinline__ int row(int lk, int g) { return lk + 4 * g; }
};

// The products with A are bound by load round trips, not bytes: one
// workgroup streams A (m x n FP64) through its CU, and a load of A takes
// hundreds of cycles to return. Each thread therefore issues the loads of
// a whole chunk (a column's rows, or a run of a row's columns) before it
// uses any of them, so that one round trip serves many loads. Over the
// column-major A a group of 16 lanes takes a column (rows r, r + 16, ...):
// the residual products and the normal-matrix product below.
constexpr int kDotLanes = 16;
constexpr int kDotGroups = kThreads / kDotLanes;
constexpr int kDotRows = (kMaxRows + kDotLanes - 1) / kDotLanes;  // rows of a column per lane

// out_i = base_i + (A vec)_i for i < m (base may be null), over the
// row-major FP64 copy of A: 16 lanes per row (columns r, r + 16, ...),
// kChunkCols of them a load round. vec (n, global) is staged in LDS first
// when it fits there.
constexpr int kChunkCols = 12;
__device__ void gemv_n(const Lp& L, const double* vec, const double* base, double* out, double* stage) {
  const int tid = threadIdx.x;
  const int m = L.m, n = L.n, ldr = L.ldr;
  const double* v = vec;
  if (n <= lay::kStage) {
    __syncthreads();
    for (int j = tid; j < n; j += kThreads) stage[j] = vec[j];
    __syncthreads();
    v = stage;
  }
  const int g = tid / kDotLanes, r = tid % kDotLanes;
  for (int i = g; i < m; i += kDotGroups) {  // uniform within a group of 16 lanes
    const double* row = L.Ar + static_cast<long long>(i) * ldr;
    double s0 = 0.0, s1 = 0.0;
    for (int j0 = r; j0 < n; j0 += kDotLanes * kChunkCols) {
      double av[kChunkCols];
#pragma unroll
      for (int u = 0; u < kChunkCols; ++u) {
        const int j = j0 + kDotLanes * u;
        av[u] = row[j < n ? j : 0];
      }
#pragma unroll
      for (int u = 0; u < kChunkCols; u += 2) {
        const int j = j0 + kDotLanes * u;
        s0 = fma(av[u], j < n ? v[j] : 0.0, s0);
        s1 = fma(av[u + 1], j + kDotLanes < n ? v[j + kDotLanes] : 0.0, s1);
      }
    }
    const double t = group_sum<kDotLanes>(s0 + s1);
    if (r == 0) out[i] = (base ? base[i] : 0.0) + t;
  }
  __syncthreads();
}

// Jacobi-scaled FP32 normal matrix in LDS (packed lower):
//   Mt = S A D A^T S,  S = diag(sc),  sc_i = (A D A^T)_ii^(-1/2),
// on MFMA from the operands A(i, k) sqrt(D_k) (the row-major FP32 copy of A
// and sqrt(D) in FP32, both read 16 bytes a lane: lane group g takes
// k = k0 + 4 g .. k0 + 4 g + 3, the same k for both operands), S applied to
// each tile's sums. sc comes from a row-wise FP32 sum of the same operands,
// one wavefront a row. Each wavefront takes up to four tiles of one block
// row at a time; two 16-column chunks are loaded ahead of their MFMAs.
__device__ void assemble32(const Lp& L, float* M, double reg) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int m = L.m, n = L.n, ldr = L.ldr;
  for (int k = tid; k < ldr; k += kThreads) L.sd[k] = k < n ? static_cast<float>(sqrt(fmax(L.D[k], 0.0))) : 0.f;
  __syncthreads();
  for (int i = wv; i < m; i += kWaves) {
    const float* ar = L.Af + static_cast<long long>(i) * ldr;
    float acc = 0.f;
    for (int k = 4 * lane; k < ldr; k += 256) {
      const flt4_t a = *reinterpret_cast<const flt4_t*>(ar + k);
      const flt4_t d = *reinterpret_cast<const flt4_t*>(L.sd + k);
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += (a[u] * d[u]) * (a[u] * d[u]);
    }
    const double dii = static_cast<double>(wave_sum(static_cast<double>(acc)));
    if (lane == 0) L.sc[i] = dii + reg > 0.0 ? 1.0 / sqrt(dii + reg) : 1.0;
  }
  __syncthreads();
  // Work items: block row I with up to four of its tiles (I, J0 .. J0 + 3),
  // J0 + 3 <= I; the row block's operands are loaded and scaled once per
  // k-chunk for all of them.
  const int nt = (m + 15) / 16;
  const int nitems = lay::assembly_items(nt);
  const flt4_t zero = {0.f, 0.f, 0.f, 0.f};
  for (int q = wv; q < nitems; q += kWaves) {
    int I, J0;
    lay::assembly_item(q, I, J0);
    const int nj = min(4, I + 1 - J0);
    const int ra = 16 * I + lrow;
    const bool va = ra < m;
    const float* pa = L.Af + static_cast<long long>(va ? ra : 0) * ldr + 4 * lk;
    const float* pd = L.sd + 4 * lk;
    const float* pb[4];
    bool vb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int cb = 16 * (J0 + c) + lrow;
      vb[c] = c < nj && cb < m;
      pb[c] = L.Af + static_cast<long long>(vb[c] ? cb : 0) * ldr + 4 * lk;
    }
    flt4_t acc[4] = {zero, zero, zero, zero};
    // One 32-column chunk's operands in registers; the next chunk's loads
    // are issued before this chunk's MFMAs, so their latency overlaps them.
    flt4_t d0, d1, a0, a1, b0[4], b1[4];
    auto load = [&](int k0) {
      d0 = *reinterpret_cast<const flt4_t*>(pd + k0);
      d1 = *reinterpret_cast<const flt4_t*>(pd + k0 + 16);
      a0 = va ? *reinterpret_cast<const flt4_t*>(pa + k0) : zero;
      a1 = va ? *reinterpret_cast<const flt4_t*>(pa + k0 + 16) : zero;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        b0[c] = vb[c] ? *reinterpret_cast<const flt4_t*>(pb[c] + k0) : zero;
        b1[c] = vb[c] ? *reinterpret_cast<const flt4_t*>(pb[c] + k0 + 16) : zero;
      }
    };
    load(0);
    for (int k0 = 0; k0 < ldr; k0 += 32) {  // ldr is a multiple of 32
      flt4_t sa0 = a0, sa1 = a1, sb0[4], sb1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        sa0[u] *= d0[u];
        sa1[u] *= d1[u];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          sb0[c][u] = b0[c][u] * d0[u];
          sb1[c][u] = b1[c][u] * d1[u];
        }
      }
      if (k0 + 32 < ldr) load(k0 + 32);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c < nj) {
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[c] = Prec<float>::mfma(sa0[u], sb0[c][u], acc[c]);
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[c] = Prec<float>::mfma(sa1[u], sb1[c][u], acc[c]);
        }
      }
    }
    // lane holds column 16 (J0 + c) + lrow, rows 16 I + Prec<float>::row(lk, g)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = 16 * (J0 + c) + lrow;
      if (c >= nj || j >= m) continue;
      const double sj = L.sc[j];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = 16 * I + Prec<float>::row(lk, g);
        if (i < m && i >= j) {
          const double si = L.sc[i];
          M[packed_index(i, j, m)] =
              static_cast<float>(si * sj * (static_cast<double>(acc[c][g]) + (i == j ? reg : 0.0)));
        }
      }
    }
  }
  __syncthreads();
}

// FP64 Jacobi-scaled normal matrix (packed lower, in global memory) of an LP
// that has switched to FP64 factorizations:
//   Mt = S A D A^T S,  S = diag(sc),  sc_i = (A D A^T)_ii^(-1/2),
// so Mt has a unit diagonal whatever the spread of D. Same work items as the
// FP32 assembly above, over the row-major FP64 copy of A (zero beyond column
// n): a lane's k-slot lk of MFMA step u is column k0 + 4 lk + u, both
// operands use that map, and A(i, k) D_k A(j, k) is formed from the operands
// A(i, k) D_k and A(j, k); S is applied to the FP64 sums when they are stored.
__device__ void assemble64(const Lp& L, double* M, double reg) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int m = L.m, n = L.n, ldr = L.ldr;
  for (int i = wv; i < m; i += kWaves) {  // diagonal: a wavefront per row
    const double* ar = L.Ar + static_cast<long long>(i) * ldr;
    double acc = 0.0;
    for (int k = lane; k < n; k += 64) {
      const double a = ar[k];
      acc = fma(a * a, fmax(L.D[k], 0.0), acc);
    }
    const double dii = wave_sum(acc);
    if (lane == 0) L.sc[i] = dii + reg > 0.0 ? 1.0 / sqrt(dii + reg) : 1.0;
  }
  __syncthreads();
  const int nt = (m + 15) / 16;
  const int nitems = lay::assembly_items(nt);
  const dbl4_t zero = {0.0, 0.0, 0.0, 0.0};
  for (int q = wv; q < nitems; q += kWaves) {
    int I, J0;
    lay::assembly_item(q, I, J0);
    const int nj = min(4, I + 1 - J0);
    const int ra = 16 * I + lrow;
    const bool va = ra < m;
    const double* pa = L.Ar + static_cast<long long>(va ? ra : 0) * ldr + 4 * lk;
    const double* pb[4];
    bool vb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int cb = 16 * (J0 + c) + lrow;
      vb[c] = c < nj && cb < m;
      pb[c] = L.Ar + static_cast<long long>(vb[c] ? cb : 0) * ldr + 4 * lk;
    }
    dbl4_t acc[4] = {zero, zero, zero, zero};
    // One 16-column step's operands in registers; the next step's loads are
    // issued before this step's MFMAs.
    dbl4_t a, d, b[4];
    auto load = [&](int k0) {
      a = va ? *reinterpret_cast<const dbl4_t*>(pa + k0) : zero;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * lk + u;
        d[u] = k < n ? fmax(L.D[k], 0.0) : 0.0;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = vb[c] ? *reinterpret_cast<const dbl4_t*>(pb[c] + k0) : zero;
    };
    load(0);
    for (int k0 = 0; k0 < ldr; k0 += 16) {  // ldr is a multiple of 32
      dbl4_t sa = a * d, sb[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) sb[c] = b[c];
      if (k0 + 16 < ldr) load(k0 + 16);
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < nj) {
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[c] = Prec<double>::mfma(sa[u], sb[c][u], acc[c]);
        }
    }
    // lane holds column 16 (J0 + c) + lrow, rows 16 I + Prec<double>::row(lk, g)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = 16 * (J0 + c) + lrow;
      if (c >= nj || j >= m) continue;
      const double sj = L.sc[j];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = 16 * I + Prec<double>::row(lk, g);
        if (i < m && i >= j) M[packed_index(i, j, m)] = L.sc[i] * sj * (acc[c][g] + (i == j ? reg : 0.0));
      }
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

// Blocked FP64 Cholesky of the packed lower matrix in global memory, for an
// LP that has switched to FP64 factorizations: 16-column panels staged in
// LDS (P: the rows k0..m-1 of the panel, row-major with leading dimension
// 16, in the LDS region the FP32 matrix no longer needs), factored there
// column by column with the pivot rule of cholesky() above, written back,
// and the trailing lower triangle updated on FP64 MFMA from the panel (one
// read-modify-write of the trailing matrix per 16 columns instead of one
// per column). Returns the number of dropped pivots.
__device__ int cholesky64_blocked(const Lp& L, double* M, double* P) {
  constexpr int kB = lay::kPanel64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int m = L.m;
  int dropped = 0;
  for (int k0 = 0; k0 < m; k0 += kB) {
    const int kb = min(kB, m - k0), rows = m - k0;
    __syncthreads();
    for (int e = tid; e < rows * kB; e += kThreads) {
      const int r = e / kB, c = e % kB;
      P[e] = (c < kb && r >= c) ? M[packed_index(k0 + r, k0 + c, m)] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
      const double dk = P[j * kB + j];
      const bool drop = !(dk > 1e-13) && isfinite(dk);
      const double piv = drop ? Prec<double>::drop_pivot : (isfinite(dk) ? sqrt(dk) : 1.0);
      dropped += drop ? 1 : 0;
      for (int r = j + 1 + tid; r < rows; r += kThreads) P[r * kB + j] /= piv;
      __syncthreads();
      if (tid == 0) P[j * kB + j] = piv;
      for (int e = tid; e < rows * kB; e += kThreads) {
        const int r = e / kB, c = e % kB;
        if (c > j && c < kb && r >= c) P[e] -= P[r * kB + j] * P[c * kB + j];
      }
      __syncthreads();
    }
    for (int e = tid; e < rows * kB; e += kThreads) {
      const int r = e / kB, c = e % kB;
      if (c < kb && r >= c) M[packed_index(k0 + r, k0 + c, m)] = P[e];
    }
    // trailing update M(i, j) -= sum_c P(i, c) P(j, c), k0 + kb <= j <= i < m
    const int tr = rows - kb;
    const int nt = (tr + 15) / 16;
    for (int t = wv; t < nt * (nt + 1) / 2; t += kWaves) {
      int ti = 0;
      while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
      const int tj = t - ti * (ti + 1) / 2;
      const int ra = kb + 16 * ti + lrow, cb = kb + 16 * tj + lrow;  // panel rows of the two operands
      dbl4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int u = 0; u < kB / 4; ++u) {
        const int kk = 4 * u + lk;
        const double av = (ra < rows && kk < kb) ? P[ra * kB + kk] : 0.0;
        const double bv = (cb < rows && kk < kb) ? P[cb * kB + kk] : 0.0;
        acc = Prec<double>::mfma(av, bv, acc);
      }
      // lane holds column cb, rows kb + 16 ti + Prec<double>::row(lk, g)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = k0 + kb + 16 * ti + Prec<double>::row(lk, g), j = k0 + cb;
        if (i < m && j < m && i >= j) M[packed_index(i, j, m)] -= acc[g];
      }
    }
  }
  __syncthreads();
  return dropped;
}

// In place: the packed FP64 factor L in global memory becomes X = L^-1,
// block column by block column from the last (16 x 16 blocks):
//   X(J, J) = L(J, J)^-1 (in LDS, one column of it per thread),
//   T(I) = sum_{K = J+1 .. I} X(I, K) L(K, J) for every I > J, on FP64 MFMA,
//   kept in LDS (T, 16 doubles per row below block J) until every L(., J)
//   has been read, then X(I, J) = -T(I) X(J, J).
// D and XJ: 16 x 16 LDS scratch each (the diagonal block and its inverse).
__device__ void invert64_blocked(const Lp& L, double* M, double* T, double* D, double* XJ) {
  constexpr int kB = lay::kPanel64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int m = L.m, nbk = (m + kB - 1) / kB;
  // element (i, k) of the block (I, K) of the lower triangle, zero outside it
  auto at = [&](int i, int k) -> double { return (i < m && k < m && i >= k) ? M[packed_index(i, k, m)] : 0.0; };
  for (int J = nbk - 1; J >= 0; --J) {
    const int j0 = kB * J, jb = min(kB, m - j0);
    __syncthreads();
    for (int e = tid; e < kB * kB; e += kThreads) {
      const int i = e / kB, k = e % kB;
      D[e] = (i < jb && k < jb) ? at(j0 + i, j0 + k) : (i == k ? 1.0 : 0.0);
    }
    __syncthreads();
    if (tid < kB) {  // column tid of the inverse of the lower triangular D
      const int c = tid;
      for (int i = 0; i < kB; ++i) {
        double v;
        if (i < c) {
          v = 0.0;
        } else {
          double s = i == c ? 1.0 : 0.0;
          for (int k = c; k < i; ++k) s -= D[i * kB + k] * XJ[k * kB + c];
          v = s / D[i * kB + i];
        }
        XJ[i * kB + c] = v;
      }
    }
    __syncthreads();
    // T(I) for the block rows below J
    const int nrow = nbk - 1 - J;
    for (int q = wv; q < nrow; q += kWaves) {
      const int I = J + 1 + q, i0 = kB * I;
      dbl4_t acc = {0.0, 0.0, 0.0, 0.0};
      for (int K = J + 1; K <= I; ++K) {
        const int k0 = kB * K;
        double av[kB / 4], bv[kB / 4];
#pragma unroll
        for (int u = 0; u < kB / 4; ++u) {
          const int kk = 4 * u + lk;
          av[u] = at(i0 + lrow, k0 + kk);  // X(I, K), zero above the diagonal of X(I, I)
          bv[u] = at(k0 + kk, j0 + lrow);  // L(K, J)
        }
#pragma unroll
        for (int u = 0; u < kB / 4; ++u) acc = Prec<double>::mfma(av[u], bv[u], acc);
      }
      // lane holds column lrow of T(I), rows Prec<double>::row(lk, g)
#pragma unroll
      for (int g = 0; g < 4; ++g) T[(q * kB + Prec<double>::row(lk, g)) * kB + lrow] = acc[g];
    }
    __syncthreads();
    // X(I, J) = -T(I) X(J, J), and X(J, J) itself
    for (int q = wv; q < nrow; q += kWaves) {
      const int i0 = kB * (J + 1 + q);
      double av[kB / 4], bv[kB / 4];
#pragma unroll
      for (int u = 0; u < kB / 4; ++u) {
        const int kk = 4 * u + lk;
        av[u] = T[(q * kB + lrow) * kB + kk];
        bv[u] = XJ[kk * kB + lrow];
      }
      dbl4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int u = 0; u < kB / 4; ++u) acc = Prec<double>::mfma(av[u], bv[u], acc);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = i0 + Prec<double>::row(lk, g), j = j0 + lrow;
        if (i < m && lrow < jb) M[packed_index(i, j, m)] = -acc[g];
      }
    }
    for (int e = tid; e < kB * kB; e += kThreads) {
      const int i = e / kB, k = e % kB;
      if (i < jb && k <= i) M[packed_index(j0 + i, j0 + k, m)] = XJ[e];
    }
  }
  __syncthreads();
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

// Orders one wavefront's LDS accesses (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

using lay::kBlk;  // block width of the blocked LDS factorization and inversion
// the inversion gives each block row below a block column a wavefront slot
constexpr int kRowsPerWave = ((kMaxRows + kBlk - 1) / kBlk - 1 + kWaves - 1) / kWaves;

// Element (i, k) of the packed lower triangle as an MFMA operand: zero above
// the diagonal and outside the matrix.
__device__ __forceinline__ float packed_at(const float* M, int i, int k, int m) {
  return (i < m && k <= i) ? M[packed_index(i, k, m)] : 0.f;
}

// Factor L and inverse X = L^-1 of the kb x kb (kb <= 16) diagonal block at
// (k0, k0) of the packed matrix, by one wavefront in registers. Lane l holds
// column c = l & 15, rows r_g = (l >> 4) + 4 g, g < 4, of the block (f) and of
// X (x). Step j takes the pivot (the pivot rule of cholesky() above), scales
// column j of L and row j of X, then applies f -= L(:, j) L(:, j)^T and
// x -= L(:, j) X(j, :) below row j; the factors come from the lanes that hold
// them (ds_bpermute). Writes L back into the packed block, X (lower, zero
// elsewhere, 16 x 16 row-major, leading dimension 17) to sx and, 16 x 16
// row-major, to xg. Returns the dropped pivots.
__device__ int leaf16(float* M, int k0, int kb, int m, float* sx, float* xg, int lane) {
  const int c = lane & 15, rg = lane >> 4;
  float f[4], x[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int r = rg + 4 * g;
    f[g] = (r < kb && c <= r) ? M[packed_index(k0 + r, k0 + c, m)] : 0.f;
    x[g] = (r == c && r < kb) ? 1.f : 0.f;
  }
  int dropped = 0;
#pragma unroll
  for (int j = 0; j < kBlk; ++j) {
    if (j < kb) {
      const int jg = j >> 2, jl = j & 3;  // (j, j) sits on lane j + 16 jl, register jg
      const double dk = static_cast<double>(__shfl(f[jg], j + 16 * jl, 64));
      const bool drop = !(dk > 1e-13) && isfinite(dk);
      const float piv = drop ? Prec<float>::drop_pivot : (isfinite(dk) ? static_cast<float>(sqrt(dk)) : 1.f);
      const float rp = 1.f / piv;
      dropped += drop ? 1 : 0;
      if (c == j) {  // column j of L: scaled below the diagonal, the pivot on it, zero above
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int r = rg + 4 * g;
          f[g] = r > j ? f[g] * rp : (r == j ? piv : 0.f);
        }
      }
      if (rg == jl) x[jg] *= rp;  // row j of X
      float lr[4], t[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) lr[g] = __shfl(f[g], j + 16 * rg, 64);  // L(r_g, j)
#pragma unroll
      for (int g = 0; g < 4; ++g) t[g] = __shfl(f[g], j + 16 * (c & 3), 64);
      const int cg = c >> 2;
      const float lc = cg == 0 ? t[0] : cg == 1 ? t[1] : cg == 2 ? t[2] : t[3];  // L(c, j)
      const float xr = __shfl(x[jg], c + 16 * jl, 64);                           // X(j, c)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int r = rg + 4 * g;
        if (r > j) {
          if (c > j) f[g] -= lr[g] * lc;
          x[g] -= lr[g] * xr;
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int r = rg + 4 * g;
    const bool low = r < kb && c <= r;
    if (low) M[packed_index(k0 + r, k0 + c, m)] = f[g];
    sx[r * (kBlk + 1) + c] = low ? x[g] : 0.f;
    xg[r * kBlk + c] = low ? x[g] : 0.f;
  }
  return dropped;
}

// Blocked right-looking Cholesky of the FP32 packed lower triangle in LDS,
// 16 columns a step: (1) wavefront 0 factors and inverts the 16 x 16
// diagonal block in registers (leaf16); (2) the panel below it is
// L(I, K) = A(I, K) X(K, K)^T, one 16-row MFMA tile per wavefront; (3) the
// trailing lower triangle takes the panel's rank-16 update on MFMA
// (v_mfma_f32_16x16x4f32), one 16 x 16 tile per wavefront at a time, except
// that wavefront 0 takes the next diagonal block's tile and then factors
// that block (step (1) of the next block, overlapping the rest of the
// update). Two barriers a block. The diagonal blocks' inverses go to xg
// (16 x 16 each, row-major) for invert_lower_blocked. Returns the dropped
// pivots (counted on wavefront 0).
__device__ int cholesky_blocked(const Lp& L, float* M, float* sx) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int m = L.m;
  int dropped = 0;
  if (wv == 0) dropped += leaf16(M, 0, min(kBlk, m), m, sx, L.xs, lane);
  __syncthreads();
  for (int k0 = 0; k0 < m; k0 += kBlk) {
    const int kb = min(kBlk, m - k0);
    const int r0 = k0 + kb;  // first row below the block
    if (r0 >= m) break;      // (only the last block is narrower than kBlk)
    for (int pt = wv; r0 + kBlk * pt < m; pt += kWaves) {  // (2) rows r0 + 16 pt .. of the panel
      const int ra = r0 + kBlk * pt + lrow;
      float av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        av[u] = packed_at(M, ra, k0 + 4 * u + lk, m);  // A(ra, k0 + q)
        bv[u] = sx[lrow * (kBlk + 1) + 4 * u + lk];    // X(p, q), p = lrow
      }
      flt4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = Prec<float>::mfma(av[u], bv[u], acc);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = r0 + kBlk * pt + Prec<float>::row(lk, g);
        if (i < m) M[packed_index(i, k0 + lrow, m)] = acc[g];
      }
    }
    __syncthreads();
    // (3) trailing update, tiles (I, J), J <= I, of the rows and columns from
    // r0. Look-ahead: wavefront 0 updates tile (0, 0), the next diagonal
    // block, and factors it at once (the leaf reads and writes that tile
    // only) while the other wavefronts update the rest.
    const int nt = (m - r0 + kBlk - 1) / kBlk;
    const int tiles = nt * (nt + 1) / 2;
    auto update_tile = [&](int t) {
      int ti = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
      while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
      while (ti * (ti + 1) / 2 > t) --ti;
      const int tj = t - ti * (ti + 1) / 2;
      const int ra = r0 + kBlk * ti + lrow, cb = r0 + kBlk * tj + lrow;
      float av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u + lk;
        av[u] = packed_at(M, ra, k, m);
        bv[u] = packed_at(M, cb, k, m);
      }
      flt4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = Prec<float>::mfma(av[u], bv[u], acc);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = r0 + kBlk * ti + Prec<float>::row(lk, g), j = cb;
        if (i < m && j < m && i >= j) M[packed_index(i, j, m)] -= acc[g];
      }
    };
    if (wv == 0) {
      update_tile(0);
      wave_sync();
      dropped += leaf16(M, r0, min(kBlk, m - r0), m, sx, L.xs + (r0 / kBlk) * kBlk * kBlk, lane);
    } else {
      for (int t = wv; t < tiles; t += kWaves - 1) update_tile(t);
    }
    __syncthreads();
  }
  return dropped;
}

// In place: the FP32 packed factor L in LDS becomes X = L^-1, block column
// by block column from the last. For block column K: (a) the diagonal
// block's inverse, formed by cholesky_blocked, is copied in from xg; (b) for
// each block row I > K (spread over the wavefronts,
// block rows per wavefront at most kRowsPerWave)
// T = sum_{J=K+1..I} X(I, J) L(J, K) on MFMA, kept in registers until every
// wavefront has read L(., K); (c) T goes into the place of
// L(I, K) and X(I, K) = -T X(K, K) on MFMA.
__device__ void invert_lower_blocked(const Lp& L, float* M) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int m = L.m;
  const int nbk = (m + kBlk - 1) / kBlk;
  for (int K = nbk - 1; K >= 0; --K) {
    const int k0 = kBlk * K, kb = min(kBlk, m - k0);
    if (tid < kBlk * kBlk) {  // (a)
      const int r = tid / kBlk, c = tid % kBlk;
      if (r < kb && c <= r) M[packed_index(k0 + r, k0 + c, m)] = L.xs[K * kBlk * kBlk + tid];
    }
    __syncthreads();
    flt4_t acc[kRowsPerWave];  // (b) block rows K + 1 + wv + kWaves q of this wavefront
#pragma unroll
    for (int q = 0; q < kRowsPerWave; ++q) {
      acc[q] = flt4_t{0.f, 0.f, 0.f, 0.f};
      const int I = K + 1 + wv + kWaves * q;
      if (I >= nbk) continue;
      const int r0 = kBlk * I;
      for (int J = K + 1; J <= I; ++J) {
        const int j0 = kBlk * J;
        float av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int kk = j0 + 4 * u + lk;
          av[u] = packed_at(M, r0 + lrow, kk, m);  // X(I, J): zero above the diagonal
          bv[u] = packed_at(M, kk, k0 + lrow, m);  // L(J, K)
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[q] = Prec<float>::mfma(av[u], bv[u], acc[q]);
      }
    }
    __syncthreads();  // every L(., K) read
#pragma unroll
    for (int q = 0; q < kRowsPerWave; ++q) {  // (c)
      const int I = K + 1 + wv + kWaves * q;
      if (I >= nbk) continue;
      const int r0 = kBlk * I;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = r0 + Prec<float>::row(lk, g), j = k0 + lrow;
        if (i < m && j < k0 + kb) M[packed_index(i, j, m)] = acc[q][g];
      }
      wave_sync();
      float av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kk = k0 + 4 * u + lk;
        av[u] = (kk < k0 + kb) ? packed_at(M, r0 + lrow, kk, m) : 0.f;                // T(I, K)
        bv[u] = (kk < k0 + kb && lrow < kb) ? packed_at(M, kk, k0 + lrow, m) : 0.f;  // X(K, K)
      }
      flt4_t x = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; ++u) x = Prec<float>::mfma(av[u], bv[u], x);
      wave_sync();
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = r0 + Prec<float>::row(lk, g), j = k0 + lrow;
        if (i < m && j < k0 + kb) M[packed_index(i, j, m)] = -x[g];
      }
    }
    __syncthreads();
  }
}

// q = A D A^T p (FP64, matrix-free); p, q global (m); atp (n, global, or
// null) receives A^T p.
// One pass over the column-major A: for each of its columns j a group of 16
// lanes forms t_j = A(:, j)' p (p staged in LDS), then adds A(:, j) D_j t_j
// into its rows of q, held in registers across the group's columns. The
// four groups of a wavefront then combine by shuffles, and the wavefronts
// add into LDS one after another, so the sum order is fixed.
static_assert(2 * kMaxRows <= kStage, "normal_matvec stages p and q");
__device__ void normal_matvec(const Lp& L, const double* p, double* q, double* stage, double* atp) {
  const int tid = threadIdx.x, m = L.m, n = L.n;
  double* pv = stage;             // p (m)
  double* qs = stage + kMaxRows;  // q being summed (m)
  __syncthreads();
  for (int i = tid; i < m; i += kThreads) {
    pv[i] = p[i];
    qs[i] = 0.0;
  }
  __syncthreads();
  const int g = tid / kDotLanes, r = tid % kDotLanes;
  double qa[kDotRows], pr[kDotRows];  // pr: the lane's rows of p, the same for every column
#pragma unroll
  for (int k = 0; k < kDotRows; ++k) {
    const int i = r + kDotLanes * k;
    qa[k] = 0.0;
    pr[k] = i < m ? pv[i] : 0.0;
  }
  const double* A = L.A;
  const double* D = L.D;
  for (int j = g; j < n; j += kDotGroups) {  // uniform within a group of 16 lanes
    const double* c1 = A + static_cast<long long>(j) * m;
    double a1[kDotRows];
#pragma unroll
    for (int k = 0; k < kDotRows; ++k) {
      const int i = r + kDotLanes * k;
      a1[k] = c1[i < m ? i : 0];
    }
    const double dj = D[j];
    double s1 = 0.0;
#pragma unroll
    for (int k = 0; k < kDotRows; ++k) s1 = fma(a1[k], pr[k], s1);
    const double tj = group_sum<kDotLanes>(s1);
    if (atp && r == 0) atp[j] = tj;
    const double f = tj * dj;
#pragma unroll
    for (int k = 0; k < kDotRows; ++k) qa[k] = fma(a1[k], f, qa[k]);
  }
#pragma unroll
  for (int k = 0; k < kDotRows; ++k) {
    qa[k] += __shfl_xor(qa[k], 16, 64);
    qa[k] += __shfl_xor(qa[k], 32, 64);
  }
  const int lane = tid & 63, wave = tid >> 6;
  for (int w = 0; w < kWaves; ++w) {
    if (wave == w && lane < kDotLanes) {
#pragma unroll
      for (int k = 0; k < kDotRows; ++k) {
        const int i = r + kDotLanes * k;
        if (i < m) qs[i] += qa[k];
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < m; i += kThreads) q[i] = qs[i];
  __syncthreads();
}

// The residual products in one pass over the column-major A, as in
// normal_matvec: rb = b - A x (m) and t = A^T y (n). Each 16-lane group
// forms t_j = A(:, j)'y (y staged in LDS) and adds A(:, j) x_j into its rows
// of A x, held in registers; the wavefronts add into LDS in a fixed order.
__device__ void residual_products(const Lp& L, double* stage) {
  const int tid = threadIdx.x, m = L.m, n = L.n;
  double* yv = stage;             // y (m)
  double* qs = stage + kMaxRows;  // A x being summed (m)
  __syncthreads();
  for (int i = tid; i < m; i += kThreads) {
    yv[i] = L.y[i];
    qs[i] = 0.0;
  }
  __syncthreads();
  const int g = tid / kDotLanes, r = tid % kDotLanes;
  double qa[kDotRows];
#pragma unroll
  for (int k = 0; k < kDotRows; ++k) qa[k] = 0.0;
  for (int j = g; j < n; j += kDotGroups) {  // uniform within a group of 16 lanes
    const double* c1 = L.A + static_cast<long long>(j) * m;
    double a1[kDotRows];
#pragma unroll
    for (int k = 0; k < kDotRows; ++k) {
      const int i = r + kDotLanes * k;
      a1[k] = c1[i < m ? i : 0];
    }
    const double xj = L.x[j];
    double s1 = 0.0;
#pragma unroll
    for (int k = 0; k < kDotRows; ++k) {
      const int i = r + kDotLanes * k;
      s1 = fma(a1[k], i < m ? yv[i] : 0.0, s1);
    }
    const double tj = group_sum<kDotLanes>(s1);
    if (r == 0) L.t[j] = tj;
#pragma unroll
    for (int k = 0; k < kDotRows; ++k) qa[k] = fma(a1[k], xj, qa[k]);
  }
#pragma unroll
  for (int k = 0; k < kDotRows; ++k) {
    qa[k] += __shfl_xor(qa[k], 16, 64);
    qa[k] += __shfl_xor(qa[k], 32, 64);
  }
  const int lane = tid & 63, wave = tid >> 6;
  for (int w = 0; w < kWaves; ++w) {
    if (wave == w && lane < kDotLanes) {
#pragma unroll
      for (int k = 0; k < kDotRows; ++k) {
        const int i = r + kDotLanes * k;
        if (i < m) qs[i] += qa[k];
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < m; i += kThreads) L.rb[i] = L.b[i] - qs[i];
  __syncthreads();
}

// dy = M^-1 rhs by PCG. Returns iterations. aty (n) receives A^T dy, kept
// alongside dy from the products the iteration forms anyway: A^T x0 from the
// first residual's product, then aty += alpha A^T p with A^T p (atp, n
// scratch) from each step's product.
template <typename T>
__device__ int solve_normal(const Lp& L, const T* M, double* stage, double* red, double tol, int max_it, double* aty,
                            double* atp) {
  const int tid = threadIdx.x, m = L.m;
  double* pv = stage + lay::kStagePv;  // m doubles: preconditioner vector (LDS)
  double* tv = stage + lay::kStageTv;  // m doubles: preconditioner temporary (LDS)
  // x0 = P rhs
  for (int i = tid; i < m; i += kThreads) pv[i] = L.rhs[i];
  precondition(L, M, pv, tv);
  for (int i = tid; i < m; i += kThreads) L.dy[i] = pv[i];
  __syncthreads();
  normal_matvec(L, L.dy, L.q, stage, aty);
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
    normal_matvec(L, L.p, L.q, stage, atp);
    double pq = 0.0;
    for (int i = tid; i < m; i += kThreads) pq += L.p[i] * L.q[i];
    pq = bred<0>(pq, red);
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    for (int j = tid; j < L.n; j += kThreads) aty[j] = fma(alpha, atp[j], aty[j]);
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
  // t moves to dx, where the last loop reads it element by element
  for (int j = tid; j < L.n; j += kThreads) dx[j] = L.t[j];
  __syncthreads();
  // A^T dy comes back in dw (dv: scratch), both written below only after
  // their element has been read
  const int its = solve_normal(L, M, stage, red, pcg_tol, o.pcg_max, dw, dv);
  __syncthreads();
  for (int j = tid; j < L.n; j += kThreads) {
    const double atdy = dw[j];
    if (!active(L, j)) {
      dx[j] = dw[j] = dz[j] = dv[j] = 0.0;
      continue;
    }
    const double dxj = L.D[j] * atdy - dx[j];
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
                                                             double* trace, int trace_lp,
                                                             const double* __restrict__ state_in, double* state_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* stage = reinterpret_cast<double*>(smem);                    // kStage doubles
  double* red = reinterpret_cast<double*>(smem + lay::lds_off_red());  // kRed doubles
  float* M = reinterpret_cast<float*>(smem + lay::lds_off_matrix());   // packed lower triangle
  __shared__ int s_lp;
  __shared__ int s_drop;  // pivots the FP32 factorization dropped
  __shared__ Lp s_L;
  __shared__ float s_xkk[kBlk * (kBlk + 1)];  // inverse of the diagonal block being factored
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
      W.Ar = mat.a64r;
      W.ldr = mat.ldr;
      W.b = b_all + dsc.b_off;
      W.c = c_all + dsc.x_off;
      W.u = u_all + dsc.x_off;
      double* wsp = ws + dsc.ws_off;
      double** nv[lay::kNVectors] = {&W.x, &W.w, &W.z, &W.v, &W.rc, &W.ru, &W.D, &W.dx, &W.dw, &W.dz, &W.dv,
                                     &W.dxa, &W.dwa, &W.dza, &W.dva, &W.t};
      for (int k = 0; k < lay::kNVectors; ++k) *nv[k] = wsp + lay::n_vector(k, W.n);
      double** mv[lay::kMVectors] = {&W.y, &W.rb, &W.dy, &W.r, &W.p, &W.q, &W.zz, &W.rhs, &W.sc};
      for (int k = 0; k < lay::kMVectors; ++k) *mv[k] = wsp + lay::m_vector(k, W.m, W.n);
      W.M64 = wsp + lay::off_m64(W.m, W.n);
      W.xs = reinterpret_cast<float*>(wsp + lay::off_xs(W.m, W.n));
      W.sd = reinterpret_cast<float*>(
          lay::sd_address(reinterpret_cast<std::uintptr_t>(wsp + lay::off_xs(W.m, W.n) + lay::xs_doubles(W.m))));
    }
    __syncthreads();
    const Lp& L = s_L;
    const int m = L.m, n = L.n;
    // Starting point: x mid-box (or 1), w = u - x, z = v = 1, y = 0; or the
    // given warm start (x, w, z, v: n each, then y: m), kept strictly
    // inside the bounds (every component of an active column at least
    // 1e-10, w = u - x on bounded ones).
    double nb = 0.0;
    const double* st = (state_in && dsc.in_off >= 0) ? state_in + dsc.in_off : nullptr;
    for (int j = tid; j < n; j += kThreads) {
      const double uj = L.u[j];
      double xj = 0.0, wj = 0.0, vj = 0.0, zj = 0.0;
      if (uj > 0.0) {
        const bool bnd = uj < 1e29;
        if (st) {
          xj = bnd ? fmin(fmax(st[j], 1e-10 * uj), uj - 1e-10 * uj) : fmax(st[j], 1e-10);
          wj = bnd ? uj - xj : 0.0;
          zj = fmax(st[2 * n + j], 1e-10);
          vj = bnd ? fmax(st[3 * n + j], 1e-10) : 0.0;
        } else {
          xj = bnd ? 0.5 * uj : 1.0;
          wj = bnd ? uj - xj : 0.0;
          zj = 1.0;
          vj = bnd ? 1.0 : 0.0;
        }
        nb += bnd ? 1.0 : 0.0;
      }
      L.x[j] = xj;
      L.w[j] = wj;
      L.z[j] = zj;
      L.v[j] = vj;
    }
    for (int i = tid; i < m; i += kThreads) L.y[i] = st ? st[4 * n + i] : 0.0;
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
    bool fp64 = false;  // FP64 factorizations (global memory): the FP32 factor failed
    int fp64_from = -1;
    double best = 1e300;
    int best_it = 0;
    bool saved = false;  // the warm-start iterate has been written
    for (int it = 0; it < o.max_iter; ++it) {
      __syncthreads();
      long long t0 = clock64();
      // residuals: rb = b - A x, t = A^T y
      residual_products(L, stage);
      double rbmax = 0.0, rcmax = 0.0, rumax = 0.0, cx = 0.0, by = 0.0, uv = 0.0, xz = 0.0;
      for (int i = tid; i < m; i += kThreads) {
        rbmax = fmax(rbmax, fabs(L.rb[i]));
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
      if (state_out && dsc.out_off >= 0 && !saved && mu < o.save_mu) {  // a centred point for warm starts
        double* so = state_out + dsc.out_off;
        for (int j = tid; j < n; j += kThreads) {
          so[j] = L.x[j];
          so[n + j] = L.w[j];
          so[2 * n + j] = L.z[j];
          so[3 * n + j] = L.v[j];
        }
        for (int i = tid; i < m; i += kThreads) so[4 * n + i] = L.y[i];
        saved = true;
      }
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
        if (fp64) assemble64(L, L.M64, 0.0);
        else assemble32(L, M, 0.0);
        t0 = clock64();
        note(it, 10, static_cast<double>(t0 - t1));
        // an LP in FP64 mode stages its panels in the LDS region the FP32
        // matrix used, when they fit there
        const bool blocked64 = lay::fp64_panel_fits(L.m, o.lds_matrix_bytes);
        double* P64 = reinterpret_cast<double*>(M);
        const int dropped = fp64 ? (blocked64 ? cholesky64_blocked(L, L.M64, P64) : cholesky(L, L.M64))
                                 : cholesky_blocked(L, M, s_xkk);
        note(it, 15, dropped);
        t1 = clock64();
        note(it, 11, static_cast<double>(t1 - t0));
        // An FP32 factor that dropped a pivot preconditions the solves badly
        // (they then take tens of CG steps, each dearer than the FP64
        // factorization): the LP switches to FP64 and redoes the iteration.
        if (!fp64) {
          if (tid == 0) s_drop = dropped;  // counted on wavefront 0
          __syncthreads();
          if (s_drop > 0) {
            fp64 = true;
            fp64_from = it;
            stalled = 0;
            best_it = it;
            continue;
          }
        }
        if (fp64) {
          if (blocked64) invert64_blocked(L, L.M64, P64, stage, stage + lay::kPanel64 * lay::kPanel64);
          else invert_lower(L, L.M64, stage);
        } else {
          invert_lower_blocked(L, M);
        }
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
    if (state_out && dsc.out_off >= 0 && !saved) {
      double* so = state_out + dsc.out_off;
      for (int j = tid; j < n; j += kThreads) {
        so[j] = L.x[j];
        so[n + j] = L.w[j];
        so[2 * n + j] = L.z[j];
        so[3 * n + j] = L.v[j];
      }
      for (int i = tid; i < m; i += kThreads) so[4 * n + i] = L.y[i];
    }
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

using lay::lds_bytes;

template <typename T>
T* dev_alloc(sov_batch* h, std::size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<std::size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
  h->owned.push_back(p);
  return static_cast<T*>(p);
}

// Row-major copies of the column-major m x n matrix a, leading dimension
// ldr >= n, zero in columns n .. ldr - 1: FP64 and FP32 (rounded to nearest).
__global__ void row_major_copies_kernel(int m, int n, int ldr, const double* __restrict__ a, float* __restrict__ a32,
                                        double* __restrict__ a64r) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(m) * ldr) return;
  const int i = static_cast<int>(idx / ldr), j = static_cast<int>(idx % ldr);
  const double v = j < n ? a[i + static_cast<long long>(j) * m] : 0.0;
  a32[idx] = static_cast<float>(v);
  a64r[idx] = v;
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
  const int ldr = row_ld(n);
  const std::size_t cntr = static_cast<std::size_t>(m) * ldr;
  double* a64 = dev_alloc<double>(h, cnt);
  float* a32 = dev_alloc<float>(h, cntr);
  double* a64r = dev_alloc<double>(h, cntr);
  if (!a64 || !a32 || !a64r) return SOV_ERR_NOMEM;
  // the column-major copy goes up (the call returns after the caller's
  // array has been read); the row-major copies are built from it on the
  // device, ahead of any solve on the same stream
  hipStream_t s = sov::stream();
  SOV_HIP_TRY(hipMemcpyAsync(a64, a_colmajor, cnt * sizeof(double), hipMemcpyHostToDevice, s));
  row_major_copies_kernel<<<static_cast<unsigned>((cntr + 255) / 256), 256, 0, s>>>(m, n, ldr, a64, a32, a64r);
  SOV_HIP_TRY(hipGetLastError());
  SOV_HIP_TRY(hipStreamSynchronize(s));
  h->mats.push_back({m, n, a64, a32, ldr, a64r});
  *id = static_cast<int32_t>(h->mats.size() - 1);
  return SOV_OK;
}

int sov_batch_solve(sov_batch* h, int32_t count, const int32_t* matrix_id, const double* b_all,
                    const int64_t* b_off, const double* c_all, const double* u_all, const int64_t* x_off,
                    const sov_ipm_options* opt, sov_batch_result* res, double* x_all) {
  return sov_batch_solve_warm(h, count, matrix_id, b_all, b_off, c_all, u_all, x_off, opt, res, x_all, nullptr,
                              nullptr, nullptr, nullptr, 0.0);
}

int sov_batch_solve_warm(sov_batch* h, int32_t count, const int32_t* matrix_id, const double* b_all,
                         const int64_t* b_off, const double* c_all, const double* u_all, const int64_t* x_off,
                         const sov_ipm_options* opt, sov_batch_result* res, double* x_all, const double* state_in,
                         const int64_t* in_off, double* state_out, const int64_t* out_off, double save_mu) {
  if (!h || count < 0 || !matrix_id || !b_all || !b_off || !c_all || !u_all || !x_off || !opt || !res ||
      (state_in && !in_off) || (state_out && !out_off))
    return SOV_ERR_ARG;
  if (count == 0) return SOV_OK;
  hipStream_t s = sov::stream();
  std::vector<BatchDesc> descs(static_cast<std::size_t>(count));
  long long b_total = 0, x_total = 0, ws_total = 0, in_total = 0, out_total = 0;
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
    const long long io = state_in ? in_off[k] : -1, oo = state_out ? out_off[k] : -1;
    descs[k] = {id, 0, b_off[k], x_off[k], ws_total, io < 0 ? -1 : io, oo < 0 ? -1 : oo};
    b_total = std::max<long long>(b_total, b_off[k] + mt.m);
    x_total = std::max<long long>(x_total, x_off[k] + mt.n);
    if (io >= 0) in_total = std::max<long long>(in_total, io + 4LL * mt.n + mt.m);
    if (oo >= 0) out_total = std::max<long long>(out_total, oo + 4LL * mt.n + mt.m);
    ws_total += lay::ws_doubles(mt.m, mt.n);
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
  double* d_sin = in_total > 0 ? tmp.get<double>(static_cast<std::size_t>(in_total)) : nullptr;
  double* d_sout = out_total > 0 ? tmp.get<double>(static_cast<std::size_t>(out_total)) : nullptr;
  if (!d_mats || !d_desc || !d_b || !d_c || !d_u || !d_ws || (x_all && !d_x) || !d_out || !d_queue ||
      (in_total > 0 && !d_sin) || (out_total > 0 && !d_sout))
    return SOV_ERR_NOMEM;
  if (d_sin) SOV_HIP_TRY(hipMemcpyAsync(d_sin, state_in, in_total * sizeof(double), hipMemcpyHostToDevice, s));
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
  o.save_mu = save_mu;
  o.lds_matrix_bytes = lay::lds_matrix_bytes(mmax);
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
                                                          d_out, o, d_trace, trace_lp, d_sin, d_sout);
  const hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, s);
  if (le != hipSuccess) {
    sov::set_error(std::string("batch_ipm_kernel launch: ") + hipGetErrorString(le));
    return SOV_ERR_DEVICE;
  }
  std::vector<sov_batch_result_dev> host(static_cast<std::size_t>(count));
  SOV_HIP_TRY(hipMemcpyAsync(host.data(), d_out, count * sizeof(sov_batch_result_dev), hipMemcpyDeviceToHost, s));
  if (x_all) SOV_HIP_TRY(hipMemcpyAsync(x_all, d_x, x_total * sizeof(double), hipMemcpyDeviceToHost, s));
  if (d_sout) SOV_HIP_TRY(hipMemcpyAsync(state_out, d_sout, out_total * sizeof(double), hipMemcpyDeviceToHost, s));
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
