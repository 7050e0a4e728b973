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
#include <cstdlib>
#include <utility>
#include <vector>

namespace sov {
namespace {

std::atomic<int> g_timing{0};  // 0 off, 1 every kernel class, 2 the bulk SYRK only

constexpr int kPanelLd = kPanel + 1;  // LDS leading dimension of a panel block
constexpr int kPanelBuffers = 3;      // L21 buffers of panel pairs in rotation (see factor)

// ---------------------------------------------------------------------------
// Diagonal block: Cholesky of the nb x nb block (nb <= kPanel = 128) and the
// inverse X = L^-1 of its factor. The block lives in LDS as S (column-major,
// leading dimension kPanelLd): L in the lower triangle, X^T in the strict
// upper triangle (X(i, k) at S[i * kPanelLd + k] for i > k) and the diagonal
// of X in xdiag. It is processed in sub-panels of kSub = 32 columns; for
// sub-panel c:
//   A  its 32 x 32 diagonal block and that block's inverse, by wavefront 0
//      alone (two 16 x 16 register leaves joined by MFMA products);
//   B  the rows below it inside the block: L(r, c) = A(r, c) inv(L_cc)^T;
//   C  the rank-32 update of the trailing lower part of the block;
//   Y/X the inverse blocks left of its diagonal block,
//      X(c, <c) = -inv(L_cc) Y,  Y = L(c, <c) X(<c, <c).
// Look-ahead: once B(c) is stored, wavefront 0 applies C(c) to the next
// sub-panel's diagonal block only and goes straight on to A(c + 1), while
// wavefronts 1-7 apply the rest of C(c) and form X(c) and Y(c + 1). These
// are grouped by 16-column group, one wavefront per group, so that Y(c + 1)
// finds the X(c) it needs (the same columns) already written by its own
// wavefront. The critical path per sub-panel is then A, B and the small
// diagonal part of C. B, C, Y and X are 16 x 16 MFMA tiles with operands
// read from LDS.
// The pivot rule is the reference LDL^T's: a pivot <= 1e-13 x |its diagonal
// before elimination| is dropped (L_jj = DropPivot), a non-finite or
// < 1e-30 pivot is a breakdown. X goes to inv_out (leading dimension ldi,
// zero outside the nb x nb lower triangle): it makes the panel TRSM a GEMM
// and seeds the super-block inverses of the triangular solves. L11 itself is
// not written back: nothing reads it.
// ---------------------------------------------------------------------------
constexpr int kSub = 32;
constexpr int kPanelThreads = 512;
constexpr int kSubLd = kSub + 1;

template <typename T>
__device__ __forceinline__ T read_lane(T v, int lane);
template <>
__device__ __forceinline__ double read_lane<double>(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <>
__device__ __forceinline__ float read_lane<float>(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Each lane receives v from lane `src` (its own choice): ds_bpermute, which
// goes through the LDS crossbar without touching LDS memory.
__device__ __forceinline__ float lane_pull(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(4 * src, __float_as_int(v)));
}
__device__ __forceinline__ double lane_pull(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(4 * src, static_cast<int>(b));
  const int hi = __builtin_amdgcn_ds_bpermute(4 * src, static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

constexpr int kLeaf = 16;             // leaf width of the diagonal sub-block factorization
constexpr int kLeafLd = kLeaf + 1;

// Orders the LDS accesses of one wavefront (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// 1 / sqrt(d) in the factor's precision: the hardware estimate refined by
// Newton steps x <- x (1.5 - 0.5 d x^2), two in FP64, one in FP32.
template <typename T>
struct LeafMath;
template <>
struct LeafMath<double> {
  __device__ static __forceinline__ double rsqrt(double d) {
    double rs = __builtin_amdgcn_rsq(d);
    rs = rs * (1.5 - 0.5 * d * rs * rs);
    return rs * (1.5 - 0.5 * d * rs * rs);
  }
};
template <>
struct LeafMath<float> {
  __device__ static __forceinline__ float rsqrt(float d) {
    const float rs = __builtin_amdgcn_rsqf(d);
    return rs * (1.5f - 0.5f * d * rs * rs);
  }
};

template <typename T>
struct LeafRcp;
template <>
struct LeafRcp<float> {
  __device__ static __forceinline__ float rcp(float d) { return __builtin_amdgcn_rcpf(d); }
};
template <>
struct LeafRcp<double> {
  __device__ static __forceinline__ double rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
  }
};

// 1 / d_j for the update of step j. EXACT = false takes the pivot as is and
// clears `plain` when it would have needed the pivot rule; EXACT applies it
// (drop: 1 / DropPivot^2, breakdown: 1).
template <typename T, bool EXACT>
__device__ __forceinline__ T leaf_rinv(T d, int j, int h, T lim, bool& plain) {
  if constexpr (EXACT) {
    const bool active = j < h;
    const bool finite = isfinite(d);
    const bool drop = active & finite & (d <= lim);
    const bool bad = active & !drop & (!finite | (d < T(1e-30)));
    constexpr T kDrop = DropPivot<T>::value;
    return drop ? T(1) / (kDrop * kDrop) : (bad ? T(1) : LeafRcp<T>::rcp(d));
  } else {
    plain = plain & (d > lim) & (d < T(INFINITY));
    return LeafRcp<T>::rcp(d);
  }
}

// x -= f(lane J) nsx' and f -= f(lane J) nt' for slots of the leaf, each one
// instruction: the broadcast of lane J's f along its 16-lane row is the DPP
// operand of the fused multiply-add (v_fmac_*_dpp row_newbcast:J), and
// both read f before the second writes it. A DPP read of a register written
// by one of the two instructions just before it needs two wait states, so
// each statement starts and ends with s_nop 1 (the compiler's code around it
// cannot see what it reads and writes).
#define SOV_FMAC_PAIR(OP, F, X) \
  OP " " X ", " F ", %[nsx] row_newbcast:%[j] row_mask:0xf bank_mask:0xf\n\t" OP " " F ", " F \
     ", %[nt] row_newbcast:%[j] row_mask:0xf bank_mask:0xf\n\t"
template <int J, typename T>
__device__ __forceinline__ void leaf_fmac1(T& f0, T& x0, T nt, T nsx) {
  if constexpr (sizeof(T) == 4)
    asm volatile("s_nop 1\n\t" SOV_FMAC_PAIR("v_fmac_f32_dpp", "%[f0]", "%[x0]") "s_nop 1"
                 : [f0] "+v"(f0), [x0] "+v"(x0)
                 : [nt] "v"(nt), [nsx] "v"(nsx), [j] "n"(J));
  else
    asm volatile("s_nop 1\n\t" SOV_FMAC_PAIR("v_fmac_f64_dpp", "%[f0]", "%[x0]") "s_nop 1"
                 : [f0] "+v"(f0), [x0] "+v"(x0)
                 : [nt] "v"(nt), [nsx] "v"(nsx), [j] "n"(J));
}
template <int J, typename T>
__device__ __forceinline__ void leaf_fmac3(T& f0, T& x0, T& f1, T& x1, T& f2, T& x2, T nt, T nsx) {
  if constexpr (sizeof(T) == 4)
    asm volatile("s_nop 1\n\t" SOV_FMAC_PAIR("v_fmac_f32_dpp", "%[f0]", "%[x0]")
                     SOV_FMAC_PAIR("v_fmac_f32_dpp", "%[f1]", "%[x1]")
                         SOV_FMAC_PAIR("v_fmac_f32_dpp", "%[f2]", "%[x2]") "s_nop 1"
                 : [f0] "+v"(f0), [x0] "+v"(x0), [f1] "+v"(f1), [x1] "+v"(x1), [f2] "+v"(f2), [x2] "+v"(x2)
                 : [nt] "v"(nt), [nsx] "v"(nsx), [j] "n"(J));
  else
    asm volatile("s_nop 1\n\t" SOV_FMAC_PAIR("v_fmac_f64_dpp", "%[f0]", "%[x0]")
                     SOV_FMAC_PAIR("v_fmac_f64_dpp", "%[f1]", "%[x1]")
                         SOV_FMAC_PAIR("v_fmac_f64_dpp", "%[f2]", "%[x2]") "s_nop 1"
                 : [f0] "+v"(f0), [x0] "+v"(x0), [f1] "+v"(f1), [x1] "+v"(x1), [f2] "+v"(f2), [x2] "+v"(x2)
                 : [nt] "v"(nt), [nsx] "v"(nsx), [j] "n"(J));
}
#undef SOV_FMAC_PAIR

// Step J of the leaf (see leaf_factor). On entry nt = -f(J, i) / d_J (0 for
// i <= J) and nsx = -Xu(J, i) / d_J, already on every row group, and d =
// d_J. The step zeroes row J right of the diagonal (spent: no later step
// may read it), updates the slot holding row J + 1 first and starts the
// next step's chain from it (row J + 1 pulled to all groups, the pivot
// d_{J+1} read from its lane, its reciprocal), then updates the other three
// slots, which hides that chain, and finally forms the next multipliers.
template <typename T, bool EXACT, int J>
__device__ __forceinline__ void leaf_slot_step(T (&f)[4], T (&x)[4], int g, int i, int h, T lim, bool& plain, T& nt,
                                               T& nsx, T& d, T& myd) {
  constexpr int G = J >> 2, S = J & 3;
  myd = i == J ? d : myd;
  if constexpr (J + 1 < kLeaf) {
    constexpr int S1 = (J + 1) & 3, G1 = (J + 1) >> 2;
    f[S] = (g == G && i >= J) ? T(0) : f[S];
    leaf_fmac1<J>(f[S1], x[S1], nt, nsx);
    const T rf = lane_pull(f[S1], 16 * G1 + i), rx = lane_pull(x[S1], 16 * G1 + i);
    d = read_lane(f[S1], 16 * G1 + J + 1);
    const T rinv = leaf_rinv<T, EXACT>(d, J + 1, h, read_lane(lim, J + 1), plain);
    __builtin_amdgcn_sched_barrier(0);
    leaf_fmac3<J>(f[(J + 2) & 3], x[(J + 2) & 3], f[(J + 3) & 3], x[(J + 3) & 3], f[S], x[S], nt, nsx);
    nt = i > J + 1 ? -(rf * rinv) : T(0);
    nsx = -(rx * rinv);
    asm volatile("" : "+v"(nt), "+v"(nsx));
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename T, bool EXACT, int... J>
__device__ __forceinline__ void leaf_slot_steps(T (&f)[4], T (&x)[4], int g, int i, int h, T lim, bool& plain,
                                                T& myd, std::integer_sequence<int, J...>) {
  T d = read_lane(f[0], 0);
  const T rinv = leaf_rinv<T, EXACT>(d, 0, h, read_lane(lim, 0), plain);
  T nt = i > 0 ? -(lane_pull(f[0], i) * rinv) : T(0);
  T nsx = -(lane_pull(x[0], i) * rinv);
  (leaf_slot_step<T, EXACT, J>(f, x, g, i, h, lim, plain, nt, nsx, d, myd), ...);
}

// Cholesky factor L and inverse X = L^-1 of the h x h (h <= 16) block at D
// (column-major, leading dimension kPanelLd) by one wavefront. Lane
// (g, i) = (lane >> 4, lane & 15) holds rows 4g .. 4g + 3 (slots) of column i
// of the symmetric block f and of the unit-lower inverse Xu, in registers.
// Step j: every lane updates all its slots,
//   f(a, i) -= f(a, j) f(j, i) / d_j   (i > j),
//   Xu(a, i) -= f(a, j) Xu(j, i) / d_j,
// f(a, j) coming from lane j of the lane's own row group (DPP) and
// f(j, i), Xu(j, i) from row group j >> 2 (ds_bpermute); row j of f right of
// the diagonal is zeroed first, which makes the updates of rows <= j exact
// no-ops, so no step needs a row mask. A step past the block (j >= h) meets
// an identity pad and changes nothing. This is Gauss-Jordan on the
// unit-lower factor of f = L_u D L_u^T; at the end L(a, i) = f(a, i) /
// sqrt(d_i) and X(a, i) = Xu(a, i) / sqrt(d_a). The fast pass takes every
// pivot as is; if one needed the pivot rule (as described above the kernel)
// the exact pass reruns the block with it. Writes L lower with its
// diagonal, X strictly above the diagonal (X(i, k) at D[i * kPanelLd + k])
// and 1 / L(j, j) to xd[j]; counts drops and breakdowns into flags once.
// thr: lane j (< 16) holds the drop threshold of column j.
template <typename T>
__device__ __forceinline__ void leaf_factor(T* D, int h, T thr, T* xd, FactorFlags* flags, int lane) {
  const int g = lane >> 4, i = lane & 15;
  T f[4], x[4];
  auto load = [&]() {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int a = 4 * g + s;
      const bool in = i < h && a < h;
      const int off = in ? (a >= i ? i * kPanelLd + a : a * kPanelLd + i) : 0;
      const T v = D[off];
      f[s] = in ? v : (a == i ? T(1) : T(0));  // identity pad past the block
      x[s] = a == i ? T(1) : T(0);
    }
  };
  load();
  bool plain = true;
  T myd = T(1);
  const T lim = thr > T(1e-30) ? thr : T(1e-30);
  leaf_slot_steps<T, false>(f, x, g, i, h, lim, plain, myd, std::make_integer_sequence<int, kLeaf>{});
  bool exact = false;
  if (!plain) {  // uniform
    load();
    myd = T(1);
    leaf_slot_steps<T, true>(f, x, g, i, h, thr, plain, myd, std::make_integer_sequence<int, kLeaf>{});
    exact = true;
  }
  const T d = myd;
  T rs = LeafMath<T>::rsqrt(d), piv = d * rs;
  if (exact) {
    // the threshold of column i lives in lane i of row group 0 only
    const T thr_i = lane_pull(thr, i);
    const bool active = i < h;
    const bool finite = isfinite(d);
    const bool drop = active & finite & (d <= thr_i);
    const bool bad = active & !drop & (!finite | (d < T(1e-30)));
    rs = drop ? T(1) / DropPivot<T>::value : (bad ? T(1) : rs);
    piv = drop ? DropPivot<T>::value : (bad ? T(1) : piv);
    const unsigned long long nd = __ballot(drop && lane < 16), nb = __ballot(bad && lane < 16);
    if (lane == 0) {
      if (nd) atomicAdd(&flags->replaced, __popcll(nd));
      if (nb) atomicOr(&flags->status, 1);
    }
  }
  // 1 / L(a, a) of each slot's row a = 4g + s, from lane a of the own group
  T rsa[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) rsa[s] = lane_pull(rs, 20 * g + s);
  if (lane < h) xd[i] = rs;
  if (i < h) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int a = 4 * g + s;
      if (a < h && a >= i) D[i * kPanelLd + a] = a == i ? piv : f[s] * rs;  // L(a, i)
      if (a < h && a > i) D[a * kPanelLd + i] = x[s] * rsa[s];              // X(a, i) = Xu(a, i) / L(a, a)
    }
  }
}

// Phase A for sub-panel c (width sb) by one wavefront: the two 16 x 16
// leaves (leaf_factor) joined by three 16 x 16 x 16 MFMA products from LDS,
// L21 = A21 X11^T, A22 -= L21 L21^T, X21 = -X22 (L21 X11). Ys: a 16 x 16
// scratch only this wavefront uses.
template <typename T>
__device__ __forceinline__ void panel_diag(T* S, int c, int sb, T* xdiag, const double* d0, T* Ys,
                                           FactorFlags* flags, int lane) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int li = lane & 15, lk = lane >> 4;
  T* D = S + c * (kPanelLd + 1);  // element (i, k) of the diagonal block at D[k * kPanelLd + i]
  const int h1 = min(kLeaf, sb), h2 = sb - h1;
  const T thr1 = static_cast<T>(lane < h1 ? 1e-13 * fabs(d0[c + lane]) : 0.0);
  leaf_factor<T>(D, h1, thr1, xdiag + c, flags, lane);
  if (h2 <= 0) return;
  wave_lds_sync();
  {  // L21 = A21 X11^T over rows kLeaf.., columns 0..kLeaf-1
    T av[kLeaf / 4], bv[kLeaf / 4];
    const T xl = xdiag[c + li];
#pragma unroll
    for (int s = 0; s < kLeaf / 4; ++s) {
      const int kk = 4 * s + lk;
      const T a0 = D[kk * kPanelLd + kLeaf + li], b0 = D[li * kPanelLd + kk];
      av[s] = li < h2 ? a0 : T(0);
      bv[s] = li > kk ? b0 : (li == kk ? xl : T(0));
    }
    __builtin_amdgcn_sched_barrier(0);
    acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < kLeaf / 4; ++s) acc = Mfma<T>::step(av[s], bv[s], acc);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int i = Mfma<T>::row(lk, g);
      if (i < h2) D[li * kPanelLd + kLeaf + i] = acc[g];
    }
  }
  wave_lds_sync();
  {  // A22 -= L21 L21^T (lower part) and Y = L21 X11
    T lv[kLeaf / 4], xv[kLeaf / 4];
    const T xl = xdiag[c + li];
#pragma unroll
    for (int s = 0; s < kLeaf / 4; ++s) {
      const int kk = 4 * s + lk;
      const T l0 = D[kk * kPanelLd + kLeaf + li], x0 = D[kk * kPanelLd + li];
      lv[s] = li < h2 ? l0 : T(0);                    // L21(li, kk)
      xv[s] = kk > li ? x0 : (kk == li ? xl : T(0));  // X11(kk, li)
    }
    __builtin_amdgcn_sched_barrier(0);
    acc_t acc = acc_t{0, 0, 0, 0}, yacc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < kLeaf / 4; ++s) {
      acc = Mfma<T>::step(lv[s], lv[s], acc);
      yacc = Mfma<T>::step(lv[s], xv[s], yacc);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int i = Mfma<T>::row(lk, g);
      if (i < h2 && i >= li) D[(kLeaf + li) * kPanelLd + kLeaf + i] -= acc[g];
      Ys[i * kLeafLd + li] = yacc[g];
    }
  }
  wave_lds_sync();
  T* D2 = D + kLeaf * (kPanelLd + 1);
  const T thr2 = static_cast<T>(lane < h2 ? 1e-13 * fabs(d0[c + kLeaf + lane]) : 0.0);
  leaf_factor<T>(D2, h2, thr2, xdiag + c + kLeaf, flags, lane);
  wave_lds_sync();
  {  // X21 = -X22 Y
    T av[kLeaf / 4], bv[kLeaf / 4];
    const T xl = xdiag[c + kLeaf + li];
#pragma unroll
    for (int s = 0; s < kLeaf / 4; ++s) {
      const int kk = 4 * s + lk;
      const T a0 = D2[li * kPanelLd + kk];
      av[s] = li > kk ? a0 : (li == kk ? xl : T(0));
      bv[s] = Ys[kk * kLeafLd + li];
    }
    __builtin_amdgcn_sched_barrier(0);
    acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < kLeaf / 4; ++s) acc = Mfma<T>::step(av[s], bv[s], acc);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int i = Mfma<T>::row(lk, g);
      if (i < h2) D[(kLeaf + i) * kPanelLd + li] = -acc[g];
    }
  }
}

// In the tile helpers below, each tile's operands are read from LDS at
// in-range addresses without branching (values outside the tile's triangle
// or rows are selected to zero afterwards), all before the tile's MFMA
// chain, so that the reads overlap instead of each waiting in front of its
// instruction.

// X tile of sub-panel c (width sb): rows 16 half.. of the sub-panel, the 16
// columns of group g (< c / 16): X(c + i, x) = -sum_v inv(L_cc)(i, v) Y(v, x),
// with Y(v, x) at Yt[x * kSubLd + v].
template <typename T>
__device__ __forceinline__ void panel_x_tile(T* S, const T* Yt, const T* xdiag, int c, int sb, int half, int g,
                                             int li, int lk) {
  typedef typename Mfma<T>::acc_t acc_t;
  const T* D = S + c * (kPanelLd + 1);
  const int i = 16 * half + li;  // row of inv(L_cc) of this lane
  const int x0 = 16 * g;
  const T xi = xdiag[c + i];
  T av[kSub / 4], bv[kSub / 4];
#pragma unroll
  for (int s2 = 0; s2 < kSub / 4; ++s2) {
    const int v = 4 * s2 + lk;
    const T a0 = D[i * kPanelLd + v];
    av[s2] = i < sb ? (v < i ? a0 : (v == i ? xi : T(0))) : T(0);
    bv[s2] = Yt[(x0 + li) * kSubLd + v];
  }
  __builtin_amdgcn_sched_barrier(0);
  acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int s2 = 0; s2 < kSub / 4; ++s2) acc = Mfma<T>::step(av[s2], bv[s2], acc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ii = 16 * half + Mfma<T>::row(lk, q);
    if (ii < sb) S[(c + ii) * kPanelLd + x0 + li] = -acc[q];
  }
}

// Y tile of sub-panel c (width sb): rows 16 half.. of the sub-panel, the 16
// columns x of group g (x0 = 16 g < c):
// Y(v, x) = sum_{x0 <= j < c} L(c + v, j) X(j, x) into Yt[x * kSubLd + v]
// (X(j, x) = 0 for j < x0; blocks of 16 j's, the next block's operands read
// while the current one's MFMAs run).
template <typename T>
__device__ __forceinline__ void panel_y_tile(const T* S, T* Yt, const T* xdiag, int c, int sb, int half, int g,
                                             int li, int lk) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int v = 16 * half + li;  // row of L(c + v, .) of this lane
  const int x0 = 16 * g;
  const int kx = x0 + li;        // column of X of this lane
  const T xk = xdiag[kx];
  auto read_block = [&](int j0, T* ao, T* bo) {
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      const int jj = j0 + 4 * s2 + lk;
      const T a0 = S[jj * kPanelLd + c + v], b0 = S[jj * kPanelLd + kx];
      ao[s2] = v < sb ? a0 : T(0);
      bo[s2] = jj > kx ? b0 : (jj == kx ? xk : T(0));
    }
  };
  acc_t acc = acc_t{0, 0, 0, 0};
  T av[4], bv[4];
  read_block(x0, av, bv);
  for (int j0 = x0; j0 < c; j0 += 16) {
    T an[4] = {T(0), T(0), T(0), T(0)}, bn[4] = {T(0), T(0), T(0), T(0)};
    if (j0 + 16 < c) read_block(j0 + 16, an, bn);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) acc = Mfma<T>::step(av[s2], bv[s2], acc);
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      av[s2] = an[s2];
      bv[s2] = bn[s2];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) Yt[(x0 + li) * kSubLd + 16 * half + Mfma<T>::row(lk, q)] = acc[q];
}

// Lower tile t (row-major order over the lower triangle of 16 x 16 tiles) of
// the rank-32 update by sub-panel c (width kSub) of the block's trailing part
// (rows and columns from c + kSub, below nb).
template <typename T>
__device__ __forceinline__ void panel_c_tile(T* S, int c, int nb, int t, int li, int lk) {
  typedef typename Mfma<T>::acc_t acc_t;
  int ti = 0;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int r = c + kSub + 16 * ti + li, q = c + kSub + 16 * tj + li;
  const int rr = r < nb ? r : 0, qq = q < nb ? q : 0;
  T av[kSub / 4], bv[kSub / 4];
#pragma unroll
  for (int s2 = 0; s2 < kSub / 4; ++s2) {
    const T* col = S + (c + 4 * s2 + lk) * kPanelLd;
    const T a0 = col[rr], b0 = col[qq];
    av[s2] = r < nb ? a0 : T(0);
    bv[s2] = q < nb ? b0 : T(0);
  }
  __builtin_amdgcn_sched_barrier(0);
  acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int s2 = 0; s2 < kSub / 4; ++s2) acc = Mfma<T>::step(av[s2], bv[s2], acc);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int rr2 = c + kSub + 16 * ti + Mfma<T>::row(lk, g);
    if (rr2 < nb && q < nb && rr2 >= q) S[q * kPanelLd + rr2] -= acc[g];
  }
}

template <typename T>
__global__ __launch_bounds__(kPanelThreads) void potrf_panel_kernel(T* __restrict__ a, long long lda, int k0, int nb,
                                                                    const double* __restrict__ diag0,
                                                                    FactorFlags* __restrict__ flags,
                                                                    T* __restrict__ inv_out, long long ldi) {
  typedef typename Mfma<T>::acc_t acc_t;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);                   // [kPanel][kPanelLd]
  T* Yt = S + kPanel * kPanelLd;                           // [kPanel - kSub][kSubLd], Y(v, x) at x * kSubLd + v
  T* xdiag = Yt + (kPanel - kSub) * kSubLd;                // [kPanel]
  T* Ys = xdiag + kPanel;                                  // [kLeaf][kLeafLd], wavefront 0's scratch
  double* d0 = reinterpret_cast<double*>(Ys + kLeaf * kLeafLd);  // [kPanel]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;  // MFMA operand row / k lane group
  constexpr int kWaves = kPanelThreads / 64;
  {
    const int i = tid & (kPanel - 1), k1 = tid >> 7;
    T v[kPanel / 4];
#pragma unroll
    for (int q = 0; q < kPanel / 4; ++q) {
      const int k = k1 + 4 * q;
      v[q] = (k < nb && i >= k && i < nb) ? a[(k0 + i) + static_cast<long long>(k0 + k) * lda] : T(0);
    }
#pragma unroll
    for (int q = 0; q < kPanel / 4; ++q) S[(k1 + 4 * q) * kPanelLd + i] = v[q];
    if (tid < nb) d0[tid] = diag0[k0 + tid];
  }
  __syncthreads();

  // Iteration c: B and the rest of C for sub-panel c (none for c = -kSub),
  // and the diagonal block of sub-panel c + kSub, so that the diagonal-block
  // code exists once.
  int c = -kSub;
  for (;;) {
    const int cn = c + kSub, sbn = min(kSub, nb - cn);
    int nct = 0;
    if (c >= 0) {
      const T* D = S + c * (kPanelLd + 1);  // sub-panel c's diagonal block (sb = kSub: rows follow)
      const int rb = nb - cn;               // rows below it
      // ---- B: L(r, c) = A(r, c) inv(L_cc)^T, 16 x 16 tiles (computed, then
      // stored after a barrier: the operands are read in place)
      const int nrt = (rb + 15) >> 4;  // 16-row tiles below the sub-panel
      const int nbt = 2 * nrt;         // B tiles (two 16-column tiles each row; <= 12)
      nct = nrt * (nrt + 1) / 2;
      acc_t bacc[2];
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const int t = wave + n * kWaves;
        if (t >= nbt) break;
        const int r = cn + 16 * (t >> 1) + li;  // A row of this lane
        const int q = 16 * (t & 1) + li;        // B column of this lane
        const int rr = r < nb ? r : 0;
        const T xq = xdiag[c + q];
        T av[kSub / 4], bv[kSub / 4];
#pragma unroll
        for (int s2 = 0; s2 < kSub / 4; ++s2) {
          const int kk = 4 * s2 + lk;
          const T a0 = S[(c + kk) * kPanelLd + rr], b0 = D[q * kPanelLd + kk];
          av[s2] = r < nb ? a0 : T(0);
          bv[s2] = kk < q ? b0 : (kk == q ? xq : T(0));
        }
        __builtin_amdgcn_sched_barrier(0);
        acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int s2 = 0; s2 < kSub / 4; ++s2) acc = Mfma<T>::step(av[s2], bv[s2], acc);
        bacc[n] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const int t = wave + n * kWaves;
        if (t >= nbt) break;
        const int q = 16 * (t & 1) + li;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int r = cn + 16 * (t >> 1) + Mfma<T>::row(lk, g);
          if (r < nb) S[(c + q) * kPanelLd + r] = bacc[n][g];
        }
      }
      __syncthreads();
    }

    // ---- look-ahead (see above the kernel): wavefront 0 updates the next
    // diagonal block (the C tiles of its two tile rows) and factors it; the
    // helper wavefronts take the column groups (X(c), then Y(c + 1)) and the
    // remaining C tiles. Wavefront 4, which shares wavefront 0's SIMD, is not
    // a helper, so that the latency-bound factorization keeps its SIMD.
    const int ndiag = min(nct, 3);
    if (wave == 0) {
      for (int t = 0; t < ndiag; ++t) panel_c_tile<T>(S, c, nb, t, li, lk);
      wave_lds_sync();
      panel_diag<T>(S, cn, sbn, xdiag, d0, Ys, flags, lane);
    } else if (c >= 0 && wave != kWaves / 2) {
      constexpr int kHelpers = kWaves - 2;
      const int hw = wave - 1 - (wave > kWaves / 2 ? 1 : 0);  // 0 .. kHelpers - 1
      for (int g = hw; g < (cn >> 4); g += kHelpers) {
        if (g < (c >> 4)) {
          panel_x_tile<T>(S, Yt, xdiag, c, kSub, 0, g, li, lk);
          panel_x_tile<T>(S, Yt, xdiag, c, kSub, 1, g, li, lk);
          wave_lds_sync();
        }
        panel_y_tile<T>(S, Yt, xdiag, cn, sbn, 0, g, li, lk);
        panel_y_tile<T>(S, Yt, xdiag, cn, sbn, 1, g, li, lk);
      }
      for (int t = ndiag + hw; t < nct; t += kHelpers) panel_c_tile<T>(S, c, nb, t, li, lk);
    }
    __syncthreads();
    c = cn;
    if (c + kSub >= nb) break;  // no rows below sub-panel c
  }
  // X of the last sub-panel (its Y was formed in the last look-ahead)
  for (int u = wave; u < 2 * (c >> 4); u += kWaves) panel_x_tile<T>(S, Yt, xdiag, c, nb - c, u & 1, u >> 1, li, lk);
  __syncthreads();

  // X to the inverse buffer, zero outside the nb x nb lower triangle. (L11
  // itself is not stored back: the solves and the panel TRSM use only X.)
  {
    const int i = tid & (kPanel - 1), k1 = tid >> 7;
#pragma unroll 8
    for (int q = 0; q < kPanel / 4; ++q) {
      const int k = k1 + 4 * q;
      const bool in = i < nb && k < nb;
      inv_out[i + k * ldi] = in ? (i > k ? S[i * kPanelLd + k] : (i == k ? xdiag[i] : T(0))) : T(0);
    }
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
  return sizeof(T) * (kPanel * kPanelLd + (kPanel - kSub) * kSubLd + kPanel + kLeaf * kLeafLd) +
         sizeof(double) * kPanel;
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

}  // namespace

// Event pairs recorded around kernels while timing is on. Recording does
// not wait; the elapsed times are collected (after one stream sync) by
// sov_kernel_times, so timing can stay on inside a timed region.
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

// A pair for one timed kernel launch whose own dispatch records both events
// (hipExtLaunchKernel's start and stop events: no markers around it), when
// timing is on for `kind` (nullptr otherwise).
// Level 2 times one launch in kLevel2Stride of each kind; the others run
// without events (events on every diagonal-block launch lengthened the
// m = 8192 solve by about 5%). The stride is prime to the 64 diagonal
// blocks and the 31 bulk updates of an m = 8192 factorization, so the timed
// launches move through every position (a stride of 8 timed only the first
// panel of each pair, 20% below the mean).
constexpr unsigned kLevel2Stride = 7;
std::atomic<unsigned> g_level2_seen[kTimeSpmv + 1];

TimedPair* timed_kernel(int kind) {
  const int level = g_timing.load();
  // level 2 keeps to the kernels whose launches record their own events:
  // the bulk SYRK and the diagonal-block kernel
  if (level == 0 || (level == 2 && kind != kTimeSyrk && kind != kTimeSyrk32 && kind != kTimePotrf64 &&
                     kind != kTimePotrf32))
    return nullptr;
  if (level == 2 && g_level2_seen[kind].fetch_add(1u) % kLevel2Stride != 0) return nullptr;
  TimedPair* p = pool().next();
  if (!p) return nullptr;
  p->kind = kind;
  p->flops = 0.0;
  p->launches = 0;
  return p;
}

// Opens a timed pair on `s` when timing is on (nullptr otherwise).
TimedPair* timed_begin(hipStream_t s, int kind) {
  TimedPair* p = timed_kernel(kind);
  if (p) (void)hipEventRecord(p->a, s);
  return p;
}
void timed_end(TimedPair* p, hipStream_t s, int launches, double flops) {
  if (!p) return;
  p->launches = launches;
  p->flops = flops;
  (void)hipEventRecord(p->b, s);
}


void set_kernel_timing(int level) { g_timing.store(level); }
bool kernel_timing() { return g_timing.load() == 1; }
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
    } else if (e.kind == kTimeStrip) {
      t.strip_ms += ms;
    } else if (e.kind == kTimeInverse) {
      t.inverse_ms += ms;
    } else if (e.kind == kTimeAssemble) {
      t.assemble_ms += ms;
    } else if (e.kind == kTimeSyrk32) {
      t.syrk32_ms += ms;
      t.syrk32_launches += e.launches;
      t.syrk32_flops += e.flops;
    } else if (e.kind == kTimePotrf64) {
      t.potrf64_ms += ms;
      t.potrf64_launches += e.launches;
      t.potrf64_flops += e.flops;
    } else if (e.kind == kTimePotrf32) {
      t.potrf32_ms += ms;
      t.potrf32_launches += e.launches;
      t.potrf32_flops += e.flops;
    } else if (e.kind == kTimeSpmv) {
      t.spmv_ms += ms;
      t.spmv_launches += e.launches;
      t.spmv_bytes += e.flops;
    } else {
      t.solve_ms += ms;
      t.solve_pairs += 1.0;
      t.solve_bytes += e.flops;
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
  for (void* p : {static_cast<void*>(panel_), static_cast<void*>(sup_inv_), static_cast<void*>(tmp_),
                  static_cast<void*>(work_), static_cast<void*>(flags_)})
    if (p) (void)hipFree(p);
}

template <typename T>
int DenseCholesky<T>::reserve(int m) {
  if (m <= m_ && flags_) return SOV_OK;
  for (void* p : {static_cast<void*>(panel_), static_cast<void*>(sup_inv_), static_cast<void*>(tmp_),
                  static_cast<void*>(work_), static_cast<void*>(flags_)})
    if (p) (void)hipFree(p);
  panel_ = sup_inv_ = tmp_ = nullptr;
  work_ = nullptr;
  flags_ = nullptr;
  solve_graphs_.clear();  // they hold the old buffers' addresses
  SOV_HIP_TRY(hipMalloc(&panel_, sizeof(T) * static_cast<std::size_t>(m) * 2 * kPanel * kPanelBuffers));
  const std::size_t nsb = static_cast<std::size_t>((m + kSuper - 1) / kSuper);
  const std::size_t sq = static_cast<std::size_t>(kSuper) * kSuper;
  // Blocks above the diagonal of each super-block stay zero: nothing writes them.
  SOV_HIP_TRY(hipMalloc(&sup_inv_, sizeof(T) * nsb * sq));
  SOV_HIP_TRY(hipMemset(sup_inv_, 0, sizeof(T) * nsb * sq));
  SOV_HIP_TRY(hipMalloc(&tmp_, sizeof(T) * nsb * sq / 4));
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


// ---------------------------------------------------------------------------
// Tile-shape autotuner for the two MFMA products of the factorization: the
// bulk trailing SYRK (lower, square tiles: 64x64 or 128x128) and the panel
// products (TRSM as a GEMM and the strip update, M x 128 x 128: 64x64,
// 128x64, 128x128 or 32x32). On the first factorization in a precision each
// shape runs a product of the class at a representative size (SYRK 4096 x
// 4096 x 256, a panel pair; panel 4096 x 128 x 128, the middle of the factorization's
// 8064 .. 128 rows), one warm-up and five timed launches (HIP events, the
// fastest counts), and the fastest shape is kept for the process.
// SOVEREIGN_GEMM_SHAPE = 0 / 1 / 2 / 3 forces one shape (128x64 falls back
// to 64x64 for the SYRK).
// ---------------------------------------------------------------------------
struct TunedShapes {
  std::atomic<int> syrk{-1}, panel{-1};
  double syrk_ms[kGemmShapes] = {0, 0, 0, 0}, panel_ms[kGemmShapes] = {0, 0, 0, 0};
};
template <typename T>
TunedShapes& tuned() {
  static TunedShapes t;
  return t;
}

int forced_shape() {
  const char* e = std::getenv("SOVEREIGN_GEMM_SHAPE");
  if (!e || !*e) return -1;
  const int v = std::atoi(e);
  return v >= 0 && v < kGemmShapes ? v : -1;
}

template <typename T>
int autotune_shapes(hipStream_t s) {
  TunedShapes& t = tuned<T>();
  if (t.syrk.load() >= 0) return SOV_OK;
  const int ms = 4096, mp = 8192, k = kPanel;
  T *a = nullptr, *c = nullptr;
  const std::size_t na = static_cast<std::size_t>(mp) * k, nc = static_cast<std::size_t>(ms) * ms;
  SOV_HIP_TRY(hipMalloc(&a, sizeof(T) * na));
  if (hipMalloc(&c, sizeof(T) * nc) != hipSuccess) {
    (void)hipFree(a);
    set_error("autotune: out of device memory");
    return SOV_ERR_NOMEM;
  }
  (void)hipMemsetAsync(a, 0, sizeof(T) * na, s);
  (void)hipMemsetAsync(c, 0, sizeof(T) * nc, s);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto time_it = [&](auto&& launch) {  // fastest of five launches after a warm-up
    launch();
    double best = 1e30;
    for (int r = 0; r < 5; ++r) {
      (void)hipEventRecord(e0, s);
      launch();
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms_ = 0.f;
      (void)hipEventElapsedTime(&ms_, e0, e1);
      best = std::min(best, static_cast<double>(ms_));
    }
    return best;
  };
  int best_s = kGemm64x64, best_p = kGemm64x64;
  // the bulk update applies a panel pair: K = 2 kPanel (a holds 4096 x 256)
  for (int shape : {kGemm64x64, kGemm128x128}) {
    t.syrk_ms[shape] = time_it([&] {
      (void)gemm_update_shape<T, false, true>(shape, ms, ms, 2 * k, a, ms, a, ms, c, ms, T(-1), s);
    });
    if (t.syrk_ms[shape] < t.syrk_ms[best_s]) best_s = shape;
  }
  for (int shape : {kGemm64x64, kGemm128x64, kGemm128x128, kGemm32x32}) {
    t.panel_ms[shape] = time_it([&] {
      (void)gemm_update_shape<T, false, false>(shape, mp / 2, k, k, a, mp, a, mp, c, mp, T(-1), s);
    });
    if (t.panel_ms[shape] < t.panel_ms[best_p]) best_p = shape;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(c);
  t.panel.store(best_p);
  t.syrk.store(best_s);
  return SOV_OK;
}

// Launches the diagonal-block kernel of the block at k0 (width nb); its
// inverse goes to diagonal block (k0 % kSuper) / kPanel of super-block
// k0 / kSuper. Returns that inverse.
template <typename T>
T* launch_panel(T* a, long long lda, int k0, int nb, const double* diag0, FactorFlags* flags, T* sup_inv,
                hipStream_t s) {
  T* pinv = sup_inv + static_cast<long long>(k0 / kSuper) * kSuper * kSuper +
            static_cast<long long>(k0 % kSuper) * (kSuper + 1);
  TimedPair* tk = timed_kernel(sizeof(T) == 8 ? kTimePotrf64 : kTimePotrf32);
  if (tk) {  // the launch records its own start and stop events
    tk->launches = 1;
    tk->flops = 2.0 * nb * static_cast<double>(nb) * nb / 3.0;  // factor nb^3/3 + inverse nb^3/3
    hipExtLaunchKernelGGL(potrf_panel_kernel<T>, dim3(1), dim3(kPanelThreads), panel_lds_bytes<T>(), s, tk->a, tk->b, 0,
                          a, lda, k0, nb, diag0, flags, pinv, kSuper);
  } else {
    potrf_panel_kernel<T><<<1, kPanelThreads, panel_lds_bytes<T>(), s>>>(a, lda, k0, nb, diag0, flags, pinv, kSuper);
  }
  return pinv;
}

// Right-looking factorization by panel pairs. Pair q holds the column
// blocks A = [k0, k0 + 128) and B = [k0 + 128, k0 + 256), k0 = 256 q:
//   on s (the critical path): the diagonal block of A, the TRSM of A
//     (L21 of A into the pair's buffer), strip 1 (block B updated by A),
//     the diagonal block of B, the TRSM of B, then strip 2: the next pair's
//     two column blocks [k0 + 256, k0 + 512) updated by both panels at once
//     (K = 256);
//   on side_: L21 of both panels copied back into the factor, then the bulk
//     update of the columns from k0 + 512 by both panels (K = 256), which
//     overlaps the next pair's chain.
// Applying two panels per pass halves the reads and writes of the trailing
// matrix (at K = 128 they cost about as much as the MFMA work). Each column
// block is updated in panel order: strip 2 of pair q waits for the bulk of
// pair q - 1 (the columns from k0 + 256); strip 1 follows strip 2 of the
// previous pair on s. The buffer of pair q is written again by the TRSM of
// pair q + 3, after strip 2 of pair q + 2, which waited for the bulk of
// pair q + 1, queued on side_ after everything of pair q.
template <typename T>
int DenseCholesky<T>::factor(T* a, int m, const double* diag0, hipStream_t s) {
  int rc = reserve(m);
  if (rc != SOV_OK) return rc;
  if (!side_) {
    SOV_HIP_TRY(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    SOV_HIP_TRY(hipEventCreateWithFlags(&ev_panel_, hipEventDisableTiming));
    SOV_HIP_TRY(hipEventCreateWithFlags(&ev_bulk_, hipEventDisableTiming));
  }
  if ((rc = autotune_shapes<T>(s)) != SOV_OK) return rc;
  const int forced = forced_shape();
  const int syrk_shape = forced >= 0 ? forced : tuned<T>().syrk.load();
  const int panel_shape = forced >= 0 ? forced : tuned<T>().panel.load();
  reset_flags_kernel<<<1, 1, 0, s>>>(flags_);
  SOV_LAUNCH_CHECK();
  // The side stream starts after everything already queued on s.
  SOV_HIP_TRY(hipEventRecord(ev_bulk_, s));
  SOV_HIP_TRY(hipStreamWaitEvent(side_, ev_bulk_, 0));
  hipEvent_t bulk_done = ev_bulk_;  // completion of the last bulk update queued on the side stream
  const long long lda = m;
  const long long bstride = 2LL * kPanel * m;  // one pair buffer: m rows x 256 columns, leading dimension m
  for (int k0 = 0, q = 0; k0 < m; k0 += 2 * kPanel, ++q) {
    // buffer row r holds global row k0 + 128 + r; columns 0..127 L21 of A,
    // 128..255 L21 of B (from global row k0 + 256, buffer row 128)
    T* buf = panel_ + (q % kPanelBuffers) * bstride;
    const int nbA = std::min(kPanel, m - k0);
    const int belowA = m - k0 - nbA;
    TimedPair* tp = timed_begin(s, kTimePanel);
    T* pinvA = launch_panel<T>(a, lda, k0, nbA, diag0, flags_, sup_inv_, s);
    SOV_LAUNCH_CHECK();
    if (belowA > 0) {
      // L21 = A21 inv(L11)^T on MFMA, out of place (the updates read it in
      // the buffer; the copy back into the factor runs on the side stream)
      const T* a21 = a + (k0 + nbA) + static_cast<long long>(k0) * lda;
      SOV_HIP_TRY((gemm_update_shape<T, false, false, true>(panel_shape, belowA, nbA, nbA, a21, lda, pinvA, kSuper,
                                                            buf, lda, T(1), s)));
    }
    timed_end(tp, s, belowA > 0 ? 2 : 1, 0.0);
    if (belowA <= 0) break;
    // strip 1: block B by panel A (it also writes the strictly upper part of
    // B's diagonal block, which nothing reads)
    const int nbB = std::min(kPanel, belowA);
    const int belowB = belowA - nbB;
    {
      TimedPair* ts = timed_begin(s, kTimeStrip);
      T* a_b = a + (k0 + nbA) + static_cast<long long>(k0 + nbA) * lda;
      SOV_HIP_TRY((gemm_update_shape<T, false, false>(panel_shape, belowA, nbB, nbA, buf, lda, buf, lda, a_b, lda,
                                                      T(-1), s)));
      timed_end(ts, s, 1, 0.0);
    }
    const int kB = k0 + nbA;
    tp = timed_begin(s, kTimePanel);
    T* pinvB = launch_panel<T>(a, lda, kB, nbB, diag0, flags_, sup_inv_, s);
    SOV_LAUNCH_CHECK();
    T* bufB = buf + kPanel + kPanel * lda;  // column 128, row 128 (global row kB + nbB = k0 + 256)
    if (belowB > 0) {
      const T* a21 = a + (kB + nbB) + static_cast<long long>(kB) * lda;
      SOV_HIP_TRY((gemm_update_shape<T, false, false, true>(panel_shape, belowB, nbB, nbB, a21, lda, pinvB, kSuper,
                                                            bufB, lda, T(1), s)));
    }
    timed_end(tp, s, belowB > 0 ? 2 : 1, 0.0);
    // strip 2 on s: the next pair's columns by both panels (K = 256), after
    // the previous pair's bulk update; its launch records ev_panel_, which
    // starts the side stream's work of this pair (released earlier, the bulk
    // update's workgroups take every CU and the chain waits behind them)
    const int w2 = std::min(2 * kPanel, belowB);
    const int rest = belowB - w2;
    const T* lp = buf + kPanel;  // rows from global k0 + 256, K = nbA + nbB = 256 columns
    if (belowB > 0) {
      SOV_HIP_TRY(hipStreamWaitEvent(s, bulk_done, 0));
      TimedPair* ts = timed_begin(s, kTimeStrip);
      T* a_c = a + (kB + nbB) + static_cast<long long>(kB + nbB) * lda;
      SOV_HIP_TRY((gemm_update_shape<T, false, false>(panel_shape, belowB, w2, nbA + nbB, lp, lda, lp, lda, a_c, lda,
                                                      T(-1), s, ts ? nullptr : ev_panel_)));
      timed_end(ts, s, 1, 0.0);
      if (ts) SOV_HIP_TRY(hipEventRecord(ev_panel_, s));
    } else {
      SOV_HIP_TRY(hipEventRecord(ev_panel_, s));
    }
    // side stream: both L21 back into the factor (only the triangular solves
    // read them there), then the bulk update, which records its completion
    // itself (bulk_done: ev_bulk_, or the stop event of its timed pair)
    SOV_HIP_TRY(hipStreamWaitEvent(side_, ev_panel_, 0));
    copy_panel_kernel<T><<<blocks_for(static_cast<long long>(belowA) * nbA, 256), 256, 0, side_>>>(buf, a, lda, k0 + nbA,
                                                                                                    k0, belowA, nbA);
    SOV_LAUNCH_CHECK();
    if (belowB > 0) {
      copy_panel_kernel<T><<<blocks_for(static_cast<long long>(belowB) * nbB, 256), 256, 0, side_>>>(
          bufB, a, lda, kB + nbB, kB, belowB, nbB);
      SOV_LAUNCH_CHECK();
    }
    if (rest > 0) {
      T* a22 = a + (kB + nbB + w2) + static_cast<long long>(kB + nbB + w2) * lda;
      const T* l3 = lp + w2;
      TimedPair* tb = timed_kernel(sizeof(T) == 8 ? kTimeSyrk : kTimeSyrk32);
      bulk_done = tb ? tb->b : ev_bulk_;
      SOV_HIP_TRY((gemm_update_shape<T, false, true>(syrk_shape, rest, rest, nbA + nbB, l3, lda, l3, lda, a22, lda,
                                                     T(-1), side_, bulk_done, tb ? tb->a : nullptr)));
      if (tb) {
        tb->launches = 1;
        tb->flops = static_cast<double>(rest) * (rest + 1) * (nbA + nbB);
      }
    } else {
      SOV_HIP_TRY(hipEventRecord(ev_bulk_, side_));
      bulk_done = ev_bulk_;
    }
  }
  SOV_HIP_TRY(hipEventRecord(ev_bulk_, side_));
  SOV_HIP_TRY(hipStreamWaitEvent(s, ev_bulk_, 0));
  return build_super_inverses(a, m, s);
}

// Inverse of each kSuper x kSuper diagonal block of L from its panel
// inverses, by recursive doubling: for adjacent diagonal blocks A (size b,
// inverse X_A) and B below it (inverse X_B), the block of the inverse under
// X_A is X_BA = -X_B (L_BA X_A). One batched GEMM pair per level covers all
// full super-blocks; a partial last super-block gets its own launches.
template <typename T>
int DenseCholesky<T>::build_super_inverses(const T* a, int m, hipStream_t s) {
  const long long lda = m, ks = kSuper, sq = ks * ks;
  const int nfull = m / kSuper, rem = m - nfull * kSuper;
  TimedPair* tp = timed_begin(s, kTimeInverse);
  int launches = 0;
  for (long long b = kPanel; b < kSuper; b *= 2) {
    const int npairs = static_cast<int>(kSuper / (2 * b));
    if (nfull > 0) {
      GemmBatch g1;  // tmp = L_BA X_A
      g1.a_y = 2 * b * (1 + lda);
      g1.a_z = ks * (1 + lda);
      g1.b_y = 2 * b * (1 + ks);
      g1.b_z = sq;
      g1.c_y = b * b;
      g1.c_z = npairs * b * b;
      SOV_HIP_TRY((gemm_update<T, true, false, true>(static_cast<int>(b), static_cast<int>(b), static_cast<int>(b),
                                                     a + b, lda, sup_inv_, ks, tmp_, b, T(1), s, npairs, nfull, g1)));
      GemmBatch g2;  // X_BA = -X_B tmp
      g2.a_y = 2 * b * (1 + ks);
      g2.a_z = sq;
      g2.b_y = b * b;
      g2.b_z = npairs * b * b;
      g2.c_y = 2 * b * (1 + ks);
      g2.c_z = sq;
      SOV_HIP_TRY((gemm_update<T, true, false, true>(static_cast<int>(b), static_cast<int>(b), static_cast<int>(b),
                                                     sup_inv_ + b * (1 + ks), ks, tmp_, b, sup_inv_ + b, ks, T(-1), s,
                                                     npairs, nfull, g2)));
      launches += 2;
    }
    if (rem > 0) {
      const long long g0 = static_cast<long long>(nfull) * kSuper;
      T* X = sup_inv_ + nfull * sq;
      for (long long i0 = 0; i0 + b < rem; i0 += 2 * b) {
        const int bb = static_cast<int>(std::min(b, rem - i0 - b));
        const T* lba = a + (g0 + i0 + b) + (g0 + i0) * lda;
        SOV_HIP_TRY((gemm_update<T, true, false, true>(bb, static_cast<int>(b), static_cast<int>(b), lba, lda,
                                                       X + i0 * (1 + ks), ks, tmp_, bb, T(1), s)));
        SOV_HIP_TRY((gemm_update<T, true, false, true>(bb, static_cast<int>(b), bb, X + (i0 + b) * (1 + ks), ks, tmp_,
                                                       bb, X + (i0 + b) + i0 * ks, ks, T(-1), s)));
        launches += 2;
      }
    }
  }
  timed_end(tp, s, launches, 0.0);
  return SOV_OK;
}

template <typename T>
int DenseCholesky<T>::solve(const T* a, int m, double* x, double* /*work*/, hipStream_t s) {
  if (m > m_) return SOV_ERR_ARG;
  const long long lda = m, sq = static_cast<long long>(kSuper) * kSuper;
  const int nsb = (m + kSuper - 1) / kSuper;
  TimedPair* tp = timed_begin(s, kTimeSolve);
  const int rc = solve_graphs_.run(a, x, nullptr, m, s, [&]() -> int {
    for (int P = 0; P < nsb; ++P) {  // L y = x: y into work_, x updated in place
      const int base = P * kSuper, sz = std::min(kSuper, m - base), below = m - base - sz;
      gemv_n_kernel<T, false, true><<<blocks_for(sz, kGemvRows), kGemvRows * kGemvPhases, 0, s>>>(
          sz, sz, sup_inv_ + P * sq, kSuper, x + base, work_ + base);
      if (below > 0)
        gemv_n_kernel<T, true, false><<<blocks_for(below, kGemvRows), kGemvRows * kGemvPhases, 0, s>>>(
            below, sz, a + (base + sz) + base * lda, lda, work_ + base, x + base + sz);
    }
    SOV_LAUNCH_CHECK();
    for (int P = nsb - 1; P >= 0; --P) {  // L^T x = y: x back into x, work_ updated in place
      const int base = P * kSuper, sz = std::min(kSuper, m - base);
      gemv_t_kernel<T, false, true><<<blocks_for(sz, 4), 256, 0, s>>>(sz, sz, sup_inv_ + P * sq, kSuper,
                                                                        work_ + base, x + base);
      if (base > 0)
        gemv_t_kernel<T, true, false><<<blocks_for(base, 4), 256, 0, s>>>(sz, base, a + base, lda, x + base, work_);
    }
    SOV_LAUNCH_CHECK();
    return SOV_OK;
  });
  // algorithmic bytes: the factor's lower triangle read by the forward and the backward pass
  timed_end(tp, s, 4 * nsb, static_cast<double>(m) * (m + 1) * sizeof(T));
  return rc;
}

template class DenseCholesky<double>;
template class DenseCholesky<float>;

}  // namespace sov

extern "C" {

int sov_gemm_shapes(int32_t* shapes4, double* times_ms12) {
  if (!shapes4) return SOV_ERR_ARG;
  shapes4[0] = sov::tuned<double>().syrk.load();
  shapes4[1] = sov::tuned<double>().panel.load();
  shapes4[2] = sov::tuned<float>().syrk.load();
  shapes4[3] = sov::tuned<float>().panel.load();
  if (times_ms12) {  // f64 SYRK 64x64, 128x128; f64 panel 64x64, 128x64, 128x128, 32x32; the same for f32
    const sov::TunedShapes* t[2] = {&sov::tuned<double>(), &sov::tuned<float>()};
    for (int p = 0; p < 2; ++p) {
      times_ms12[6 * p + 0] = t[p]->syrk_ms[sov::kGemm64x64];
      times_ms12[6 * p + 1] = t[p]->syrk_ms[sov::kGemm128x128];
      for (int k = 0; k < 4; ++k) times_ms12[6 * p + 2 + k] = t[p]->panel_ms[k];
    }
  }
  return SOV_OK;
}

void sov_set_kernel_timing(int32_t on) { sov::set_kernel_timing(on == 2 ? 2 : (on != 0 ? 1 : 0)); }

int sov_kernel_times(double* out) {
  if (!out) return SOV_ERR_ARG;
  if (sov::device_ready()) sov::collect_kernel_times(sov::stream());
  sov::KernelTimes& t = sov::kernel_times();
  const double v[24] = {t.syrk_ms,        t.syrk_launches,    t.syrk_flops,      t.panel_ms,       t.solve_ms,
                        t.assemble_ms,    t.panel_launches,   t.strip_ms,        t.inverse_ms,     t.syrk32_ms,
                        t.syrk32_launches, t.syrk32_flops,    t.potrf64_ms,      t.potrf64_launches, t.potrf64_flops,
                        t.potrf32_ms,     t.potrf32_launches, t.potrf32_flops,   t.spmv_ms,        t.spmv_launches,
                        t.spmv_bytes,     t.solve_pairs,      t.solve_bytes,
                        sov::g_timing.load() == 2 ? static_cast<double>(sov::kLevel2Stride) : 1.0};
  for (int i = 0; i < 24; ++i) out[i] = v[i];
  t = sov::KernelTimes{};
  return SOV_OK;
}

}  // extern "C"
