// Separable convex QP with second-difference constraints (the reference's
// LISWET structure, detect_liswet_structured, solver/qp/src/
// qp_interior_point.cpp:892-967):
//
//     min 0.5 ||x||^2 - target' x   s.t.  x_r - 2 x_{r+1} + x_{r+2} >= 0,  r < n-2.
//
// The interior-point iterations run on the GPU in one persistent workgroup
// (the whole Mehrotra loop in one launch: the problem is a few 20k-vectors,
// so one CU's bandwidth and a barrier per phase beat a launch per phase).
// The method is the reference's x-space structured IPM
// (solve_liswet_structured_ipm, :2378-2826): start x = target, s = max(1,
// A x), z = 1/s; normal matrix M = I + A' diag(z/s) A (pentadiagonal);
// sigma = clamp((mu_aff/mu)^3, 0.01, 0.99); common step with tau = 0.995.
// M is Jacobi-scaled (S = sqrt(max(1, diag M)), as the reference's banded
// Cholesky) and solved by block cyclic reduction on its 2 x 2-block
// tridiagonal form: log2(n/2) levels, each a parallel sweep of 2 x 2 block
// operations with one barrier. Mixed precision: the reduction is factored in
// FP32 and its solves are refined against the FP64 bands (at most eight
// corrections); when the refinement stalls the factor is redone in FP64 for
// this and all later iterations. The iterations stop at the structured
// tolerance (1e-10, 1e-12 for m <= 50) or when the method cannot progress;
// the host then polishes in knot space from the last iterate
// (polish_liswet_knots :2269-2377, solve_liswet_knot_projection :1509-1580)
// and certifies (certify_liswet :1410-1480), in long double.

#include "internal.hpp"

#include "qp_polish.hpp"
#include "sovereign_qp.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

namespace {

constexpr int kT = 1024;  // threads of the workgroup
constexpr int kMaxLevels = 40;

struct QpOpts {
  double tol;
  int max_iter;
  int precision;  // SOV_PREC_FP32: mixed (FP32 factor while refinement converges), else FP64
};

// Device-side result record.
struct QpDev {
  int iterations;
  int status;  // 0 converged, 1 stopped (no progress / iteration cap), 6 numerical failure
  int fp32_factors, fp64_factors;
  int refinements;
  int pad;
  double pres, dres, cres, mu;
  long long t_factor, t_solve;  // wall-clock ticks spent in factorizations and in solves
};

// Reduction blocks of one factor (global memory), arithmetic type T.
template <typename T>
struct Fac {
  T* A;   // 3 per block, all levels
  T* B;   // 4 per block, all levels
  T* Lf;  // 3 per block: Cholesky factor of the odd blocks and of the top block
};

// Workspace pointers (global memory).
struct Ws {
  double *x, *s, *z, *ax, *atz, *rp, *rd, *d, *diag, *s1, *s2, *sc;
  double *rhs, *dx, *ds, *dz, *dxa, *dsa, *dza, *rc, *res, *corr, *wt;
  double *brhs;  // reduction right-hand sides, 2 per block over all levels
  double *bx;    // block solution, 2 per level-0 block
};

struct Levels {
  int count;
  int size[kMaxLevels];
  long long off[kMaxLevels];  // block offset of level l in the level-stacked arrays
};

__device__ __forceinline__ double bsum(double v, double* sh) { return block_reduce<0>(v, sh); }
__device__ __forceinline__ double bmax(double v, double* sh) { return block_reduce<1>(v, sh); }

// (A x)_r = x_r - 2 x_{r+1} + x_{r+2}
__device__ void apply_a(int n, const double* x, double* out) {
  for (int r = threadIdx.x; r < n - 2; r += kT) out[r] = x[r] - 2.0 * x[r + 1] + x[r + 2];
}
// (A' z)_j = z_j - 2 z_{j-1} + z_{j-2} (terms within 0..m-1)
__device__ void apply_at(int n, const double* z, double* out) {
  const int m = n - 2;
  for (int j = threadIdx.x; j < n; j += kT) {
    double v = 0.0;
    if (j < m) v += z[j];
    if (j >= 1 && j - 1 < m) v -= 2.0 * z[j - 1];
    if (j >= 2 && j - 2 < m) v += z[j - 2];
    out[j] = v;
  }
}

// ---- block cyclic reduction of the scaled normal matrix -------------------
//
// Mt = S^-1 (M + reg I) S^-1 is block tridiagonal with 2 x 2 blocks: block k
// holds rows 2k, 2k+1; A_k is its diagonal block (symmetric, stored a00, a10,
// a11) and B_k couples block k+1 to block k (rows 2k+2, 2k+3 x columns 2k,
// 2k+1; stored b00, b01, b10, b11). Level l+1 keeps the even blocks of level
// l; eliminating the odd ones with the Cholesky factor L_j of A_j gives
//   A'_k = A_k - W'W - V'V,   W = L_{k-1}^-1 B_{k-1}',  V = L_{k+1}^-1 B_k,
//   B'_k = -U'V,              U = L_{k+1}^-1 B_{k+1}'
// (block Cholesky of the odd-even permuted matrix, so as stable as the
// banded Cholesky for any condition number the arithmetic can hold). The
// factor keeps every level's B and the odd blocks' L (and the last level's).
// Arithmetic type T: float for the mixed path, double otherwise.

// 2 x 2 Cholesky: (a00, a10, a11) -> (l00, l10, l11); false on a pivot at or
// below `tiny` times the diagonal entry.
template <typename T>
__device__ __forceinline__ bool chol2(T a00, T a10, T a11, T tiny, T* l) {
  const bool ok0 = a00 > tiny * fmax(T(1), a00);
  const T l00 = sqrt(ok0 ? a00 : T(1));
  const T l10 = a10 / l00;
  const T p = a11 - l10 * l10;
  const bool ok1 = p > tiny * fmax(T(1), a11) && isfinite(p);
  l[0] = l00;
  l[1] = l10;
  l[2] = sqrt(ok1 ? p : T(1));
  return ok0 && ok1;
}
// y = L^-1 v, y = L^-T v, y = L^-T L^-1 v
template <typename T>
__device__ __forceinline__ void lsolve(const T* l, T v0, T v1, T& y0, T& y1) {
  y0 = v0 / l[0];
  y1 = (v1 - l[1] * y0) / l[2];
}
template <typename T>
__device__ __forceinline__ void ltsolve(const T* l, T v0, T v1, T& y0, T& y1) {
  y1 = v1 / l[2];
  y0 = (v0 - l[1] * y1) / l[0];
}
template <typename T>
__device__ __forceinline__ void blocksolve(const T* l, T v0, T v1, T& y0, T& y1) {
  T t0, t1;
  lsolve(l, v0, v1, t0, t1);
  ltsolve(l, t0, t1, y0, y1);
}

template <typename T>
__device__ constexpr T pivot_tiny() {
  // FP64: the reference's pivot test (1e-18 of the row, :1098-1105). FP32:
  // a pivot that lost all but ~4 digits to cancellation hands the system to
  // the FP64 factor rather than to a refinement that would stall.
  return sizeof(T) == 4 ? T(2e-6) : T(1e-18);
}

template <typename T>
__device__ void level0(int n, const Ws& w, double reg, const Fac<T>& f) {
  const int N = (n + 1) / 2;
  for (int k = threadIdx.x; k < N; k += kT) {
    const int i0 = 2 * k, i1 = 2 * k + 1;
    auto mt_d = [&](int i) { return i < n ? (w.diag[i] + reg) / (w.sc[i] * w.sc[i]) : 1.0; };
    auto mt_1 = [&](int i) { return (i < n && i >= 1) ? w.s1[i] / (w.sc[i] * w.sc[i - 1]) : 0.0; };
    auto mt_2 = [&](int i) { return (i < n && i >= 2) ? w.s2[i] / (w.sc[i] * w.sc[i - 2]) : 0.0; };
    f.A[3LL * k + 0] = static_cast<T>(mt_d(i0));
    f.A[3LL * k + 1] = static_cast<T>(mt_1(i1));
    f.A[3LL * k + 2] = static_cast<T>(mt_d(i1));
    f.B[4LL * k + 0] = static_cast<T>(mt_2(i0 + 2));
    f.B[4LL * k + 1] = static_cast<T>(mt_1(i0 + 2));
    f.B[4LL * k + 2] = T(0);
    f.B[4LL * k + 3] = static_cast<T>(mt_2(i1 + 2));
  }
}

// Level l -> l+1. Even block k = 2k' reduces its odd neighbours; the left
// neighbour's factor was written by block k - 2's pass, the right one's
// (k + 1) is written here, so every odd block's L is written exactly once.
template <typename T>
__device__ void reduce_level(const Levels& L, int l, const Fac<T>& f, int* ok) {
  const int Nl = L.size[l], Nn = L.size[l + 1];
  const T* Al = f.A + 3 * L.off[l];
  const T* Bl = f.B + 4 * L.off[l];
  T* Ll = f.Lf + 3 * L.off[l];
  T* An = f.A + 3 * L.off[l + 1];
  T* Bn = f.B + 4 * L.off[l + 1];
  constexpr T tiny = pivot_tiny<T>();
  for (int kp = threadIdx.x; kp < Nn; kp += kT) {
    const int k = 2 * kp;
    T a0 = Al[3LL * k], a1 = Al[3LL * k + 1], a2 = Al[3LL * k + 2];
    T bn[4] = {T(0), T(0), T(0), T(0)};
    if (k >= 1) {  // W = L_{k-1}^-1 B_{k-1}'; A -= W'W
      T lf[3];
      const T* aj = Al + 3LL * (k - 1);
      if (!chol2(aj[0], aj[1], aj[2], tiny, lf)) *ok = 0;
      const T* b = Bl + 4LL * (k - 1);
      T w00, w10, w01, w11;
      lsolve(lf, b[0], b[1], w00, w10);  // column 0 of B' = row 0 of B
      lsolve(lf, b[2], b[3], w01, w11);
      a0 -= w00 * w00 + w10 * w10;
      a1 -= w01 * w00 + w11 * w10;
      a2 -= w01 * w01 + w11 * w11;
    }
    if (k + 1 < Nl) {  // V = L_{k+1}^-1 B_k; A -= V'V; B' = -U'V, U = L_{k+1}^-1 B_{k+1}'
      T lf[3];
      const T* aj = Al + 3LL * (k + 1);
      if (!chol2(aj[0], aj[1], aj[2], tiny, lf)) *ok = 0;
      Ll[3LL * (k + 1)] = lf[0];
      Ll[3LL * (k + 1) + 1] = lf[1];
      Ll[3LL * (k + 1) + 2] = lf[2];
      const T* b = Bl + 4LL * k;
      T v00, v10, v01, v11;
      lsolve(lf, b[0], b[2], v00, v10);  // column 0 of B
      lsolve(lf, b[1], b[3], v01, v11);
      a0 -= v00 * v00 + v10 * v10;
      a1 -= v01 * v00 + v11 * v10;
      a2 -= v01 * v01 + v11 * v11;
      if (k + 2 < Nl) {
        const T* cb = Bl + 4LL * (k + 1);
        T u00, u10, u01, u11;
        lsolve(lf, cb[0], cb[1], u00, u10);
        lsolve(lf, cb[2], cb[3], u01, u11);
        bn[0] = -(u00 * v00 + u10 * v10);
        bn[1] = -(u00 * v01 + u10 * v11);
        bn[2] = -(u01 * v00 + u11 * v10);
        bn[3] = -(u01 * v01 + u11 * v11);
      }
    }
    An[3LL * kp + 0] = a0;
    An[3LL * kp + 1] = a1;
    An[3LL * kp + 2] = a2;
    for (int q = 0; q < 4; ++q) Bn[4LL * kp + q] = bn[q];
  }
}

// Factorization: level 0 from the bands, every reduction level, then the
// top block. Returns false on a failed pivot (block-uniform).
template <typename T>
__device__ bool factor(int n, const Ws& w, const Levels& L, double reg, const Fac<T>& f, int* ok_sh) {
  if (threadIdx.x == 0) *ok_sh = 1;
  level0<T>(n, w, reg, f);
  __syncthreads();
  for (int l = 0; l + 1 < L.count; ++l) {
    reduce_level<T>(L, l, f, ok_sh);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const long long top = L.off[L.count - 1];
    const T* a = f.A + 3 * top;
    if (!chol2(a[0], a[1], a[2], pivot_tiny<T>(), f.Lf + 3 * top)) *ok_sh = 0;
  }
  __syncthreads();
  const bool ok = *ok_sh != 0;
  __syncthreads();
  return ok;
}

// out = (M + reg I)^-1 v with the factor (v, out: n FP64 vectors, distinct).
// Right-hand sides and block solutions are kept in FP64, block arithmetic is
// in T.
template <typename T>
__device__ void solve(int n, const Ws& w, const Levels& L, const Fac<T>& f, const double* v, double* out) {
  const int N0 = L.size[0];
  double* R = w.brhs;
  for (int k = threadIdx.x; k < N0; k += kT) {  // level-0 right-hand side, scaled by S^-1
    const int i0 = 2 * k, i1 = 2 * k + 1;
    R[2LL * k] = v[i0] / w.sc[i0];
    R[2LL * k + 1] = i1 < n ? v[i1] / w.sc[i1] : 0.0;
  }
  __syncthreads();
  // forward: r'_k = r_k - B_{k-1} A_{k-1}^-1 r_{k-1} - B_k' A_{k+1}^-1 r_{k+1}
  for (int l = 0; l + 1 < L.count; ++l) {
    const int Nl = L.size[l], Nn = L.size[l + 1];
    const T* Bl = f.B + 4 * L.off[l];
    const T* Ll = f.Lf + 3 * L.off[l];
    const double* Rl = R + 2 * L.off[l];
    double* Rn = R + 2 * L.off[l + 1];
    for (int kp = threadIdx.x; kp < Nn; kp += kT) {
      const int k = 2 * kp;
      T r0 = static_cast<T>(Rl[2LL * k]), r1 = static_cast<T>(Rl[2LL * k + 1]);
      if (k >= 1) {
        T t0, t1;
        blocksolve(Ll + 3LL * (k - 1), static_cast<T>(Rl[2LL * (k - 1)]), static_cast<T>(Rl[2LL * (k - 1) + 1]), t0,
                   t1);
        const T* b = Bl + 4LL * (k - 1);
        r0 -= b[0] * t0 + b[1] * t1;
        r1 -= b[2] * t0 + b[3] * t1;
      }
      if (k + 1 < Nl) {
        T t0, t1;
        blocksolve(Ll + 3LL * (k + 1), static_cast<T>(Rl[2LL * (k + 1)]), static_cast<T>(Rl[2LL * (k + 1) + 1]), t0,
                   t1);
        const T* b = Bl + 4LL * k;
        r0 -= b[0] * t0 + b[2] * t1;
        r1 -= b[1] * t0 + b[3] * t1;
      }
      Rn[2LL * kp] = r0;
      Rn[2LL * kp + 1] = r1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {  // the top block sits at level-0 index 0
    const long long top = L.off[L.count - 1];
    T y0, y1;
    blocksolve(f.Lf + 3 * top, static_cast<T>(R[2 * top]), static_cast<T>(R[2 * top + 1]), y0, y1);
    w.bx[0] = y0;
    w.bx[1] = y1;
  }
  __syncthreads();
  // backward: odd blocks of level l, x_k = A_k^-1 (r_k - B_{k-1} x_{k-1} - B_k' x_{k+1})
  for (int l = L.count - 2; l >= 0; --l) {
    const int Nl = L.size[l], stride = 1 << l;
    const T* Bl = f.B + 4 * L.off[l];
    const T* Ll = f.Lf + 3 * L.off[l];
    const double* Rl = R + 2 * L.off[l];
    for (int k = 2 * threadIdx.x + 1; k < Nl; k += 2 * kT) {
      T r0 = static_cast<T>(Rl[2LL * k]), r1 = static_cast<T>(Rl[2LL * k + 1]);
      const double* xl = w.bx + 2LL * (k - 1) * stride;
      const T x0 = static_cast<T>(xl[0]), x1 = static_cast<T>(xl[1]);
      const T* b = Bl + 4LL * (k - 1);
      r0 -= b[0] * x0 + b[1] * x1;
      r1 -= b[2] * x0 + b[3] * x1;
      if (k + 1 < Nl) {
        const double* xr = w.bx + 2LL * (k + 1) * stride;
        const T y0 = static_cast<T>(xr[0]), y1 = static_cast<T>(xr[1]);
        const T* c = Bl + 4LL * k;
        r0 -= c[0] * y0 + c[2] * y1;
        r1 -= c[1] * y0 + c[3] * y1;
      }
      T z0, z1;
      blocksolve(Ll + 3LL * k, r0, r1, z0, z1);
      w.bx[2LL * k * stride] = z0;
      w.bx[2LL * k * stride + 1] = z1;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += kT) out[i] = w.bx[i] / w.sc[i];
  __syncthreads();
}

// out = (M + reg I) v from the FP64 bands.
__device__ void matvec(int n, const Ws& w, double reg, const double* v, double* out) {
  for (int i = threadIdx.x; i < n; i += kT) {
    double y = (w.diag[i] + reg) * v[i];
    if (i >= 1) y += w.s1[i] * v[i - 1];
    if (i + 1 < n) y += w.s1[i + 1] * v[i + 1];
    if (i >= 2) y += w.s2[i] * v[i - 2];
    if (i + 2 < n) y += w.s2[i + 2] * v[i + 2];
    out[i] = y;
  }
}

// Per-launch state shared by the workgroup (LDS): workspace pointers, the
// level table and the result record live here so that no thread keeps a
// private copy of them.
struct Ctx {
  Ws w;
  Fac<float> f32;
  Fac<double> f64;
  Levels L;
  QpDev rec;
  double sh[kT / 64];
  double reg;
  int prec;  // current factor precision
  int ok;
};

// Factor solve with its time booked to the result record (thread 0).
__device__ void timed_solve(int n, Ctx& c, const double* v, double* out) {
  const long long t0 = wall_clock64();
  if (c.prec == SOV_PREC_FP32) solve<float>(n, c.w, c.L, c.f32, v, out);
  else solve<double>(n, c.w, c.L, c.f64, v, out);
  if (threadIdx.x == 0) c.rec.t_solve += wall_clock64() - t0;
}

template <typename T>
__device__ bool timed_factor(int n, Ctx& c, double reg, const Fac<T>& f) {
  const long long t0 = wall_clock64();
  const bool ok = factor<T>(n, c.w, c.L, reg, f, &c.ok);
  if (threadIdx.x == 0) c.rec.t_factor += wall_clock64() - t0;
  return ok;
}

// Solves (M + reg I) dx = rhs with the current factor, refining against the
// FP64 bands. FP64 factor: at most two corrections, as the reference. FP32
// factor: at most eight; false when the residual is not below 1e-10 x |rhs|
// by then (the caller refactors in FP64).
__device__ bool refined_solve(int n, Ctx& c, double* dx) {
  const Ws& w = c.w;
  const int tid = threadIdx.x;
  const bool fp32 = c.prec == SOV_PREC_FP32;
  timed_solve(n, c, w.rhs, dx);
  double rn0 = 0.0;
  for (int j = tid; j < n; j += kT) rn0 = fmax(rn0, fabs(w.rhs[j]));
  rn0 = fmax(bmax(rn0, c.sh), 1e-300);
  const int max_ref = fp32 ? 8 : 2;
  for (int k = 0; k <= max_ref; ++k) {
    matvec(n, w, c.reg, dx, w.res);
    __syncthreads();
    double rn = 0.0;
    for (int j = tid; j < n; j += kT) {
      const double r = w.rhs[j] - w.res[j];
      w.res[j] = r;
      rn = fmax(rn, fabs(r));
    }
    rn = bmax(rn, c.sh);
    if (!isfinite(rn)) return false;
    if (rn <= 1e-14 * rn0) return true;
    if (k == max_ref) return !fp32 || rn <= 1e-10 * rn0;
    timed_solve(n, c, w.res, w.corr);
    for (int j = tid; j < n; j += kT) dx[j] += w.corr[j];
    if (tid == 0) ++c.rec.refinements;
    __syncthreads();
  }
  return true;
}

// Newton direction for the complementarity right-hand side w.rc:
//   (I + A' D A) dx = -rd - A' ((rc + z rp) / s),  ds = A dx + rp,
//   dz = (-rc - z ds) / s.
__device__ bool direction(int n, Ctx& c, double* dx, double* ds, double* dz) {
  const Ws& w = c.w;
  const int m = n - 2, tid = threadIdx.x;
  for (int r = tid; r < m; r += kT) w.wt[r] = (w.rc[r] + w.z[r] * w.rp[r]) / w.s[r];
  __syncthreads();
  apply_at(n, w.wt, w.corr);
  __syncthreads();
  for (int j = tid; j < n; j += kT) w.rhs[j] = -w.rd[j] - w.corr[j];
  __syncthreads();
  if (!refined_solve(n, c, dx)) return false;
  __syncthreads();
  apply_a(n, dx, w.res);
  __syncthreads();
  for (int r = tid; r < m; r += kT) {
    const double v = w.res[r] + w.rp[r];
    ds[r] = v;
    dz[r] = (-w.rc[r] - w.z[r] * v) / w.s[r];
  }
  __syncthreads();
  return true;
}

// As direction(); when the FP32 refinement stalls, the factor is redone in
// FP64, which then serves this and every later iteration.
__device__ bool direction_safe(int n, Ctx& c, double* dx, double* ds, double* dz) {
  if (direction(n, c, dx, ds, dz)) return true;
  if (c.prec != SOV_PREC_FP32) return false;
  __syncthreads();
  if (threadIdx.x == 0) c.prec = SOV_PREC_FP64;
  __syncthreads();
  if (!timed_factor<double>(n, c, c.reg, c.f64)) return false;
  if (threadIdx.x == 0) ++c.rec.fp64_factors;
  return direction(n, c, dx, ds, dz);
}

// Largest step in [0, 1] keeping v + a d >= 0 (block-wide).
__device__ double max_step(int m, Ctx& c, const double* v, const double* d) {
  double a = 1.0;
  for (int r = threadIdx.x; r < m; r += kT)
    if (d[r] < 0.0) a = fmin(a, -v[r] / d[r]);
  a = block_reduce<2>(a, c.sh);
  return isfinite(a) ? fmax(0.0, a) : 0.0;
}

__global__ __launch_bounds__(kT) void liswet_ipm_kernel(int n, const double* __restrict__ target, Ws ws,
                                                        Fac<float> f32, Fac<double> f64, Levels levels, QpOpts o,
                                                        QpDev* out) {
  __shared__ Ctx c;
  const int m = n - 2, tid = threadIdx.x;
  if (tid == 0) {
    c.w = ws;
    c.f32 = f32;
    c.f64 = f64;
    c.L = levels;
    c.rec = QpDev{};
    c.rec.status = 1;
    c.prec = o.precision;
    c.reg = 0.0;
  }
  __syncthreads();
  const Ws& w = c.w;
  for (int j = tid; j < n; j += kT) w.x[j] = target[j];
  __syncthreads();
  apply_a(n, w.x, w.s);
  __syncthreads();
  for (int r = tid; r < m; r += kT) {
    w.s[r] = fmax(1.0, w.s[r]);
    w.z[r] = 1.0 / w.s[r];
  }
  double tmax = 0.0;
  for (int j = tid; j < n; j += kT) tmax = fmax(tmax, fabs(target[j]));
  tmax = bmax(tmax, c.sh);
  for (int it = 0; it < o.max_iter; ++it) {
    apply_a(n, w.x, w.ax);
    apply_at(n, w.z, w.atz);
    __syncthreads();
    double mu = 0.0, maxc = 0.0, rpm = 0.0, axm = 0.0, rdm = 0.0, atzm = 0.0;
    for (int r = tid; r < m; r += kT) {
      const double rp = w.ax[r] - w.s[r];
      w.rp[r] = rp;
      const double sz = w.s[r] * w.z[r];
      mu += sz;
      maxc = fmax(maxc, fabs(sz));
      rpm = fmax(rpm, fabs(rp));
      axm = fmax(axm, fabs(w.ax[r]));
    }
    for (int j = tid; j < n; j += kT) {
      const double rd = w.x[j] - target[j] - w.atz[j];
      w.rd[j] = rd;
      rdm = fmax(rdm, fabs(rd));
      atzm = fmax(atzm, fabs(w.atz[j]));
    }
    mu = bsum(mu, c.sh) / m;
    maxc = bmax(maxc, c.sh);
    rpm = bmax(rpm, c.sh);
    axm = bmax(axm, c.sh);
    rdm = bmax(rdm, c.sh);
    atzm = bmax(atzm, c.sh);
    const double pres = rpm / fmax(1.0, axm);
    const double dres = rdm / fmax(1.0, tmax + atzm);
    const double cres = maxc / fmax(1.0, tmax);
    const bool finite = isfinite(mu) && isfinite(pres) && isfinite(dres) && isfinite(cres);
    const bool done = pres <= o.tol && dres <= o.tol && cres <= o.tol;
    if (tid == 0) {
      c.rec.pres = pres;
      c.rec.dres = dres;
      c.rec.cres = cres;
      c.rec.mu = mu;
      c.rec.iterations = it;
      if (!finite) c.rec.status = 6;
      else if (done) c.rec.status = 0;
    }
    if (!finite || done) break;
    // normal matrix bands: row r adds d_r a a' on rows/columns r..r+2, a = (1, -2, 1)
    for (int r = tid; r < m; r += kT) w.d[r] = w.z[r] / w.s[r];
    __syncthreads();
    double dmax = 0.0;
    for (int i = tid; i < n; i += kT) {
      const double d0 = i < m ? w.d[i] : 0.0;
      const double d1 = (i >= 1 && i - 1 < m) ? w.d[i - 1] : 0.0;
      const double d2 = (i >= 2 && i - 2 < m) ? w.d[i - 2] : 0.0;
      const double dg = 1.0 + d0 + 4.0 * d1 + d2;
      w.diag[i] = dg;
      w.s1[i] = -2.0 * d1 - 2.0 * d2;
      w.s2[i] = d2;
      w.sc[i] = sqrt(fmax(1.0, dg));
      dmax = fmax(dmax, dg);
    }
    dmax = bmax(dmax, c.sh);
    // factor: FP32 while the mixed path holds, else FP64; on a failed pivot
    // one retry with regularization 1e-12 max diag (at most 1e-6)
    bool fok = false;
    for (int attempt = 0; attempt < 2 && !fok; ++attempt) {
      const double reg = attempt == 0 ? 0.0 : fmin(1e-12 * fmax(1.0, dmax), 1e-6);
      __syncthreads();
      if (tid == 0) c.reg = reg;
      __syncthreads();
      if (c.prec == SOV_PREC_FP32) {
        fok = timed_factor<float>(n, c, reg, c.f32);
        if (fok) {
          if (tid == 0) ++c.rec.fp32_factors;
          break;
        }
        __syncthreads();
        if (tid == 0) c.prec = SOV_PREC_FP64;
        __syncthreads();
      }
      fok = timed_factor<double>(n, c, reg, c.f64);
      if (fok && tid == 0) ++c.rec.fp64_factors;
    }
    if (!fok) {
      if (tid == 0) c.rec.status = 6;
      break;
    }
    // affine predictor: rc = s z
    for (int r = tid; r < m; r += kT) w.rc[r] = w.s[r] * w.z[r];
    __syncthreads();
    if (!direction_safe(n, c, w.dxa, w.dsa, w.dza)) {
      if (tid == 0) c.rec.status = 6;
      break;
    }
    const double as = max_step(m, c, w.s, w.dsa), az = max_step(m, c, w.z, w.dza);
    double mua = 0.0;
    for (int r = tid; r < m; r += kT) mua += (w.s[r] + as * w.dsa[r]) * (w.z[r] + az * w.dza[r]);
    mua = bsum(mua, c.sh) / m;
    const double ratio = mu > 0.0 ? fmax(0.0, mua / mu) : 0.0;
    const double sigma = fmax(0.01, fmin(0.99, ratio * ratio * ratio));
    // corrector: rc = s z - sigma mu + ds_aff dz_aff
    for (int r = tid; r < m; r += kT) w.rc[r] = w.s[r] * w.z[r] - sigma * mu + w.dsa[r] * w.dza[r];
    __syncthreads();
    if (!direction_safe(n, c, w.dx, w.ds, w.dz)) {
      if (tid == 0) c.rec.status = 6;
      break;
    }
    const double ps = max_step(m, c, w.s, w.ds), pz = max_step(m, c, w.z, w.dz);
    const double alpha = fmin(fmin(1.0, 0.995 * ps), fmin(1.0, 0.995 * pz));
    if (!(alpha > 0.0) || !isfinite(alpha)) {
      if (tid == 0) c.rec.status = 6;
      break;
    }
    for (int j = tid; j < n; j += kT) w.x[j] += alpha * w.dx[j];
    for (int r = tid; r < m; r += kT) {
      w.s[r] += alpha * w.ds[r];
      w.z[r] += alpha * w.dz[r];
    }
    __syncthreads();
    if (tid == 0) c.rec.iterations = it + 1;
  }
  __syncthreads();
  if (tid == 0) *out = c.rec;
}

// ---- host -------------------------------------------------------------------

int levels_for(int n, Levels* L) {
  int N = (n + 1) / 2;
  L->count = 0;
  long long off = 0;
  while (true) {
    if (L->count >= kMaxLevels) return SOV_ERR_ARG;
    L->size[L->count] = N;
    L->off[L->count] = off;
    ++L->count;
    off += N;
    if (N == 1) break;
    N = (N + 1) / 2;
  }
  return SOV_OK;
}

}  // namespace

using sovqp::Certificate;
using sovqp::ld;

extern "C" {

void sov_qp_default_options(sov_qp_options* o) {
  if (!o) return;
  o->max_iterations = 200;
  o->precision = SOV_PREC_FP32;
  o->tolerance = 0.0;  // 0: the structured tolerance (1e-10, 1e-12 for m <= 50)
  o->polish = 1;
}

int sov_qp_liswet_solve(int32_t n, const double* target, const sov_qp_options* opt, sov_qp_result* res,
                        double* x_out, double* lambda_out) {
  if (n < 3 || !target || !res) return SOV_ERR_ARG;
  *res = sov_qp_result{};
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  sov_qp_options o;
  sov_qp_default_options(&o);
  if (opt) o = *opt;
  const auto t0 = std::chrono::steady_clock::now();
  const int m = n - 2;
  Levels L;
  if (levels_for(n, &L) != SOV_OK) return SOV_ERR_ARG;
  const long long blocks = L.off[L.count - 1] + 1;
  const std::size_t nb = static_cast<std::size_t>(blocks);
  const std::size_t nd = 12 * static_cast<std::size_t>(n) + 11 * static_cast<std::size_t>(m) + 2 * nb +
                         2 * static_cast<std::size_t>(L.size[0]) + 10 * nb;
  const std::size_t nf = 10 * nb;
  double* buf = nullptr;
  float* fbuf = nullptr;
  double* dtarget = nullptr;
  QpDev* dout = nullptr;
  SOV_HIP_TRY(hipMalloc(&buf, nd * sizeof(double)));
  struct Free {
    void* p[4];
    ~Free() {
      for (void* q : p)
        if (q) (void)hipFree(q);
    }
  } fr{{buf, nullptr, nullptr, nullptr}};
  SOV_HIP_TRY(hipMalloc(&fbuf, nf * sizeof(float)));
  fr.p[1] = fbuf;
  SOV_HIP_TRY(hipMalloc(&dtarget, n * sizeof(double)));
  fr.p[2] = dtarget;
  SOV_HIP_TRY(hipMalloc(&dout, sizeof(QpDev)));
  fr.p[3] = dout;
  Ws w;
  double* p = buf;
  for (double** v : {&w.x, &w.atz, &w.rd, &w.diag, &w.s1, &w.s2, &w.sc, &w.rhs, &w.dx, &w.dxa, &w.res, &w.corr}) {
    *v = p;
    p += n;
  }
  for (double** v : {&w.s, &w.z, &w.ax, &w.rp, &w.d, &w.ds, &w.dz, &w.dsa, &w.dza, &w.rc, &w.wt}) {
    *v = p;
    p += m;
  }
  w.brhs = p;
  p += 2 * nb;
  w.bx = p;
  p += 2 * static_cast<std::size_t>(L.size[0]);
  const Fac<double> f64{p, p + 3 * nb, p + 7 * nb};
  const Fac<float> f32{fbuf, fbuf + 3 * nb, fbuf + 7 * nb};
  hipStream_t s = sov::stream();
  SOV_HIP_TRY(hipMemcpyAsync(dtarget, target, n * sizeof(double), hipMemcpyHostToDevice, s));
  QpOpts qo;
  qo.tol = o.tolerance > 0 ? o.tolerance : (m <= 50 ? 1e-12 : 1e-10);
  qo.max_iter = std::max(1, o.max_iterations);
  qo.precision = o.precision == SOV_PREC_FP32 ? SOV_PREC_FP32 : SOV_PREC_FP64;
  hipEvent_t e0, e1;
  SOV_HIP_TRY(hipEventCreate(&e0));
  SOV_HIP_TRY(hipEventCreate(&e1));
  (void)hipEventRecord(e0, s);
  liswet_ipm_kernel<<<1, kT, 0, s>>>(n, dtarget, w, f32, f64, L, qo, dout);
  const hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, s);
  if (le != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    sov::set_error(std::string("liswet_ipm_kernel launch: ") + hipGetErrorString(le));
    return SOV_ERR_DEVICE;
  }
  QpDev rec{};
  std::vector<double> xh(n), sh(m), zh(m);
  SOV_HIP_TRY(hipMemcpyAsync(&rec, dout, sizeof(QpDev), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipMemcpyAsync(xh.data(), w.x, n * sizeof(double), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipMemcpyAsync(sh.data(), w.s, m * sizeof(double), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipMemcpyAsync(zh.data(), w.z, m * sizeof(double), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  float kms = 0.f;
  (void)hipEventElapsedTime(&kms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->iterations = rec.iterations;
  res->ipm_status = rec.status;
  res->fp32_factorizations = rec.fp32_factors;
  res->fp64_factorizations = rec.fp64_factors;
  res->refinements = rec.refinements;
  res->ipm_primal = rec.pres;
  res->ipm_dual = rec.dres;
  res->ipm_complementarity = rec.cres;
  res->kernel_ms = kms;
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0) {
      res->factor_ms = static_cast<double>(rec.t_factor) / khz;
      res->solve_ms = static_cast<double>(rec.t_solve) / khz;
    }
  }
  // knot-space polish and certificate on the host
  std::vector<ld> t(target, target + n), x(xh.begin(), xh.end());
  Certificate cert;
  std::vector<ld> lam, slack;
  if (o.polish) {
    std::vector<char> guess(m);
    for (int r = 0; r < m; ++r) guess[r] = sh[r] < zh[r];
    std::vector<ld> xp;
    sovqp::PolishStats st;
    if (sovqp::active_set_polish(t, guess, xp, cert, st)) x.swap(xp);
    res->polish_fits = st.fits;
  }
  cert = sovqp::certify(t, x, lam, slack);
  ld obj = 0.0L;
  for (int j = 0; j < n; ++j) obj += 0.5L * x[j] * x[j] - t[j] * x[j];
  res->objective = static_cast<double>(obj);
  res->primal = static_cast<double>(cert.primal);
  res->stationarity = static_cast<double>(cert.stationarity);
  res->complementarity = static_cast<double>(cert.complementarity);
  res->min_lambda = static_cast<double>(cert.min_lambda);
  res->consistency = static_cast<double>(cert.consistency);
  res->status = cert.ok ? 0 : 6;
  if (x_out)
    for (int j = 0; j < n; ++j) x_out[j] = static_cast<double>(x[j]);
  if (lambda_out)
    for (int r = 0; r < m; ++r) lambda_out[r] = static_cast<double>(lam[r]);
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  sov::count_operation();
  return SOV_OK;
}

int sov_qp_liswet_polish(int32_t n, const double* target, const double* s, const double* z, sov_qp_result* res,
                         double* x_out, double* lambda_out) {
  if (n < 3 || !target || !s || !z || !res) return SOV_ERR_ARG;
  *res = sov_qp_result{};
  const auto t0 = std::chrono::steady_clock::now();
  const int m = n - 2;
  std::vector<ld> t(target, target + n), x, lam, slack;
  std::vector<char> guess(m);
  for (int r = 0; r < m; ++r) guess[r] = s[r] < z[r];
  Certificate cert;
  sovqp::PolishStats st;
  if (!sovqp::active_set_polish(t, guess, x, cert, st)) x.assign(t.begin(), t.end());
  cert = sovqp::certify(t, x, lam, slack);
  ld obj = 0.0L;
  for (int j = 0; j < n; ++j) obj += 0.5L * x[j] * x[j] - t[j] * x[j];
  res->status = cert.ok ? 0 : 6;
  res->polish_fits = st.fits;
  res->objective = static_cast<double>(obj);
  res->primal = static_cast<double>(cert.primal);
  res->stationarity = static_cast<double>(cert.stationarity);
  res->complementarity = static_cast<double>(cert.complementarity);
  res->min_lambda = static_cast<double>(cert.min_lambda);
  res->consistency = static_cast<double>(cert.consistency);
  if (x_out)
    for (int j = 0; j < n; ++j) x_out[j] = static_cast<double>(x[j]);
  if (lambda_out)
    for (int r = 0; r < m; ++r) lambda_out[r] = static_cast<double>(lam[r]);
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return SOV_OK;
}

}  // extern "C"
