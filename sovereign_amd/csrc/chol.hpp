// Blocked dense Cholesky of the normal equations and its triangular solves.
#pragma once

#include "common.hpp"
#include "graphs.hpp"

namespace sov {

// Replacement value for a dropped pivot: L(k,k) = kDropPivot<T>, so the
// direction is scaled out (the reference's LDL^T stores d_k = 1e128).
template <typename T>
struct DropPivot;
template <>
struct DropPivot<double> {
  static constexpr double value = 1e64;
};
template <>
struct DropPivot<float> {
  static constexpr float value = 1e18f;
};

constexpr int kPanel = 128;   // factorization panel width (K of the trailing update)
constexpr int kSuper = 1024;  // diagonal blocks whose inverses the triangular solves use

// Per-factorization device state shared by the kernels.
struct FactorFlags {
  int status;    // 0 ok, 1 breakdown (non-finite or < 1e-30 pivot)
  int replaced;  // dropped pivots
};

// Owns the factor workspace for an m x m matrix stored column-major in `a`
// (lower triangle used). All work goes to `s`.
template <typename T>
class DenseCholesky {
 public:
  ~DenseCholesky();
  int reserve(int m);  // allocates the panel inverses, scratch and flags
  // In-place factorization of a (lower); diag0 holds the diagonal before
  // elimination (for the drop test). Asynchronous; status in flags().
  int factor(T* a, int m, const double* diag0, hipStream_t s);
  // x <- L^-T L^-1 x on device vector x (FP64), using the factor in a.
  // The launch sequence is replayed as a graph per (a, x, m).
  int solve(const T* a, int m, double* x, double* work, hipStream_t s);
  FactorFlags* flags() const { return flags_; }
  // Inverses of the factor's kSuper x kSuper diagonal blocks (block P at
  // P * kSuper * kSuper, leading dimension kSuper, zero above the diagonal).
  const T* super_inverses() const { return sup_inv_; }

 private:
  int build_super_inverses(const T* a, int m, hipStream_t s);

  int m_ = 0;
  T* panel_ = nullptr;      // L21 of two consecutive panels (2 x m x kPanel, leading dimension m)
  hipStream_t side_ = nullptr;          // trailing updates run here, overlapping the next panel
  hipEvent_t ev_panel_ = nullptr, ev_bulk_ = nullptr;
  T* sup_inv_ = nullptr;  // inverses of the kSuper x kSuper diagonal blocks of L (leading dimension kSuper)
  T* tmp_ = nullptr;      // products of the recursive doubling that builds them
  double* work_ = nullptr;  // intermediate vector of the solves (m)
  FactorFlags* flags_ = nullptr;
  LaunchGraphs solve_graphs_;  // solve sequences replayed per (factor, vector, m)
};

// Global switch for per-kernel event timing (adds synchronisation) and the
// totals it accumulates (sov_kernel_times).
struct KernelTimes {
  double syrk_ms = 0.0, syrk_launches = 0.0, syrk_flops = 0.0;
  double panel_ms = 0.0, solve_ms = 0.0, assemble_ms = 0.0, panel_launches = 0.0;
  double strip_ms = 0.0, inverse_ms = 0.0;
  double syrk32_ms = 0.0, syrk32_launches = 0.0, syrk32_flops = 0.0;  // FP32 factorizations' SYRK
  // the diagonal-block (panel) kernel alone, per precision: time, launches, flops (factor + inverse)
  double potrf64_ms = 0.0, potrf64_launches = 0.0, potrf64_flops = 0.0;
  double potrf32_ms = 0.0, potrf32_launches = 0.0, potrf32_flops = 0.0;
  double spmv_ms = 0.0, spmv_launches = 0.0, spmv_bytes = 0.0;  // standalone CSR SpMV launches
  double solve_pairs = 0.0, solve_bytes = 0.0;                  // triangular solve pairs, factor bytes read
};
void set_kernel_timing(int level);
bool kernel_timing();
KernelTimes& kernel_times();
void collect_kernel_times(hipStream_t s);

// Event pair around work on a stream (nullptr when timing is off).
enum TimedKind {
  kTimeSyrk = 0,  // FP64 bulk trailing update
  kTimePanel = 1,
  kTimeSolve = 2,
  kTimeStrip = 3,
  kTimeInverse = 4,
  kTimeAssemble = 5,
  kTimeSyrk32 = 6,   // FP32 bulk trailing update
  kTimePotrf64 = 7,  // FP64 diagonal-block kernel
  kTimePotrf32 = 8,  // FP32 diagonal-block kernel
  kTimeSpmv = 9      // standalone CSR SpMV
};
// Event pair of one timed launch (or a run of launches), with its launch
// count and its algorithmic work (flops, or bytes for the memory-bound
// kinds).
struct TimedPair {
  hipEvent_t a = nullptr, b = nullptr;
  int kind = 0;
  double flops = 0.0;
  int launches = 0;
};
// A pair for a kernel launch that records both events itself
// (hipExtLaunchKernel), or nullptr when timing is off for `kind`.
TimedPair* timed_kernel(int kind);
TimedPair* timed_begin(hipStream_t s, int kind);
void timed_end(TimedPair* p, hipStream_t s, int launches, double flops);

}  // namespace sov
