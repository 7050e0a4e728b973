// Normal equations M = A diag(d) A^T + reg I on the device: assembly plan,
// factorization (FP64, or FP32 with FP64 preconditioned CG) and solves.
#pragma once

#include "chol.hpp"
#include "internal.hpp"
#include "sparse_chol.hpp"

#include <memory>

namespace sov {

constexpr double kNormalReg = 1e-12;  // interior_point.cpp:99,103

// Scalars of the preconditioned CG kept on the device.
struct PcgScalars {
  double rz, pq, alpha, beta, rnorm, bnorm;
};

class NormalSolver {
 public:
  ~NormalSolver();
  // path: SOV_NE_AUTO (sparse when the ordered factor stays sparse; the
  // SOVEREIGN_NE_PATH environment variable, dense / sparse, overrides it),
  // SOV_NE_DENSE or SOV_NE_SPARSE.
  int init(sov_mat* a, int path = SOV_NE_AUTO);
  // factor(d) with d on the device. Asynchronous; check breakdown().
  int factor(const double* d_dev, int precision, hipStream_t s);
  // Waits for the last factorization and reports 0 ok, 1 breakdown.
  int breakdown(hipStream_t s);
  // x <- M^-1 x in place (device vector of nrows). For an FP32 factor this
  // runs PCG against the FP64 matrix-free product until the relative
  // residual is below tol or max_iter is reached. Returns SOV_OK, or
  // SOV_SINGULAR when PCG did not reach tol.
  int solve(double* x_dev, hipStream_t s, double tol, int max_iter);

  int precision() const { return prec_; }
  int last_pcg_iterations() const { return last_iters_; }
  double last_residual() const { return last_res_; }
  int replaced_pivots() const { return replaced_; }
  int dense_mode() const { return dense_a_ != nullptr; }
  // The sparse factorization is in use (every factorization FP64, direct solves).
  bool sparse_mode() const { return sp_ != nullptr; }
  const SparseSymbolic* sparse_symbolic() const { return sp_ ? &sp_->sym : nullptr; }
  DenseCholesky<double> ch64;
  DenseCholesky<float> ch32;
  double assemble_ms = 0.0;

 private:
  template <typename T>
  int assemble(const double* d_dev, T* M, hipStream_t s);
  int matvec(const double* p, double* q, hipStream_t s);  // q = M p in FP64, matrix-free
  int init_sparse(const std::vector<long long>& eoff, int path);

  sov_mat* A_ = nullptr;
  int m_ = 0, n_ = 0;
  const double* d_ = nullptr;  // d of the current factorization (device, caller-owned)
  // Sparse assembly plan: entry e (a lower-triangle position of M) sums
  // val[ta[t]] * d[tj[t]] * val[tb[t]] over t in [eptr[e], eptr[e+1]).
  long long entries_ = 0, triples_ = 0;
  int32_t* eptr_ = nullptr;
  int32_t* ta_ = nullptr;
  int32_t* tb_ = nullptr;
  int32_t* tj_ = nullptr;
  long long* eoff_ = nullptr;
  int32_t* ediag_ = nullptr;  // row index for diagonal entries, -1 otherwise
  // Sparse factorization: each plan entry's slot in the fronts, and the
  // permuted row of diagonal entries (-1 otherwise).
  std::unique_ptr<SparseCholesky> sp_;
  long long* slot_ = nullptr;
  int32_t* ediag_perm_ = nullptr;
  // Dense-A plan (when the sparse plan would exceed the dense matrix):
  double* dense_a_ = nullptr;  // m x n column-major
  void* wbuf_ = nullptr;       // A diag(sqrt d), T = factor precision
  double* M64_ = nullptr;
  float* M32_ = nullptr;
  double* diag0_ = nullptr;
  // PCG workspace
  double *r_ = nullptr, *z_ = nullptr, *p_ = nullptr, *q_ = nullptr, *t_ = nullptr, *b_ = nullptr;
  double* partial_ = nullptr;
  PcgScalars* sc_ = nullptr;
  PcgScalars* sc_host_ = nullptr;
  LaunchGraphs pcg_graphs_;  // one CG iteration's launches, replayed per (solution vector, d)
  int prec_ = 0;
  int last_iters_ = 0;
  double last_res_ = 0.0;
  int replaced_ = 0;
};

}  // namespace sov
