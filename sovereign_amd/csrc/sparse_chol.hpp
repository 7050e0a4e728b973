// Sparse Cholesky of the normal matrix M = A diag(d) A^T + reg I for LPs whose
// factor stays sparse (staircase and other banded-block models), the GPU
// counterpart of the reference's AMD-ordered up-looking LDL^T
// (solver/numerical/src/sparse_ldlt.cpp:19-223, amd_ordering.cpp:12-130).
//
// Host, once per matrix (symbolic):
//   minimum-degree ordering of the pattern (the reference's rule), the
//   elimination tree, its postorder and column counts, fundamental supernodes
//   merged into small relaxed ones, each supernode's row structure, the
//   supernodal tree cut into levels (leaves first), the maps that place a
//   child's update rows in its parent's front, and for every entry of M its
//   slot in a front.
// Device, per factorization (numeric, multifrontal):
//   every front is a dense column-major f x f block in one HBM buffer; the
//   entries of M are scattered into the fronts, then one launch per level
//   gives each front of the level a workgroup that adds its children's update
//   matrices, factors its leading columns with the reference's pivot rule
//   (a pivot <= 1e-13 |original diagonal| is replaced by 1e128, the
//   reference's drop, a non-finite or < 1e-30 one fails) and forms its own
//   update matrix. The fronts' leading columns are the factor L (Cholesky
//   form: L_kk = sqrt(d_k)).
// Solves walk the same levels: forward (leaves first, each front sums its
// children's update vectors) and backward (root first, each front gathers
// its ancestors' solution entries). Every sum runs in a fixed order, so
// results do not change from run to run.
#pragma once

#include "chol.hpp"
#include "internal.hpp"
#include "sparse_symbolic.hpp"

#include <cstdint>
#include <vector>

namespace sov {

class SparseCholesky {
 public:
  ~SparseCholesky();
  SparseSymbolic sym;
  int upload();                                  // device copies of the symbolic data, buffers
  double* fronts() { return fronts_; }           // filled by the caller's scatter of M's entries
  double* diag0() { return diag0_; }             // original diagonal (permuted order), set by the scatter
  int zero_fronts(hipStream_t s);
  int factor(hipStream_t s);                     // asynchronous; flags() report breakdown / replaced pivots
  int solve(double* x_dev, hipStream_t s);       // x <- M^-1 x (original order, device)
  FactorFlags* flags() { return flags_; }

 private:
  double* fronts_ = nullptr;
  double* diag0_ = nullptr;
  double* vf_ = nullptr;   // per-front vectors of the solves
  double* xp_ = nullptr;   // permuted right-hand side / solution
  FactorFlags* flags_ = nullptr;
  int32_t *d_perm_ = nullptr, *d_first_ = nullptr, *d_ncol_ = nullptr, *d_front_ = nullptr, *d_rows_ = nullptr;
  int32_t *d_child_ptr_ = nullptr, *d_child_idx_ = nullptr, *d_maps_ = nullptr, *d_level_idx_ = nullptr;
  int64_t *d_front_base_ = nullptr, *d_vec_base_ = nullptr, *d_map_base_ = nullptr;
};

}  // namespace sov
