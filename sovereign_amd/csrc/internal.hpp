// Cross-file declarations of the HIP library (not part of the C ABI).
#pragma once

#include "common.hpp"

#include <vector>

// Device-resident sparse matrix. y = A x runs over the CSR of A (one output
// per row), y = A^T x over the CSC arrays read as the CSR of A^T (one output
// per column). `group_*` is the number of lanes cooperating on one output,
// the power of two nearest the mean entries per output (capped at 64).
struct sov_mat {
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  int32_t* r_ptr = nullptr;  // CSR of A
  int32_t* r_col = nullptr;
  double* r_val = nullptr;
  int32_t* c_ptr = nullptr;  // CSC of A
  int32_t* c_row = nullptr;
  double* c_val = nullptr;
  int group_rows = 1, group_cols = 1;
  double* x_buf = nullptr;  // staging for host-vector products
  double* y_buf = nullptr;
  // Host copy of the CSC arrays (normal-equation planning, content checks).
  std::vector<int32_t> h_cptr, h_crow;
  std::vector<double> h_cval;
};

namespace sov {

int mat_upload(int32_t nrows, int32_t ncols, const int32_t* col_ptr, const int32_t* row_idx,
               const double* values, sov_mat** out);
void mat_release(sov_mat* a);

// y = op(A) x on device vectors, on `s`.
int spmv_launch(const sov_mat* a, bool transpose, const double* x, double* y, hipStream_t s);

// Dense LU of a host column-major matrix (sov_dense_lu without the probe gate).
int dense_lu_run(int64_t n, double* a, int32_t* piv);

}  // namespace sov
