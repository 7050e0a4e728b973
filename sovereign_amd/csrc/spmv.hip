// Sparse matrix-vector products y = A x and y = A^T x, FP64, int32 indices.
//
// Replaces the reference's `row_dot` kernel (one thread per output row,
// solver/gpu/src/cuda_driver.cpp:81-133) and its per-call host copies
// (cuda_driver.cpp:584-613): the matrix lives in HBM behind a handle, lanes
// cooperate per output so loads are coalesced, and the device-vector entry
// point lets the interior-point loop keep every vector resident.

#include "chol.hpp"
#include "internal.hpp"
#include "spmv_device.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <mutex>

namespace {

constexpr int kBlock = 256;

template <int G>
__global__ __launch_bounds__(kBlock) void csr_spmv_kernel(int nrows, const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ idx,
                                                          const double* __restrict__ val,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ y) {
  const int sub = threadIdx.x & (G - 1);
  const int row = (blockIdx.x * kBlock + threadIdx.x) / G;
  const double v = csr_row_dot<G>(ptr, idx, val, x, row, nrows, sub);
  if (sub == 0 && row < nrows) y[row] = v;
}

int group_for(double mean) {
  int g = 1;
  while (g < 64 && g < mean) g <<= 1;
  return g;
}

template <typename T>
bool dev_copy(T** dst, const T* src, std::size_t count) {
  if (hipMalloc(reinterpret_cast<void**>(dst), std::max<std::size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    *dst = nullptr;
    return false;
  }
  if (count == 0) return true;
  return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

}  // namespace

namespace sov {

int mat_upload(int32_t nrows, int32_t ncols, const int32_t* col_ptr, const int32_t* row_idx,
               const double* values, sov_mat** out) {
  *out = nullptr;
  if (nrows < 0 || ncols < 0 || !col_ptr) {
    set_error("sov_mat_upload_csc: bad dimensions");
    return SOV_ERR_ARG;
  }
  const int64_t nnz = col_ptr[ncols];
  if (col_ptr[0] != 0 || nnz < 0) {
    set_error("sov_mat_upload_csc: col_ptr must start at 0");
    return SOV_ERR_ARG;
  }
  for (int j = 0; j < ncols; ++j) {
    if (col_ptr[j + 1] < col_ptr[j]) {
      set_error("sov_mat_upload_csc: col_ptr is not monotone");
      return SOV_ERR_ARG;
    }
  }
  for (int64_t p = 0; p < nnz; ++p) {
    if (row_idx[p] < 0 || row_idx[p] >= nrows) {
      set_error("sov_mat_upload_csc: row index out of range");
      return SOV_ERR_ARG;
    }
  }
  auto* m = new sov_mat;
  m->nrows = nrows;
  m->ncols = ncols;
  m->nnz = nnz;
  m->h_cptr.assign(col_ptr, col_ptr + ncols + 1);
  m->h_crow.assign(row_idx, row_idx + nnz);
  m->h_cval.assign(values, values + nnz);
  // CSR of A by a counting sort of the entries on their row (stable in j).
  std::vector<int32_t> rptr(static_cast<std::size_t>(nrows) + 1, 0), rcol(static_cast<std::size_t>(nnz));
  std::vector<double> rval(static_cast<std::size_t>(nnz));
  for (int64_t p = 0; p < nnz; ++p) ++rptr[row_idx[p] + 1];
  for (int r = 0; r < nrows; ++r) rptr[r + 1] += rptr[r];
  {
    std::vector<int32_t> next(rptr.begin(), rptr.end() - 1);
    for (int j = 0; j < ncols; ++j) {
      for (int p = col_ptr[j]; p < col_ptr[j + 1]; ++p) {
        const int32_t q = next[row_idx[p]]++;
        rcol[q] = j;
        rval[q] = values[p];
      }
    }
  }
  m->group_rows = group_for(nrows ? static_cast<double>(nnz) / nrows : 1.0);
  m->group_cols = group_for(ncols ? static_cast<double>(nnz) / ncols : 1.0);
  const std::size_t vec = static_cast<std::size_t>(std::max(nrows, ncols));
  bool ok = dev_copy(&m->r_ptr, rptr.data(), rptr.size()) && dev_copy(&m->r_col, rcol.data(), rcol.size()) &&
            dev_copy(&m->r_val, rval.data(), rval.size()) &&
            dev_copy(&m->c_ptr, col_ptr, static_cast<std::size_t>(ncols) + 1) &&
            dev_copy(&m->c_row, row_idx, static_cast<std::size_t>(nnz)) &&
            dev_copy(&m->c_val, values, static_cast<std::size_t>(nnz)) &&
            hipMalloc(&m->x_buf, std::max<std::size_t>(vec, 1) * sizeof(double)) == hipSuccess &&
            hipMalloc(&m->y_buf, std::max<std::size_t>(vec, 1) * sizeof(double)) == hipSuccess;
  if (!ok) {
    mat_release(m);
    set_error("sov_mat_upload_csc: device allocation or copy failed");
    return SOV_ERR_NOMEM;
  }
  *out = m;
  return SOV_OK;
}

void mat_release(sov_mat* m) {
  if (!m) return;
  void* ptrs[] = {m->r_ptr, m->r_col, m->r_val, m->c_ptr, m->c_row, m->c_val, m->x_buf, m->y_buf};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m;
}

int spmv_launch(const sov_mat* a, bool transpose, const double* x, double* y, hipStream_t s) {
  const int out = transpose ? a->ncols : a->nrows;
  if (out == 0) return SOV_OK;
  const int32_t* ptr = transpose ? a->c_ptr : a->r_ptr;
  const int32_t* idx = transpose ? a->c_row : a->r_col;
  const double* val = transpose ? a->c_val : a->r_val;
  const int g = transpose ? a->group_cols : a->group_rows;
  TimedPair* tk = timed_kernel(kTimeSpmv);
  if (tk) {  // algorithmic bytes: values and indices, row pointers, x and y once each
    const long long nnz = a->nnz;
    const int in = transpose ? a->nrows : a->ncols;
    tk->launches = 1;
    tk->flops = 12.0 * nnz + 4.0 * (out + 1) + 8.0 * (in + out);
  }
  SOV_DISPATCH_GROUP(g, {
    const unsigned grid = blocks_for(static_cast<long long>(out) * G, kBlock);
    if (tk)
      hipExtLaunchKernelGGL(csr_spmv_kernel<G>, dim3(grid), dim3(kBlock), 0, s, tk->a, tk->b, 0, out, ptr, idx, val, x,
                            y);
    else
      csr_spmv_kernel<G><<<grid, kBlock, 0, s>>>(out, ptr, idx, val, x, y);
  });
  SOV_LAUNCH_CHECK();
  return SOV_OK;
}

}  // namespace sov

namespace {

// One resident copy per thread and direction for the one-shot entry point,
// replaced when the contents change (cuda_driver.cpp:509-582 semantics).
// Never freed at thread or process exit: the HIP runtime may already be shut
// down then, and the driver reclaims the memory with the process.
struct OneShotCache {
  sov_mat* mat = nullptr;
};

bool same_contents(const sov_mat* m, int32_t nrows, int32_t ncols, const int32_t* cp, const int32_t* ri,
                   const double* v) {
  if (!m || m->nrows != nrows || m->ncols != ncols || m->h_cptr[ncols] != cp[ncols]) return false;
  const std::size_t nnz = static_cast<std::size_t>(cp[ncols]);
  return std::memcmp(m->h_cptr.data(), cp, (static_cast<std::size_t>(ncols) + 1) * sizeof(int32_t)) == 0 &&
         std::memcmp(m->h_crow.data(), ri, nnz * sizeof(int32_t)) == 0 &&
         std::memcmp(m->h_cval.data(), v, nnz * sizeof(double)) == 0;
}

}  // namespace

extern "C" {

int sov_mat_upload_csc(int32_t nrows, int32_t ncols, const int32_t* col_ptr, const int32_t* row_idx,
                       const double* values, sov_mat** out) {
  if (!out) return SOV_ERR_ARG;
  *out = nullptr;
  if (!sov::device_ready()) {
    sov::set_error("no usable GPU");
    return SOV_ERR_NODEV;
  }
  return sov::mat_upload(nrows, ncols, col_ptr, row_idx, values, out);
}

int sov_mat_free(sov_mat* a) {
  sov::mat_release(a);
  return SOV_OK;
}

int sov_spmv_dev(sov_mat* a, int32_t transpose, const double* x_dev, double* y_dev) {
  if (!a || !x_dev || !y_dev) return SOV_ERR_ARG;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  const int rc = sov::spmv_launch(a, transpose != 0, x_dev, y_dev, sov::stream());
  if (rc == SOV_OK) sov::count_operation();
  return rc;
}

int sov_spmv(sov_mat* a, int32_t transpose, const double* x_host, double* y_host) {
  if (!a || !x_host || !y_host) return SOV_ERR_ARG;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  const std::size_t in = static_cast<std::size_t>(transpose ? a->nrows : a->ncols);
  const std::size_t out = static_cast<std::size_t>(transpose ? a->ncols : a->nrows);
  hipStream_t s = sov::stream();
  if (in) SOV_HIP_TRY(hipMemcpyAsync(a->x_buf, x_host, in * sizeof(double), hipMemcpyHostToDevice, s));
  const int rc = sov::spmv_launch(a, transpose != 0, a->x_buf, a->y_buf, s);
  if (rc != SOV_OK) return rc;
  if (out) SOV_HIP_TRY(hipMemcpyAsync(y_host, a->y_buf, out * sizeof(double), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  sov::count_operation();
  return SOV_OK;
}

int sov_spmv_csc(int32_t nrows, int32_t ncols, const int32_t* col_ptr, const int32_t* row_idx,
                 const double* values, const double* x_host, double* y_host, int32_t transpose) {
  if (!col_ptr || !x_host || !y_host) return SOV_ERR_ARG;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  static thread_local OneShotCache caches[2];
  OneShotCache& c = caches[transpose ? 1 : 0];
  if (!same_contents(c.mat, nrows, ncols, col_ptr, row_idx, values)) {
    sov::mat_release(c.mat);
    c.mat = nullptr;
    const int rc = sov::mat_upload(nrows, ncols, col_ptr, row_idx, values, &c.mat);
    if (rc != SOV_OK) return rc;
  }
  return sov_spmv(c.mat, transpose, x_host, y_host);
}

}  // extern "C"
