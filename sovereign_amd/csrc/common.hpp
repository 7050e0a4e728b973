// Shared definitions of the HIP library: error plumbing, the library stream,
// launch helpers and wavefront reductions for gfx950 (64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "sovereign_hip.h"

namespace sov {

// Thread-local text of the last failure, returned by sov_last_error().
void set_error(const std::string& text);
const char* last_error();

// Library stream (non-blocking, default priority) of the current device.
hipStream_t stream();

// True once sov_dev_probe has found a usable device.
bool device_ready();

void count_operation();
void count_factorization();

constexpr int kWave = 64;

inline unsigned blocks_for(long long count, int per_block) {
  return static_cast<unsigned>((count + per_block - 1) / per_block);
}

}  // namespace sov

#define SOV_HIP_TRY(expr)                                                              \
  do {                                                                                 \
    hipError_t sov_err_ = (expr);                                                      \
    if (sov_err_ != hipSuccess) {                                                      \
      sov::set_error(std::string(#expr) + ": " + hipGetErrorString(sov_err_));         \
      return SOV_ERR_DEVICE;                                                           \
    }                                                                                  \
  } while (0)

#define SOV_LAUNCH_CHECK()                                                             \
  do {                                                                                 \
    hipError_t sov_err_ = hipGetLastError();                                           \
    if (sov_err_ != hipSuccess) {                                                      \
      sov::set_error(std::string("kernel launch: ") + hipGetErrorString(sov_err_));    \
      return SOV_ERR_DEVICE;                                                           \
    }                                                                                  \
  } while (0)

// ---------------------------------------------------------------------------
// Wavefront reductions (64 lanes) through cross-lane shuffles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
// Sum over aligned groups of G lanes (G a power of two <= 64).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide reduction of one value per thread; result valid in thread 0.
// `scratch` holds blockDim.x / 64 doubles. op: 0 sum, 1 max, 2 min.
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  v = OP == 0 ? wave_sum(v) : OP == 1 ? wave_max(v) : wave_min(v);
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = scratch[0];
    for (int w = 1; w < nw; ++w) r = OP == 0 ? r + scratch[w] : OP == 1 ? fmax(r, scratch[w]) : fmin(r, scratch[w]);
    scratch[0] = r;
  }
  __syncthreads();
  return scratch[0];
}
