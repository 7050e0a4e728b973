// Device-side CSR row products shared by the SpMV and IPM kernels.
//
// G lanes of a wavefront cooperate on one output: they stride over the row's
// entries G apart (so a group reads G consecutive values and indices, one
// coalesced segment), then combine their partial sums with cross-lane
// shuffles. Summation order is fixed, so results are deterministic.
#pragma once

#include "common.hpp"

// Partial sum of lane `sub` (0 <= sub < G) of row `row`, then reduced over the
// group; every lane of the group returns the row's value. Rows >= nrows
// contribute 0 but still take part in the shuffles.
template <int G>
__device__ __forceinline__ double csr_row_dot(const int32_t* __restrict__ ptr,
                                              const int32_t* __restrict__ idx,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, int row, int nrows,
                                              int sub) {
  double acc = 0.0;
  if (row < nrows) {
    const int end = ptr[row + 1];
    for (int p = ptr[row] + sub; p < end; p += G) acc = fma(val[p], x[idx[p]], acc);
  }
  if (G > 1) acc = group_sum<G>(acc);
  return acc;
}

// Dispatch a kernel template on the group size (1..64, a power of two).
#define SOV_DISPATCH_GROUP(G_VAR, ...)        \
  switch (G_VAR) {                            \
    case 1: { constexpr int G = 1; __VA_ARGS__; break; }   \
    case 2: { constexpr int G = 2; __VA_ARGS__; break; }   \
    case 4: { constexpr int G = 4; __VA_ARGS__; break; }   \
    case 8: { constexpr int G = 8; __VA_ARGS__; break; }   \
    case 16: { constexpr int G = 16; __VA_ARGS__; break; } \
    case 32: { constexpr int G = 32; __VA_ARGS__; break; } \
    default: { constexpr int G = 64; __VA_ARGS__; break; } \
  }
