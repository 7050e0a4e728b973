// Memory layout of the batched child-LP kernel (batch_ipm.hip), shared by
// the kernel, which carves each LP's workspace and LDS with these functions,
// the host code that sizes the buffers, and a host-only audit
// (tests/batch_layout_audit.cpp) that walks every vector access of the FP32
// assembly against the real allocation ends. Plain C++ apart from the
// __host__ __device__ markers, so that g++ compiles it too.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define SOV_HD __host__ __device__
#else
#define SOV_HD
#endif

namespace sov_batch_layout {

constexpr int kMaxRows = 272;    // rows of the largest LP: its packed FP32 normal matrix fills the LDS
constexpr int kStage = 1024;     // doubles of LDS staging (vectors, partial sums)
constexpr int kRed = 64;         // doubles of LDS for block reductions
constexpr int kBlk = 16;         // block width of the LDS factorization and inversion
constexpr int kNVectors = 16;    // n-vectors of the workspace: x w z v rc ru D dx dw dz dv dxa dwa dza dva t
constexpr int kMVectors = 9;     // m-vectors: y rb dy r p q zz rhs sc
constexpr int kBlockFloats = kBlk * kBlk;  // inverse of one 16 x 16 diagonal block

// Leading dimension of the row-major copies of A: n rounded up to 32 (the
// assembly reads 32 columns a step, 16 bytes a lane).
SOV_HD constexpr int row_ld(int n) { return (n + 31) / 32 * 32; }
SOV_HD constexpr int diag_blocks(int m) { return (m + kBlk - 1) / kBlk; }

// Per-LP workspace, in doubles from the LP's start (8-byte aligned):
//   [0, 16 n)                 the n-vectors, vector k at k n
//   [16 n, 16 n + 9 m)        the m-vectors, vector k at 16 n + k m
//   M64                       packed lower FP64 normal matrix, m (m + 1) / 2
//   xs                        256 floats per diagonal block (128 doubles)
//   sd                        ldr floats (sqrt(D) in FP32), from the first
//                             16-byte boundary at or after the end of xs
// The last two doubles are the alignment slack of sd (at most 8 bytes are
// skipped, since the start is 8-byte aligned) and one spare.
SOV_HD constexpr long long n_vector(int k, int n) { return static_cast<long long>(k) * n; }
SOV_HD constexpr long long m_vector(int k, int m, int n) {
  return static_cast<long long>(kNVectors) * n + static_cast<long long>(k) * m;
}
SOV_HD constexpr long long off_m64(int m, int n) { return m_vector(kMVectors, m, n); }
SOV_HD constexpr long long off_xs(int m, int n) { return off_m64(m, n) + static_cast<long long>(m) * (m + 1) / 2; }
SOV_HD constexpr long long xs_doubles(int m) { return static_cast<long long>(diag_blocks(m)) * kBlockFloats / 2; }
SOV_HD constexpr long long ws_doubles(int m, int n) { return off_xs(m, n) + xs_doubles(m) + row_ld(n) / 2 + 2; }
// Address of sd given the address just past xs.
SOV_HD inline std::uintptr_t sd_address(std::uintptr_t xs_end) {
  return (xs_end + 15) & ~static_cast<std::uintptr_t>(15);
}

// Dynamic LDS, in bytes from its base: stage (kStage doubles), red (kRed
// doubles), then the packed FP32 normal matrix of the largest LP in the launch.
SOV_HD constexpr std::size_t lds_off_red() { return sizeof(double) * kStage; }
SOV_HD constexpr std::size_t lds_off_matrix() { return sizeof(double) * (kStage + kRed); }
SOV_HD constexpr std::size_t lds_bytes(int mmax) {
  return lds_off_matrix() + sizeof(float) * static_cast<std::size_t>(mmax) * (mmax + 1) / 2;
}
// Bytes of the LDS matrix region (the packed FP32 normal matrix of the
// launch's largest LP, mmax rows).
SOV_HD constexpr long long lds_matrix_bytes(int mmax) {
  return static_cast<long long>(sizeof(float)) * mmax * (mmax + 1) / 2;
}
// An LP that has switched to FP64 factorizations no longer uses that region:
// its blocked FP64 Cholesky and inverse stage a 16-column panel of m rows
// there (16 m doubles), when it fits.
constexpr int kPanel64 = 16;
SOV_HD constexpr bool fp64_panel_fits(int m, long long matrix_bytes) {
  return static_cast<long long>(sizeof(double)) * kPanel64 * m <= matrix_bytes;
}
// Staging offsets (doubles) used by the preconditioner: vector and temporary.
constexpr int kStagePv = 0, kStageTv = 512;
static_assert(kStageTv >= kMaxRows && kStageTv + kMaxRows <= kStage, "preconditioner staging overlaps");

// Work items of the FP32 assembly: block row I (of nt) with up to four of its
// tiles (I, J0 .. J0 + 3), J0 a multiple of 4 and J0 <= I.
SOV_HD inline int assembly_items(int nt) {
  int q = 0;
  for (int I = 0; I < nt; ++I) q += I / 4 + 1;
  return q;
}
SOV_HD inline void assembly_item(int q, int& I, int& J0) {
  int i = 0, rem = q;
  while (rem >= i / 4 + 1) {
    rem -= i / 4 + 1;
    ++i;
  }
  I = i;
  J0 = 4 * rem;
}

}  // namespace sov_batch_layout
