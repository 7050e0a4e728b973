// Host-only audit of the batched child-LP kernel's memory layout
// (sovereign_amd/csrc/batch_layout.hpp, the functions the kernel carves its
// workspace and LDS with and the host sizes its buffers with). For batches of
// LPs of every row count 1..272 and many column counts, laid out one after
// another as the host does, it checks at byte granularity:
//   * each workspace region (16 n-vectors, 9 m-vectors, the packed FP64
//     matrix, the diagonal-block inverses, sqrt(D)) lies inside the LP's
//     share of the workspace, after the one before it; sqrt(D) is 16-byte
//     aligned and all ldr floats of it fit;
//   * every 16-byte load of the FP32 assembly (row sums and tiles, every lane
//     of every work item, whether its predicate is on or off) lies inside the
//     row-major FP32 copy of A (m x ldr floats) or inside sqrt(D), 16-byte
//     aligned, and every 32-byte load of the FP64 assembly's tiles inside the
//     row-major FP64 copy (m x ldr doubles), 32-byte aligned;
//   * the work items cover every lower 16 x 16 tile exactly once, and every
//     packed-matrix store lies inside the LDS bytes the launch reserves;
//   * the dynamic LDS of the largest LP, with room for the kernel's static
//     LDS, fits the 160 KiB of a workgroup.
// Prints "ok <checks>" and exits 0, or prints the first failures and exits 1.
// Built with g++ by tests/test_batch_layout.py.
#include "batch_layout.hpp"

#include <cstdint>
#include <cstdio>
#include <vector>

namespace L = sov_batch_layout;

namespace {

long long checks = 0;
int failures = 0;
// Planted mistakes: with argument "w" each LP's workspace ends 16 bytes
// early (the spare double and the alignment slack), with "a" the FP32 copy
// of A ends one float early, with "d" the FP64 copy one double early; the
// audit must then fail.
std::uintptr_t plant_ws = 0, plant_a32 = 0, plant_a64 = 0;

void fail(const char* what, int m, int n, long long detail) {
  if (failures < 20) std::printf("FAIL %s m=%d n=%d detail=%lld\n", what, m, n, detail);
  ++failures;
}

// [a, a + bytes) inside [lo, hi) and a a multiple of align.
void in_range(std::uintptr_t a, std::size_t bytes, std::uintptr_t lo, std::uintptr_t hi, std::size_t align,
              const char* what, int m, int n) {
  ++checks;
  if (a < lo || a + bytes > hi) fail(what, m, n, static_cast<long long>(a - lo));
  if (a % align != 0) fail(what, m, n, static_cast<long long>(a % align));
}

void audit_lp(int m, int n, std::uintptr_t wsp, int mmax) {
  const int ldr = L::row_ld(n);
  const std::uintptr_t end = wsp + 8 * static_cast<std::uintptr_t>(L::ws_doubles(m, n)) - plant_ws;
  // workspace regions, in order
  std::uintptr_t prev = wsp;
  for (int k = 0; k < L::kNVectors; ++k) {
    const std::uintptr_t a = wsp + 8 * L::n_vector(k, n);
    ++checks;
    if (a < prev) fail("n-vector overlaps", m, n, k);
    in_range(a, 8 * static_cast<std::size_t>(n), wsp, end, 8, "n-vector", m, n);
    prev = a + 8 * n;
  }
  for (int k = 0; k < L::kMVectors; ++k) {
    const std::uintptr_t a = wsp + 8 * L::m_vector(k, m, n);
    ++checks;
    if (a < prev) fail("m-vector overlaps", m, n, k);
    in_range(a, 8 * static_cast<std::size_t>(m), wsp, end, 8, "m-vector", m, n);
    prev = a + 8 * m;
  }
  const std::uintptr_t m64 = wsp + 8 * L::off_m64(m, n);
  ++checks;
  if (m64 < prev) fail("M64 overlaps", m, n, 0);
  in_range(m64, 8 * static_cast<std::size_t>(m) * (m + 1) / 2, wsp, end, 8, "M64", m, n);
  prev = m64 + 8 * static_cast<std::uintptr_t>(m) * (m + 1) / 2;
  const std::uintptr_t xs = wsp + 8 * L::off_xs(m, n);
  ++checks;
  if (xs < prev) fail("xs overlaps", m, n, 0);
  // the factorization writes 256 floats per diagonal block
  in_range(xs, 4 * static_cast<std::size_t>(L::kBlockFloats) * L::diag_blocks(m), wsp, end, 4, "xs", m, n);
  const std::uintptr_t xs_end = xs + 8 * L::xs_doubles(m);
  const std::uintptr_t sd = L::sd_address(xs_end);
  ++checks;
  if (sd < xs_end) fail("sd overlaps xs", m, n, 0);
  in_range(sd, 4 * static_cast<std::size_t>(ldr), wsp, end, 16, "sd", m, n);

  // the FP32 copy of A: its own allocation (256-byte aligned), m x ldr floats
  const std::uintptr_t a32 = 0x40000000;
  const std::uintptr_t a32_end = a32 + 4 * static_cast<std::uintptr_t>(m) * ldr - plant_a32;
  const std::uintptr_t sd_end = sd + 4 * static_cast<std::uintptr_t>(ldr);
  const std::uintptr_t a64 = 0x80000000;
  const std::uintptr_t a64_end = a64 + 8 * static_cast<std::uintptr_t>(m) * ldr - plant_a64;
  // row sums: wavefront per row, lane takes k = 4 lane + 256 t
  for (int i = 0; i < m; ++i)
    for (int lane = 0; lane < 64; ++lane)
      for (int k = 4 * lane; k < ldr; k += 256) {
        in_range(a32 + 4 * (static_cast<std::uintptr_t>(i) * ldr + k), 16, a32, a32_end, 16, "rowsum A", m, n);
        in_range(sd + 4 * static_cast<std::uintptr_t>(k), 16, sd, sd_end, 16, "rowsum sd", m, n);
      }
  // FP64 mode (an LP whose FP32 solves stalled): the blocked Cholesky stages
  // a 16-column panel of up to m rows, and the blocked inverse the products
  // T(I) of the block rows below J, in the LDS matrix region; the diagonal
  // block and its inverse (2 x 256 doubles) in the staging area.
  if (L::fp64_panel_fits(m, L::lds_matrix_bytes(mmax))) {
    const std::uintptr_t reg = L::lds_off_matrix(), reg_end = L::lds_bytes(mmax);
    ++checks;
    if (static_cast<long long>(reg_end - reg) != L::lds_matrix_bytes(mmax)) fail("matrix region size", m, n, 0);
    for (int k0 = 0; k0 < m; k0 += L::kPanel64) {
      const int rows = m - k0;
      in_range(reg, 8 * static_cast<std::size_t>(rows) * L::kPanel64, reg, reg_end, 8, "FP64 panel", m, n);
    }
    const int nbk = (m + L::kPanel64 - 1) / L::kPanel64;
    for (int J = 0; J < nbk; ++J) {
      const int nrow = nbk - 1 - J;
      if (nrow > 0)
        in_range(reg, 8 * static_cast<std::size_t>(nrow) * L::kPanel64 * L::kPanel64, reg, reg_end, 8, "FP64 T", m,
                 n);
    }
    in_range(0, 8 * 2 * L::kPanel64 * L::kPanel64, 0, 8 * static_cast<std::uintptr_t>(L::kStage), 8, "FP64 D, XJ",
             m, n);
  }
  // tiles
  const int nt = (m + 15) / 16;
  const int items = L::assembly_items(nt);
  std::vector<int> seen(static_cast<std::size_t>(nt) * nt, 0);
  const std::uintptr_t lds_end = L::lds_bytes(mmax);
  for (int q = 0; q < items; ++q) {
    int I, J0;
    L::assembly_item(q, I, J0);
    ++checks;
    if (I >= nt || J0 > I || J0 % 4 != 0) {
      fail("work item", m, n, q);
      continue;
    }
    const int nj = (I + 1 - J0) < 4 ? (I + 1 - J0) : 4;
    for (int c = 0; c < nj; ++c) ++seen[static_cast<std::size_t>(I) * nt + J0 + c];
    for (int lane = 0; lane < 64; ++lane) {
      const int lrow = lane & 15, lk = lane >> 4;
      const int ra = 16 * I + lrow;
      const std::uintptr_t pa = a32 + 4 * (static_cast<std::uintptr_t>(ra < m ? ra : 0) * ldr + 4 * lk);
      const std::uintptr_t pd = sd + 4 * static_cast<std::uintptr_t>(4 * lk);
      std::uintptr_t pb[4];
      for (int c = 0; c < 4; ++c) {
        const int cb = 16 * (J0 + c) + lrow;
        const bool vb = c < nj && cb < m;
        pb[c] = a32 + 4 * (static_cast<std::uintptr_t>(vb ? cb : 0) * ldr + 4 * lk);
      }
      for (int k0 = 0; k0 < ldr; k0 += 32)
        for (int h = 0; h < 2; ++h) {
          const std::uintptr_t off = 4 * static_cast<std::uintptr_t>(k0 + 16 * h);
          in_range(pd + off, 16, sd, sd_end, 16, "tile sd", m, n);
          in_range(pa + off, 16, a32, a32_end, 16, "tile A row", m, n);
          for (int c = 0; c < 4; ++c) in_range(pb[c] + off, 16, a32, a32_end, 16, "tile A column", m, n);
        }
      // FP64 mode: 32-byte loads of the row-major FP64 copy of A (m x ldr
      // doubles, its own 256-byte aligned allocation), 16 columns a step
      const std::uintptr_t pa64 = a64 + 8 * (static_cast<std::uintptr_t>(ra < m ? ra : 0) * ldr + 4 * lk);
      for (int k0 = 0; k0 < ldr; k0 += 16) {
        in_range(pa64 + 8 * static_cast<std::uintptr_t>(k0), 32, a64, a64_end, 32, "FP64 tile A row", m, n);
        for (int c = 0; c < 4; ++c) {
          const int cb = 16 * (J0 + c) + lrow;
          const bool vb = c < nj && cb < m;
          const std::uintptr_t p = a64 + 8 * (static_cast<std::uintptr_t>(vb ? cb : 0) * ldr + 4 * lk + k0);
          in_range(p, 32, a64, a64_end, 32, "FP64 tile A column", m, n);
        }
      }
      // stores of the packed matrix: lane column 16 (J0 + c) + lrow, rows 16 I + 4 lk + g
      for (int c = 0; c < nj; ++c) {
        const int j = 16 * (J0 + c) + lrow;
        if (j >= m) continue;
        for (int g = 0; g < 4; ++g) {
          const int i = 16 * I + 4 * lk + g;
          if (i >= m || i < j) continue;
          const long long idx = static_cast<long long>(j) * (2 * m - j + 1) / 2 + (i - j);
          in_range(L::lds_off_matrix() + 4 * static_cast<std::uintptr_t>(idx), 4, L::lds_off_matrix(), lds_end, 4,
                   "packed store", m, n);
        }
      }
    }
  }
  for (int I = 0; I < nt; ++I)
    for (int J = 0; J < nt; ++J) {
      ++checks;
      const int want = J <= I ? 1 : 0;
      if (seen[static_cast<std::size_t>(I) * nt + J] != want) fail("tile coverage", m, n, I * 1000 + J);
    }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1 && argv[1][0] == 'w') plant_ws = 8 + 8;  // past the spare double and the alignment slack
  if (argc > 1 && argv[1][0] == 'a') plant_a32 = 4;
  if (argc > 1 && argv[1][0] == 'd') plant_a64 = 8;
  // static LDS of the kernel (queue slot, LP descriptor, a 16 x 17 float block) is
  // under 2 KiB; the dynamic part of the largest LP must leave room for it
  ++checks;
  if (L::lds_bytes(L::kMaxRows) + 2048 > 160 * 1024) fail("LDS budget", L::kMaxRows, 0, L::lds_bytes(L::kMaxRows));
  const int ns[] = {1, 2, 3, 7, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 300,
                    511, 512, 513, 700, 1000};
  // one batch per column count: every row count 1..272, in order, workspaces
  // one after another from a 256-byte aligned base, as the host lays them out
  for (int n : ns) {
    const std::uintptr_t base = 0x10000000;
    long long ws_off = 0;
    for (int m = 1; m <= L::kMaxRows; ++m) {
      audit_lp(m, n, base + 8 * static_cast<std::uintptr_t>(ws_off), L::kMaxRows);
      ws_off += L::ws_doubles(m, n);
    }
  }
  // the smallest launch that holds each m alone (mmax = m)
  for (int m = 1; m <= L::kMaxRows; ++m) audit_lp(m, 700, 0x10000000, m);
  if (failures) {
    std::printf("%d failures in %lld checks\n", failures, checks);
    return 1;
  }
  std::printf("ok %lld\n", checks);
  return 0;
}
