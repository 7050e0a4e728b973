// Host-only check of the sparse Cholesky's symbolic analysis
// (sovereign_amd/csrc/sparse_symbolic.cpp, compiled in) and of the index
// arithmetic its GPU kernels use (sovereign_amd/csrc/sparse_chol.hip):
//   * structural invariants: rows of every front ascending, its own columns
//     first; child update rows inside the parent's front (maps in range and
//     increasing); fronts and row vectors inside their buffers; children on
//     lower levels than their parents; every lower entry of M in a front;
//   * the multifrontal factorization and both solve sweeps, run
//     sequentially with the kernels' loop structure over the same arrays
//     (every index bounds-checked), on the given matrix and right-hand side.
// Reads from stdin: m nnz, the lower CSC (col_ptr m+1, row_idx nnz, values
// nnz, row >= col), then m right-hand-side values. Writes to stdout one line
// "ok supernodes levels max_front nnz_l front_total replaced" and the m
// solution values, or "FAIL <what>" and exits 1.
// Built with g++ by tests/test_sparse_chol.py.
#include "sparse_symbolic.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

[[noreturn]] void fail(const char* what, long long a = 0, long long b = 0) {
  std::printf("FAIL %s %lld %lld\n", what, a, b);
  std::exit(1);
}

// Bounds-checked element access.
template <typename T>
T& at(std::vector<T>& v, long long i, const char* what) {
  if (i < 0 || i >= static_cast<long long>(v.size())) fail(what, i, static_cast<long long>(v.size()));
  return v[static_cast<std::size_t>(i)];
}

}  // namespace

int main() {
  int m = 0;
  long long nnz = 0;
  if (std::scanf("%d %lld", &m, &nnz) != 2) fail("input");
  std::vector<int32_t> lcp(static_cast<std::size_t>(m) + 1), lri(static_cast<std::size_t>(nnz));
  std::vector<double> val(static_cast<std::size_t>(nnz)), rhs(static_cast<std::size_t>(m));
  for (auto& v : lcp)
    if (std::scanf("%d", &v) != 1) fail("col_ptr");
  for (auto& v : lri)
    if (std::scanf("%d", &v) != 1) fail("row_idx");
  for (auto& v : val)
    if (std::scanf("%lf", &v) != 1) fail("values");
  for (auto& v : rhs)
    if (std::scanf("%lf", &v) != 1) fail("rhs");

  sov::SparseSymbolic y;
  if (y.analyse(m, lcp, lri, 0) != SOV_OK) fail("analyse");
  const int ns = static_cast<int>(y.first.size());
  // structure
  for (int s = 0; s < ns; ++s) {
    const int f = y.front[s], nc = y.ncol[s];
    if (nc < 1 || f < nc) fail("front size", s, f);
    if (y.vec_base[s] < 0 || y.vec_base[s] + f > y.vec_total) fail("vec range", s);
    if (y.front_base[s] < 0 || y.front_base[s] + static_cast<long long>(f) * f > y.front_total) fail("front range", s);
    for (int t = 0; t < f; ++t) {
      const int r = at(y.rows, y.vec_base[s] + t, "rows");
      if (t < nc && r != y.first[s] + t) fail("own columns first", s, t);
      if (t > 0 && r <= y.rows[y.vec_base[s] + t - 1]) fail("rows ascending", s, t);
      if (r < 0 || r >= m) fail("row range", s, r);
    }
    const int p = y.sparent[s];
    if (p >= 0) {
      for (int t = nc; t < f; ++t) {
        const int pos = at(y.maps, y.map_base[s] + (t - nc), "maps");
        if (pos < 0 || pos >= y.front[p]) fail("map range", s, pos);
        if (y.rows[y.vec_base[p] + pos] != y.rows[y.vec_base[s] + t]) fail("map row", s, t);
        if (t > nc && pos <= y.maps[y.map_base[s] + (t - nc - 1)]) fail("map increasing", s, t);
      }
    }
  }
  std::vector<int> lev(static_cast<std::size_t>(ns), -1);
  const int nlev = static_cast<int>(y.level_ptr.size()) - 1;
  for (int l = 0; l < nlev; ++l)
    for (int q = y.level_ptr[l]; q < y.level_ptr[l + 1]; ++q) lev[at(y.level_idx, q, "level_idx")] = l;
  for (int s = 0; s < ns; ++s) {
    if (lev[s] < 0) fail("supernode without level", s);
    for (int q = y.child_ptr[s]; q < y.child_ptr[s + 1]; ++q)
      if (lev[y.child_idx[q]] >= lev[s]) fail("child level", s, y.child_idx[q]);
  }

  // numeric: scatter M into the fronts (the plan kernel's slots)
  std::vector<double> fronts(static_cast<std::size_t>(y.front_total), 0.0), diag0(static_cast<std::size_t>(m), 0.0);
  for (int j = 0; j < m; ++j)
    for (int p = lcp[j]; p < lcp[j + 1]; ++p) {
      const int i = lri[p];
      const long long sl = y.slot(i, j);
      if (sl < 0) fail("entry without slot", i, j);
      at(fronts, sl, "slot") = val[p];
      if (i == j) at(diag0, y.iperm[i], "diag0") = val[p];
    }
  // front_factor_kernel, level by level
  int replaced = 0;
  for (int l = 0; l < nlev; ++l)
    for (int q = y.level_ptr[l]; q < y.level_ptr[l + 1]; ++q) {
      const int s = y.level_idx[q];
      const long long f = y.front[s];
      const int nc = y.ncol[s], c0 = y.first[s];
      const long long G = y.front_base[s];
      for (int ci = y.child_ptr[s]; ci < y.child_ptr[s + 1]; ++ci) {
        const int c = y.child_idx[ci];
        const long long fc = y.front[c], ncc = y.ncol[c], u = fc - ncc;
        const long long U = y.front_base[c] + ncc * fc + ncc;
        for (long long j = 0; j < u; ++j)
          for (long long i = j; i < u; ++i) {
            const long long mi = at(y.maps, y.map_base[c] + i, "map i"), mj = at(y.maps, y.map_base[c] + j, "map j");
            at(fronts, G + mi + mj * f, "extend-add dst") += at(fronts, U + i + j * fc, "extend-add src");
          }
      }
      for (int k = 0; k < nc; ++k) {
        double& dk = at(fronts, G + k + k * f, "pivot");
        const double thr = 1e-13 * std::fabs(at(diag0, c0 + k, "diag0 read"));
        double piv;
        if (std::isfinite(dk) && dk <= thr) {
          piv = 1e64;
          ++replaced;
        } else if (!std::isfinite(dk) || dk < 1e-30) {
          fail("breakdown", s, k);
        } else {
          piv = std::sqrt(dk);
        }
        dk = piv;
        for (long long i = k + 1; i < f; ++i) at(fronts, G + i + k * f, "scale") /= piv;
        for (long long j = k + 1; j < nc; ++j) {
          const double ljk = fronts[G + j + k * f];
          for (long long i = j; i < f; ++i) at(fronts, G + i + j * f, "block update") -= fronts[G + i + k * f] * ljk;
        }
      }
      for (long long j = nc; j < f; ++j)
        for (long long i = j; i < f; ++i) {
          double acc = 0.0;
          for (int k = 0; k < nc; ++k) acc = std::fma(fronts[G + i + k * f], fronts[G + j + k * f], acc);
          at(fronts, G + i + j * f, "schur") -= acc;
        }
    }
  // solves: permute in, forward (leaves first), backward (root first), permute out
  std::vector<double> xp(static_cast<std::size_t>(m)), vf(static_cast<std::size_t>(y.vec_total), 0.0);
  for (int k = 0; k < m; ++k) xp[k] = rhs[at(y.perm, k, "perm")];
  for (int l = 0; l < nlev; ++l)
    for (int q = y.level_ptr[l]; q < y.level_ptr[l + 1]; ++q) {
      const int s = y.level_idx[q];
      const long long f = y.front[s], G = y.front_base[s], W = y.vec_base[s];
      const int nc = y.ncol[s], c0 = y.first[s];
      for (long long p = 0; p < f; ++p) at(vf, W + p, "w init") = p < nc ? at(xp, c0 + p, "xp") : 0.0;
      for (int ci = y.child_ptr[s]; ci < y.child_ptr[s + 1]; ++ci) {
        const int c = y.child_idx[ci];
        const long long ncc = y.ncol[c], u = y.front[c] - ncc;
        for (long long t = 0; t < u; ++t)
          at(vf, W + at(y.maps, y.map_base[c] + t, "fw map"), "fw dst") += at(vf, y.vec_base[c] + ncc + t, "fw src");
      }
      for (int k = 0; k < nc; ++k) {
        const double yk = vf[W + k] / fronts[G + k + k * f];
        for (long long p = k + 1; p < f; ++p) vf[W + p] -= fronts[G + p + k * f] * yk;
        xp[c0 + k] = yk;
      }
    }
  for (int l = nlev - 1; l >= 0; --l)
    for (int q = y.level_ptr[l]; q < y.level_ptr[l + 1]; ++q) {
      const int s = y.level_idx[q];
      const long long f = y.front[s], G = y.front_base[s], W = y.vec_base[s];
      const int nc = y.ncol[s], c0 = y.first[s];
      for (int k = 0; k < nc; ++k) {
        double acc = 0.0;
        for (long long p = nc; p < f; ++p)
          acc += fronts[G + p + k * f] * at(xp, at(y.rows, W + p, "bw rows"), "bw xp");
        vf[W + k] = xp[c0 + k] - acc;
      }
      for (int k = nc - 1; k >= 0; --k) {
        const double xk = vf[W + k] / fronts[G + k + k * f];
        for (int qq = 0; qq < k; ++qq) vf[W + qq] -= fronts[G + k + qq * f] * xk;
        xp[c0 + k] = xk;
      }
    }
  std::vector<double> x(static_cast<std::size_t>(m));
  for (int k = 0; k < m; ++k) x[y.perm[k]] = xp[k];
  std::printf("ok %d %d %d %lld %lld %d\n", ns, nlev, y.max_front, static_cast<long long>(y.nnz_l),
              static_cast<long long>(y.front_total), replaced);
  for (int k = 0; k < m; ++k) std::printf("%.17g\n", x[k]);
  return 0;
}
