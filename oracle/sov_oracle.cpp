// CPU oracle: a restatement of the reference solver's interior-point path.
//
// TEST INFRASTRUCTURE ONLY (see sov_oracle.h). Nothing in the shipped HIP
// library links or calls this file; tests compare the library against it.
//
// Each piece cites the reference function it restates. Where the reference's
// arithmetic order decides rounding (pivot choice, summation order inside a
// column) the same order is kept, so the oracle reproduces the reference to the
// last bit on the same inputs when both are built without FMA contraction.

#include "sov_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <set>
#include <stdexcept>
#include <utility>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

struct Csc {
  int m = 0, n = 0;
  std::vector<int> ptr, idx;
  std::vector<double> val;
};

// ---------------------------------------------------------------------------
// SpMV. gpu_spmv.cpp:36-51 (y = A x, CSC scatter that skips x_j == 0) and
// gpu_spmv.cpp:54-67 (y = A^T x, one dot product per column).
// ---------------------------------------------------------------------------
void csc_times(const Csc& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(static_cast<std::size_t>(A.m), 0.0);
  for (int j = 0; j < A.n; ++j) {
    const double xj = x[j];
    if (xj == 0.0) continue;
    for (int p = A.ptr[j]; p < A.ptr[j + 1]; ++p) y[A.idx[p]] += A.val[p] * xj;
  }
}

void csc_transpose_times(const Csc& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(static_cast<std::size_t>(A.n), 0.0);
  for (int j = 0; j < A.n; ++j) {
    double acc = 0.0;
    for (int p = A.ptr[j]; p < A.ptr[j + 1]; ++p) acc += A.val[p] * x[A.idx[p]];
    y[j] = acc;
  }
}

// ---------------------------------------------------------------------------
// DenseLU (dense_lu.cpp:9-108): right-looking LU with partial pivoting on a
// column-major matrix. The pivot is the FIRST row holding the largest |a(i,k)|
// (strict >), and a largest magnitude below 1e-14 reports singular.
// ---------------------------------------------------------------------------
struct Lu {
  int n = 0;
  bool ok = false;
  std::vector<double> a;
  std::vector<int> perm;

  bool factor(std::vector<double> mat, int order) {
    n = order;
    ok = false;
    a = std::move(mat);
    perm.resize(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) perm[i] = i;
    const std::size_t N = static_cast<std::size_t>(n);
    for (int k = 0; k < n; ++k) {
      double* colk = &a[static_cast<std::size_t>(k) * N];
      int prow = k;
      double pmax = std::abs(colk[k]);
      for (int i = k + 1; i < n; ++i) {
        const double v = std::abs(colk[i]);
        if (v > pmax) {
          pmax = v;
          prow = i;
        }
      }
      if (pmax < 1e-14) return false;
      if (prow != k) {
        for (int j = 0; j < n; ++j) std::swap(a[j * N + k], a[j * N + prow]);
        std::swap(perm[k], perm[prow]);
      }
      const double pivot = colk[k];
      for (int i = k + 1; i < n; ++i) {
        colk[i] /= pivot;
        const double l = colk[i];
        for (int j = k + 1; j < n; ++j) a[j * N + i] -= l * a[j * N + k];
      }
    }
    ok = true;
    return true;
  }

  bool solve(std::vector<double>& x) const {
    if (!ok || static_cast<int>(x.size()) != n) return false;
    const std::size_t N = static_cast<std::size_t>(n);
    std::vector<double> w(N);
    for (int i = 0; i < n; ++i) w[i] = x[perm[i]];
    for (int i = 0; i < n; ++i) {
      double acc = w[i];
      for (int j = 0; j < i; ++j) acc -= a[j * N + i] * w[j];
      w[i] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
      double acc = w[i];
      for (int j = i + 1; j < n; ++j) acc -= a[j * N + i] * w[j];
      w[i] = acc / a[i * N + i];
    }
    x.swap(w);
    return true;
  }

  bool solve_transpose(std::vector<double>& x) const {
    if (!ok || static_cast<int>(x.size()) != n) return false;
    const std::size_t N = static_cast<std::size_t>(n);
    for (int i = 0; i < n; ++i) {  // U^T w = x
      double acc = x[i];
      for (int j = 0; j < i; ++j) acc -= a[i * N + j] * x[j];
      x[i] = acc / a[i * N + i];
    }
    for (int i = n - 1; i >= 0; --i) {  // L^T z = w
      double acc = x[i];
      for (int j = i + 1; j < n; ++j) acc -= a[i * N + j] * x[j];
      x[i] = acc;
    }
    std::vector<double> z(N, 0.0);
    for (int i = 0; i < n; ++i) z[perm[i]] = x[i];
    x.swap(z);
    return true;
  }
};

// ---------------------------------------------------------------------------
// Normal equations, dense: build_normal_eq (interior_point.cpp:74-101).
// Full (both triangles) column-major M = A diag(d) A^T, then +1e-12 on the
// diagonal. Columns with d_j == 0 are skipped.
// ---------------------------------------------------------------------------
void normal_dense(const Csc& A, const std::vector<double>& d, std::vector<double>& M) {
  const std::size_t m = static_cast<std::size_t>(A.m);
  M.assign(m * m, 0.0);
  for (int j = 0; j < A.n; ++j) {
    const double dj = d[j];
    if (dj == 0.0) continue;
    for (int p = A.ptr[j]; p < A.ptr[j + 1]; ++p) {
      const double left = A.val[p] * dj;
      const std::size_t r = static_cast<std::size_t>(A.idx[p]);
      for (int q = A.ptr[j]; q < A.ptr[j + 1]; ++q) {
        M[static_cast<std::size_t>(A.idx[q]) * m + r] += left * A.val[q];
      }
    }
  }
  for (std::size_t i = 0; i < m; ++i) M[i * m + i] += 1e-12;
}

// ---------------------------------------------------------------------------
// Normal equations, sparse pattern and values (sparse_symmetric.cpp:39-169).
// Lower triangle in CSC, rows sorted, diagonal always present.
// ---------------------------------------------------------------------------
struct SymPattern {
  int n = 0;
  std::vector<int> ptr, idx;
};

SymPattern normal_pattern(const Csc& A) {
  SymPattern P;
  P.n = A.m;
  P.ptr.assign(static_cast<std::size_t>(A.m) + 1, 0);
  // Row lists of A (the transpose of its column structure).
  std::vector<int> rstart(static_cast<std::size_t>(A.m) + 1, 0);
  for (int p = 0; p < A.ptr[A.n]; ++p) ++rstart[A.idx[p] + 1];
  for (int r = 0; r < A.m; ++r) rstart[r + 1] += rstart[r];
  std::vector<int> rcols(static_cast<std::size_t>(rstart[A.m]));
  {
    std::vector<int> fill(rstart.begin(), rstart.end() - 1);
    for (int j = 0; j < A.n; ++j)
      for (int p = A.ptr[j]; p < A.ptr[j + 1]; ++p) rcols[fill[A.idx[p]]++] = j;
  }
  std::vector<int> seen(static_cast<std::size_t>(A.m), -1);
  std::vector<int> rows;
  for (int k = 0; k < A.m; ++k) {
    P.ptr[k] = static_cast<int>(P.idx.size());
    rows.clear();
    rows.push_back(k);
    seen[k] = k;
    for (int q = rstart[k]; q < rstart[k + 1]; ++q) {
      const int j = rcols[q];
      for (int p = A.ptr[j]; p < A.ptr[j + 1]; ++p) {
        const int i = A.idx[p];
        if (i > k && seen[i] != k) {
          seen[i] = k;
          rows.push_back(i);
        }
      }
    }
    std::sort(rows.begin(), rows.end());
    P.idx.insert(P.idx.end(), rows.begin(), rows.end());
  }
  P.ptr[A.m] = static_cast<int>(P.idx.size());
  return P;
}

void normal_values(const Csc& A, const std::vector<double>& d, const SymPattern& P,
                   std::vector<double>& out) {
  out.assign(P.idx.size(), 0.0);
  for (int j = 0; j < A.n; ++j) {
    const double dj = d[j];
    if (dj == 0.0) continue;
    for (int a = A.ptr[j]; a < A.ptr[j + 1]; ++a) {
      const int i = A.idx[a];
      const double aij = A.val[a];
      for (int b = A.ptr[j]; b < A.ptr[j + 1]; ++b) {
        const int k = A.idx[b];
        if (i < k) continue;
        const auto first = P.idx.begin() + P.ptr[k];
        const auto last = P.idx.begin() + P.ptr[k + 1];
        const auto hit = std::lower_bound(first, last, i);
        if (hit != last && *hit == i) out[static_cast<std::size_t>(hit - P.idx.begin())] += aij * dj * A.val[b];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Minimum-degree ordering (amd_ordering.cpp:12-130): greedy exact minimum
// degree on an explicit adjacency graph. Nodes with degree above
// max(16, 10 sqrt(n)) are removed and ordered last by (degree, index); the
// rest are eliminated lowest (degree, index) first, each elimination joining
// the eliminated node's neighbours into a clique. The running count of clique
// sizes is compared against the fill limit (0 = none).
// ---------------------------------------------------------------------------
bool min_degree_order(const SymPattern& P, std::size_t fill_limit, std::vector<int>& perm,
                      std::vector<int>& iperm, bool* limit_hit) {
  const int n = P.n;
  *limit_hit = false;
  perm.assign(static_cast<std::size_t>(n), -1);
  iperm.assign(static_cast<std::size_t>(n), -1);
  if (n == 0) return true;
  std::vector<std::vector<int>> nbr(static_cast<std::size_t>(n));
  for (int j = 0; j < n; ++j) {
    for (int p = P.ptr[j]; p < P.ptr[j + 1]; ++p) {
      const int i = P.idx[p];
      if (i == j) continue;
      nbr[j].push_back(i);
      nbr[i].push_back(j);
    }
  }
  for (auto& v : nbr) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  const double dense_cut = std::max(16.0, 10.0 * std::sqrt(static_cast<double>(n)));
  std::vector<char> is_dense(static_cast<std::size_t>(n), 0);
  std::vector<std::pair<int, int>> dense_list;
  for (int i = 0; i < n; ++i) {
    if (static_cast<double>(nbr[i].size()) > dense_cut) {
      is_dense[i] = 1;
      dense_list.emplace_back(static_cast<int>(nbr[i].size()), i);
    }
  }
  if (!dense_list.empty()) {
    for (int i = 0; i < n; ++i) {
      if (is_dense[i]) {
        std::vector<int>().swap(nbr[i]);
        continue;
      }
      auto& v = nbr[i];
      v.erase(std::remove_if(v.begin(), v.end(), [&](int w) { return is_dense[w] != 0; }), v.end());
    }
    std::sort(dense_list.begin(), dense_list.end());
  }
  const int n_sparse = n - static_cast<int>(dense_list.size());
  std::set<std::pair<int, int>> ready;
  for (int i = 0; i < n; ++i)
    if (!is_dense[i]) ready.emplace(static_cast<int>(nbr[i].size()), i);

  const std::size_t nd = dense_list.size();
  std::size_t fill = nd * (nd - (nd > 0 ? 1 : 0)) / 2;
  for (int k = n_sparse; k < n; ++k) {
    const int v = dense_list[static_cast<std::size_t>(k - n_sparse)].second;
    perm[k] = v;
    iperm[v] = k;
  }
  std::vector<int> stamp_of(static_cast<std::size_t>(n), -1);
  int stamp = 0;
  for (int k = 0; k < n_sparse; ++k) {
    const int v = ready.begin()->second;
    ready.erase(ready.begin());
    perm[k] = v;
    iperm[v] = k;
    std::vector<int> clique;
    clique.swap(nbr[v]);
    fill += clique.size();
    if (fill_limit > 0 && fill > fill_limit) {
      *limit_hit = true;
      return false;
    }
    for (int u : clique) {
      auto& list = nbr[u];
      ready.erase({static_cast<int>(list.size()), u});
      ++stamp;
      std::size_t kept = 0;
      for (int w : list) {
        if (w == v) continue;
        stamp_of[w] = stamp;
        list[kept++] = w;
      }
      list.resize(kept);
      stamp_of[u] = stamp;
      for (int w : clique)
        if (stamp_of[w] != stamp) list.push_back(w);
      ready.emplace(static_cast<int>(list.size()), u);
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Up-looking sparse LDL^T (sparse_ldlt.cpp:19-223, Davis TOMS 849): permuted
// upper triangle, elimination tree, row-by-row numeric factorization with
// regularization added to the diagonal, tiny-pivot replacement (d_k <=
// threshold * |diagonal| -> 1e128) and failure on a non-finite or < 1e-30 pivot.
// ---------------------------------------------------------------------------
struct Ldlt {
  int n = 0;
  bool analysed = false, factored = false;
  std::vector<int> perm, iperm;
  std::vector<int> cptr, cidx, cmap;
  std::vector<double> cval;
  std::vector<int> parent, flag, llen, lptr, lidx, stack;
  std::vector<double> lval, dg, work;
  std::size_t in_nnz = 0;
  double tiny = 0.0;
  int replaced = 0;

  bool analyse(const SymPattern& P, std::size_t limit) {
    analysed = factored = false;
    n = P.n;
    in_nnz = P.idx.size();
    if (n > 0) {
      bool hit = false;
      if (!min_degree_order(P, limit, perm, iperm, &hit)) return false;
    }
    cptr.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int j = 0; j < n; ++j)
      for (int p = P.ptr[j]; p < P.ptr[j + 1]; ++p)
        ++cptr[std::max(iperm[P.idx[p]], iperm[j]) + 1];
    for (int k = 0; k < n; ++k) cptr[k + 1] += cptr[k];
    cidx.resize(in_nnz);
    cmap.resize(in_nnz);
    cval.assign(in_nnz, 0.0);
    {
      std::vector<int> fill(cptr.begin(), cptr.end() - 1);
      for (int j = 0; j < n; ++j) {
        for (int p = P.ptr[j]; p < P.ptr[j + 1]; ++p) {
          const int a = iperm[P.idx[p]], b = iperm[j];
          const int q = fill[std::max(a, b)]++;
          cidx[q] = std::min(a, b);
          cmap[p] = q;
        }
      }
    }
    parent.assign(static_cast<std::size_t>(n), -1);
    flag.assign(static_cast<std::size_t>(n), -1);
    llen.assign(static_cast<std::size_t>(n), 0);
    for (int k = 0; k < n; ++k) {
      flag[k] = k;
      for (int p = cptr[k]; p < cptr[k + 1]; ++p) {
        for (int i = cidx[p]; i < k && flag[i] != k; i = parent[i]) {
          if (parent[i] == -1) parent[i] = k;
          ++llen[i];
          flag[i] = k;
        }
      }
    }
    std::size_t lnnz = 0;
    for (int k = 0; k < n; ++k) lnnz += static_cast<std::size_t>(llen[k]);
    if ((limit > 0 && lnnz > limit) || lnnz > static_cast<std::size_t>(std::numeric_limits<int>::max()))
      return false;
    lptr.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int k = 0; k < n; ++k) lptr[k + 1] = lptr[k] + llen[k];
    lidx.assign(static_cast<std::size_t>(lptr[n]), 0);
    lval.assign(lidx.size(), 0.0);
    dg.assign(static_cast<std::size_t>(n), 0.0);
    work.assign(static_cast<std::size_t>(n), 0.0);
    stack.assign(static_cast<std::size_t>(n), 0);
    analysed = true;
    return true;
  }

  bool factor(const std::vector<double>& values, double reg) {
    factored = false;
    if (!analysed || values.size() != in_nnz) return false;
    if (n == 0) {
      factored = true;
      return true;
    }
    std::fill(cval.begin(), cval.end(), 0.0);
    for (std::size_t p = 0; p < values.size(); ++p) cval[cmap[p]] += values[p];
    replaced = 0;
    std::fill(flag.begin(), flag.end(), -1);
    std::fill(llen.begin(), llen.end(), 0);
    std::vector<double>& y = work;
    for (int k = 0; k < n; ++k) {
      y[k] = 0.0;
      int top = n;
      flag[k] = k;
      for (int p = cptr[k]; p < cptr[k + 1]; ++p) {
        int i = cidx[p];
        y[i] += cval[p];
        int len = 0;
        for (; flag[i] != k; i = parent[i]) {
          stack[len++] = i;
          flag[i] = k;
        }
        while (len > 0) stack[--top] = stack[--len];
      }
      const double diag = y[k] + reg;
      double dk = diag;
      y[k] = 0.0;
      for (; top < n; ++top) {
        const int i = stack[top];
        const double yi = y[i];
        y[i] = 0.0;
        const int p0 = lptr[i], p1 = lptr[i] + llen[i];
        for (int p = p0; p < p1; ++p) y[lidx[p]] -= lval[p] * yi;
        const double lki = yi / dg[i];
        dk -= lki * yi;
        lidx[p1] = k;
        lval[p1] = lki;
        ++llen[i];
      }
      if (tiny > 0.0 && std::isfinite(dk) && dk <= tiny * std::abs(diag)) {
        dg[k] = 1e128;
        ++replaced;
        continue;
      }
      if (!std::isfinite(dk) || dk < 1e-30) {
        std::fill(y.begin(), y.end(), 0.0);
        return false;
      }
      dg[k] = dk;
    }
    factored = true;
    return true;
  }

  bool solve(std::vector<double>& x) const {
    if (!factored || static_cast<int>(x.size()) != n) return false;
    if (n == 0) return true;
    std::vector<double> y(static_cast<std::size_t>(n));
    for (int k = 0; k < n; ++k) y[k] = x[perm[k]];
    for (int j = 0; j < n; ++j) {
      const double yj = y[j];
      if (yj == 0.0) continue;
      for (int p = lptr[j]; p < lptr[j + 1]; ++p) y[lidx[p]] -= lval[p] * yj;
    }
    for (int k = 0; k < n; ++k) y[k] /= dg[k];
    for (int j = n - 1; j >= 0; --j) {
      double acc = y[j];
      for (int p = lptr[j]; p < lptr[j + 1]; ++p) acc -= lval[p] * y[lidx[p]];
      y[j] = acc;
    }
    for (int k = 0; k < n; ++k) x[perm[k]] = y[k];
    return true;
  }
};

// ---------------------------------------------------------------------------
// NormalEquations (interior_point.cpp:129-212) with the CPU device policy
// (SOVEREIGN_DEVICE=cpu, gpu_spmv.cpp:123-126): dense LU below 200 rows or
// when the column pairs exceed 4x the dense triangle; otherwise sparse LDL^T
// with a fill limit of 0.9x the dense triangle while m <= 8000 (a failed
// analysis falls back to dense LU), and unlimited sparse above 8000 rows.
// A failed sparse numeric factorization retries dense LU while m <= 8000.
// ---------------------------------------------------------------------------
constexpr double kReg = 1e-12;        // interior_point.cpp:103
constexpr int kDenseMaxRows = 8000;   // interior_point.cpp:106

}  // namespace

struct orc_ne {
  const Csc* A = nullptr;
  Csc own;
  bool sparse = false, last_sparse = false;
  int sparse_count = 0, dense_count = 0, max_replaced = 0;
  SymPattern pat;
  Ldlt ldlt;
  std::vector<double> vals;
  Lu lu;
  double factor_seconds = 0.0;

  void init(int mode) {  // mode: 0 auto, 1 dense, 2 sparse
    const int m = A->m;
    if (mode == 1 || m == 0) return;
    const double tri = 0.5 * static_cast<double>(m) * static_cast<double>(m + 1);
    std::size_t limit = 0;
    if (mode != 2 && m <= kDenseMaxRows) {
      if (m < 200) return;
      double pairs = 0.0;
      for (int j = 0; j < A->n; ++j) {
        const double k = static_cast<double>(A->ptr[j + 1] - A->ptr[j]);
        pairs += 0.5 * k * (k + 1.0);
      }
      if (pairs > 4.0 * tri) return;
      limit = static_cast<std::size_t>(0.9 * tri);
    }
    pat = normal_pattern(*A);
    sparse = ldlt.analyse(pat, limit);
    ldlt.tiny = 1e-13;
  }

  bool factor(const std::vector<double>& d) {
    const auto t0 = Clock::now();
    last_sparse = false;
    if (sparse) {
      normal_values(*A, d, pat, vals);
      if (ldlt.factor(vals, kReg)) {
        last_sparse = true;
        ++sparse_count;
        max_replaced = std::max(max_replaced, ldlt.replaced);
        factor_seconds += seconds_since(t0);
        return true;
      }
      if (A->m > kDenseMaxRows) {
        factor_seconds += seconds_since(t0);
        return false;
      }
    }
    std::vector<double> M;
    normal_dense(*A, d, M);
    ++dense_count;
    const bool ok = lu.factor(std::move(M), A->m);
    factor_seconds += seconds_since(t0);
    return ok;
  }

  bool solve(std::vector<double>& rhs) const { return last_sparse ? ldlt.solve(rhs) : lu.solve(rhs); }
};

namespace {

// ---------------------------------------------------------------------------
// build_ipm_form (interior_point.cpp:224-409).
// ---------------------------------------------------------------------------
struct Form {
  Csc A;
  std::vector<double> b, c, shift, col_scale, row_scale;
  std::vector<int> row_origin;
  int n_struct = 0;
};

Form make_form(const orc_lp_model& mdl, bool scaling) {
  Form F;
  const int nv = mdl.n_vars, m0 = mdl.n_rows;
  F.n_struct = nv;
  F.shift.assign(static_cast<std::size_t>(nv), 0.0);
  std::vector<std::pair<int, double>> upper;  // (variable, ub - lb)
  for (int i = 0; i < nv; ++i) {
    F.shift[i] = mdl.lb[i];
    const double ub = mdl.ub[i];
    if (std::isfinite(ub) && ub < 1e29) {
      const double width = ub - mdl.lb[i];
      if (width < -1e-12) throw std::runtime_error("inconsistent bounds");
      upper.emplace_back(i, width);
    }
  }
  const int mt = m0 + static_cast<int>(upper.size());
  std::vector<double> rhs(static_cast<std::size_t>(mt), 0.0);
  std::vector<std::vector<std::pair<int, double>>> rows(static_cast<std::size_t>(mt));
  F.row_scale.assign(static_cast<std::size_t>(mt), 1.0);
  F.row_origin.assign(static_cast<std::size_t>(mt), -1);
  for (int r = 0; r < m0; ++r) F.row_origin[r] = r;

  // Constraint rows, gathered from the column-major input. The rhs absorbs
  // the lower-bound shift of every variable in the row.
  for (int r = 0; r < m0; ++r) rhs[r] = mdl.rhs[r];
  for (int j = 0; j < nv; ++j) {
    for (int p = mdl.col_ptr[j]; p < mdl.col_ptr[j + 1]; ++p) {
      const int r = mdl.row_idx[p];
      if (r < 0 || r >= m0) throw std::runtime_error("row index out of range");
      const double v = mdl.values[p];
      if (v != 0.0) rows[r].emplace_back(j, v);
      rhs[r] -= v * F.shift[j];
    }
  }
  for (std::size_t k = 0; k < upper.size(); ++k) {
    const int r = m0 + static_cast<int>(k);
    rows[r].emplace_back(upper[k].first, 1.0);
    rhs[r] = upper[k].second;
  }
  std::vector<char> kind(static_cast<std::size_t>(mt), 'L');
  for (int r = 0; r < m0; ++r) {
    const char s = mdl.row_sense[r];
    kind[r] = (s == 'G' || s == 'g') ? 'G' : (s == 'E' || s == 'e') ? 'E' : 'L';
  }
  int slacks = 0;
  for (int r = 0; r < mt; ++r)
    if (kind[r] != 'E') ++slacks;
  const int nt = nv + slacks;
  F.A.m = mt;
  F.A.n = nt;
  F.b = rhs;
  F.c.assign(static_cast<std::size_t>(nt), 0.0);
  const double sgn = mdl.maximize ? -1.0 : 1.0;
  for (int i = 0; i < nv; ++i) F.c[i] = sgn * mdl.obj[i];

  std::vector<std::vector<std::pair<int, double>>> cols(static_cast<std::size_t>(nt));
  for (int r = 0; r < mt; ++r)
    for (const auto& e : rows[r]) cols[e.first].emplace_back(r, e.second);
  int sc = nv;
  for (int r = 0; r < mt; ++r) {
    if (kind[r] == 'L') cols[sc++].emplace_back(r, 1.0);
    else if (kind[r] == 'G') cols[sc++].emplace_back(r, -1.0);
  }
  F.A.ptr.assign(static_cast<std::size_t>(nt) + 1, 0);
  for (int j = 0; j < nt; ++j) {
    auto& e = cols[j];
    std::sort(e.begin(), e.end());
    std::vector<std::pair<int, double>> merged;
    for (const auto& t : e) {
      if (!merged.empty() && merged.back().first == t.first) merged.back().second += t.second;
      else merged.push_back(t);
    }
    for (const auto& t : merged) {
      if (std::abs(t.second) > 0.0) {
        F.A.idx.push_back(t.first);
        F.A.val.push_back(t.second);
      }
    }
    F.A.ptr[j + 1] = static_cast<int>(F.A.val.size());
  }

  F.col_scale.assign(static_cast<std::size_t>(nt), 1.0);
  if (scaling && mt > 0 && nt > 0) {
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> rmax(static_cast<std::size_t>(mt), 0.0);
      for (std::size_t p = 0; p < F.A.val.size(); ++p)
        rmax[F.A.idx[p]] = std::max(rmax[F.A.idx[p]], std::abs(F.A.val[p]));
      std::vector<double> rs(static_cast<std::size_t>(mt), 1.0);
      for (int i = 0; i < mt; ++i) {
        if (rmax[i] > 0.0) rs[i] = 1.0 / std::sqrt(rmax[i]);
        F.row_scale[i] *= rs[i];
        F.b[i] *= rs[i];
      }
      for (std::size_t p = 0; p < F.A.val.size(); ++p) F.A.val[p] *= rs[F.A.idx[p]];
      for (int j = 0; j < nt; ++j) {
        double cmax = 0.0;
        for (int p = F.A.ptr[j]; p < F.A.ptr[j + 1]; ++p) cmax = std::max(cmax, std::abs(F.A.val[p]));
        const double s = cmax > 0.0 ? 1.0 / std::sqrt(cmax) : 1.0;
        for (int p = F.A.ptr[j]; p < F.A.ptr[j + 1]; ++p) F.A.val[p] *= s;
        F.c[j] *= s;
        F.col_scale[j] *= s;
      }
    }
  }
  return F;
}

// ---------------------------------------------------------------------------
// solve_ipm (interior_point.cpp:454-740) with solve_newton (:418-452),
// scaling_diagonal (:411-415) and step_to_bound (:214-222).
// ---------------------------------------------------------------------------
double inf_norm(const std::vector<double>& v) {
  double r = 0.0;
  for (double t : v) r = std::max(r, std::abs(t));
  return r;
}
double dotp(const std::vector<double>& a, const std::vector<double>& b) {
  double r = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i) r += a[i] * b[i];
  return r;
}
double max_step(const std::vector<double>& v, const std::vector<double>& dv) {
  double a = 1.0;
  for (std::size_t i = 0; i < v.size(); ++i)
    if (dv[i] < 0.0) a = std::min(a, -v[i] / dv[i]);
  return a;
}

bool newton(const Csc& A, const orc_ne& ne, const std::vector<double>& d, const std::vector<double>& s,
            const std::vector<double>& rp, const std::vector<double>& rd, const std::vector<double>& rxs,
            std::vector<double>& dx, std::vector<double>& dy, std::vector<double>& ds) {
  const int n = A.n;
  std::vector<double> t(static_cast<std::size_t>(n));
  for (int j = 0; j < n; ++j) t[j] = d[j] * rd[j] - rxs[j] / std::max(s[j], 1e-16);
  std::vector<double> At;
  csc_times(A, t, At);
  dy = rp;
  for (std::size_t i = 0; i < dy.size(); ++i) dy[i] += At[i];
  if (!ne.solve(dy)) return false;
  std::vector<double> Atdy;
  csc_transpose_times(A, dy, Atdy);
  dx.assign(static_cast<std::size_t>(n), 0.0);
  ds.assign(static_cast<std::size_t>(n), 0.0);
  for (int j = 0; j < n; ++j) {
    dx[j] = d[j] * (Atdy[j] - rd[j]) + rxs[j] / std::max(s[j], 1e-16);
    ds[j] = rd[j] - Atdy[j];
  }
  return true;
}

struct Iterate {
  std::vector<double> x, y, s;
  int status = 7, iterations = 0;
  double gap = 0, pres = 0, dres = 0;
};

Iterate mehrotra(const Csc& A, const std::vector<double>& b, const std::vector<double>& c,
                 const orc_ipm_options& opt, orc_ne& ne) {
  const int m = A.m, n = A.n;
  Iterate out;
  std::vector<double> x(static_cast<std::size_t>(n), 1.0), s(static_cast<std::size_t>(n), 1.0),
      y(static_cast<std::size_t>(m), 0.0);
  {  // starting point (:488-506)
    std::vector<double> ones(static_cast<std::size_t>(n), 1.0);
    if (ne.factor(ones)) {
      std::vector<double> Ax;
      csc_times(A, ones, Ax);
      std::vector<double> dy = b;
      for (int i = 0; i < m; ++i) dy[i] -= Ax[i];
      if (ne.solve(dy)) {
        std::vector<double> Atdy;
        csc_transpose_times(A, dy, Atdy);
        for (int j = 0; j < n; ++j) {
          x[j] = std::max(1.0, std::abs(Atdy[j]));
          s[j] = std::max(1.0, std::abs(c[j] - Atdy[j]));
        }
        y = dy;
      }
    }
  }
  const double bnorm = std::max(1.0, inf_norm(b));
  const double cnorm = std::max(1.0, inf_norm(c));
  const double tau = opt.fraction_to_boundary;
  Iterate best;
  best.x = x;
  best.y = y;
  best.s = s;
  double best_metric = std::numeric_limits<double>::infinity();
  double best_gap = best_metric, best_p = best_metric, best_d = best_metric;
  int best_it = 0;

  auto give_best = [&]() {
    out.x = best.x;
    out.y = best.y;
    out.s = best.s;
    out.status = best_p <= opt.feasibility_tol ? 1 : 6;
    out.iterations = best_it;
    out.gap = best_gap;
    out.pres = best_p;
    out.dres = best_d;
  };

  for (int it = 0; it < opt.max_iterations; ++it) {
    std::vector<double> Ax, Aty;
    csc_times(A, x, Ax);
    std::vector<double> rp = b;
    for (int i = 0; i < m; ++i) rp[i] -= Ax[i];
    csc_transpose_times(A, y, Aty);
    std::vector<double> rd = c;
    for (int j = 0; j < n; ++j) rd[j] -= Aty[j] + s[j];
    const double xs = dotp(x, s);
    const double mu = xs / static_cast<double>(n);
    const double pres = inf_norm(rp) / bnorm;
    const double dres = inf_norm(rd) / cnorm;
    const double gap = std::abs(xs) / (1.0 + std::abs(dotp(c, x)));
    const double metric = std::max(pres, std::max(dres, gap));
    if (metric < best_metric) {
      best_metric = metric;
      best.x = x;
      best.y = y;
      best.s = s;
      best_gap = gap;
      best_p = pres;
      best_d = dres;
      best_it = it + 1;
    }
    if (gap < best_gap) best_gap = gap;
    if (pres < opt.feasibility_tol && dres < opt.feasibility_tol && gap < opt.optimality_tol) {
      out.x = x;
      out.y = y;
      out.s = s;
      out.status = 0;
      out.iterations = it + 1;
      out.gap = gap;
      out.pres = pres;
      out.dres = dres;
      return out;
    }
    std::vector<double> rxs(static_cast<std::size_t>(n));
    for (int j = 0; j < n; ++j) rxs[j] = -x[j] * s[j];
    std::vector<double> d(static_cast<std::size_t>(n));
    for (int j = 0; j < n; ++j) d[j] = std::max(x[j], 1e-16) / std::max(s[j], 1e-16);
    std::vector<double> dxa, dya, dsa;
    if (!ne.factor(d) || !newton(A, ne, d, s, rp, rd, rxs, dxa, dya, dsa)) {
      give_best();
      return out;
    }
    const double ap_aff = max_step(x, dxa);
    const double ad_aff = max_step(s, dsa);
    double mu_aff = 0.0;
    for (int j = 0; j < n; ++j) mu_aff += (x[j] + ap_aff * dxa[j]) * (s[j] + ad_aff * dsa[j]);
    mu_aff /= static_cast<double>(n);
    const double sigma = mu > 0.0 ? std::min(1.0, std::pow(mu_aff / mu, 3.0)) : 0.0;
    for (int j = 0; j < n; ++j) rxs[j] = -x[j] * s[j] - dxa[j] * dsa[j] + sigma * mu;
    std::vector<double> dx, dy, ds;
    if (!newton(A, ne, d, s, rp, rd, rxs, dx, dy, ds)) {
      give_best();
      return out;
    }
    const double ap = std::min(1.0, tau * max_step(x, dx));
    const double ad = std::min(1.0, tau * max_step(s, ds));
    for (int j = 0; j < n; ++j) {
      x[j] += ap * dx[j];
      s[j] += ad * ds[j];
      x[j] = std::max(x[j], 1e-14);
      s[j] = std::max(s[j], 1e-14);
    }
    for (int i = 0; i < m; ++i) y[i] += ad * dy[i];
  }
  give_best();
  return out;
}

Csc csc_from(int m, int n, const int32_t* cp, const int32_t* ri, const double* v) {
  Csc A;
  A.m = m;
  A.n = n;
  A.ptr.assign(cp, cp + n + 1);
  const int nnz = cp[n];
  A.idx.assign(ri, ri + nnz);
  A.val.assign(v, v + nnz);
  return A;
}

void fill_result(orc_ipm_result* r, const Iterate& it, const orc_ne& ne) {
  r->status = it.status;
  r->iterations = it.iterations;
  r->duality_gap = it.gap;
  r->primal_residual = it.pres;
  r->dual_residual = it.dres;
  r->factorizations = ne.sparse_count + ne.dense_count;
  r->used_sparse = ne.last_sparse ? 1 : 0;
  r->factor_seconds = ne.factor_seconds;
}

}  // namespace

struct orc_form {
  Form f;
};

extern "C" {

void orc_default_options(orc_ipm_options* o) {
  o->max_iterations = 250;
  o->feasibility_tol = 1e-9;
  o->optimality_tol = 1e-9;
  o->fraction_to_boundary = 0.999;
  o->enable_scaling = 1;
  o->normal_mode = 0;
}

int orc_lp_solve(const orc_lp_model* model, const orc_ipm_options* opt, orc_ipm_result* res,
                 double* primal, double* row_dual, double* var_dual) {
  const auto t0 = Clock::now();
  *res = orc_ipm_result{};
  res->status = 7;
  try {
    Form F = make_form(*model, opt->enable_scaling != 0);
    if (F.A.n <= 0) return -1;
    if (F.A.m <= 0) {  // interior_point.cpp:464-478
      double obj = model->obj_constant;
      for (int i = 0; i < F.n_struct; ++i) {
        if (primal) primal[i] = F.shift[i];
        obj += model->obj[i] * F.shift[i];
      }
      res->status = 0;
      res->objective = obj;
      res->total_seconds = seconds_since(t0);
      return 0;
    }
    orc_ne ne;
    ne.A = &F.A;
    ne.init(opt->normal_mode);
    const Iterate it = mehrotra(F.A, F.b, F.c, *opt, ne);
    fill_result(res, it, ne);
    // publish_iterate (:521-569): undo the column scaling and the bound shift.
    double obj = model->obj_constant;
    for (int j = 0; j < F.n_struct; ++j) {
      const double v = F.shift[j] + F.col_scale[j] * it.x[j];
      if (primal) primal[j] = v;
      if (var_dual) var_dual[j] = it.s[j] / F.col_scale[j];
      obj += model->obj[j] * v;
    }
    if (row_dual)
      for (int i = 0; i < F.A.m; ++i)
        if (F.row_origin[i] >= 0) row_dual[F.row_origin[i]] = F.row_scale[i] * it.y[i];
    res->objective = obj;
    res->total_seconds = seconds_since(t0);
    return 0;
  } catch (const std::exception&) {
    res->status = 7;
    return -1;
  }
}

orc_form* orc_form_build(const orc_lp_model* model, int enable_scaling) {
  try {
    auto* h = new orc_form;
    h->f = make_form(*model, enable_scaling != 0);
    return h;
  } catch (const std::exception&) {
    return nullptr;
  }
}

void orc_form_dims(const orc_form* h, int32_t* m, int32_t* n, int32_t* nnz, int32_t* ns) {
  *m = h->f.A.m;
  *n = h->f.A.n;
  *nnz = static_cast<int32_t>(h->f.A.val.size());
  *ns = h->f.n_struct;
}

void orc_form_get(const orc_form* h, int32_t* cp, int32_t* ri, double* v, double* b, double* c,
                  double* cs, double* rs, double* sh) {
  const Form& F = h->f;
  std::copy(F.A.ptr.begin(), F.A.ptr.end(), cp);
  std::copy(F.A.idx.begin(), F.A.idx.end(), ri);
  std::copy(F.A.val.begin(), F.A.val.end(), v);
  std::copy(F.b.begin(), F.b.end(), b);
  std::copy(F.c.begin(), F.c.end(), c);
  std::copy(F.col_scale.begin(), F.col_scale.end(), cs);
  std::copy(F.row_scale.begin(), F.row_scale.end(), rs);
  std::copy(F.shift.begin(), F.shift.end(), sh);
}

void orc_form_free(orc_form* h) { delete h; }

int orc_solve_ipm_form(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri, const double* v,
                       const double* b, const double* c, const orc_ipm_options* opt,
                       orc_ipm_result* res, double* x, double* y, double* s) {
  const auto t0 = Clock::now();
  *res = orc_ipm_result{};
  res->status = 7;
  if (n <= 0 || m <= 0) return -1;
  const Csc A = csc_from(m, n, cp, ri, v);
  orc_ne ne;
  ne.A = &A;
  ne.init(opt->normal_mode);
  const std::vector<double> bv(b, b + m), cv(c, c + n);
  const Iterate it = mehrotra(A, bv, cv, *opt, ne);
  fill_result(res, it, ne);
  res->objective = dotp(cv, it.x);
  if (x) std::copy(it.x.begin(), it.x.end(), x);
  if (y) std::copy(it.y.begin(), it.y.end(), y);
  if (s) std::copy(it.s.begin(), it.s.end(), s);
  res->total_seconds = seconds_since(t0);
  return 0;
}

void orc_spmv_csc(int32_t nrows, int32_t ncols, const int32_t* cp, const int32_t* ri, const double* v,
                  const double* x, double* y, int transpose) {
  const Csc A = csc_from(nrows, ncols, cp, ri, v);
  std::vector<double> xv(x, x + (transpose ? nrows : ncols)), yv;
  if (transpose) csc_transpose_times(A, xv, yv);
  else csc_times(A, xv, yv);
  std::copy(yv.begin(), yv.end(), y);
}

int orc_dense_lu_factor(int32_t n, double* a, int32_t* piv) {
  Lu lu;
  const std::size_t N = static_cast<std::size_t>(n);
  const bool ok = lu.factor(std::vector<double>(a, a + N * N), n);
  std::copy(lu.a.begin(), lu.a.end(), a);
  for (int i = 0; i < n; ++i) piv[i] = lu.perm[i];
  return ok ? 0 : 1;
}

void orc_dense_lu_solve(int32_t n, const double* a, const int32_t* piv, double* x) {
  Lu lu;
  lu.n = n;
  lu.ok = true;
  lu.a.assign(a, a + static_cast<std::size_t>(n) * n);
  lu.perm.assign(piv, piv + n);
  std::vector<double> v(x, x + n);
  lu.solve(v);
  std::copy(v.begin(), v.end(), x);
}

void orc_dense_lu_solve_transpose(int32_t n, const double* a, const int32_t* piv, double* x) {
  Lu lu;
  lu.n = n;
  lu.ok = true;
  lu.a.assign(a, a + static_cast<std::size_t>(n) * n);
  lu.perm.assign(piv, piv + n);
  std::vector<double> v(x, x + n);
  lu.solve_transpose(v);
  std::copy(v.begin(), v.end(), x);
}

void orc_normal_eq_dense(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri, const double* v,
                         const double* d, double* M) {
  const Csc A = csc_from(m, n, cp, ri, v);
  std::vector<double> dv(d, d + n), out;
  normal_dense(A, dv, out);
  std::copy(out.begin(), out.end(), M);
}

int64_t orc_normal_eq_pattern_nnz(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri) {
  Csc A;
  A.m = m;
  A.n = n;
  A.ptr.assign(cp, cp + n + 1);
  A.idx.assign(ri, ri + cp[n]);
  return static_cast<int64_t>(normal_pattern(A).idx.size());
}

void orc_normal_eq_pattern(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri, int32_t* pcp,
                           int32_t* pri) {
  Csc A;
  A.m = m;
  A.n = n;
  A.ptr.assign(cp, cp + n + 1);
  A.idx.assign(ri, ri + cp[n]);
  const SymPattern P = normal_pattern(A);
  std::copy(P.ptr.begin(), P.ptr.end(), pcp);
  std::copy(P.idx.begin(), P.idx.end(), pri);
}

orc_ne* orc_ne_create(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri, const double* v,
                      int32_t mode) {
  auto* h = new orc_ne;
  h->own = csc_from(m, n, cp, ri, v);
  h->A = &h->own;
  h->init(mode);
  return h;
}

int orc_ne_factor(orc_ne* h, const double* d) {
  return h->factor(std::vector<double>(d, d + h->A->n)) ? 1 : 0;
}

int orc_ne_solve(const orc_ne* h, double* rhs) {
  std::vector<double> v(rhs, rhs + h->A->m);
  if (!h->solve(v)) return 0;
  std::copy(v.begin(), v.end(), rhs);
  return 1;
}

int orc_ne_is_sparse(const orc_ne* h) { return h->sparse ? 1 : 0; }
int64_t orc_ne_factor_nnz(const orc_ne* h) {
  return h->sparse ? static_cast<int64_t>(h->ldlt.lidx.size()) : 0;
}
int32_t orc_ne_replaced_pivots(const orc_ne* h) { return h->max_replaced; }
void orc_ne_free(orc_ne* h) { delete h; }

int orc_ldlt_factor_solve(int32_t n, const int32_t* pcp, const int32_t* pri, const double* vals, double reg,
                          double tiny, double* rhs, int32_t* replaced, int64_t* factor_nnz, int32_t* perm) {
  SymPattern P;
  P.n = n;
  P.ptr.assign(pcp, pcp + n + 1);
  P.idx.assign(pri, pri + pcp[n]);
  Ldlt f;
  if (!f.analyse(P, 0)) return -1;
  if (factor_nnz) *factor_nnz = static_cast<int64_t>(f.lidx.size());
  if (perm) std::copy(f.perm.begin(), f.perm.end(), perm);
  f.tiny = tiny;
  if (!f.factor(std::vector<double>(vals, vals + pcp[n]), reg)) return 1;
  if (replaced) *replaced = f.replaced;
  if (rhs) {
    std::vector<double> x(rhs, rhs + n);
    if (!f.solve(x)) return 1;
    std::copy(x.begin(), x.end(), rhs);
  }
  return 0;
}

int orc_amd_order(int32_t n, const int32_t* pcp, const int32_t* pri, int64_t fill_limit, int32_t* perm,
                  int32_t* iperm) {
  SymPattern P;
  P.n = n;
  P.ptr.assign(pcp, pcp + n + 1);
  P.idx.assign(pri, pri + pcp[n]);
  std::vector<int> pv, ipv;
  bool hit = false;
  const bool ok = min_degree_order(P, static_cast<std::size_t>(fill_limit), pv, ipv, &hit);
  std::copy(pv.begin(), pv.end(), perm);
  std::copy(ipv.begin(), ipv.end(), iperm);
  return ok ? 1 : 0;
}

}  // extern "C"
