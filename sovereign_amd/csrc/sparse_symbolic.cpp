// Host side of the sparse Cholesky: minimum-degree ordering and the symbolic
// analysis of the supernodal multifrontal factorization (see sparse_chol.hpp).
#include "sparse_symbolic.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <queue>
#include <utility>

namespace sov {

bool min_degree_order(int n, const int32_t* col_ptr, const int32_t* row_idx, int64_t fill_limit,
                      std::vector<int>& perm, std::vector<int>& iperm, int64_t* predicted_nnz) {
  perm.assign(static_cast<std::size_t>(n), -1);
  iperm.assign(static_cast<std::size_t>(n), -1);
  if (predicted_nnz) *predicted_nnz = 0;
  if (n <= 0) return true;
  // Neighbour lists in both directions, without self loops or repeats.
  std::vector<std::vector<int>> nb(static_cast<std::size_t>(n));
  for (int j = 0; j < n; ++j)
    for (int p = col_ptr[j]; p < col_ptr[j + 1]; ++p) {
      const int i = row_idx[p];
      if (i == j) continue;
      nb[j].push_back(i);
      nb[i].push_back(j);
    }
  for (auto& l : nb) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  // Nodes of high degree leave the graph and come last, by (degree, index).
  const double dense_degree = std::max(16.0, 10.0 * std::sqrt(static_cast<double>(n)));
  std::vector<char> late(static_cast<std::size_t>(n), 0);
  std::vector<std::pair<int, int>> late_nodes;
  for (int i = 0; i < n; ++i)
    if (static_cast<double>(nb[i].size()) > dense_degree) {
      late[i] = 1;
      late_nodes.push_back({static_cast<int>(nb[i].size()), i});
    }
  if (!late_nodes.empty()) {
    for (int i = 0; i < n; ++i) {
      if (late[i]) {
        std::vector<int>().swap(nb[i]);
        continue;
      }
      auto& l = nb[i];
      l.erase(std::remove_if(l.begin(), l.end(), [&](int w) { return late[w] != 0; }), l.end());
    }
    std::sort(late_nodes.begin(), late_nodes.end());
  }
  const int n_early = n - static_cast<int>(late_nodes.size());
  for (std::size_t t = 0; t < late_nodes.size(); ++t) {
    perm[n_early + static_cast<int>(t)] = late_nodes[t].second;
    iperm[late_nodes[t].second] = n_early + static_cast<int>(t);
  }
  const int64_t nl = static_cast<int64_t>(late_nodes.size());
  int64_t fill = nl > 0 ? nl * (nl - 1) / 2 : 0;
  // Lazy heap of (degree, node): a node's entry is current while its degree
  // equals the length of its list; older entries are skipped when popped.
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<std::pair<int, int>>> heap;
  for (int i = 0; i < n; ++i)
    if (!late[i]) heap.push({static_cast<int>(nb[i].size()), i});
  std::vector<char> gone(static_cast<std::size_t>(n), 0);
  std::vector<int> seen(static_cast<std::size_t>(n), -1);
  int stamp = 0;
  for (int k = 0; k < n_early; ++k) {
    int v = -1;
    while (!heap.empty()) {
      const std::pair<int, int> top = heap.top();
      heap.pop();
      if (!gone[top.second] && top.first == static_cast<int>(nb[top.second].size())) {
        v = top.second;
        break;
      }
    }
    if (v < 0) return false;  // cannot happen: every remaining node has a current entry
    gone[v] = 1;
    perm[k] = v;
    iperm[v] = k;
    std::vector<int> clique;
    clique.swap(nb[v]);
    fill += static_cast<int64_t>(clique.size());
    if (fill_limit > 0 && fill > fill_limit) {
      if (predicted_nnz) *predicted_nnz = fill;
      return false;
    }
    // v's neighbours become a clique: each drops v and gains the others.
    for (int u : clique) {
      std::vector<int>& l = nb[u];
      ++stamp;
      std::size_t keep = 0;
      for (int w : l) {
        if (w == v) continue;
        seen[w] = stamp;
        l[keep++] = w;
      }
      l.resize(keep);
      seen[u] = stamp;
      for (int w : clique)
        if (seen[w] != stamp) l.push_back(w);
      heap.push({static_cast<int>(l.size()), u});
    }
  }
  if (predicted_nnz) *predicted_nnz = fill;
  return true;
}

int SparseSymbolic::analyse(int mm, const std::vector<int32_t>& lcp, const std::vector<int32_t>& lri,
                            int64_t fill_limit) {
  m = mm;
  std::vector<int> p0, ip0;
  int64_t predicted = 0;
  if (!min_degree_order(m, lcp.data(), lri.data(), fill_limit, p0, ip0, &predicted)) return SOV_SINGULAR;

  // Upper-triangle column lists of P M P^T: entry (i, j) goes to column
  // max(pi, pj) with row min(pi, pj).
  auto upper_lists = [&](const std::vector<int>& ip, std::vector<int>& uptr, std::vector<int>& uidx) {
    uptr.assign(static_cast<std::size_t>(m) + 1, 0);
    for (int j = 0; j < m; ++j)
      for (int p = lcp[j]; p < lcp[j + 1]; ++p) {
        const int a = ip[lri[p]], b = ip[j];
        if (a != b) ++uptr[std::max(a, b) + 1];
      }
    for (int k = 0; k < m; ++k) uptr[k + 1] += uptr[k];
    uidx.assign(static_cast<std::size_t>(uptr[m]), 0);
    std::vector<int> next(uptr.begin(), uptr.end() - 1);
    for (int j = 0; j < m; ++j)
      for (int p = lcp[j]; p < lcp[j + 1]; ++p) {
        const int a = ip[lri[p]], b = ip[j];
        if (a != b) uidx[next[std::max(a, b)]++] = std::min(a, b);
      }
  };
  std::vector<int> uptr, uidx;
  upper_lists(ip0, uptr, uidx);
  // Elimination tree (ancestor links with path compression).
  std::vector<int> par(static_cast<std::size_t>(m), -1), anc(static_cast<std::size_t>(m), -1);
  for (int k = 0; k < m; ++k)
    for (int p = uptr[k]; p < uptr[k + 1]; ++p)
      for (int i = uidx[p]; i != -1 && i < k;) {
        const int up = anc[i];
        anc[i] = k;
        if (up == -1) par[i] = k;
        i = up;
      }
  // Postorder (children in increasing order, roots in increasing order):
  // an equivalent ordering whose subtrees are contiguous column ranges.
  std::vector<int> head(static_cast<std::size_t>(m), -1), nxt(static_cast<std::size_t>(m), -1);
  for (int i = m - 1; i >= 0; --i)
    if (par[i] >= 0) {
      nxt[i] = head[par[i]];
      head[par[i]] = i;
    }
  std::vector<int> post;
  post.reserve(static_cast<std::size_t>(m));
  std::vector<int> stack;
  for (int r = 0; r < m; ++r) {
    if (par[r] != -1) continue;
    stack.push_back(r);
    while (!stack.empty()) {
      const int v = stack.back();
      if (head[v] != -1) {  // descend into the next child
        const int c = head[v];
        head[v] = nxt[c];
        stack.push_back(c);
      } else {
        post.push_back(v);
        stack.pop_back();
      }
    }
  }
  std::vector<int> newlabel(static_cast<std::size_t>(m));
  for (int k = 0; k < m; ++k) newlabel[post[k]] = k;
  perm.assign(static_cast<std::size_t>(m), 0);
  iperm.assign(static_cast<std::size_t>(m), 0);
  for (int k = 0; k < m; ++k) {
    perm[k] = p0[post[k]];
    iperm[perm[k]] = k;
  }
  parent.assign(static_cast<std::size_t>(m), -1);
  for (int i = 0; i < m; ++i) parent[newlabel[i]] = par[i] < 0 ? -1 : newlabel[par[i]];
  upper_lists(iperm, uptr, uidx);
  // Column counts: row k of L is the union of the tree paths from the
  // entries of upper column k up to k.
  std::vector<int> below(static_cast<std::size_t>(m), 0), flag(static_cast<std::size_t>(m), -1);
  for (int k = 0; k < m; ++k) {
    flag[k] = k;
    for (int p = uptr[k]; p < uptr[k + 1]; ++p)
      for (int i = uidx[p]; flag[i] != k; i = parent[i]) {
        ++below[i];
        flag[i] = k;
      }
  }
  colcount.assign(static_cast<std::size_t>(m), 0);
  nnz_l = 0;
  flops = 0.0;
  for (int k = 0; k < m; ++k) {
    colcount[k] = below[k] + 1;
    nnz_l += colcount[k];
    flops += static_cast<double>(below[k]) * below[k];
  }
  // Fundamental supernodes: column j joins j - 1's supernode when j is its
  // only child's parent and the column structures nest exactly.
  std::vector<int> nchild(static_cast<std::size_t>(m), 0);
  for (int k = 0; k < m; ++k)
    if (parent[k] >= 0) ++nchild[parent[k]];
  first.clear();
  ncol.clear();
  snode_of.assign(static_cast<std::size_t>(m), -1);
  for (int j = 0; j < m; ++j) {
    const bool join = j > 0 && parent[j - 1] == j && colcount[j - 1] == colcount[j] + 1 && nchild[j] == 1;
    if (!join) {
      first.push_back(j);
      ncol.push_back(0);
    }
    ++ncol.back();
    snode_of[j] = static_cast<int>(first.size()) - 1;
  }
  const int ns = static_cast<int>(first.size());
  sparent.assign(static_cast<std::size_t>(ns), -1);
  for (int s = 0; s < ns; ++s) {
    const int last = first[s] + ncol[s] - 1;
    sparent[s] = parent[last] < 0 ? -1 : snode_of[parent[last]];
  }
  child_ptr.assign(static_cast<std::size_t>(ns) + 1, 0);
  for (int s = 0; s < ns; ++s)
    if (sparent[s] >= 0) ++child_ptr[sparent[s] + 1];
  for (int s = 0; s < ns; ++s) child_ptr[s + 1] += child_ptr[s];
  child_idx.assign(static_cast<std::size_t>(child_ptr[ns]), 0);
  {
    std::vector<int> fillp(child_ptr.begin(), child_ptr.end() - 1);
    for (int s = 0; s < ns; ++s)
      if (sparent[s] >= 0) child_idx[fillp[sparent[s]]++] = s;
  }
  // Lower column lists of the permuted matrix (rows below the diagonal).
  std::vector<int> lptr(static_cast<std::size_t>(m) + 1, 0), lidx;
  for (int k = 0; k < m; ++k)
    for (int p = uptr[k]; p < uptr[k + 1]; ++p) ++lptr[uidx[p] + 1];
  for (int k = 0; k < m; ++k) lptr[k + 1] += lptr[k];
  lidx.assign(static_cast<std::size_t>(lptr[m]), 0);
  {
    std::vector<int> fillp(lptr.begin(), lptr.end() - 1);
    for (int k = 0; k < m; ++k)
      for (int p = uptr[k]; p < uptr[k + 1]; ++p) lidx[fillp[uidx[p]]++] = k;
  }
  // Row structure of each front: its own columns, then the rows below them
  // from M's columns and from the children's fronts, ascending.
  front.assign(static_cast<std::size_t>(ns), 0);
  vec_base.assign(static_cast<std::size_t>(ns), 0);
  rows.clear();
  std::vector<int> mark(static_cast<std::size_t>(m), -1), tail;
  for (int s = 0; s < ns; ++s) {
    const int c0 = first[s], last = c0 + ncol[s] - 1;
    vec_base[s] = static_cast<int64_t>(rows.size());
    for (int j = c0; j <= last; ++j) {
      rows.push_back(j);
      mark[j] = s;
    }
    tail.clear();
    for (int j = c0; j <= last; ++j)
      for (int p = lptr[j]; p < lptr[j + 1]; ++p) {
        const int r = lidx[p];
        if (r > last && mark[r] != s) {
          mark[r] = s;
          tail.push_back(r);
        }
      }
    for (int q = child_ptr[s]; q < child_ptr[s + 1]; ++q) {
      const int c = child_idx[q];
      for (int t = ncol[c]; t < front[c]; ++t) {
        const int r = rows[vec_base[c] + t];
        if (r > last && mark[r] != s) {
          mark[r] = s;
          tail.push_back(r);
        }
      }
    }
    std::sort(tail.begin(), tail.end());
    rows.insert(rows.end(), tail.begin(), tail.end());
    front[s] = ncol[s] + static_cast<int>(tail.size());
  }
  vec_total = static_cast<int64_t>(rows.size());
  // Where each update row of a front sits in its parent's front.
  map_base.assign(static_cast<std::size_t>(ns), 0);
  maps.clear();
  for (int s = 0; s < ns; ++s) {
    map_base[s] = static_cast<int64_t>(maps.size());
    const int p = sparent[s];
    if (p < 0) continue;
    const int* pr = rows.data() + vec_base[p];
    for (int t = ncol[s]; t < front[s]; ++t) {
      const int r = rows[vec_base[s] + t];
      const int* at = std::lower_bound(pr, pr + front[p], r);
      if (at == pr + front[p] || *at != r) return SOV_ERR_ARG;  // an update row outside the parent: a bug
      maps.push_back(static_cast<int>(at - pr));
    }
  }
  // Fronts in the buffer, and levels of the supernodal tree.
  front_base.assign(static_cast<std::size_t>(ns), 0);
  front_total = 0;
  max_front = 0;
  for (int s = 0; s < ns; ++s) {
    front_base[s] = front_total;
    front_total += static_cast<int64_t>(front[s]) * front[s];
    max_front = std::max(max_front, front[s]);
  }
  std::vector<int> level(static_cast<std::size_t>(ns), 0);
  int nlev = 0;
  for (int s = 0; s < ns; ++s) {  // children come before their parents
    for (int q = child_ptr[s]; q < child_ptr[s + 1]; ++q) level[s] = std::max(level[s], level[child_idx[q]] + 1);
    nlev = std::max(nlev, level[s] + 1);
  }
  level_ptr.assign(static_cast<std::size_t>(nlev) + 1, 0);
  for (int s = 0; s < ns; ++s) ++level_ptr[level[s] + 1];
  for (int l = 0; l < nlev; ++l) level_ptr[l + 1] += level_ptr[l];
  level_idx.assign(static_cast<std::size_t>(ns), 0);
  {
    std::vector<int> fillp(level_ptr.begin(), level_ptr.end() - 1);
    for (int s = 0; s < ns; ++s) level_idx[fillp[level[s]]++] = s;
  }
  return SOV_OK;
}

int64_t SparseSymbolic::slot(int row, int col) const {
  const int a = iperm[row], b = iperm[col];
  const int r = std::max(a, b), c = std::min(a, b);
  const int s = snode_of[c];
  const int* rs = rows.data() + vec_base[s];
  const int* at = std::lower_bound(rs, rs + front[s], r);
  if (at == rs + front[s] || *at != r) return -1;
  return front_base[s] + static_cast<int64_t>(c - first[s]) * front[s] + (at - rs);
}

}  // namespace sov
