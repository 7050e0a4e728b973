// Symbolic side of the sparse Cholesky (sparse_chol.hpp): the ordering and
// the supernodal analysis, plain C++ (no HIP), so that host-only builds can
// use it too.
#pragma once

#include "sovereign_hip.h"

#include <cstdint>
#include <vector>

namespace sov {

// Minimum-degree ordering (AMDOrdering::compute, amd_ordering.cpp:12-130):
// greedy exact minimum degree on the explicit graph of a symmetric pattern
// (CSC, either triangle or both; the diagonal is ignored), the node of least
// (degree, index) first; nodes of degree above max(16, 10 sqrt(n)) are taken
// out and ordered last by (degree, index). perm[k] is the node eliminated
// k-th, iperm its inverse. predicted_nnz receives the fill the elimination
// predicts (a lower bound on nnz(L)). Returns false, with perm incomplete,
// when fill_limit > 0 and the prediction passes it.
bool min_degree_order(int n, const int32_t* col_ptr, const int32_t* row_idx, int64_t fill_limit,
                      std::vector<int>& perm, std::vector<int>& iperm, int64_t* predicted_nnz);

struct SparseSymbolic {
  int m = 0;
  std::vector<int> perm, iperm;  // elimination order (postordered): perm[k] = original index
  std::vector<int> parent;       // elimination tree on the permuted columns (-1: root)
  std::vector<int> colcount;     // entries of column k of L, diagonal included
  int64_t nnz_l = 0;             // sum of colcount
  double flops = 0.0;            // sum of colcount^2 (the reference's factor_flops)
  // supernodes (contiguous permuted columns first[s] .. first[s] + ncol[s] - 1)
  std::vector<int> first, ncol, front;  // front: rows of the supernode's front (its columns first)
  std::vector<int64_t> front_base;      // offset of the front in the front buffer (doubles)
  std::vector<int64_t> vec_base;        // offset of the front's rows in the row buffer
  std::vector<int> rows;                // permuted row indices of every front, ascending
  std::vector<int> sparent;             // supernodal tree
  std::vector<int> child_ptr, child_idx;
  std::vector<int64_t> map_base;        // per supernode: where its update rows' parent positions start
  std::vector<int> maps;                // parent front position of each update row
  std::vector<int> level_ptr, level_idx;  // supernodes of each level, leaves first
  std::vector<int> snode_of;              // supernode of each permuted column
  int64_t front_total = 0, vec_total = 0;
  int max_front = 0;

  // Runs the symbolic analysis of the lower pattern (CSC of the m x m
  // matrix, row >= col, diagonal present or not). Returns SOV_OK, or
  // SOV_SINGULAR when the ordering's fill prediction passes fill_limit.
  int analyse(int m, const std::vector<int32_t>& lcp, const std::vector<int32_t>& lri, int64_t fill_limit);
  // Slot of the entry (row, col) of the original matrix in the front buffer.
  int64_t slot(int row, int col) const;
};

}  // namespace sov
