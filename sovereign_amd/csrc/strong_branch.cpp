// Strong branching and a best-bound branch-and-bound over batched child LPs.
//
// Scoring follows pick_strong_branch (solver/milp/src/branch_and_bound.cpp:
// 726-793): candidates are the fractional integer variables ordered by
// closeness of their fractional part to 1/2, the first k are branched down
// (ub = floor x) and up (lb = ceil x), a child's degradation is
// |child objective - parent objective| (1e6 when the child is infeasible),
// and the score is min(down, up) + 1e-4 max(down, up); the first best score
// wins. The 2k child LPs go to the GPU together (sov_batch_solve) instead of
// one warm-started dual simplex after another.
//
// The tree search is the core of BranchAndBoundSolver::solve
// (branch_and_bound.cpp:906-1523) without cuts, heuristics or presolve:
// best-bound node selection, integrality check before pruning, and children
// taken from the strong-branching solves of the chosen variable.

#include "internal.hpp"
#include "sovereign_bnb.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <memory>
#include <queue>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
constexpr double kInfeasibleDegradation = 1e6;  // branch_and_bound.cpp:784

// The node LP as a dense bounded-variable form: min c'x, A x = b, lb <= x <= ub
// (structural columns first, then one slack per inequality row), rows,
// columns and objective scaled for the IPM. The IPM's variable j is x_j /
// colscale[j]; lb and ub stay in the model's units.
struct DenseForm {
  int m = 0, n = 0, nv = 0;
  std::vector<double> A, b, c, lb, ub, colscale;
  double sense = 1.0;   // -1 when the model maximizes
  double cscale = 1.0;  // objective divided by this
  double constant = 0.0;
};

DenseForm dense_form(const sov_lp_model& mdl, const double* lb, const double* ub) {
  DenseForm F;
  F.nv = mdl.n_vars;
  F.m = mdl.n_rows;
  int ns = 0;
  for (int r = 0; r < F.m; ++r)
    if (mdl.row_sense[r] != 'E' && mdl.row_sense[r] != 'e') ++ns;
  F.n = F.nv + ns;
  F.A.assign(static_cast<std::size_t>(F.m) * F.n, 0.0);
  for (int j = 0; j < F.nv; ++j)
    for (int p = mdl.col_ptr[j]; p < mdl.col_ptr[j + 1]; ++p)
      F.A[mdl.row_idx[p] + static_cast<std::size_t>(j) * F.m] += mdl.values[p];
  int sc = F.nv;
  for (int r = 0; r < F.m; ++r) {
    const char s = mdl.row_sense[r];
    if (s == 'L' || s == 'l') F.A[r + static_cast<std::size_t>(sc++) * F.m] = 1.0;
    else if (s == 'G' || s == 'g') F.A[r + static_cast<std::size_t>(sc++) * F.m] = -1.0;
  }
  F.b.assign(mdl.rhs, mdl.rhs + F.m);
  F.sense = mdl.maximize ? -1.0 : 1.0;
  F.constant = mdl.obj_constant;
  F.c.assign(static_cast<std::size_t>(F.n), 0.0);
  for (int j = 0; j < F.nv; ++j) F.c[j] = F.sense * mdl.obj[j];
  F.lb.assign(static_cast<std::size_t>(F.n), 0.0);
  F.ub.assign(static_cast<std::size_t>(F.n), 1e30);
  for (int j = 0; j < F.nv; ++j) {
    F.lb[j] = lb[j];
    F.ub[j] = ub[j];
  }
  // Row and column equilibration as build_ipm_form's (solver/lp/src/
  // interior_point.cpp:370-405): two passes of row scaling by 1 / sqrt(row
  // max) and column scaling by 1 / sqrt(column max). With the rows scaled
  // alone the slack columns sit two orders of magnitude below the
  // structural ones, and the batched IPM needed about twice the iterations
  // of the reference's form on the knapsack children.
  F.colscale.assign(static_cast<std::size_t>(F.n), 1.0);
  for (int pass = 0; pass < 2; ++pass) {
    for (int r = 0; r < F.m; ++r) {
      double mx = 0.0;
      for (int j = 0; j < F.n; ++j) mx = std::max(mx, std::abs(F.A[r + static_cast<std::size_t>(j) * F.m]));
      if (mx > 0.0) {
        const double sr = 1.0 / std::sqrt(mx);
        for (int j = 0; j < F.n; ++j) F.A[r + static_cast<std::size_t>(j) * F.m] *= sr;
        F.b[r] *= sr;
      }
    }
    for (int j = 0; j < F.n; ++j) {
      double* col = F.A.data() + static_cast<std::size_t>(j) * F.m;
      double mx = 0.0;
      for (int r = 0; r < F.m; ++r) mx = std::max(mx, std::abs(col[r]));
      if (mx > 0.0) {
        const double sc = 1.0 / std::sqrt(mx);
        for (int r = 0; r < F.m; ++r) col[r] *= sc;
        F.c[j] *= sc;
        F.colscale[j] *= sc;
      }
    }
  }
  double cm = 0.0;
  for (double v : F.c) cm = std::max(cm, std::abs(v));
  F.cscale = std::max(1.0, cm);
  for (double& v : F.c) v /= F.cscale;
  return F;
}

// One child (or node) LP in shifted form 0 <= (x - lb) / colscale <= (ub -
// lb) / colscale (lb, ub in the model's units). Returns false when the box
// is empty.
bool shifted_child(const DenseForm& F, const std::vector<double>& lb, const std::vector<double>& ub, double* b,
                   double* u, double* shift_obj) {
  for (int i = 0; i < F.m; ++i) b[i] = F.b[i];
  double k = 0.0;
  for (int j = 0; j < F.n; ++j) {
    if (ub[j] < lb[j] - 1e-9) return false;
    const double l = lb[j] / F.colscale[j];  // in the IPM's units
    u[j] = ub[j] >= 1e29 ? 1e30 : std::max(0.0, ub[j] - lb[j]) / F.colscale[j];
    if (l != 0.0) {
      for (int i = 0; i < F.m; ++i) b[i] -= F.A[i + static_cast<std::size_t>(j) * F.m] * l;
      k += F.c[j] * l;
    }
  }
  *shift_obj = k;
  return true;
}

double model_objective(const DenseForm& F, double scaled_min_obj) {
  return F.sense * (scaled_min_obj * F.cscale) + F.constant;
}

struct ChildSpec {
  std::vector<double> lb, ub;
  bool empty = false;
};

struct ChildOut {
  int status = 7;
  int iterations = 0;
  double objective = 0.0;  // model sense
  std::vector<double> x;   // structural values (unshifted)
};

// Solves a set of children of one node in one batched launch.
int solve_children(sov_batch* batch, int mat_id, const DenseForm& F, const std::vector<ChildSpec>& kids,
                   const sov_ipm_options& opt, std::vector<ChildOut>& out, double* kernel_ms, int* lp_solves) {
  const int K = static_cast<int>(kids.size());
  out.assign(static_cast<std::size_t>(K), ChildOut{});
  std::vector<int> live;
  std::vector<double> b_all, c_all, u_all, shift_obj;
  std::vector<int64_t> b_off, x_off;
  for (int k = 0; k < K; ++k) {
    std::vector<double> b(static_cast<std::size_t>(F.m)), u(static_cast<std::size_t>(F.n));
    double so = 0.0;
    if (kids[k].empty || !shifted_child(F, kids[k].lb, kids[k].ub, b.data(), u.data(), &so)) {
      out[k].status = 2;  // Infeasible: empty box
      continue;
    }
    live.push_back(k);
    b_off.push_back(static_cast<int64_t>(b_all.size()));
    x_off.push_back(static_cast<int64_t>(c_all.size()));
    b_all.insert(b_all.end(), b.begin(), b.end());
    c_all.insert(c_all.end(), F.c.begin(), F.c.end());
    u_all.insert(u_all.end(), u.begin(), u.end());
    shift_obj.push_back(so);
  }
  if (live.empty()) return SOV_OK;
  const int L = static_cast<int>(live.size());
  std::vector<int32_t> mats(static_cast<std::size_t>(L), mat_id);
  std::vector<sov_batch_result> res(static_cast<std::size_t>(L));
  std::vector<double> x_all(c_all.size());
  const int rc = sov_batch_solve(batch, L, mats.data(), b_all.data(), b_off.data(), c_all.data(), u_all.data(),
                                 x_off.data(), &opt, res.data(), x_all.data());
  if (rc != SOV_OK) return rc;
  if (kernel_ms) *kernel_ms += sov_batch_last_kernel_ms(batch);
  if (lp_solves) *lp_solves += L;
  for (int q = 0; q < L; ++q) {
    ChildOut& o = out[live[q]];
    o.status = res[q].status;
    o.iterations = res[q].iterations;
    o.objective = model_objective(F, res[q].objective + shift_obj[q]);
    o.x.resize(static_cast<std::size_t>(F.nv));
    const ChildSpec& kid = kids[live[q]];
    for (int j = 0; j < F.nv; ++j) o.x[j] = F.colscale[j] * x_all[x_off[q] + j] + kid.lb[j];
  }
  return SOV_OK;
}

// Per-LP comparator: every child through the plain FP64 single-LP GPU path
// (sov_lp_solve with the reference's upper-bound rows).
int solve_children_one_by_one(const sov_lp_model& mdl, const std::vector<ChildSpec>& kids, const DenseForm& F,
                              const sov_ipm_options& opt, std::vector<ChildOut>& out, int* lp_solves) {
  const int K = static_cast<int>(kids.size());
  out.assign(static_cast<std::size_t>(K), ChildOut{});
  for (int k = 0; k < K; ++k) {
    if (kids[k].empty) {
      out[k].status = 2;
      continue;
    }
    bool emptybox = false;
    for (int j = 0; j < F.nv; ++j) emptybox = emptybox || kids[k].ub[j] < kids[k].lb[j] - 1e-9;
    if (emptybox) {
      out[k].status = 2;
      continue;
    }
    sov_lp_model child = mdl;
    child.lb = kids[k].lb.data();
    child.ub = kids[k].ub.data();
    sov_ipm_result r{};
    out[k].x.assign(static_cast<std::size_t>(F.nv), 0.0);
    const int rc = sov_lp_solve(&child, &opt, &r, out[k].x.data(), nullptr, nullptr);
    if (rc < 0) return rc;
    out[k].status = r.status;
    out[k].iterations = r.iterations;
    out[k].objective = r.objective;
    if (lp_solves) ++*lp_solves;
  }
  return SOV_OK;
}

// Children the batched kernel could not finish (status 5 or 6: it stops early
// when its FP32-preconditioned solves stall) go through the per-LP path,
// whose late iterations factor in FP64.
int finish_unconverged(const sov_lp_model& mdl, const std::vector<ChildSpec>& kids, const DenseForm& F,
                       const sov_ipm_options& opt, std::vector<ChildOut>& out, int* lp_solves) {
  std::vector<ChildSpec> redo;
  std::vector<int> where;
  for (std::size_t k = 0; k < out.size(); ++k)
    if (out[k].status == 5 || out[k].status == 6) {
      redo.push_back(kids[k]);
      where.push_back(static_cast<int>(k));
    }
  if (redo.empty()) return SOV_OK;
  if (const char* tr = std::getenv("SOVEREIGN_SB_TRACE"); tr && tr[0] == '1') {
    std::fprintf(stderr, "[strong-branch] %zu of %zu children left unconverged by the batch:", redo.size(), out.size());
    for (int k : where) std::fprintf(stderr, " %d(st%d,it%d)", k, out[k].status, out[k].iterations);
    std::fprintf(stderr, "\n");
  }
  std::vector<ChildOut> again;
  const int rc = solve_children_one_by_one(mdl, redo, F, opt, again, lp_solves);
  if (rc != SOV_OK) return rc;
  for (std::size_t q = 0; q < where.size(); ++q) out[where[q]] = std::move(again[q]);
  return SOV_OK;
}

double frac_part(double v) { return std::min(v - std::floor(v), std::ceil(v) - v); }

// fractional_vars + the closeness-to-1/2 order of branch_and_bound.cpp:731-741.
std::vector<int> fractional_candidates(const int8_t* is_int, const double* x, int nv, double tol) {
  std::vector<int> c;
  for (int j = 0; j < nv; ++j)
    if (is_int[j] && frac_part(x[j]) > tol) c.push_back(j);
  std::stable_sort(c.begin(), c.end(), [&](int a, int b) {
    return std::abs(frac_part(x[a]) - 0.5) < std::abs(frac_part(x[b]) - 0.5);
  });
  return c;
}

void children_specs(const std::vector<double>& lb, const std::vector<double>& ub, const std::vector<int>& cand,
                    const double* x, int nslack_total, std::vector<ChildSpec>& kids) {
  kids.clear();
  for (int v : cand) {
    ChildSpec down, up;
    down.lb = lb;
    down.ub = ub;
    up.lb = lb;
    up.ub = ub;
    down.ub[v] = std::min(ub[v], std::floor(x[v]));
    up.lb[v] = std::max(lb[v], std::ceil(x[v]));
    for (ChildSpec* s : {&down, &up}) {
      s->lb.resize(lb.size() + nslack_total, 0.0);
      s->ub.resize(ub.size() + nslack_total, 1e30);
    }
    kids.push_back(std::move(down));
    kids.push_back(std::move(up));
  }
}

double degradation(const ChildOut& c, double parent) {
  if (c.status == 2) return kInfeasibleDegradation;
  const double obj = c.status == 0 ? c.objective : parent;
  return std::max(0.0, std::abs(obj - parent));
}

}  // namespace

extern "C" {

void sov_bnb_default_options(sov_bnb_options* o) {
  o->max_nodes = 100000;
  o->strong_branch_candidates = 64;
  o->mip_gap = 1e-6;
  o->integer_tol = 1e-6;
  o->child_solver = SOV_CHILD_BATCHED;
  o->time_limit_seconds = 0.0;
  o->candidate_order = nullptr;
  o->candidate_order_user = nullptr;
}

int sov_strong_branch(const sov_lp_model* node, const double* node_lb, const double* node_ub, const int8_t* is_int,
                      const double* x_node, double parent_obj, int32_t max_candidates, int32_t child_solver,
                      const sov_ipm_options* opt, sov_sb_child* out, int32_t* ncand_out, int32_t* best,
                      double* kernel_ms) {
  if (!node || !node_lb || !node_ub || !is_int || !x_node || !opt || !out || !ncand_out || !best) return SOV_ERR_ARG;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  *best = -1;
  std::vector<int> cand = fractional_candidates(is_int, x_node, node->n_vars, 1e-6);
  if (static_cast<int>(cand.size()) > max_candidates) cand.resize(static_cast<std::size_t>(max_candidates));
  *ncand_out = static_cast<int32_t>(cand.size());
  if (cand.empty()) return SOV_OK;
  const DenseForm F = dense_form(*node, node_lb, node_ub);
  std::vector<double> lb(node_lb, node_lb + node->n_vars), ub(node_ub, node_ub + node->n_vars);
  std::vector<ChildSpec> kids;
  children_specs(lb, ub, cand, x_node, F.n - F.nv, kids);
  std::vector<ChildOut> res;
  int rc;
  if (child_solver == SOV_CHILD_BATCHED) {
    sov_batch* batch = nullptr;
    if ((rc = sov_batch_create(&batch)) != SOV_OK) return rc;
    int32_t mid = 0;
    rc = sov_batch_add_matrix(batch, F.m, F.n, F.A.data(), &mid);
    if (rc == SOV_OK) rc = solve_children(batch, mid, F, kids, *opt, res, kernel_ms, nullptr);
    sov_batch_free(batch);
    if (rc == SOV_OK) rc = finish_unconverged(*node, kids, F, *opt, res, nullptr);
  } else {
    rc = solve_children_one_by_one(*node, kids, F, *opt, res, nullptr);
  }
  if (rc != SOV_OK) return rc;
  double best_score = -1.0;
  for (std::size_t k = 0; k < cand.size(); ++k) {
    const ChildOut& d = res[2 * k];
    const ChildOut& u = res[2 * k + 1];
    sov_sb_child& o = out[k];
    o.var = cand[k];
    o.status_down = d.status;
    o.status_up = u.status;
    o.iters_down = d.iterations;
    o.iters_up = u.iterations;
    o.obj_down = d.status == 0 ? d.objective : parent_obj;
    o.obj_up = u.status == 0 ? u.objective : parent_obj;
    const double dd = degradation(d, parent_obj), du = degradation(u, parent_obj);
    o.score = std::min(dd, du) + 1e-4 * std::max(dd, du);
    if (o.score > best_score) {
      best_score = o.score;
      *best = static_cast<int32_t>(k);
    }
  }
  return SOV_OK;
}

int sov_bnb_solve(const sov_lp_model* milp, const int8_t* is_int, const sov_bnb_options* bo,
                  const sov_ipm_options* opt, sov_bnb_result* res, double* x_best) {
  if (!milp || !is_int || !bo || !opt || !res) return SOV_ERR_ARG;
  *res = sov_bnb_result{};
  res->status = 7;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  const auto t0 = Clock::now();
  const int nv = milp->n_vars;
  std::vector<double> lb0(milp->lb, milp->lb + nv), ub0(milp->ub, milp->ub + nv);
  const DenseForm F = dense_form(*milp, lb0.data(), ub0.data());
  const int nslack = F.n - F.nv;
  if (F.m > 272 && bo->child_solver == SOV_CHILD_BATCHED) {
    sov::set_error("sov_bnb_solve: batched child LPs need at most 272 rows");
    return SOV_ERR_ARG;
  }
  std::unique_ptr<sov_batch, int (*)(sov_batch*)> batch(nullptr, sov_batch_free);
  int32_t mat_id = 0;
  if (bo->child_solver == SOV_CHILD_BATCHED) {
    sov_batch* b = nullptr;
    int rc = sov_batch_create(&b);
    if (rc != SOV_OK) return rc;
    batch.reset(b);
    rc = sov_batch_add_matrix(b, F.m, F.n, F.A.data(), &mat_id);
    if (rc != SOV_OK) return rc;
  }
  double kernel_ms = 0.0;
  int lp_solves = 0;
  double sb_seconds = 0.0;
  auto solve_set = [&](const std::vector<ChildSpec>& kids, std::vector<ChildOut>& outv) {
    if (bo->child_solver == SOV_CHILD_BATCHED) {
      const int rc = solve_children(batch.get(), mat_id, F, kids, *opt, outv, &kernel_ms, &lp_solves);
      return rc != SOV_OK ? rc : finish_unconverged(*milp, kids, F, *opt, outv, &lp_solves);
    }
    return solve_children_one_by_one(*milp, kids, F, *opt, outv, &lp_solves);
  };
  // Minimisation view: bounds and incumbent in F.sense * objective.
  struct Node {
    double bound;  // min-sense LP bound
    long long id;
    std::vector<double> lb, ub, x;
  };
  auto worse = [](const Node& a, const Node& b) { return a.bound != b.bound ? a.bound > b.bound : a.id > b.id; };
  std::priority_queue<Node, std::vector<Node>, decltype(worse)> open(worse);
  long long next_id = 0;
  {
    ChildSpec root;
    root.lb = lb0;
    root.ub = ub0;
    root.lb.resize(static_cast<std::size_t>(F.n), 0.0);
    root.ub.resize(static_cast<std::size_t>(F.n), 1e30);
    std::vector<ChildOut> r;
    const int rc = solve_set({root}, r);
    if (rc != SOV_OK) return rc;
    if (r[0].status == 2) {
      res->status = 2;
      return SOV_OK;
    }
    if (r[0].status != 0) {
      res->status = 6;
      return SOV_OK;
    }
    res->root_bound = r[0].objective;
    open.push({F.sense * r[0].objective, next_id++, lb0, ub0, r[0].x});
  }
  double incumbent = std::numeric_limits<double>::infinity();  // min sense
  std::vector<double> best_x;
  long long nodes = 0;
  bool limit_hit = false;
  while (!open.empty()) {
    if (nodes >= bo->max_nodes ||
        (bo->time_limit_seconds > 0 &&
         std::chrono::duration<double>(Clock::now() - t0).count() > bo->time_limit_seconds)) {
      limit_hit = true;
      break;
    }
    Node node = open.top();
    open.pop();
    const double tol = bo->mip_gap * std::max(1.0, std::abs(incumbent));
    if (node.bound >= incumbent - tol) continue;  // pruned by bound
    ++nodes;
    std::vector<int> cand = fractional_candidates(is_int, node.x.data(), nv, bo->integer_tol);
    if (cand.empty()) {  // integral: new incumbent
      incumbent = node.bound;
      best_x = node.x;
      for (int j = 0; j < nv; ++j)
        if (is_int[j]) best_x[j] = std::round(best_x[j]);
      continue;
    }
    // Candidate order: most fractional (the reference's FullStrong rule), or
    // a caller-supplied ranking (the learned surrogate) over the same set.
    if (bo->candidate_order) {
      std::vector<double> key(static_cast<std::size_t>(nv), 0.0);
      bo->candidate_order(nv, node.x.data(), node.lb.data(), node.ub.data(), key.data(), bo->candidate_order_user);
      std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return key[a] > key[b]; });
    }
    const int k = std::min<int>(static_cast<int>(cand.size()), bo->strong_branch_candidates);
    cand.resize(static_cast<std::size_t>(k));
    std::vector<ChildSpec> kids;
    children_specs(node.lb, node.ub, cand, node.x.data(), nslack, kids);
    std::vector<ChildOut> outs;
    const auto ts = Clock::now();
    const int rc = solve_set(kids, outs);
    sb_seconds += std::chrono::duration<double>(Clock::now() - ts).count();
    if (rc != SOV_OK) return rc;
    const double parent = F.sense * node.bound;  // model sense
    int best = 0;
    double best_score = -1.0;
    for (int q = 0; q < k; ++q) {
      const double dd = degradation(outs[2 * q], parent), du = degradation(outs[2 * q + 1], parent);
      const double sc = std::min(dd, du) + 1e-4 * std::max(dd, du);
      if (sc > best_score) {
        best_score = sc;
        best = q;
      }
    }
    for (int side = 0; side < 2; ++side) {
      const ChildOut& ch = outs[2 * best + side];
      if (ch.status != 0) continue;  // infeasible (or unresolved) child: no subtree
      const double b = F.sense * ch.objective;
      if (b >= incumbent - bo->mip_gap * std::max(1.0, std::abs(incumbent))) continue;
      std::vector<double> clb(kids[2 * best + side].lb.begin(), kids[2 * best + side].lb.begin() + nv);
      std::vector<double> cub(kids[2 * best + side].ub.begin(), kids[2 * best + side].ub.begin() + nv);
      open.push({b, next_id++, std::move(clb), std::move(cub), ch.x});
    }
  }
  res->nodes = nodes;
  res->lp_solves = lp_solves;
  res->strong_branch_seconds = sb_seconds;
  res->kernel_ms = kernel_ms;
  res->seconds = std::chrono::duration<double>(Clock::now() - t0).count();
  if (best_x.empty()) {
    res->status = limit_hit ? 5 : 2;
    return SOV_OK;
  }
  res->status = limit_hit ? 1 : 0;
  res->objective = F.sense * incumbent;
  if (x_best)
    for (int j = 0; j < nv; ++j) x_best[j] = best_x[j];
  return SOV_OK;
}

}  // extern "C"
