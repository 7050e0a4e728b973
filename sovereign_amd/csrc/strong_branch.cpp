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
// (branch_and_bound.cpp:906-1523) without primal heuristics or model
// presolve: root cut rounds (cuts.hpp), best-bound node selection with
// plunging, the reference's branching rules and pseudocost learning, a node
// presolve infeasibility test, integrality checked before pruning, children
// taken from the strong-branching solves of the chosen variable, and child
// LPs warm-started from their parent's iterate (sovereign_bnb.h).

#include "internal.hpp"
#include "cuts.hpp"
#include "sovereign_bnb.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <memory>
#include <queue>
#include <system_error>
#include <thread>
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

// A batched LP's iterate (x, w, z, v: n each, y: m, in the form's units)
// and the lower bounds it was shifted by (model units, n): the warm start
// of its children.
struct WarmState {
  std::vector<double> v, lb;
};

struct ChildSpec {
  std::vector<double> lb, ub;
  bool empty = false;
  std::shared_ptr<const WarmState> warm;  // the parent LP's iterate, or none (default start)
};

struct WarmOpts {
  bool on = false;
  double save_mu = 0.1;
};

// Node presolve: proves a node LP infeasible from its rows' activity ranges
// under the node's bounds, with activity-based bound propagation between
// passes (the infeasibility tests of Presolver, solver/presolve/src/
// presolve.cpp:186-246 and 560-660, with its margin 1e-6 (1 + |scale|) and
// |candidate| <= 1e9). The node's branch-and-bound descends into LPs that an
// interior point cannot certify infeasible (it only stalls); this test
// catches those that bound propagation proves empty. The propagated bounds
// serve only the test (no integer rounding, so what it proves holds for the
// LP relaxation); the LP itself is solved on the node's own bounds.
class NodePresolve {
 public:
  explicit NodePresolve(const sov_lp_model& mdl)
      : m_(mdl.n_rows), n_(mdl.n_vars), sense_(mdl.row_sense, mdl.row_sense + mdl.n_rows),
        rhs_(mdl.rhs, mdl.rhs + mdl.n_rows), cptr_(mdl.col_ptr, mdl.col_ptr + mdl.n_vars + 1),
        crow_(mdl.row_idx, mdl.row_idx + mdl.col_ptr[mdl.n_vars]), cval_(mdl.values, mdl.values + mdl.col_ptr[mdl.n_vars]) {
    ptr_.assign(static_cast<std::size_t>(m_) + 1, 0);
    for (int p = 0; p < cptr_[n_]; ++p) ++ptr_[crow_[p] + 1];
    for (int r = 0; r < m_; ++r) ptr_[r + 1] += ptr_[r];
    col_.resize(static_cast<std::size_t>(ptr_[m_]));
    val_.resize(col_.size());
    std::vector<int> fill(ptr_.begin(), ptr_.end() - 1);
    for (int j = 0; j < n_; ++j)
      for (int p = cptr_[j]; p < cptr_[j + 1]; ++p) {
        const int q = fill[crow_[p]]++;
        col_[q] = j;
        val_[q] = cval_[p];
      }
  }

  // Full test: activity ranges and bound propagation to a fixed point (at
  // most 8 passes). lb, ub: at least n_vars entries (structural first).
  bool infeasible(std::vector<double> lb, std::vector<double> ub) const {
    for (int j = 0; j < n_; ++j)
      if (crossed(lb[j], ub[j])) return true;
    for (int pass = 0; pass < 8; ++pass) {
      bool changed = false;
      for (int r = 0; r < m_; ++r) {
        const bool le = is_le(sense_[r]), ge = is_ge(sense_[r]);
        Activity act;
        for (int p = ptr_[r]; p < ptr_[r + 1]; ++p) act.add(val_[p], lb[col_[p]], ub[col_[p]], 1.0);
        if (violates(act, r)) return true;
        const double mn = act.mn, mx = act.mx;
        // a_j x_j <= rhs - (min activity of the rest), >= rhs - (max of the rest)
        for (int p = ptr_[r]; p < ptr_[r + 1]; ++p) {
          const double a = val_[p];
          const int j = col_[p];
          const double lo = a > 0 ? lb[j] : ub[j], hi = a > 0 ? ub[j] : lb[j];
          const bool lo_inf = std::abs(lo) >= kInf, hi_inf = std::abs(hi) >= kInf;
          if (le && act.mn_inf - (lo_inf ? 1 : 0) == 0) {
            const double cand = (rhs_[r] - (mn - (lo_inf ? 0.0 : a * lo))) / a;
            if (std::abs(cand) <= 1e9) {
              double& bnd = a > 0 ? ub[j] : lb[j];
              if (a > 0 ? cand < bnd - margin(cand) : cand > bnd + margin(cand)) {
                bnd = cand;
                changed = true;
              }
            }
          }
          if (ge && act.mx_inf - (hi_inf ? 1 : 0) == 0) {
            const double cand = (rhs_[r] - (mx - (hi_inf ? 0.0 : a * hi))) / a;
            if (std::abs(cand) <= 1e9) {
              double& bnd = a > 0 ? lb[j] : ub[j];
              if (a > 0 ? cand > bnd + margin(cand) : cand < bnd - margin(cand)) {
                bnd = cand;
                changed = true;
              }
            }
          }
          if (crossed(lb[j], ub[j])) return true;
        }
      }
      if (!changed) break;
    }
    return false;
  }

  // Quick test of a set of children of one node (their bounds differ in a
  // few variables): each row's activity range under the first child's
  // bounds, updated column by column for the others. Marks the children it
  // proves infeasible.
  template <typename Spec>
  void mark(std::vector<Spec>& kids) const {
    const std::vector<char> bad = infeasible(kids);
    for (std::size_t k = 0; k < kids.size(); ++k)
      if (bad[k]) kids[k].empty = true;
  }

  // The test of mark() without marking: entry k is 1 when child k's box is
  // crossed or a row it touches cannot be met. Reads kids only.
  template <typename Spec>
  std::vector<char> infeasible(const std::vector<Spec>& kids) const {
    std::vector<char> out(kids.size(), 0);
    if (kids.empty() || m_ == 0) return out;
    const Spec& k0 = kids[0];
    std::vector<Activity> base(static_cast<std::size_t>(m_));
    for (int j = 0; j < n_; ++j)
      for (int p = cptr_[j]; p < cptr_[j + 1]; ++p) base[crow_[p]].add(cval_[p], k0.lb[j], k0.ub[j], 1.0);
    std::vector<Activity> act;
    std::vector<int> rows;
    for (std::size_t q = 0; q < kids.size(); ++q) {
      const Spec& k = kids[q];
      if (k.empty) continue;
      bool bad = false;
      rows.clear();
      act = base;
      for (int j = 0; j < n_ && !bad; ++j) {
        if (k.lb[j] == k0.lb[j] && k.ub[j] == k0.ub[j]) continue;
        bad = crossed(k.lb[j], k.ub[j]);
        for (int p = cptr_[j]; p < cptr_[j + 1]; ++p) {
          act[crow_[p]].add(cval_[p], k0.lb[j], k0.ub[j], -1.0);
          act[crow_[p]].add(cval_[p], k.lb[j], k.ub[j], 1.0);
          rows.push_back(crow_[p]);
        }
      }
      const bool first = &k == &kids[0];
      if (first)
        for (int r = 0; r < m_ && !bad; ++r) bad = violates(act[r], r);
      else
        for (int r : rows)
          if (!bad) bad = violates(act[r], r);
      if (bad) out[q] = 1;
    }
    return out;
  }

 private:
  static constexpr double kInf = 1e29;
  static double margin(double scale) { return 1e-6 * (1.0 + std::abs(scale)); }
  static bool crossed(double lo, double hi) { return lo - hi > margin(std::max(std::abs(lo), std::abs(hi))); }
  static bool is_le(char c) { return c == 'L' || c == 'l' || c == 'E' || c == 'e'; }
  static bool is_ge(char c) { return c == 'G' || c == 'g' || c == 'E' || c == 'e'; }

  // A row's activity range: finite parts, counts of infinite terms, and the
  // magnitude that scales its margin.
  struct Activity {
    double mn = 0.0, mx = 0.0, mag = 0.0;
    int mn_inf = 0, mx_inf = 0;
    void add(double a, double lb, double ub, double w) {
      const double lo = a > 0 ? lb : ub, hi = a > 0 ? ub : lb;
      if (std::abs(lo) >= kInf) mn_inf += static_cast<int>(w);
      else {
        mn += w * a * lo;
        mag += w * std::abs(a * lo);
      }
      if (std::abs(hi) >= kInf) mx_inf += static_cast<int>(w);
      else {
        mx += w * a * hi;
        mag += w * std::abs(a * hi);
      }
    }
  };
  bool violates(const Activity& act, int r) const {
    const double mg = margin(act.mag);
    if (is_le(sense_[r]) && act.mn_inf == 0 && act.mn > rhs_[r] + mg) return true;
    if (is_ge(sense_[r]) && act.mx_inf == 0 && act.mx < rhs_[r] - mg) return true;
    return false;
  }

  int m_, n_;
  std::vector<char> sense_;
  std::vector<double> rhs_;
  std::vector<int> cptr_, crow_;
  std::vector<double> cval_;
  std::vector<int> ptr_, col_;
  std::vector<double> val_;
};

struct ChildOut {
  int status = 7;
  int iterations = 0;
  double objective = 0.0;  // model sense
  std::vector<double> x;   // structural values (unshifted)
  std::shared_ptr<const WarmState> state;  // batched solves with warm starts on: the iterate its children start from
};

// The parent's iterate mapped into a child's shifted box: x moves by the
// change of lower bound, components pushed at least `lift` (or half the
// box) inside their bounds, their multipliers raised so that each
// complementarity product is at least save_mu; w, z, v, y carried over.
void warm_start_for(const DenseForm& F, const WarmState& W, const ChildSpec& kid, const std::vector<double>& u,
                    double save_mu, double* out) {
  const int n = F.n, m = F.m;
  const double lift = std::sqrt(save_mu);
  for (int j = 0; j < n; ++j) {
    double x = W.v[j] + (W.lb[j] - kid.lb[j]) / F.colscale[j];
    double w = 0.0, z = W.v[2 * n + j], v = W.v[3 * n + j];
    const double uj = u[j];
    if (uj <= 0.0) {
      x = w = z = v = 0.0;
    } else if (uj < 1e29) {
      const double d = std::min(lift, 0.5 * uj);
      x = std::min(std::max(x, d), uj - d);
      w = uj - x;
      z = std::max(z, save_mu / x);
      v = std::max(v, save_mu / w);
    } else {
      x = std::max(x, lift);
      z = std::max(z, save_mu / x);
      v = 0.0;
    }
    out[j] = x;
    out[n + j] = w;
    out[2 * n + j] = z;
    out[3 * n + j] = v;
  }
  for (int i = 0; i < m; ++i) out[4 * n + i] = W.v[4 * n + i];
}

// Solves a set of children of one node in one batched launch.
int solve_children(sov_batch* batch, int mat_id, const DenseForm& F, const std::vector<ChildSpec>& kids,
                   const sov_ipm_options& opt, std::vector<ChildOut>& out, double* kernel_ms, int* lp_solves,
                   const WarmOpts& warm = WarmOpts{}) {
  const int K = static_cast<int>(kids.size());
  out.assign(static_cast<std::size_t>(K), ChildOut{});
  std::vector<int> live;
  std::vector<double> b_all, c_all, u_all, shift_obj, s_in;
  std::vector<int64_t> b_off, x_off, in_off, out_off;
  const int state_len = 4 * F.n + F.m;
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
    if (warm.on) {
      out_off.push_back(static_cast<int64_t>(out_off.size()) * state_len);
      if (kids[k].warm) {
        in_off.push_back(static_cast<int64_t>(s_in.size()));
        s_in.resize(s_in.size() + static_cast<std::size_t>(state_len));
        warm_start_for(F, *kids[k].warm, kids[k], u, warm.save_mu, s_in.data() + in_off.back());
      } else {
        in_off.push_back(-1);
      }
    }
  }
  if (live.empty()) return SOV_OK;
  const int L = static_cast<int>(live.size());
  std::vector<int32_t> mats(static_cast<std::size_t>(L), mat_id);
  std::vector<sov_batch_result> res(static_cast<std::size_t>(L));
  std::vector<double> x_all(c_all.size());
  std::vector<double> s_out(warm.on ? static_cast<std::size_t>(L) * state_len : 0);
  const bool any_in = !s_in.empty();
  const int rc = sov_batch_solve_warm(batch, L, mats.data(), b_all.data(), b_off.data(), c_all.data(), u_all.data(),
                                      x_off.data(), &opt, res.data(), x_all.data(), any_in ? s_in.data() : nullptr,
                                      any_in ? in_off.data() : nullptr, warm.on ? s_out.data() : nullptr,
                                      warm.on ? out_off.data() : nullptr, warm.save_mu);
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
    if (warm.on) {
      auto st = std::make_shared<WarmState>();
      st->v.assign(s_out.begin() + out_off[q], s_out.begin() + out_off[q] + state_len);
      st->lb = kid.lb;
      o.state = std::move(st);
    }
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

// An LP the interior point left unconverged may be infeasible, which the
// interior point cannot certify: the full presolve test settles the ones
// that bound propagation proves empty.
int prove_infeasible(const NodePresolve& pre, const std::vector<ChildSpec>& kids, std::vector<ChildOut>& out) {
  for (std::size_t k = 0; k < out.size(); ++k)
    if (out[k].status != 0 && out[k].status != 2 && pre.infeasible(kids[k].lb, kids[k].ub)) out[k].status = 2;
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

// A variable with lb <= -1e29 has no finite shift into the batched child form
// 0 <= x <= u; such models go through the per-LP path.
bool batched_form_ok(const double* lb, int nv) {
  for (int j = 0; j < nv; ++j)
    if (lb[j] <= -1e29) return false;
  return true;
}

// PseudoCostStats (branch_and_bound.cpp:74-79), indexed by variable. A
// variable is "known" to the reference once either count is positive: every
// site that creates its map entry also increments a count.
struct PseudoCost {
  double down_sum = 0.0, up_sum = 0.0;
  int down_count = 0, up_count = 0;
};

// product_score (branch_and_bound.cpp:795-797).
double product_score(double down_gain, double up_gain) {
  return std::max(down_gain, 1e-6) * std::max(up_gain, 1e-6);
}

// The branching decision of one node, with the two children's LP results
// when strong branching solved them (the node LP of that child is then not
// solved again: it is the same LP).
struct Decision {
  int var = -1;
  bool have_children = false;
  ChildOut down, up;
};

// Solves the down/up children of `vars` at a node as one set (one batched
// launch, or one per-LP solve each); out[2q], out[2q+1] are var q's down, up.
using PairSolver = std::function<int(const std::vector<int>& vars, std::vector<ChildOut>& out)>;

// pick_most_fractional (branch_and_bound.cpp:144-159): the largest
// fractional part, the first such variable on ties.
int pick_most_fractional(const std::vector<int>& cands, const double* x) {
  int best = -1;
  double best_score = -1.0;
  for (int i : cands) {
    const double score = frac_part(x[i]);
    if (score > best_score) {
      best_score = score;
      best = i;
    }
  }
  return best;
}

// pick_pseudo_cost (branch_and_bound.cpp:161-189).
int pick_pseudo_cost(const std::vector<int>& cands, const double* x, const std::vector<PseudoCost>& st) {
  int best = -1;
  double best_score = -1.0;
  for (int i : cands) {
    const double fdown = x[i] - std::floor(x[i]), fup = std::ceil(x[i]) - x[i];
    double down_est = fdown, up_est = fup;
    if (st[i].down_count > 0) down_est = fdown * (st[i].down_sum / st[i].down_count);
    if (st[i].up_count > 0) up_est = fup * (st[i].up_sum / st[i].up_count);
    const double score = std::min(down_est, up_est) + 1e-4 * std::max(down_est, up_est);
    if (score > best_score) {
      best_score = score;
      best = i;
    }
  }
  return best >= 0 ? best : pick_most_fractional(cands, x);
}

// pick_strong_branch (branch_and_bound.cpp:726-793) with every candidate's
// two children in one set. `order` (may be empty) is a ranking key per
// variable applied, stably, after the closeness order (the learned
// surrogate); the first k candidates are branched.
Decision pick_full_strong(const std::vector<int>& frac, const double* x, double parent, double tol, int k,
                          const std::vector<double>& order, std::vector<PseudoCost>& st, const PairSolver& solve,
                          int* rc, long long* solved) {
  Decision d;
  std::vector<int> cands = frac;
  std::sort(cands.begin(), cands.end(),
            [&](int a, int b) { return std::abs(frac_part(x[a]) - 0.5) < std::abs(frac_part(x[b]) - 0.5); });
  if (!order.empty())
    std::stable_sort(cands.begin(), cands.end(), [&](int a, int b) { return order[a] > order[b]; });
  if (static_cast<int>(cands.size()) > k) cands.resize(static_cast<std::size_t>(std::max(k, 0)));
  if (cands.empty()) {
    d.var = pick_most_fractional(frac, x);
    return d;
  }
  std::vector<ChildOut> outs;
  if ((*rc = solve(cands, outs)) != SOV_OK) return d;
  *solved += 2 * static_cast<long long>(cands.size());
  int best = -1;
  double best_score = -1.0;
  for (std::size_t q = 0; q < cands.size(); ++q) {
    const int i = cands[q];
    const ChildOut& dn = outs[2 * q];
    const ChildOut& up = outs[2 * q + 1];
    const double down_obj = dn.status == 0 ? dn.objective : parent;
    const double up_obj = up.status == 0 ? up.objective : parent;
    const double fdown = x[i] - std::floor(x[i]), fup = std::ceil(x[i]) - x[i];
    if (fdown > tol) {
      st[i].down_sum += std::abs(down_obj - parent) / fdown;
      ++st[i].down_count;
    }
    if (fup > tol) {
      st[i].up_sum += std::abs(up_obj - parent) / fup;
      ++st[i].up_count;
    }
    const double dd = degradation(dn, parent), du = degradation(up, parent);
    const double score = std::min(dd, du) + 1e-4 * std::max(dd, du);
    if (score > best_score) {
      best_score = score;
      best = static_cast<int>(q);
    }
  }
  if (best < 0) {
    d.var = pick_most_fractional(frac, x);
    return d;
  }
  d.var = cands[best];
  d.have_children = true;
  d.down = std::move(outs[2 * best]);
  d.up = std::move(outs[2 * best + 1]);
  return d;
}

// pick_reliability_branch (branch_and_bound.cpp:799-899). The reference
// strong-branches unreliable candidates one at a time, in pseudocost order,
// until max_strong are done or kLookahead in a row fail to improve the best
// score. Every candidate that loop can reach is among the first max_strong
// unreliable ones, so their children are solved together in one set first;
// the sequential rule is then replayed over those results, and only what it
// reaches updates the pseudocosts or decides, exactly as in the reference.
Decision pick_reliability(const std::vector<int>& frac, const double* x, double parent, double tol, int reliability,
                          int max_strong, std::vector<PseudoCost>& st, const PairSolver& solve, int* rc,
                          long long* solved, long long* used) {
  Decision d;
  if (frac.empty()) return d;
  double down_sum = 0.0, up_sum = 0.0;
  int down_n = 0, up_n = 0;
  for (const PseudoCost& p : st) {
    if (p.down_count > 0) {
      down_sum += p.down_sum / p.down_count;
      ++down_n;
    }
    if (p.up_count > 0) {
      up_sum += p.up_sum / p.up_count;
      ++up_n;
    }
  }
  const double avg_down = down_n > 0 ? down_sum / down_n : 1.0;
  const double avg_up = up_n > 0 ? up_sum / up_n : 1.0;
  struct Scored {
    int var;
    double score;
    bool reliable;
  };
  std::vector<Scored> scored;
  scored.reserve(frac.size());
  for (int i : frac) {
    const double fdown = x[i] - std::floor(x[i]), fup = std::ceil(x[i]) - x[i];
    const PseudoCost& p = st[i];
    const bool known = p.down_count > 0 || p.up_count > 0;
    const double unit_down = p.down_count > 0 ? p.down_sum / p.down_count : avg_down;
    const double unit_up = p.up_count > 0 ? p.up_sum / p.up_count : avg_up;
    const bool reliable = known && std::min(p.down_count, p.up_count) >= reliability;
    scored.push_back({i, product_score(fdown * unit_down, fup * unit_up), reliable});
  }
  std::sort(scored.begin(), scored.end(), [](const Scored& a, const Scored& b) { return a.score > b.score; });
  std::vector<int> spec;  // the candidates the sequential loop can reach, in its order
  for (const Scored& s : scored)
    if (!s.reliable && static_cast<int>(spec.size()) < max_strong) spec.push_back(s.var);
  std::vector<ChildOut> outs;
  if (!spec.empty()) {
    if ((*rc = solve(spec, outs)) != SOV_OK) return d;
    *solved += 2 * static_cast<long long>(spec.size());
  }
  constexpr int kLookahead = 4;
  int best = scored.front().var, best_q = -1;
  double best_score = -1.0;
  int strong_done = 0, no_improve = 0;
  for (const Scored& s : scored) {
    double score = s.score;
    int q = -1;
    if (!s.reliable && strong_done < max_strong && no_improve < kLookahead) {
      q = strong_done++;  // spec[q] == s.var: the q-th unreliable candidate in this order
      *used += 2;
      const ChildOut& dn = outs[2 * q];
      const ChildOut& up = outs[2 * q + 1];
      const bool down_inf = dn.status == 2, up_inf = up.status == 2;
      if (down_inf && up_inf) {  // both children die; any choice prunes the node
        d.var = s.var;
        d.have_children = true;
        d.down = dn;
        d.up = up;
        return d;
      }
      const double fdown = x[s.var] - std::floor(x[s.var]), fup = std::ceil(x[s.var]) - x[s.var];
      double down_deg = down_inf ? kInfeasibleDegradation : 0.0, up_deg = up_inf ? kInfeasibleDegradation : 0.0;
      if (!down_inf && dn.status == 0) {
        down_deg = std::abs(dn.objective - parent);
        st[s.var].down_sum += down_deg / std::max(fdown, tol);
        ++st[s.var].down_count;
      }
      if (!up_inf && up.status == 0) {
        up_deg = std::abs(up.objective - parent);
        st[s.var].up_sum += up_deg / std::max(fup, tol);
        ++st[s.var].up_count;
      }
      score = product_score(down_deg, up_deg);
      no_improve = score > best_score ? 0 : no_improve + 1;
    }
    if (score > best_score) {
      best_score = score;
      best = s.var;
      best_q = q;
    }
  }
  d.var = best;
  if (best_q >= 0) {
    d.have_children = true;
    d.down = std::move(outs[2 * best_q]);
    d.up = std::move(outs[2 * best_q + 1]);
  }
  return d;
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
  o->branch_rule = SOV_BRANCH_RELIABILITY;
  o->reliability_threshold = 4;
  o->max_strong_per_node = 8;
  o->plunging = 1;
  o->max_cut_rounds = 5;
  o->max_cuts_per_node = 20;
  o->cmir_at_root = 0;
  o->warm_start = 1;
  o->warm_mu = 0.1;
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
  if (child_solver == SOV_CHILD_BATCHED && F.m > 272) {
    sov::set_error("sov_strong_branch: batched child LPs need at most 272 rows");
    return SOV_ERR_ARG;
  }
  std::vector<double> lb(node_lb, node_lb + node->n_vars), ub(node_ub, node_ub + node->n_vars);
  std::vector<ChildSpec> kids;
  children_specs(lb, ub, cand, x_node, F.n - F.nv, kids);
  const NodePresolve presolve(*node);
  std::vector<ChildOut> res;
  int rc;
  if (child_solver == SOV_CHILD_BATCHED && batched_form_ok(node_lb, node->n_vars)) {
    // The presolve test runs on a host thread while the GPU solves every
    // child; a child it proves infeasible then gets the result it gets when
    // marked before the solve (status 2), whatever its LP returned.
    std::vector<char> pre_bad;
    std::thread pre;
    try {
      pre = std::thread([&] { pre_bad = presolve.infeasible(kids); });
    } catch (const std::system_error&) {
      pre_bad = presolve.infeasible(kids);  // no thread to be had: the test runs here
    }
    sov_batch* batch = nullptr;
    rc = sov_batch_create(&batch);
    int32_t mid = 0;
    if (rc == SOV_OK) rc = sov_batch_add_matrix(batch, F.m, F.n, F.A.data(), &mid);
    if (rc == SOV_OK) rc = solve_children(batch, mid, F, kids, *opt, res, kernel_ms, nullptr);
    if (batch) sov_batch_free(batch);
    if (pre.joinable()) pre.join();
    if (rc == SOV_OK)
      for (std::size_t k = 0; k < kids.size(); ++k)
        if (pre_bad[k]) {
          kids[k].empty = true;
          res[k] = ChildOut{};
          res[k].status = 2;
        }
    if (rc == SOV_OK) rc = prove_infeasible(presolve, kids, res);
    if (rc == SOV_OK) rc = finish_unconverged(*node, kids, F, *opt, res, nullptr);
  } else {
    presolve.mark(kids);
    rc = solve_children_one_by_one(*node, kids, F, *opt, res, nullptr);
    if (rc == SOV_OK) rc = prove_infeasible(presolve, kids, res);
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
  if (bo->branch_rule < SOV_BRANCH_MOST_FRACTIONAL || bo->branch_rule > SOV_BRANCH_FULL_STRONG) {
    sov::set_error("sov_bnb_solve: unknown branch_rule");
    return SOV_ERR_ARG;
  }
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  const auto t0 = Clock::now();
  const int nv = milp->n_vars;
  const std::vector<double> lb0(milp->lb, milp->lb + nv), ub0(milp->ub, milp->ub + nv);
  const bool batched = bo->child_solver == SOV_CHILD_BATCHED && batched_form_ok(lb0.data(), nv);
  int lp_solves = 0;

  // Root cut rounds (apply_cuts at depth 0, branch_and_bound.cpp:1214-1273):
  // separate at the root LP's point, gate, append, re-solve; a round whose
  // LP does not solve is taken back and ends the loop (the reference keeps
  // the pre-cut LP there). No incumbent exists yet, so the gate checks the
  // bound box only. The tree then searches the model with the cuts.
  res->root_bound_before_cuts = std::numeric_limits<double>::quiet_NaN();
  sov::OwnedLpModel with_cuts;
  const sov_lp_model* mdl = milp;
  if (bo->max_cut_rounds > 0) {
    std::vector<int8_t> kind(static_cast<std::size_t>(nv));
    for (int j = 0; j < nv; ++j)
      kind[j] = !is_int[j] ? sov::kContinuous : (lb0[j] >= 0.0 && ub0[j] <= 1.0 ? sov::kBinary : sov::kInteger);
    sov::CutModel cm;
    int rc = sov::cut_model_from(*milp, kind.data(), cm);
    if (rc != SOV_OK) {
      sov::set_error("sov_bnb_solve: malformed model");
      return rc;
    }
    const std::size_t base_rows = cm.rows.size();
    const std::size_t row_cap = batched ? 272 : std::numeric_limits<std::size_t>::max();
    std::vector<double> x(static_cast<std::size_t>(nv));
    sov_ipm_result r{};
    if ((rc = sov_lp_solve(milp, opt, &r, x.data(), nullptr, nullptr)) < 0) return rc;
    ++lp_solves;
    if (r.status == 0) {
      res->root_bound_before_cuts = r.objective;
      std::vector<int> kept;  // family of each cut in the model
      for (int round = 0; round < bo->max_cut_rounds && cm.rows.size() < row_cap; ++round) {
        const std::size_t before = cm.rows.size();
        std::vector<int> fam;
        int rejected = 0;
        sov::apply_cuts(cm, x.data(), bo->integer_tol, bo->max_cuts_per_node, {}, base_rows, bo->cmir_at_root != 0,
                        &rejected, &fam);
        res->cut_rejections += rejected;
        if (cm.rows.size() > row_cap) {
          cm.rows.resize(row_cap);
          fam.resize(row_cap - before);
        }
        if (cm.rows.size() == before) break;
        with_cuts.build(*milp, cm);
        std::vector<double> xc(static_cast<std::size_t>(nv));
        sov_ipm_result rc2{};
        if ((rc = sov_lp_solve(&with_cuts.model, opt, &rc2, xc.data(), nullptr, nullptr)) < 0) return rc;
        ++lp_solves;
        if (rc2.status != 0) {
          cm.rows.resize(before);
          break;
        }
        kept.insert(kept.end(), fam.begin(), fam.end());
        ++res->cut_rounds;
        x = std::move(xc);
      }
      for (int f : kept) {
        ++res->cuts_added;
        if (f == sov::kCutCover) ++res->cuts_cover;
        else if (f == sov::kCutGomory) ++res->cuts_gomory;
        else ++res->cuts_cmir;
      }
      if (cm.rows.size() > base_rows) {
        with_cuts.build(*milp, cm);
        mdl = &with_cuts.model;
      }
    }
  }
  res->root_rows = mdl->n_rows;
  const DenseForm F = dense_form(*mdl, lb0.data(), ub0.data());
  const int nslack = F.n - F.nv;
  if (batched && F.m > 272) {
    sov::set_error("sov_bnb_solve: batched child LPs need at most 272 rows");
    return SOV_ERR_ARG;
  }
  std::unique_ptr<sov_batch, int (*)(sov_batch*)> batch(nullptr, sov_batch_free);
  int32_t mat_id = 0;
  if (batched) {
    sov_batch* b = nullptr;
    int rc = sov_batch_create(&b);
    if (rc != SOV_OK) return rc;
    batch.reset(b);
    rc = sov_batch_add_matrix(b, F.m, F.n, F.A.data(), &mat_id);
    if (rc != SOV_OK) return rc;
  }
  double kernel_ms = 0.0;
  const NodePresolve presolve(*mdl);
  WarmOpts warm;
  warm.on = bo->warm_start != 0;
  warm.save_mu = bo->warm_mu > 0.0 ? bo->warm_mu : 0.1;
  long long lp_iterations = 0;
  auto solve_set = [&](std::vector<ChildSpec> kids, std::vector<ChildOut>& outv) {
    int rc;
    if (batched) {
      // presolve test on a host thread beside the GPU solve (as in
      // sov_strong_branch); a child it proves infeasible gets status 2
      std::vector<char> pre_bad;
      std::thread pre;
      try {
        pre = std::thread([&] { pre_bad = presolve.infeasible(kids); });
      } catch (const std::system_error&) {
        pre_bad = presolve.infeasible(kids);
      }
      rc = solve_children(batch.get(), mat_id, F, kids, *opt, outv, &kernel_ms, &lp_solves, warm);
      if (pre.joinable()) pre.join();
      if (rc == SOV_OK)
        for (std::size_t k = 0; k < kids.size(); ++k)
          if (pre_bad[k]) {
            kids[k].empty = true;
            outv[k] = ChildOut{};
            outv[k].status = 2;
            --lp_solves;  // settled by the presolve test, not by its LP
          }
      if (rc == SOV_OK) rc = prove_infeasible(presolve, kids, outv);
      if (rc == SOV_OK) rc = finish_unconverged(*mdl, kids, F, *opt, outv, &lp_solves);
    } else {
      presolve.mark(kids);
      rc = solve_children_one_by_one(*mdl, kids, F, *opt, outv, &lp_solves);
      if (rc == SOV_OK) rc = prove_infeasible(presolve, kids, outv);
    }
    for (const ChildOut& o : outv) lp_iterations += o.iterations;
    return rc;
  };
  auto node_spec = [&](const std::vector<double>& lb, const std::vector<double>& ub,
                       std::shared_ptr<const WarmState> from = nullptr) {
    ChildSpec s;
    s.lb = lb;
    s.ub = ub;
    s.warm = std::move(from);
    s.lb.resize(static_cast<std::size_t>(F.n), 0.0);
    s.ub.resize(static_cast<std::size_t>(F.n), 1e30);
    return s;
  };
  bool any_continuous = false;
  for (int j = 0; j < nv; ++j) any_continuous = any_continuous || !is_int[j];

  // Open nodes carry their branching bounds on top of the root's
  // (SearchNode, branch_and_bound.cpp:58-71). Bounds and objectives are kept
  // in the minimisation sense (F.sense * model objective).
  struct BoundChange {
    int var;
    double lower, upper;
  };
  struct OpenNode {
    double bound = 0.0;
    int depth = 0;
    long long id = 0;
    std::vector<BoundChange> bounds;
    int branch_var = -1;
    bool branch_up = false;
    double branch_distance = 0.0, parent_obj = 0.0;
    std::shared_ptr<const ChildOut> solved;  // the LP, when strong branching solved it
    std::shared_ptr<const WarmState> parent_state;  // the parent LP's iterate (warm start of this node's LP)
  };
  auto worse = [](const OpenNode& a, const OpenNode& b) { return a.bound != b.bound ? a.bound > b.bound : a.id > b.id; };
  std::priority_queue<OpenNode, std::vector<OpenNode>, decltype(worse)> open(worse);
  std::vector<OpenNode> plunge;  // the preferred child, popped next (branch_and_bound.cpp:1014-1048)
  long long next_id = 1;
  {
    OpenNode root;
    root.bound = -std::numeric_limits<double>::infinity();
    root.id = next_id++;
    open.push(std::move(root));
  }
  auto rel_gap = [](double bound, double inc) { return (inc - bound) / std::max(1.0, std::abs(inc)); };
  std::vector<PseudoCost> pseudo(static_cast<std::size_t>(nv));
  bool has_inc = false, any_error = false, limit_hit = false, have_root = false;
  double incumbent = 0.0;  // min sense
  std::vector<double> best_x;
  long long nodes = 0, node_lps = 0, strong_lps = 0, strong_used = 0;
  double sb_seconds = 0.0, lp_seconds = 0.0;
  int max_depth = 0, failures = 0;

  auto offer = [&](std::vector<double> x, double obj_min) -> int {  // offer_incumbent + polish_incumbent
    if (has_inc && !(obj_min < incumbent - 1e-12)) return SOV_OK;
    if (any_continuous) {
      std::vector<double> lb = lb0, ub = ub0;
      for (int j = 0; j < nv; ++j)
        if (is_int[j]) lb[j] = ub[j] = std::round(x[j]);
      std::vector<ChildOut> r;
      const int rc = solve_set({node_spec(lb, ub)}, r);
      if (rc != SOV_OK) return rc;
      if (r[0].status == 0) {
        for (int j = 0; j < nv; ++j) x[j] = is_int[j] ? lb[j] : r[0].x[j];
        obj_min = F.sense * r[0].objective;
      }
      if (has_inc && !(obj_min < incumbent - 1e-12)) return SOV_OK;
    }
    has_inc = true;
    incumbent = obj_min;
    best_x = std::move(x);
    return SOV_OK;
  };

  while (!(plunge.empty() && open.empty())) {
    if (nodes >= bo->max_nodes || (bo->time_limit_seconds > 0 &&
                                   std::chrono::duration<double>(Clock::now() - t0).count() >= bo->time_limit_seconds)) {
      limit_hit = true;
      break;
    }
    OpenNode node;
    if (!plunge.empty()) {
      node = std::move(plunge.back());
      plunge.pop_back();
    } else {
      node = open.top();
      open.pop();
    }
    ++nodes;
    max_depth = std::max(max_depth, node.depth);
    if (has_inc && rel_gap(node.bound, incumbent) <= bo->mip_gap) continue;  // objective cutoff
    std::vector<double> lb = lb0, ub = ub0;
    for (const BoundChange& bc : node.bounds) {
      lb[bc.var] = bc.lower;
      ub[bc.var] = bc.upper;
    }
    ChildOut lp;
    if (node.solved) {
      lp = *node.solved;
    } else {
      const auto tl = Clock::now();
      std::vector<ChildOut> r;
      const int rc = solve_set({node_spec(lb, ub, node.parent_state)}, r);
      if (rc != SOV_OK) return rc;
      lp = std::move(r[0]);
      ++node_lps;
      lp_seconds += std::chrono::duration<double>(Clock::now() - tl).count();
    }
    if (lp.status == 2) continue;  // infeasible
    if (!have_root) {
      have_root = true;
      res->root_bound = lp.status == 0 ? lp.objective : 0.0;
    }
    const double lp_obj = F.sense * lp.objective;  // min sense
    if (lp.status == 0 && node.branch_var >= 0 && node.branch_distance > bo->integer_tol) {
      PseudoCost& p = pseudo[node.branch_var];  // node-LP pseudocost learning (:1159-1170)
      const double unit = std::abs(lp_obj - node.parent_obj) / node.branch_distance;
      if (node.branch_up) {
        p.up_sum += unit;
        ++p.up_count;
      } else {
        p.down_sum += unit;
        ++p.down_count;
      }
    }
    if (lp.status != 0) {  // no certified bound: the subtree is dropped, optimality not certified
      any_error = true;
      ++failures;
      continue;
    }
    std::vector<int> frac;
    for (int j = 0; j < nv; ++j)
      if (is_int[j] && std::abs(lp.x[j] - std::round(lp.x[j])) > bo->integer_tol) frac.push_back(j);
    auto snapped = [&]() {  // snap_integer_primal
      std::vector<double> x = lp.x;
      for (int j = 0; j < nv; ++j)
        if (is_int[j] && std::abs(x[j] - std::round(x[j])) <= 10 * bo->integer_tol) x[j] = std::round(x[j]);
      return x;
    };
    if (frac.empty()) {  // integrality before bound pruning (:1317)
      const int rc = offer(snapped(), lp_obj);
      if (rc != SOV_OK) return rc;
      continue;
    }
    if (has_inc && rel_gap(lp_obj, incumbent) <= bo->mip_gap) continue;

    const double parent = lp.objective;  // model sense, as the reference's parent_obj
    PairSolver pairs = [&](const std::vector<int>& vars, std::vector<ChildOut>& outv) {
      std::vector<ChildSpec> kids;
      children_specs(lb, ub, vars, lp.x.data(), nslack, kids);
      for (ChildSpec& k : kids) k.warm = lp.state;
      return solve_set(kids, outv);
    };
    Decision d;
    int rc = SOV_OK;
    const auto ts = Clock::now();
    switch (bo->branch_rule) {
      case SOV_BRANCH_RELIABILITY:
        d = pick_reliability(frac, lp.x.data(), parent, bo->integer_tol, bo->reliability_threshold,
                             bo->max_strong_per_node, pseudo, pairs, &rc, &strong_lps, &strong_used);
        break;
      case SOV_BRANCH_FULL_STRONG: {
        std::vector<double> key;
        if (bo->candidate_order) {
          key.assign(static_cast<std::size_t>(nv), 0.0);
          bo->candidate_order(nv, lp.x.data(), lb.data(), ub.data(), key.data(), bo->candidate_order_user);
        }
        long long solved_now = 0;
        d = pick_full_strong(frac, lp.x.data(), parent, bo->integer_tol, bo->strong_branch_candidates, key, pseudo,
                             pairs, &rc, &solved_now);
        strong_lps += solved_now;
        strong_used += solved_now;
        break;
      }
      case SOV_BRANCH_PSEUDOCOST:
        d.var = pick_pseudo_cost(frac, lp.x.data(), pseudo);
        break;
      default:
        d.var = pick_most_fractional(frac, lp.x.data());
    }
    sb_seconds += std::chrono::duration<double>(Clock::now() - ts).count();
    if (rc != SOV_OK) return rc;
    if (d.var < 0) {
      rc = offer(snapped(), lp_obj);
      if (rc != SOV_OK) return rc;
      continue;
    }
    const int v = d.var;
    const double val = lp.x[v], fl = std::floor(val), ce = std::ceil(val);
    const bool prefer_up = val - fl >= 0.5;  // plunge toward the nearer integer
    for (const bool up : {prefer_up, !prefer_up}) {
      const double lower = up ? std::max(lb[v], ce) : lb[v];
      const double upper = up ? ub[v] : std::min(ub[v], fl);
      if (lower > upper + 1e-12) continue;
      OpenNode child;
      child.bounds = node.bounds;
      auto same = std::find_if(child.bounds.begin(), child.bounds.end(), [&](const BoundChange& b) { return b.var == v; });
      if (same != child.bounds.end()) *same = BoundChange{v, lower, upper};
      else child.bounds.push_back(BoundChange{v, lower, upper});
      child.id = next_id++;
      child.depth = node.depth + 1;
      child.bound = lp_obj;
      child.branch_var = v;
      child.branch_up = up;
      child.branch_distance = up ? ce - val : val - fl;
      child.parent_obj = lp_obj;
      if (d.have_children) child.solved = std::make_shared<const ChildOut>(up ? d.up : d.down);
      child.parent_state = lp.state;
      if (bo->plunging && up == prefer_up) plunge.push_back(std::move(child));
      else open.push(std::move(child));
    }
  }
  res->nodes = nodes;
  res->lp_solves = lp_solves;
  res->node_lp_solves = node_lps;
  res->strong_lp_solves = strong_lps;
  res->strong_lp_used = strong_used;
  res->strong_branch_seconds = sb_seconds;
  res->node_lp_seconds = lp_seconds;
  res->kernel_ms = kernel_ms;
  res->max_depth = max_depth;
  res->node_lp_failures = failures;
  res->lp_iterations = lp_iterations;
  res->seconds = std::chrono::duration<double>(Clock::now() - t0).count();
  if (!has_inc) {
    res->status = any_error ? 6 : limit_hit ? 5 : 2;
    return SOV_OK;
  }
  res->status = (limit_hit || any_error) ? 1 : 0;
  res->objective = F.sense * incumbent;
  if (x_best)
    for (int j = 0; j < nv; ++j) x_best[j] = best_x[j];
  return SOV_OK;
}

}  // extern "C"
