// Cut separation and the validity gate (see cuts.hpp). The separation rules,
// tolerances and limits are the reference's (solver/cuts/src/cuts.cpp), so a
// node separates the same cuts; the code works on index arrays instead of
// name maps. Rows are scanned in model order and each row's terms in the
// order they are stored; where the reference sorts (cover terms by
// coefficient, cMIR candidates by efficacy) ties keep that order.

#include "cuts.hpp"

#include "internal.hpp"
#include "sovereign_cuts.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace sov {
namespace {

bool is_integer(int8_t k) { return k == kInteger || k == kBinary; }

}  // namespace

std::vector<Cut> separate_covers(const CutModel& m, const double* x, double int_tol, int max_cuts) {
  std::vector<Cut> out;
  struct Item {
    int var;
    double a;
  };
  std::vector<Item> items;
  for (const CutRow& r : m.rows) {
    if (r.sense != 'L') continue;
    // knapsack: the binary terms with a > 0 against rhs minus the least
    // activity of every other term (which must be bounded on that side)
    items.clear();
    double cap = r.rhs;
    bool bounded = true;
    for (std::size_t t = 0; t < r.idx.size(); ++t) {
      const int j = r.idx[t];
      const double a = r.val[t];
      if (m.kind[j] == kBinary && a > 1e-12) {
        items.push_back({j, a});
        continue;
      }
      const double at = a > 0.0 ? m.lb[j] : m.ub[j];
      if (std::abs(at) >= 1e20) {
        bounded = false;
        break;
      }
      cap -= a * at;
    }
    if (!bounded || items.size() < 2) continue;
    std::stable_sort(items.begin(), items.end(), [](const Item& p, const Item& q) { return p.a > q.a; });
    std::size_t k = 0;
    double weight = 0.0;
    while (k < items.size()) {
      weight += items[k++].a;
      if (weight > cap + 1e-9) break;
    }
    if (weight <= cap + 1e-9 || k < 2) continue;
    double lhs = 0.0;
    for (std::size_t t = 0; t < k; ++t) lhs += x[items[t].var];
    const double bound = static_cast<double>(k) - 1.0;
    if (lhs <= bound + int_tol) continue;  // not violated
    Cut c;
    c.family = kCutCover;
    c.row.sense = 'L';
    c.row.rhs = bound;
    for (std::size_t t = 0; t < k; ++t) {
      c.row.idx.push_back(items[t].var);
      c.row.val.push_back(1.0);
    }
    out.push_back(std::move(c));
    if (static_cast<int>(out.size()) >= max_cuts) break;
  }
  return out;
}

std::vector<Cut> separate_gomory(const CutModel& m, const double* x, double int_tol, int max_cuts) {
  std::vector<Cut> out;
  for (const CutRow& r : m.rows) {
    if (r.sense != 'L') continue;
    // floor(a_j) x_j <= a_j x_j must hold over each column's range: a
    // column that may go negative disqualifies the row unless its
    // coefficient is integral and non-negative; a continuous column (which
    // is dropped, not rounded) must never contribute negatively
    bool usable = true, fractional_coef = false;
    for (std::size_t t = 0; t < r.idx.size() && usable; ++t) {
      const int j = r.idx[t];
      const double a = r.val[t];
      const double fa = std::floor(a + 1e-12);
      const bool fractional = std::abs(a - fa) > 1e-9;
      const bool may_be_negative = m.lb[j] < -1e-9;
      if (may_be_negative && (fractional || a < 0.0)) usable = false;
      else if (!is_integer(m.kind[j]) && may_be_negative) usable = false;
      else if (!is_integer(m.kind[j]) && a < 0.0) usable = false;
      fractional_coef = fractional_coef || fractional;
    }
    if (!usable) continue;
    Cut c;
    c.family = kCutGomory;
    c.row.sense = 'L';
    c.row.rhs = std::floor(r.rhs + 1e-12);
    double violation = -c.row.rhs;
    for (std::size_t t = 0; t < r.idx.size(); ++t) {
      const int j = r.idx[t];
      if (!is_integer(m.kind[j])) continue;
      const double fa = std::floor(r.val[t] + 1e-12);
      if (std::abs(fa) > 1e-12) {
        c.row.idx.push_back(j);
        c.row.val.push_back(fa);
      }
      violation += fa * x[j];
    }
    if (!fractional_coef || c.row.idx.empty() || violation <= int_tol) continue;
    out.push_back(std::move(c));
    if (static_cast<int>(out.size()) >= max_cuts) break;
  }
  return out;
}

std::vector<Cut> separate_cmir(const CutModel& m, const double* x, double int_tol, int max_cuts,
                               std::size_t base_rows) {
  constexpr double kInf = 1e20;
  // one term of the row over a non-negative substitute: x - l, or u - x when
  // the LP point is nearer the upper bound
  struct Term {
    int var;
    bool integer, upper;
    double bound, coef, point;
  };
  std::vector<Cut> found;
  std::vector<Term> terms;
  std::vector<double> g, best_g;

  auto separate = [&](const CutRow& r, double sign) {
    terms.clear();
    double b = sign * r.rhs;
    bool fractional_point = false;
    for (std::size_t t = 0; t < r.idx.size(); ++t) {
      const double a = sign * r.val[t];
      if (a == 0.0) continue;
      const int j = r.idx[t];
      const bool integer = is_integer(m.kind[j]);
      double lo = m.lb[j], hi = m.ub[j];
      if (integer) {
        lo = lo > -kInf ? std::ceil(lo - 1e-9) : -kInf;
        hi = hi < kInf ? std::floor(hi + 1e-9) : kInf;
      }
      const double xv = x[j];
      if (integer && std::abs(xv - std::round(xv)) > int_tol) fractional_point = true;
      const bool has_lo = lo > -kInf, has_hi = hi < kInf;
      if (!has_lo && !has_hi) return;
      const bool upper = !has_lo || (has_hi && hi - xv < xv - lo);
      Term tm{j, integer, upper, upper ? hi : lo, upper ? -a : a, upper ? hi - xv : xv - lo};
      b -= a * tm.bound;
      terms.push_back(tm);
    }
    if (!fractional_point) return;
    // divisors: 1 and the coefficients of integer terms off their bound
    std::vector<double> deltas{1.0};
    for (const Term& tm : terms)
      if (tm.integer && tm.point > int_tol && std::abs(tm.coef) > 1e-6) deltas.push_back(std::abs(tm.coef));
    std::sort(deltas.begin(), deltas.end());
    deltas.erase(std::unique(deltas.begin(), deltas.end(),
                             [](double p, double q) { return std::abs(p - q) <= 1e-9 * std::max(1.0, q); }),
                 deltas.end());
    if (deltas.size() > 8) deltas.resize(8);
    // the MIR inequality for divisor delta: its coefficients in g, its
    // right-hand side in rhs_out, its efficacy (or -1 when unusable)
    auto build = [&](double delta, double& rhs_out) -> double {
      const double beta = b / delta;
      const double f0 = beta - std::floor(beta);
      if (f0 < 0.01 || f0 > 0.99 || std::abs(beta) > 1e9) return -1.0;
      g.assign(terms.size(), 0.0);
      rhs_out = std::floor(beta);
      double activity = 0.0, norm = 0.0;
      for (std::size_t i = 0; i < terms.size(); ++i) {
        const Term& tm = terms[i];
        double gi = 0.0;
        if (tm.integer) {
          const double q = tm.coef / delta;
          const double fj = q - std::floor(q);
          gi = std::floor(q) + std::max(0.0, fj - f0) / (1.0 - f0);
        } else if (tm.coef < 0.0) {
          gi = tm.coef / (delta * (1.0 - f0));
        }
        g[i] = gi;
        activity += gi * tm.point;
        norm += gi * gi;
      }
      if (norm <= 1e-18) return -1.0;
      return (activity - rhs_out) / std::sqrt(norm);
    };
    double best = -1.0, best_rhs = 0.0, best_delta = 0.0;
    for (double d : deltas) {
      double rhs_d = 0.0;
      const double e = build(d, rhs_d);
      if (e > best) {
        best = e;
        best_rhs = rhs_d;
        best_delta = d;
        best_g = g;
      }
    }
    if (best_delta > 0.0) {
      const double base = best_delta;
      for (double scale : {0.5, 0.25, 0.125}) {
        double rhs_d = 0.0;
        const double e = build(base * scale, rhs_d);
        if (e > best) {
          best = e;
          best_rhs = rhs_d;
          best_g = g;
        }
      }
    }
    if (best < 1e-4) return;
    // back to the original variables
    Cut c;
    c.family = kCutCmir;
    c.efficacy = best;
    c.row.sense = 'L';
    double rhs = best_rhs, largest = 0.0, smallest = std::numeric_limits<double>::infinity();
    for (std::size_t i = 0; i < terms.size(); ++i) {
      const double gi = best_g[i];
      if (gi == 0.0) continue;
      const Term& tm = terms[i];
      const double coef = tm.upper ? -gi : gi;
      rhs += tm.upper ? -gi * tm.bound : gi * tm.bound;
      c.row.idx.push_back(tm.var);
      c.row.val.push_back(coef);
      largest = std::max(largest, std::abs(coef));
      smallest = std::min(smallest, std::abs(coef));
    }
    if (c.row.idx.empty() || largest > 1e6 * smallest || std::abs(rhs) > 1e12) return;
    c.row.rhs = rhs + 1e-9 * std::max(1.0, std::abs(rhs));  // rounding error goes to validity's side
    found.push_back(std::move(c));
  };

  const std::size_t nr = std::min(base_rows, m.rows.size());
  for (std::size_t i = 0; i < nr; ++i) {
    const CutRow& r = m.rows[i];
    if (r.sense != 'G') separate(r, 1.0);
    if (r.sense != 'L') separate(r, -1.0);
  }
  std::stable_sort(found.begin(), found.end(), [](const Cut& p, const Cut& q) { return p.efficacy > q.efficacy; });
  if (static_cast<int>(found.size()) > max_cuts) found.resize(static_cast<std::size_t>(std::max(0, max_cuts)));
  return found;
}

int cut_rejection(const CutModel& m, const CutRow& cut, const std::vector<std::vector<double>>& points, double tol) {
  // 1. no known integer-feasible point may be cut off
  for (const auto& p : points) {
    double lhs = 0.0;
    for (std::size_t t = 0; t < cut.idx.size(); ++t) lhs += cut.val[t] * p[cut.idx[t]];
    const double viol = cut.sense == 'L' ? lhs - cut.rhs : cut.sense == 'G' ? cut.rhs - lhs : std::abs(lhs - cut.rhs);
    if (viol > tol) return 1;
  }
  // 2. some point of the bound box must satisfy it
  double lo_sum = 0.0, hi_sum = 0.0;
  for (std::size_t t = 0; t < cut.idx.size(); ++t) {
    const int j = cut.idx[t];
    const double c = cut.val[t];
    const double lo = c * (c > 0.0 ? m.lb[j] : m.ub[j]);
    const double hi = c * (c > 0.0 ? m.ub[j] : m.lb[j]);
    lo_sum += std::max(lo, -1e30);
    hi_sum += std::min(hi, 1e30);
  }
  if ((cut.sense != 'G' && lo_sum > cut.rhs + tol) || (cut.sense != 'L' && hi_sum < cut.rhs - tol)) return 2;
  return 0;
}

int apply_cuts(CutModel& m, const double* x, double int_tol, int max_cuts,
               const std::vector<std::vector<double>>& points, std::size_t base_rows, bool with_cmir,
               int* rejected, std::vector<int>* added_families) {
  std::vector<Cut> cand = separate_covers(m, x, int_tol, 8);
  if (with_cmir) {
    std::vector<Cut> c = separate_cmir(m, x, int_tol, 8, base_rows);
    cand.insert(cand.end(), std::make_move_iterator(c.begin()), std::make_move_iterator(c.end()));
  }
  {
    std::vector<Cut> c = separate_gomory(m, x, int_tol, 8);
    cand.insert(cand.end(), std::make_move_iterator(c.begin()), std::make_move_iterator(c.end()));
  }
  int added = 0;
  for (Cut& c : cand) {
    if (added >= max_cuts) break;
    if (cut_rejection(m, c.row, points, std::max(int_tol, 1e-6)) != 0) {
      if (rejected) ++*rejected;
      continue;
    }
    if (added_families) added_families->push_back(c.family);
    m.rows.push_back(std::move(c.row));
    ++added;
  }
  return added;
}

// The model's rows, each row's terms in increasing variable order, a
// repeated entry summed.
int cut_model_from(const sov_lp_model& md, const int8_t* kind, CutModel& out) {
  const int nv = md.n_vars, nr = md.n_rows;
  if (nv <= 0 || nr < 0 || !md.col_ptr || (nr > 0 && (!md.row_sense || !md.rhs)) || !md.lb || !md.ub || !kind)
    return SOV_ERR_ARG;
  out.lb.assign(md.lb, md.lb + nv);
  out.ub.assign(md.ub, md.ub + nv);
  out.kind.assign(kind, kind + nv);
  out.rows.assign(static_cast<std::size_t>(nr), CutRow{});
  for (int i = 0; i < nr; ++i) {
    const char s = md.row_sense[i];
    out.rows[i].sense = (s == 'G' || s == 'g') ? 'G' : (s == 'E' || s == 'e') ? 'E' : 'L';
    out.rows[i].rhs = md.rhs[i];
  }
  for (int j = 0; j < nv; ++j)
    for (int p = md.col_ptr[j]; p < md.col_ptr[j + 1]; ++p) {
      const int i = md.row_idx[p];
      if (i < 0 || i >= nr) return SOV_ERR_ARG;
      CutRow& r = out.rows[i];
      if (!r.idx.empty() && r.idx.back() == j) r.val.back() += md.values[p];
      else {
        r.idx.push_back(j);
        r.val.push_back(md.values[p]);
      }
    }
  return SOV_OK;
}

void OwnedLpModel::build(const sov_lp_model& base, const CutModel& m) {
  const int nv = base.n_vars, nr = static_cast<int>(m.rows.size());
  col_ptr.assign(static_cast<std::size_t>(nv) + 1, 0);
  for (const CutRow& r : m.rows)
    for (int j : r.idx) ++col_ptr[j + 1];
  for (int j = 0; j < nv; ++j) col_ptr[j + 1] += col_ptr[j];
  row_idx.assign(static_cast<std::size_t>(col_ptr[nv]), 0);
  values.assign(row_idx.size(), 0.0);
  std::vector<int32_t> next(col_ptr.begin(), col_ptr.end() - 1);
  rhs.resize(static_cast<std::size_t>(nr));
  sense.assign(static_cast<std::size_t>(nr), 'L');
  for (int i = 0; i < nr; ++i) {
    const CutRow& r = m.rows[i];
    for (std::size_t t = 0; t < r.idx.size(); ++t) {
      const int32_t q = next[r.idx[t]]++;
      row_idx[q] = i;
      values[q] = r.val[t];
    }
    rhs[i] = r.rhs;
    sense[i] = r.sense;
  }
  model = base;
  model.n_rows = nr;
  model.col_ptr = col_ptr.data();
  model.row_idx = row_idx.data();
  model.values = values.data();
  model.row_sense = sense.data();
  model.rhs = rhs.data();
}

}  // namespace sov

extern "C" {

int sov_separate_cuts(const sov_lp_model* model, const int8_t* var_kind, const double* x, int32_t family,
                      double int_tol, int32_t limit, int32_t base_rows, sov_cut_list* out) {
  if (!model || !x || !out || (family != SOV_CUT_COVER && family != SOV_CUT_GOMORY && family != SOV_CUT_CMIR))
    return SOV_ERR_ARG;
  sov::CutModel m;
  int rc = sov::cut_model_from(*model, var_kind, m);
  if (rc != SOV_OK) {
    sov::set_error("sov_separate_cuts: bad model");
    return rc;
  }
  std::vector<sov::Cut> cuts;
  if (family == SOV_CUT_COVER) cuts = sov::separate_covers(m, x, int_tol, limit);
  else if (family == SOV_CUT_GOMORY) cuts = sov::separate_gomory(m, x, int_tol, limit);
  else
    cuts = sov::separate_cmir(m, x, int_tol, limit,
                              base_rows < 0 ? m.rows.size() : static_cast<std::size_t>(base_rows));
  std::size_t nnz = 0;
  for (const auto& c : cuts) nnz += c.row.idx.size();
  out->n_cuts = static_cast<int32_t>(cuts.size());
  out->nnz = static_cast<int64_t>(nnz);
  if (static_cast<int64_t>(cuts.size()) > out->max_cuts || static_cast<int64_t>(nnz) > out->max_nnz) {
    sov::set_error("sov_separate_cuts: the output list is too small (n_cuts and nnz hold the sizes needed)");
    return SOV_ERR_ARG;
  }
  if (!out->row_ptr || (nnz > 0 && (!out->col || !out->val)) ||
      (!cuts.empty() && (!out->sense || !out->rhs || !out->family || !out->efficacy)))
    return SOV_ERR_ARG;
  int64_t p = 0;
  out->row_ptr[0] = 0;
  for (std::size_t k = 0; k < cuts.size(); ++k) {
    const sov::Cut& c = cuts[k];
    for (std::size_t t = 0; t < c.row.idx.size(); ++t, ++p) {
      out->col[p] = c.row.idx[t];
      out->val[p] = c.row.val[t];
    }
    out->row_ptr[k + 1] = static_cast<int32_t>(p);
    out->sense[k] = c.row.sense;
    out->rhs[k] = c.row.rhs;
    out->family[k] = c.family;
    out->efficacy[k] = c.efficacy;
  }
  return SOV_OK;
}

int sov_cut_validity(const sov_lp_model* model, int32_t nnz, const int32_t* col, const double* val, char sense,
                     double rhs, int32_t n_points, const double* points, double tol, int32_t* verdict) {
  if (!model || !verdict || nnz < 0 || (nnz > 0 && (!col || !val)) || n_points < 0 || (n_points > 0 && !points))
    return SOV_ERR_ARG;
  const int nv = model->n_vars;
  if (nv <= 0 || !model->lb || !model->ub) return SOV_ERR_ARG;
  sov::CutModel m;
  m.lb.assign(model->lb, model->lb + nv);
  m.ub.assign(model->ub, model->ub + nv);
  m.kind.assign(static_cast<std::size_t>(nv), sov::kContinuous);
  sov::CutRow cut;
  for (int32_t t = 0; t < nnz; ++t) {
    if (col[t] < 0 || col[t] >= nv) return SOV_ERR_ARG;
    cut.idx.push_back(col[t]);
    cut.val.push_back(val[t]);
  }
  cut.sense = (sense == 'G' || sense == 'g') ? 'G' : (sense == 'E' || sense == 'e') ? 'E' : 'L';
  cut.rhs = rhs;
  std::vector<std::vector<double>> pts(static_cast<std::size_t>(n_points));
  for (int32_t k = 0; k < n_points; ++k)
    pts[k].assign(points + static_cast<std::size_t>(k) * nv, points + static_cast<std::size_t>(k + 1) * nv);
  *verdict = sov::cut_rejection(m, cut, pts, tol);
  return SOV_OK;
}

}  // extern "C"
