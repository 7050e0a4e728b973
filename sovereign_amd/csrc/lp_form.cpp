// General LP -> interior-point form, solve on the GPU, map the result back.
//
// build_ipm_form (solver/lp/src/interior_point.cpp:224-409): lower bounds are
// shifted out (b absorbs A*lb), every finite upper bound becomes an extra <=
// row, <= rows get a slack and >= rows a surplus column, duplicate entries of
// a column are merged and zeros dropped, then two passes of square-root row
// and column equilibration. publish_iterate (:521-569) undoes the scaling and
// the shift. This runs once per solve on the host; it is not on the hot path.

#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <utility>
#include <vector>

namespace sov {


namespace {

struct IpmForm {
  int m = 0, n = 0, n_struct = 0;
  std::vector<int32_t> cptr, crow;
  std::vector<double> cval, b, c, shift, col_scale, row_scale;
  std::vector<int> row_origin;
};

void to_ipm_form(const sov_lp_model& mdl, bool scaling, IpmForm& F) {
  const int nv = mdl.n_vars, m0 = mdl.n_rows;
  F.n_struct = nv;
  F.shift.assign(static_cast<std::size_t>(nv), 0.0);
  std::vector<int> ub_var;
  std::vector<double> ub_width;
  for (int v = 0; v < nv; ++v) {
    F.shift[v] = mdl.lb[v];
    if (std::isfinite(mdl.ub[v]) && mdl.ub[v] < 1e29) {
      const double w = mdl.ub[v] - mdl.lb[v];
      if (w < -1e-12) throw std::runtime_error("inconsistent bounds");
      ub_var.push_back(v);
      ub_width.push_back(w);
    }
  }
  const int mt = m0 + static_cast<int>(ub_var.size());
  F.m = mt;
  F.b.assign(static_cast<std::size_t>(mt), 0.0);
  F.row_scale.assign(static_cast<std::size_t>(mt), 1.0);
  F.row_origin.assign(static_cast<std::size_t>(mt), -1);
  for (int r = 0; r < m0; ++r) {
    F.row_origin[r] = r;
    F.b[r] = mdl.rhs[r];
  }
  // Entries per structural column (row, value), ub rows appended.
  std::vector<std::vector<std::pair<int, double>>> col(static_cast<std::size_t>(nv));
  for (int v = 0; v < nv; ++v) {
    for (int p = mdl.col_ptr[v]; p < mdl.col_ptr[v + 1]; ++p) {
      const int r = mdl.row_idx[p];
      if (r < 0 || r >= m0) throw std::runtime_error("row index out of range");
      const double a = mdl.values[p];
      if (a != 0.0) col[v].emplace_back(r, a);
      F.b[r] -= a * F.shift[v];
    }
  }
  for (std::size_t k = 0; k < ub_var.size(); ++k) {
    const int r = m0 + static_cast<int>(k);
    col[ub_var[k]].emplace_back(r, 1.0);
    F.b[r] = ub_width[k];
  }
  // Slack (<=, +1) or surplus (>=, -1) per inequality row, in row order.
  std::vector<std::pair<int, double>> slack;
  for (int r = 0; r < mt; ++r) {
    const char sense = r < m0 ? mdl.row_sense[r] : 'L';
    if (sense == 'L' || sense == 'l') slack.emplace_back(r, 1.0);
    else if (sense == 'G' || sense == 'g') slack.emplace_back(r, -1.0);
  }
  const int nt = nv + static_cast<int>(slack.size());
  F.n = nt;
  F.c.assign(static_cast<std::size_t>(nt), 0.0);
  const double sgn = mdl.maximize ? -1.0 : 1.0;
  for (int v = 0; v < nv; ++v) F.c[v] = sgn * mdl.obj[v];
  F.cptr.assign(static_cast<std::size_t>(nt) + 1, 0);
  for (int j = 0; j < nt; ++j) {
    std::vector<std::pair<int, double>> e;
    if (j < nv) e = std::move(col[j]);
    else e.push_back(slack[j - nv]);
    std::sort(e.begin(), e.end());
    for (std::size_t k = 0; k < e.size();) {
      double sum = e[k].second;
      std::size_t l = k + 1;
      while (l < e.size() && e[l].first == e[k].first) sum += e[l++].second;
      if (std::abs(sum) > 0.0) {
        F.crow.push_back(e[k].first);
        F.cval.push_back(sum);
      }
      k = l;
    }
    F.cptr[j + 1] = static_cast<int32_t>(F.cval.size());
  }
  F.col_scale.assign(static_cast<std::size_t>(nt), 1.0);
  if (!scaling || mt == 0 || nt == 0) return;
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<double> rmax(static_cast<std::size_t>(mt), 0.0);
    for (std::size_t p = 0; p < F.cval.size(); ++p) rmax[F.crow[p]] = std::max(rmax[F.crow[p]], std::abs(F.cval[p]));
    for (int r = 0; r < mt; ++r) {
      const double f = rmax[r] > 0.0 ? 1.0 / std::sqrt(rmax[r]) : 1.0;
      rmax[r] = f;
      F.row_scale[r] *= f;
      F.b[r] *= f;
    }
    for (std::size_t p = 0; p < F.cval.size(); ++p) F.cval[p] *= rmax[F.crow[p]];
    for (int j = 0; j < nt; ++j) {
      double cm = 0.0;
      for (int p = F.cptr[j]; p < F.cptr[j + 1]; ++p) cm = std::max(cm, std::abs(F.cval[p]));
      const double f = cm > 0.0 ? 1.0 / std::sqrt(cm) : 1.0;
      for (int p = F.cptr[j]; p < F.cptr[j + 1]; ++p) F.cval[p] *= f;
      F.c[j] *= f;
      F.col_scale[j] *= f;
    }
  }
}

}  // namespace
}  // namespace sov

extern "C" int sov_lp_solve(const sov_lp_model* model, const sov_ipm_options* opt, sov_ipm_result* res, double* primal,
                            double* row_dual, double* var_dual) {
  if (!model || !opt || !res) return SOV_ERR_ARG;
  *res = sov_ipm_result{};
  res->status = 7;
  if (!sov::device_ready()) {
    sov::set_error("no usable GPU");
    return SOV_ERR_NODEV;
  }
  sov::IpmForm F;
  try {
    sov::to_ipm_form(*model, opt->enable_scaling != 0, F);
  } catch (const std::exception& ex) {
    sov::set_error(std::string("IPM error: ") + ex.what());
    return SOV_ERR_ARG;
  }
  if (F.n <= 0) {
    sov::set_error("IPM: empty model");
    return SOV_ERR_ARG;
  }
  if (F.m <= 0) {  // interior_point.cpp:464-478: nothing constrains x beyond its bounds
    double obj = model->obj_constant;
    for (int v = 0; v < F.n_struct; ++v) {
      if (primal) primal[v] = F.shift[v];
      obj += model->obj[v] * F.shift[v];
    }
    res->status = 0;
    res->objective = obj;
    return SOV_OK;
  }
  sov_mat* A = nullptr;
  int rc = sov::mat_upload(F.m, F.n, F.cptr.data(), F.crow.data(), F.cval.data(), &A);
  if (rc != SOV_OK) return rc;
  std::vector<double> x(static_cast<std::size_t>(F.n)), y(static_cast<std::size_t>(F.m)), s(static_cast<std::size_t>(F.n));
  sov_ipm* h = nullptr;
  rc = sov_ipm_create_path(A, opt->normal_path, &h);
  if (rc == SOV_OK) rc = sov_ipm_solve(h, F.b.data(), F.c.data(), opt, res, x.data(), y.data(), s.data());
  sov_ipm_free(h);
  sov::mat_release(A);
  if (rc != SOV_OK) return rc;
  double obj = model->obj_constant;
  for (int v = 0; v < F.n_struct; ++v) {
    const double val = F.shift[v] + F.col_scale[v] * x[v];
    if (primal) primal[v] = val;
    if (var_dual) var_dual[v] = s[v] / F.col_scale[v];
    obj += model->obj[v] * val;
  }
  if (row_dual)
    for (int i = 0; i < F.m; ++i)
      if (F.row_origin[i] >= 0) row_dual[F.row_origin[i]] = F.row_scale[i] * y[i];
  res->objective = obj;
  return SOV_OK;
}
