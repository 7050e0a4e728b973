// C entry points over the reference solver's own CPU code, compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`).
//
// TEST INFRASTRUCTURE ONLY: used in this container to pin the oracle
// restatement (sov_oracle.cpp) and to write golden fixtures; the result lands
// in oracle/_ref/, which is never committed.
//
// Each function builds the reference's own types from plain arrays and calls
// the reference API:
//   InteriorPointSolver::solve   solver/lp/include/sovereign/interior_point.hpp:28-36
//   DenseLU                      solver/numerical/include/sovereign/dense_lu.hpp:9-23
//   spmv_csc_auto                solver/gpu/include/sovereign/gpu_spmv.hpp:22-31
//   build_normal_eq_pattern/values, SparseLDLT, AMDOrdering (numerical/)
//   QpInteriorPointSolver::solve solver/qp/include/sovereign/qp_interior_point.hpp:24-31
//   generate_cover/mir/cmir_cuts, check_cut_validity  solver/cuts/include/sovereign/cuts.hpp:17-55

#include "sov_oracle.h"

#include "sovereign/amd_ordering.hpp"
#include "sovereign/cuts.hpp"
#include "sovereign/dense_lu.hpp"
#include "sovereign/gpu_spmv.hpp"
#include "sovereign/interior_point.hpp"
#include "sovereign/qp_interior_point.hpp"
#include "sovereign/sparse_ldlt.hpp"
#include "sovereign/sparse_matrix.hpp"
#include "sovereign/sparse_symmetric.hpp"
#include "sovereign/types.hpp"

#include <cstdlib>
#include <string>
#include <vector>

using namespace sovereign;

namespace {

std::string var_name(int j) { return "v" + std::to_string(j); }

OptimizationModel to_model(const orc_lp_model& m) {
  OptimizationModel model;
  model.problem_type = ProblemType::LP;
  model.sense = m.maximize ? Sense::Maximize : Sense::Minimize;
  model.objective.constant = m.obj_constant;
  for (int j = 0; j < m.n_vars; ++j) {
    Variable v;
    v.name = var_name(j);
    v.lower_bound = m.lb[j];
    v.upper_bound = m.ub[j];
    model.variables.push_back(v);
    if (m.obj[j] != 0.0) model.objective.linear[v.name] = m.obj[j];
  }
  model.constraints.resize(static_cast<std::size_t>(m.n_rows));
  for (int r = 0; r < m.n_rows; ++r) {
    Constraint& c = model.constraints[static_cast<std::size_t>(r)];
    c.name = "r" + std::to_string(r);
    const char s = m.row_sense[r];
    c.sense = s == 'G' ? ConstraintSense::Ge : s == 'E' ? ConstraintSense::Eq : ConstraintSense::Le;
    c.rhs = m.rhs[r];
  }
  for (int j = 0; j < m.n_vars; ++j)
    for (int p = m.col_ptr[j]; p < m.col_ptr[j + 1]; ++p)
      model.constraints[static_cast<std::size_t>(m.row_idx[p])].linear[var_name(j)] += m.values[p];
  return model;
}

int status_code(SolverStatus s) { return static_cast<int>(s); }

SparseMatrixCSC to_csc(int m, int n, const int32_t* cp, const int32_t* ri, const double* v) {
  SparseMatrixCSC A;
  A.resize(static_cast<std::size_t>(m), static_cast<std::size_t>(n));
  for (int j = 0; j < n; ++j) {
    for (int p = cp[j]; p < cp[j + 1]; ++p) A.push_back(ri[p], v[p]);
    A.finish_column(static_cast<std::size_t>(j));
  }
  return A;
}

}  // namespace

extern "C" {

int ref_lp_solve(const orc_lp_model* m, const orc_ipm_options* o, orc_ipm_result* res,
                 double* primal, double* row_dual, double* var_dual) {
  setenv("SOVEREIGN_DEVICE", "cpu", 1);
  setenv("SOVEREIGN_IPM_NORMAL_EQUATIONS",
         o->normal_mode == 1 ? "dense" : o->normal_mode == 2 ? "sparse" : "auto", 1);
  InteriorPointOptions opt;
  opt.max_iterations = o->max_iterations;
  opt.feasibility_tol = o->feasibility_tol;
  opt.optimality_tol = o->optimality_tol;
  opt.fraction_to_boundary = o->fraction_to_boundary;
  opt.enable_scaling = o->enable_scaling != 0;
  const SolverResult r = InteriorPointSolver(opt).solve(to_model(*m));
  *res = orc_ipm_result{};
  res->status = status_code(r.status);
  res->iterations = static_cast<int32_t>(r.iterations);
  res->objective = r.objective_value;
  res->duality_gap = r.duality_gap;
  res->primal_residual = r.primal_residual;
  res->dual_residual = r.dual_residual;
  auto it = r.lp_diagnostics.find("ipm");
  if (it != r.lp_diagnostics.end()) res->factorizations = static_cast<int32_t>(it->second.refactorizations);
  res->used_sparse = r.message.find("sparse LDL") != std::string::npos ? 1 : 0;
  for (int j = 0; j < m->n_vars; ++j) {
    auto pv = r.primal.find(var_name(j));
    if (primal) primal[j] = pv == r.primal.end() ? 0.0 : pv->second;
    auto dv = r.dual.find(var_name(j));
    if (var_dual) var_dual[j] = dv == r.dual.end() ? 0.0 : dv->second;
  }
  if (row_dual) {
    for (int i = 0; i < m->n_rows; ++i) {
      auto dv = r.dual.find("row_" + std::to_string(i));
      row_dual[i] = dv == r.dual.end() ? 0.0 : dv->second;
    }
  }
  return 0;
}

// DenseLU::factorize + solve (dense_lu.cpp:9-80). Returns 0 ok, 1 singular.
int ref_dense_lu_solve(int32_t n, const double* a, double* x, int transpose) {
  setenv("SOVEREIGN_DEVICE", "cpu", 1);
  const std::size_t N = static_cast<std::size_t>(n);
  DenseLU lu;
  if (!lu.factorize(std::vector<double>(a, a + N * N), N)) return 1;
  std::vector<double> v(x, x + n);
  const bool ok = transpose ? lu.solve_transpose(v) : lu.solve(v);
  for (int i = 0; i < n; ++i) x[i] = v[static_cast<std::size_t>(i)];
  return ok ? 0 : 1;
}

void ref_spmv_csc(int32_t nrows, int32_t ncols, const int32_t* cp, const int32_t* ri, const double* v,
                  const double* x, double* y, int transpose) {
  setenv("SOVEREIGN_DEVICE", "cpu", 1);
  const int nnz = cp[ncols];
  std::vector<int> p(cp, cp + ncols + 1), i(ri, ri + nnz);
  std::vector<double> val(v, v + nnz), xv(x, x + (transpose ? nrows : ncols)), yv;
  spmv_csc_auto(static_cast<std::size_t>(nrows), static_cast<std::size_t>(ncols), p, i, val, xv, yv,
                transpose != 0);
  for (std::size_t k = 0; k < yv.size(); ++k) y[k] = yv[k];
}

int64_t ref_normal_eq_pattern(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri,
                              int32_t* pcp, int32_t* pri) {
  std::vector<double> ones(static_cast<std::size_t>(cp[n]), 1.0);
  const SparseSymmetricPattern P = build_normal_eq_pattern(to_csc(m, n, cp, ri, ones.data()));
  if (pcp) for (std::size_t k = 0; k < P.col_ptr.size(); ++k) pcp[k] = P.col_ptr[k];
  if (pri) for (std::size_t k = 0; k < P.row_idx.size(); ++k) pri[k] = P.row_idx[k];
  return static_cast<int64_t>(P.row_idx.size());
}

// build_normal_eq_values + SparseLDLT (AMD-ordered, tiny-pivot 1e-13, reg 1e-12)
// + solve, as NormalEquations does on its sparse path (interior_point.cpp:154-182).
int ref_sparse_normal_solve(int32_t m, int32_t n, const int32_t* cp, const int32_t* ri,
                            const double* v, const double* d, double* rhs, int64_t* lnnz) {
  const SparseMatrixCSC A = to_csc(m, n, cp, ri, v);
  const SparseSymmetricPattern P = build_normal_eq_pattern(A);
  SparseLDLT ldlt;
  if (!ldlt.symbolic_analyze(P, 0)) return -1;
  ldlt.set_tiny_pivot_threshold(1e-13);
  std::vector<double> vals;
  build_normal_eq_values(A, std::vector<double>(d, d + n), P, vals);
  if (!ldlt.numeric_factor(vals, 1e-12)) return 1;
  std::vector<double> x(rhs, rhs + m);
  if (!ldlt.solve(x)) return 1;
  for (int i = 0; i < m; ++i) rhs[i] = x[static_cast<std::size_t>(i)];
  if (lnnz) *lnnz = static_cast<int64_t>(ldlt.factor_nnz());
  return 0;
}

int ref_amd_order(int32_t nn, const int32_t* pcp, const int32_t* pri, int64_t limit, int32_t* perm,
                  int32_t* iperm) {
  SparseSymmetricPattern P;
  P.n = static_cast<std::size_t>(nn);
  P.col_ptr.assign(pcp, pcp + nn + 1);
  P.row_idx.assign(pri, pri + pcp[nn]);
  AMDOrdering amd;
  amd.set_fill_limit(static_cast<std::size_t>(limit));
  std::vector<int> pv, ipv;
  const bool ok = amd.compute(P, pv, ipv);
  for (std::size_t k = 0; k < pv.size(); ++k) {
    perm[k] = pv[k];
    iperm[k] = ipv[k];
  }
  return ok ? 1 : 0;
}

// The reference's QpInteriorPointSolver on a dense model: A (m x n,
// row-major), senses 'L'/'G'/'E', Q (n x n, row-major, symmetric) of
// 1/2 x'Qx, stored as one term per nonzero entry (a term (i, j) adds q/2 to
// Q_ij and Q_ji, so each symmetric pair is stored twice with its own value).
// out6: status, iterations, objective, duality gap, primal and dual residual.
int ref_qp_solve(int32_t n, int32_t m, const double* A, const char* senses, const double* rhs,
                 const double* lb, const double* ub, const double* c, const double* Q, int32_t maximize,
                 double constant, double* out6, double* primal) {
  OptimizationModel model;
  model.problem_type = ProblemType::QP;
  model.sense = maximize ? Sense::Maximize : Sense::Minimize;
  model.objective.constant = constant;
  for (int j = 0; j < n; ++j) {
    Variable v;
    v.name = var_name(j);
    v.lower_bound = lb[j];
    v.upper_bound = ub[j];
    model.variables.push_back(v);
    if (c[j] != 0.0) model.objective.linear[v.name] = c[j];
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double q = Q[static_cast<std::size_t>(i) * n + j];
      if (q != 0.0) model.objective.quadratic[var_name(i)][var_name(j)] = q;
    }
  for (int r = 0; r < m; ++r) {
    Constraint cons;
    cons.name = "r" + std::to_string(r);
    cons.sense = senses[r] == 'G' ? ConstraintSense::Ge : senses[r] == 'E' ? ConstraintSense::Eq : ConstraintSense::Le;
    cons.rhs = rhs[r];
    for (int j = 0; j < n; ++j) {
      const double a = A[static_cast<std::size_t>(r) * n + j];
      if (a != 0.0) cons.linear[var_name(j)] = a;
    }
    model.constraints.push_back(cons);
  }
  const SolverResult r = QpInteriorPointSolver().solve(model);
  out6[0] = static_cast<double>(status_code(r.status));
  out6[1] = static_cast<double>(r.iterations);
  out6[2] = r.has_objective_value ? r.objective_value : 0.0;
  out6[3] = r.duality_gap;
  out6[4] = r.primal_residual;
  out6[5] = r.dual_residual;
  for (int j = 0; j < n; ++j) {
    auto it = r.primal.find(var_name(j));
    primal[j] = it == r.primal.end() ? 0.0 : it->second;
  }
  return 0;
}

// The reference's cut generators on the model's rows at x. kind: 0
// continuous, 1 integer, 2 binary. family 1 cover, 2 Gomory, 4 cMIR. Cuts
// come back in CSR form over variable indices (capacities max_cuts / max_nnz;
// the sizes needed are reported either way); returns -1 when they do not fit.
int ref_separate_cuts(const orc_lp_model* m, const int8_t* kind, const double* x, int32_t family, double int_tol,
                      int32_t limit, int32_t base_rows, int32_t max_cuts, int64_t max_nnz, int32_t* n_cuts,
                      int64_t* nnz, int32_t* row_ptr, int32_t* col, double* val, double* rhs) {
  OptimizationModel model = to_model(*m);
  model.problem_type = ProblemType::MILP;
  for (int j = 0; j < m->n_vars; ++j)
    model.variables[static_cast<std::size_t>(j)].type =
        kind[j] == 2 ? VariableType::Binary : kind[j] == 1 ? VariableType::Integer : VariableType::Continuous;
  std::unordered_map<std::string, double> point;
  for (int j = 0; j < m->n_vars; ++j) point[var_name(j)] = x[j];
  std::vector<Cut> cuts;
  if (family == 1) cuts = generate_cover_cuts(model, point, int_tol, limit);
  else if (family == 2) cuts = generate_mir_cuts(model, point, int_tol, limit);
  else
    cuts = generate_cmir_cuts(model, point, int_tol, limit,
                              base_rows < 0 ? static_cast<std::size_t>(-1) : static_cast<std::size_t>(base_rows));
  int64_t total = 0;
  for (const auto& c : cuts) total += static_cast<int64_t>(c.constraint.linear.size());
  *n_cuts = static_cast<int32_t>(cuts.size());
  *nnz = total;
  if (*n_cuts > max_cuts || total > max_nnz) return -1;
  int64_t p = 0;
  row_ptr[0] = 0;
  for (std::size_t k = 0; k < cuts.size(); ++k) {
    for (const auto& kv : cuts[k].constraint.linear) {
      col[p] = std::atoi(kv.first.c_str() + 1);  // "v<j>"
      val[p] = kv.second;
      ++p;
    }
    row_ptr[k + 1] = static_cast<int32_t>(p);
    rhs[k] = cuts[k].constraint.rhs;
  }
  return 0;
}

// check_cut_validity on one cut: 0 accepted, 1 a reference point is cut
// off, 2 the bound box cannot satisfy it (the reason string's kind).
int ref_cut_validity(const orc_lp_model* m, int32_t nnz, const int32_t* col, const double* val, char sense, double rhs,
                     int32_t n_points, const double* points, double tol) {
  OptimizationModel model = to_model(*m);
  Constraint cut;
  cut.name = "c";
  cut.sense = sense == 'G' ? ConstraintSense::Ge : sense == 'E' ? ConstraintSense::Eq : ConstraintSense::Le;
  cut.rhs = rhs;
  for (int t = 0; t < nnz; ++t) cut.linear[var_name(col[t])] += val[t];
  std::vector<std::unordered_map<std::string, double>> pts(static_cast<std::size_t>(n_points));
  for (int k = 0; k < n_points; ++k)
    for (int j = 0; j < m->n_vars; ++j) pts[k][var_name(j)] = points[static_cast<std::size_t>(k) * m->n_vars + j];
  const std::string why = check_cut_validity(model, cut, pts, tol);
  if (why.empty()) return 0;
  return why.find("removes a known") != std::string::npos ? 1 : 2;
}

}  // extern "C"
