/* CPU oracle for the interior-point linear-algebra path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain C++ restatement of the reference
 * solver's CPU algorithms (Mehrotra IPM, DenseLU, CSC SpMV, normal equations,
 * minimum-degree ordering and up-looking LDL^T). It exists to check the HIP
 * library; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it, and never as the thing that is measured or shipped.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef SOV_ORACLE_H
#define SOV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A general-form LP, the array image of sovereign::OptimizationModel
 * (solver/core/include/sovereign/types.hpp:32-59) restricted to LP:
 * constraint matrix in CSC (n_rows x n_vars), one sense per row
 * ('L' <=, 'G' >=, 'E' =), bounds per variable (ub >= 1e29 or inf = none). */
typedef struct {
  int32_t n_vars;
  int32_t n_rows;
  const int32_t* col_ptr; /* n_vars + 1 */
  const int32_t* row_idx; /* nnz */
  const double* values;   /* nnz */
  const double* obj;      /* n_vars */
  double obj_constant;
  int32_t maximize;
  const double* lb; /* n_vars */
  const double* ub; /* n_vars */
  const char* row_sense; /* n_rows */
  const double* rhs;     /* n_rows */
} orc_lp_model;

typedef struct {
  int32_t max_iterations;      /* 250 */
  double feasibility_tol;      /* 1e-9 */
  double optimality_tol;       /* 1e-9 */
  double fraction_to_boundary; /* 0.999 */
  int32_t enable_scaling;      /* 1 */
  int32_t normal_mode;         /* 0 auto, 1 dense LU, 2 sparse LDL^T */
} orc_ipm_options;

/* status follows sovereign::SolverStatus (types.hpp:20-30):
 * 0 Optimal, 1 Feasible, 6 NumericalError, 7 Error. */
typedef struct {
  int32_t status;
  int32_t iterations;
  double objective;
  double duality_gap;
  double primal_residual;
  double dual_residual;
  int32_t factorizations;
  int32_t used_sparse; /* 1 when the last factorization was the sparse LDL^T */
  double factor_seconds;
  double total_seconds;
} orc_ipm_result;

void orc_default_options(orc_ipm_options* opt);

/* InteriorPointSolver::solve (interior_point.cpp:747-757): build_ipm_form then
 * solve_ipm. primal: n_vars, row_dual: n_rows, var_dual: n_vars (any may be
 * NULL). Returns 0, or -1 on a malformed model (status is then 7). */
int orc_lp_solve(const orc_lp_model* model, const orc_ipm_options* opt,
                 orc_ipm_result* res, double* primal, double* row_dual, double* var_dual);

/* build_ipm_form (interior_point.cpp:224-409) exposed as a handle. */
typedef struct orc_form orc_form;
orc_form* orc_form_build(const orc_lp_model* model, int enable_scaling);
void orc_form_dims(const orc_form* f, int32_t* m, int32_t* n, int32_t* nnz, int32_t* n_structural);
void orc_form_get(const orc_form* f, int32_t* col_ptr, int32_t* row_idx, double* values,
                  double* b, double* c, double* col_scale, double* row_scale, double* shift);
void orc_form_free(orc_form* f);

/* solve_ipm (interior_point.cpp:454-740) on an LP already in IPM form
 * min c'x, Ax = b, x >= 0. x, s: n; y: m (outputs, may be NULL). The
 * objective reported is c'x of this form. */
int orc_solve_ipm_form(int32_t m, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                       const double* values, const double* b, const double* c,
                       const orc_ipm_options* opt, orc_ipm_result* res, double* x, double* y,
                       double* s);

/* spmv_csc_cpu / spmv_csc_transpose_cpu (gpu_spmv.cpp:36-67). */
void orc_spmv_csc(int32_t nrows, int32_t ncols, const int32_t* col_ptr, const int32_t* row_idx,
                  const double* values, const double* x, double* y, int transpose);

/* DenseLU (dense_lu.cpp:9-108), column-major a[j*n+i]. factor returns 0 ok,
 * 1 singular (largest remaining pivot < 1e-14). */
int orc_dense_lu_factor(int32_t n, double* a, int32_t* piv);
void orc_dense_lu_solve(int32_t n, const double* lu, const int32_t* piv, double* x);
void orc_dense_lu_solve_transpose(int32_t n, const double* lu, const int32_t* piv, double* x);

/* build_normal_eq (interior_point.cpp:74-101): full column-major m x m
 * A diag(d) A^T + 1e-12 I. */
void orc_normal_eq_dense(int32_t m, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                         const double* values, const double* d, double* M);

/* build_normal_eq_pattern / build_normal_eq_values (sparse_symmetric.cpp:39-169)
 * as a two-call API: pattern_nnz returns the count, pattern fills it. */
int64_t orc_normal_eq_pattern_nnz(int32_t m, int32_t n, const int32_t* col_ptr,
                                  const int32_t* row_idx);
void orc_normal_eq_pattern(int32_t m, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                           int32_t* p_col_ptr, int32_t* p_row_idx);

/* NormalEquations (interior_point.cpp:129-212) on the CPU device policy. */
typedef struct orc_ne orc_ne;
orc_ne* orc_ne_create(int32_t m, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                      const double* values, int32_t normal_mode);
int orc_ne_factor(orc_ne* h, const double* d);      /* 1 ok, 0 failed */
int orc_ne_solve(const orc_ne* h, double* rhs);     /* 1 ok, 0 failed */
int orc_ne_is_sparse(const orc_ne* h);
int64_t orc_ne_factor_nnz(const orc_ne* h);
int32_t orc_ne_replaced_pivots(const orc_ne* h);
void orc_ne_free(orc_ne* h);

/* SparseLDLT (sparse_ldlt.cpp:19-223) on a lower-triangle pattern with its
 * values: symbolic analysis (unlimited fill), numeric factorization with
 * regularization `reg` and tiny-pivot threshold `tiny` (0 = off), then a solve
 * of rhs in place (rhs may be NULL). Returns 0 ok, 1 factorization failed,
 * -1 analysis failed. perm (n) is the elimination order. */
int orc_ldlt_factor_solve(int32_t n, const int32_t* p_col_ptr, const int32_t* p_row_idx,
                          const double* values, double reg, double tiny, double* rhs,
                          int32_t* replaced, int64_t* factor_nnz, int32_t* perm);

/* AMDOrdering::compute (amd_ordering.cpp:12-130) on a lower-triangle
 * pattern; returns 1 ok, 0 when the fill limit (0 = none) is exceeded. */
int orc_amd_order(int32_t n, const int32_t* p_col_ptr, const int32_t* p_row_idx,
                  int64_t fill_limit, int32_t* perm, int32_t* iperm);

/* Structured separable QP  min 0.5|x|^2 - target'x,  x_r - 2x_{r+1} + x_{r+2} >= 0
 * (liswet_oracle.cpp; solve_liswet_structured_ipm, qp_interior_point.cpp:2378-2826,
 * and certify_liswet, :1410-1480). */
int orc_liswet_ipm(int32_t n, const double* target, int32_t max_iterations, int32_t mode,
                   double* x_out);
int orc_liswet_certify(int32_t n, const double* target, const double* x, double* lambda_out, double* out7);
int orc_liswet_polish(int32_t n, const double* target, const double* s, const double* z, double* x_out);

#ifdef __cplusplus
}
#endif

#endif
