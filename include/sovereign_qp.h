/* sovereign_qp.h - the structured separable convex QP on the GPU (C ABI).
 *
 *     min 0.5 ||x||^2 - target' x   s.t.  x_r - 2 x_{r+1} + x_{r+2} >= 0,
 *                                          r = 0 .. n-3, x free
 *
 * This is the LISWET shape that the reference's QP solver detects
 * (detect_liswet_structured, solver/qp/src/qp_interior_point.cpp:892-967)
 * and solves through ConvexQpSolver::solve -> solve_liswet_structured_gi /
 * solve_liswet_structured_ipm (:3153-3161). sov_qp_liswet_solve replaces
 * that dispatch for this shape: one call takes the target vector (the
 * negated linear term of the model) and returns x, the multipliers and the
 * certificate that certify_liswet (:1410-1480) computes.
 *
 * Method: the reference's x-space interior-point iterations
 * (solve_liswet_structured_ipm, :2378-2826) in one persistent GPU workgroup,
 * the pentadiagonal normal matrix I + A' diag(z/s) A solved by block cyclic
 * reduction of its 2 x 2-block tridiagonal form, factored in FP32 with FP64
 * refinement (SOV_PREC_FP32) or in FP64 throughout (SOV_PREC_FP64), then a
 * knot-space polish on the host in long double: a primal active-set method
 * over the rows' active set, warm-started from the interior point's guess
 * {r : s_r < z_r} as polish_liswet_knots (:2269-2377) is, each candidate the
 * least-squares piecewise-linear fit of solve_liswet_knot_projection
 * (:1509-1580), and the result certified by certify_liswet's checks.
 * Status codes as in sovereign_hip.h. */
#ifndef SOVEREIGN_QP_H
#define SOVEREIGN_QP_H

#include "sovereign_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t max_iterations; /* IPM iterations (reference default 200) */
  int32_t precision;      /* SOV_PREC_FP32 (mixed) or SOV_PREC_FP64 */
  double tolerance;       /* <= 0: 1e-10, 1e-12 when n - 2 <= 50 (the reference's structured tolerances) */
  int32_t polish;         /* 1: knot-space polish from the last iterate (default) */
  int32_t pad;
} sov_qp_options;

typedef struct {
  int32_t status;      /* 0: certified optimal; 6: the certificate failed */
  int32_t ipm_status;  /* 0 reached the tolerance, 1 stopped (iteration cap), 6 numerical failure */
  int32_t iterations;
  int32_t polish_fits; /* knot fits computed by the polish */
  int32_t fp32_factorizations, fp64_factorizations;
  int32_t refinements;
  int32_t pad;
  double objective;    /* 0.5 |x|^2 - target' x */
  double primal;       /* max(0, -(A x)_r) */
  double stationarity; /* |x - target - A' lambda|_inf */
  double complementarity, min_lambda, consistency;
  double ipm_primal, ipm_dual, ipm_complementarity; /* scaled residuals of the last iterate */
  double factor_ms;    /* part of kernel_ms spent in factorizations */
  double solve_ms;     /* part of kernel_ms spent in factor solves */
  double kernel_ms;    /* GPU time of the interior-point launch */
  double seconds;      /* wall time of the call */
} sov_qp_result;

void sov_qp_default_options(sov_qp_options* opt);

/* Solves the QP for target[0..n) (n >= 3). x_out (n) and lambda_out (n - 2)
 * may be NULL. opt NULL means the defaults. */
int sov_qp_liswet_solve(int32_t n, const double* target, const sov_qp_options* opt, sov_qp_result* res,
                        double* x_out, double* lambda_out);

/* The host polish alone, from an interior point (s, z: n - 2 entries each)
 * of the same QP: the polish step of the solver above (the role of
 * polish_liswet_knots, :2269-2377), for callers that run their own
 * interior-point iterations. Needs no device. Fills status, polish_fits,
 * objective, the certificate fields and seconds of res. */
int sov_qp_liswet_polish(int32_t n, const double* target, const double* s, const double* z, sov_qp_result* res,
                         double* x_out, double* lambda_out);

/* ----------------------------------------------------------- generic QP --
 * QpInteriorPointSolver::solve (solver/qp/include/sovereign/qp_interior_point.hpp:
 * 24-31) for a convex QP that is not of the structured shape: the generic
 * branch, solve_qp_ipm (solver/qp/src/qp_interior_point.cpp:2848-3116) on
 * the form of build_qp_form (:118-247), first from the default starting
 * point, then from Mehrotra's when that fails (:3166-3180).
 *
 * Model: minimize (or maximize) 1/2 sum_t q_t x_{q_row[t]} x_{q_col[t]} +
 * obj' x + obj_constant, the quadratic terms as the reference stores them (a
 * diagonal term enters Q_ii, an off-diagonal term adds q/2 to Q_ij and to
 * Q_ji), rows of a CSC matrix with senses 'L' / 'G' / 'E', bounds (lb <=
 * -1e29: free, ub >= 1e29: none).
 *
 * Method (GPU, FP64): the Newton system [Q + X^-1 S, -A^T; A, 0] of QpKkt
 * (:398-656) in Schur-complement form: K = Q + X^-1 S + delta I by the
 * library's blocked dense Cholesky on MFMA, S = A K^-1 A^T + delta I as the
 * MFMA product V'V with V' = A L^-T (block triangular solve through the
 * factor's diagonal-block inverses), then S by the same Cholesky. delta
 * starts at 1e-9 and is raised 100-fold on a breakdown up to 1e-5 (the
 * reference's quasi-definite regularisation, :386-388); each solve is
 * refined against the unregularised system at most 3 times, a step kept
 * only when it halves the residual (:469-505). Status codes are
 * SolverStatus: 0 Optimal, 7 Error (factorization or iteration limit). */
typedef struct {
  int32_t n_vars;
  int32_t n_rows;
  const int32_t* col_ptr;
  const int32_t* row_idx;
  const double* values;
  const double* obj;
  double obj_constant;
  int32_t maximize;
  int32_t pad;
  const double* lb;
  const double* ub;
  const char* row_sense;
  const double* rhs;
  int64_t q_nnz;
  const int32_t* q_row;
  const int32_t* q_col;
  const double* q_val;
} sov_qp_model;

/* QpInteriorPointOptions (qp_interior_point.hpp:7-18). */
typedef struct {
  int32_t max_iterations;      /* 100 */
  int32_t iterative_refinement; /* 1 */
  double feasibility_tol;      /* 1e-8 */
  double optimality_tol;       /* 1e-8 */
  double fraction_to_boundary; /* 0.999 */
  int32_t common_step;         /* 1: equal primal and dual steps when Q != 0 */
  int32_t retry_mehrotra_start; /* 1 */
} sov_qp_ipm_options;

typedef struct {
  int32_t status;     /* 0 Optimal, 7 Error */
  int32_t iterations; /* both attempts when the second one ran */
  double objective;   /* of the model, constant included */
  double duality_gap;
  double primal_residual;
  double dual_residual;
  int32_t factorizations;     /* K and S factorizations */
  int32_t refinement_steps;   /* accepted refinement steps */
  int32_t mehrotra_start;     /* 1 when the answer came from the second attempt */
  int32_t n_form;             /* columns of the interior-point form (slacks included) */
  int32_t m_form;             /* rows of the form (upper-bound rows included) */
  int32_t pad;
  double regularization;      /* delta at the end */
  double seconds;
} sov_qp_ipm_result;

void sov_qp_ipm_default_options(sov_qp_ipm_options* opt);

/* primal (n_vars), row_dual (the form's m_form rows: the model's rows, then
 * one per finite upper bound) and var_dual (n_vars) may be NULL. */
int sov_qp_ipm_solve(const sov_qp_model* model, const sov_qp_ipm_options* opt, sov_qp_ipm_result* res,
                     double* primal, double* row_dual, double* var_dual);

#ifdef __cplusplus
}
#endif

#endif /* SOVEREIGN_QP_H */
