/* Cut separation of the tree search, as host entry points (no device needed).
 *
 * Replaces, for callers that build their own node models, the reference's
 * generators and validity gate (solver/cuts/include/sovereign/cuts.hpp:17-55):
 *   generate_cover_cuts  -> sov_separate_cuts(..., SOV_CUT_COVER, ...)
 *   generate_mir_cuts    -> sov_separate_cuts(..., SOV_CUT_GOMORY, ...)
 *   generate_cmir_cuts   -> sov_separate_cuts(..., SOV_CUT_CMIR, ...)
 *   check_cut_validity   -> sov_cut_validity
 * The tree search of sov_bnb_solve (sovereign_bnb.h) applies them at the
 * root through apply_cuts' gate (solver/milp/src/branch_and_bound.cpp:
 * 648-694).
 *
 * Variables are identified by index; var_kind gives each one's type as the
 * reference's VariableType: 0 continuous, 1 integer, 2 binary (covers use
 * binaries only). Status: 0 ok, < 0 error (sov_last_error()). */
#ifndef SOVEREIGN_CUTS_H
#define SOVEREIGN_CUTS_H

#include "sovereign_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SOV_CUT_COVER 1
#define SOV_CUT_GOMORY 2
#define SOV_CUT_CMIR 4

/* A list of cuts in CSR form, filled by sov_separate_cuts. The caller sets
 * max_cuts and max_nnz and provides row_ptr (max_cuts + 1), col and val
 * (max_nnz), sense, rhs, family and efficacy (max_cuts). Every cut is a
 * '<=' row; efficacy is the cMIR separation score (0 for the others). */
typedef struct {
  int32_t max_cuts;
  int32_t n_cuts; /* out */
  int64_t max_nnz;
  int64_t nnz;    /* out */
  int32_t* row_ptr;
  int32_t* col;
  double* val;
  char* sense;
  double* rhs;
  int32_t* family;
  double* efficacy;
} sov_cut_list;

/* One generator over the model's rows at the point x (n_vars): covers and
 * Gomory rounding over '<=' rows, cMIR over the first base_rows rows (-1: all)
 * on both sides of equalities. limit is the generator's max_cuts (the
 * reference's default is 8), int_tol its violation tolerance (1e-6). When
 * the list is too small, n_cuts and nnz still report the sizes needed and
 * the call returns SOV_ERR_ARG. */
int sov_separate_cuts(const sov_lp_model* model, const int8_t* var_kind, const double* x, int32_t family,
                      double int_tol, int32_t limit, int32_t base_rows, sov_cut_list* out);

/* The validity gate for one cut sum val[t] x[col[t]] (sense) rhs: *verdict =
 * 0 accepted, 1 it removes one of the known integer-feasible points (points:
 * n_points x n_vars, row-major), 2 no point within the model's variable bounds
 * satisfies it. tol as the reference's (max(integer_tol, 1e-6) in the tree). */
int sov_cut_validity(const sov_lp_model* model, int32_t nnz, const int32_t* col, const double* val, char sense,
                     double rhs, int32_t n_points, const double* points, double tol, int32_t* verdict);

#ifdef __cplusplus
}
#endif

#endif /* SOVEREIGN_CUTS_H */
