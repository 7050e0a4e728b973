/* sovereign_hip.h - C ABI of the MI355X (gfx950) linear-algebra core of the
 * Sovereign LP solver's interior-point path.
 *
 * Plain C: pointers, sizes and int status codes; no C++ or torch types cross
 * this boundary and nothing throws. Host pointers are named *_host (or are
 * host pointers when not marked); device pointers are named *_dev.
 *
 * Status codes: 0 ok, 1 singular / numerical breakdown, < 0 errors.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference checkout).
 */
#ifndef SOVEREIGN_HIP_H
#define SOVEREIGN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOV_OK 0
#define SOV_SINGULAR 1
#define SOV_ERR_ARG (-1)
#define SOV_ERR_DEVICE (-2)
#define SOV_ERR_NOMEM (-3)
#define SOV_ERR_NODEV (-4)

/* ---------------------------------------------------------------- device --
 * Replaces cuda::DeviceInfo / cuda::device_info()
 * (solver/gpu/include/sovereign/cuda_driver.hpp:14-27, probe at
 * solver/gpu/src/cuda_driver.cpp:661-750). Probes once per process: opens
 * device 0, checks the kernels on small problems (the self-test of
 * cuda_driver.cpp:616-659) and reports why the GPU is unusable otherwise.
 * SOVEREIGN_DISABLE_HIP=1 disables the device. */
typedef struct {
  int32_t available;
  char name[256];
  char arch[64];
  int32_t compute_major;
  int32_t compute_minor;
  uint64_t memory_bytes;
  int32_t driver_version;
  int32_t multiprocessors;
  char reason[256];
} sov_dev_info;

int sov_dev_probe(sov_dev_info* out);

/* Per-process counters: gpu_operations / gpu_factorizations /
 * reset_gpu_operations (solver/gpu/include/sovereign/gpu_spmv.hpp:89-92). */
uint64_t sov_gpu_operations(void);
uint64_t sov_gpu_factorizations(void);
void sov_reset_gpu_operations(void);

/* Last error text of the calling thread ("" when none). */
const char* sov_last_error(void);

/* Synchronises the library stream (all work is issued on it). */
int sov_sync(void);
/* The library's HIP stream (hipStream_t), for callers that time kernels. */
void* sov_stream(void);

/* Kernel timing: with timing on (1), the factorization, solve and SpMV
 * kernels are bracketed by HIP events on the streams they run on (recording
 * does not wait; the totals are collected after one stream
 * synchronisation); 2 times only the kernels whose own launches record the
 * events (the bulk trailing-update SYRK and the diagonal-block kernel), one
 * launch in out[23] of each, which keeps event traffic off the
 * factorization's critical path (launch counts and times are those of the
 * timed launches).
 * sov_kernel_times fills out[0..23] with the totals since the last call and
 * resets them:
 *   0 FP64 trailing-update SYRK ms (the bulk update of each panel, MFMA),
 *   1 its launches, 2 its flops (algorithmic: lower triangle, M(M+1)K),
 *   3 panel chain ms (diagonal-block kernel + panel TRSM + copy),
 *   4 triangular-solve ms, 5 normal-equation assembly ms,
 *   6 panel-chain launches, 7 look-ahead strip update ms,
 *   8 super-block inverse ms,
 *   9 FP32 trailing-update SYRK ms, 10 its launches, 11 its flops,
 *   12 FP64 diagonal-block kernel ms, 13 its launches, 14 its flops
 *      (2 nb^3 / 3 per launch: the block's factor and its inverse),
 *   15-17 the same for the FP32 diagonal-block kernel,
 *   18 standalone SpMV kernel ms, 19 its launches, 20 its algorithmic bytes
 *      (12 nnz + 4 (rows + 1) + 8 (m + n) per product),
 *   21 triangular solve pairs, 22 their algorithmic bytes (the factor's
 *      lower triangle read twice, m (m + 1) elements per pair),
 *   23 the level-2 stride (one launch in this many of a kind is timed; 1
 *      at the other levels).
 * No counterpart in the reference, which times nothing on the device. */
void sov_set_kernel_timing(int32_t on);
int sov_kernel_times(double* out24);

/* MFMA tile shapes chosen by the factorization's autotuner (0: 64x64,
 * 1: 128x64, 2: 128x128, 3: 32x32; -1 before the first factorization in
 * that precision): shapes4 = {FP64 SYRK, FP64 panel products, FP32 SYRK,
 * FP32 panel products}. times_ms12 (may be NULL) receives the tuning runs'
 * times per launch: FP64 SYRK 64x64, 128x128, FP64 panel 64x64, 128x64,
 * 128x128, 32x32, then the same six for FP32. SOVEREIGN_GEMM_SHAPE=0/1/2/3
 * overrides the choice. No counterpart in the reference. */
int sov_gemm_shapes(int32_t* shapes4, double* times_ms12);

/* -------------------------------------------------------------- dense LU --
 * Replaces cuda::dense_lu (cuda_driver.hpp:29-34; run_dense_lu at
 * cuda_driver.cpp:461-502). In-place LU with partial pivoting of an n x n
 * column-major host matrix: unit-lower L below the diagonal, U on and above
 * it, piv the row permutation in DenseLU's convention (row i of P*A is row
 * piv[i] of A; solver/numerical/src/dense_lu.cpp:19-40). Singular (1) when
 * the largest remaining pivot magnitude is below 1e-14. */
int sov_dense_lu(int64_t n, double* a_host, int32_t* piv_host);

/* Allocation-free form of sov_dense_lu, on device memory. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). a_dev (n x n, column-major) is factored in place, piv_dev (n
 * int32) receives DenseLU's permutation; results equal sov_dense_lu's on the
 * same matrix. The pivot swap list and the singular flag live in a fixed
 * device array of the library (one per GPU, held until the process ends),
 * so n is at most 32768 (larger n: SOV_ERR_ARG) and the call allocates
 * nothing. Calls that overlap on one GPU (other host threads, other
 * streams) share that array and race; calls are serialised on the library
 * stream, which this call synchronises before it returns the status
 * (0 ok, 1 singular). */
int sov_dense_lu_dev_noalloc(int64_t n, double* a_dev, int32_t* piv_dev);

/* ------------------------------------------------------------------ SpMV --
 * Replaces cuda::spmv_csc (cuda_driver.hpp:36-40) and its content-compared
 * per-thread MatrixCache (cuda_driver.cpp:509-582) by an explicit handle:
 * the matrix is uploaded once (CSR of A for y = A x, the CSC arrays as CSR
 * of A^T for y = A^T x) and stays resident in HBM. int32 indices, FP64. */
typedef struct sov_mat sov_mat;

int sov_mat_upload_csc(int32_t nrows, int32_t ncols, const int32_t* col_ptr,
                       const int32_t* row_idx, const double* values, sov_mat** out);
int sov_mat_free(sov_mat* a);
/* y = A x (transpose = 0) or y = A^T x (transpose = 1), host vectors. */
int sov_spmv(sov_mat* a, int32_t transpose, const double* x_host, double* y_host);
/* Same on device vectors, asynchronous on the library stream. */
int sov_spmv_dev(sov_mat* a, int32_t transpose, const double* x_dev, double* y_dev);

/* One-shot form of cuda::spmv_csc: keeps one resident copy per calling
 * thread and direction, replaced when the matrix contents change. */
int sov_spmv_csc(int32_t nrows, int32_t ncols, const int32_t* col_ptr, const int32_t* row_idx,
                 const double* values, const double* x_host, double* y_host, int32_t transpose);

/* ------------------------------------------------------- normal equations --
 * The IPM's KKT-solve interface: NormalEquations{factor(d); solve(rhs)}
 * (solver/lp/src/interior_point.cpp:129-212). M = A diag(d) A^T + 1e-12 I is
 * assembled in HBM and factored by a blocked dense Cholesky on MFMA. Pivots
 * follow the reference's sparse LDL^T (solver/numerical/src/sparse_ldlt.cpp:
 * 165-178): a pivot <= 1e-13 x |original diagonal| drops that direction (the
 * reference's 1e128 replacement), a non-finite or < 1e-30 pivot fails.
 * precision: 0 FP64 factor; 1 FP32 factor (FP32 MFMA, FP32 accumulation) with
 * FP64 preconditioned-CG refinement against the FP64 matrix-free product. */
typedef struct sov_ne sov_ne;

#define SOV_PREC_FP64 0
#define SOV_PREC_FP32 1

int sov_ne_create(sov_mat* a, sov_ne** out);

/* The factorization the normal equations use. The reference's selector
 * (interior_point.cpp:129-175) picks a dense LU or the AMD-ordered sparse
 * LDL^T by size and fill; here:
 *   SOV_NE_DENSE   the blocked dense Cholesky above;
 *   SOV_NE_SPARSE  the supernodal multifrontal Cholesky: the reference's
 *                  minimum-degree ordering (amd_ordering.cpp:12-130) and
 *                  elimination-tree analysis (sparse_ldlt.cpp:19-103) once,
 *                  then per factorization a level-scheduled multifrontal
 *                  factorization on the GPU with the same pivot rule, FP64
 *                  only (a FP32 request factors in FP64, solves are direct);
 *   SOV_NE_AUTO    sparse when the pattern's mean degree is at most 64 and
 *                  the ordering predicts a factor of at most a tenth of the
 *                  dense triangle, or when the dense matrix would not fit in
 *                  device memory; dense otherwise. The environment variable
 *                  SOVEREIGN_NE_PATH=dense|sparse overrides AUTO. */
#define SOV_NE_AUTO 0
#define SOV_NE_DENSE 1
#define SOV_NE_SPARSE 2
int sov_ne_create_path(sov_mat* a, int32_t path, sov_ne** out);

/* Symbolic figures of a sparse factorization (all zero for a dense one):
 * nnz(L) with the diagonal, the reference's factor_flops (sum over columns
 * of the squared count below the diagonal), supernodes, tree levels (one
 * launch each per factorization and per solve sweep), largest front. */
typedef struct {
  int32_t sparse;
  int32_t supernodes;
  int32_t levels;
  int32_t max_front;
  int64_t nnz_l;
  double factor_flops;
  int64_t front_doubles;
} sov_ne_info;
int sov_ne_get_info(const sov_ne* h, sov_ne_info* info);

/* Minimum-degree ordering alone (AMDOrdering::compute, amd_ordering.cpp:
 * 12-130): a symmetric pattern (CSC of either triangle or both, n x n), the
 * node of least (degree, index) eliminated first, nodes of degree above
 * max(16, 10 sqrt(n)) last by (degree, index). perm (n): the node eliminated
 * k-th; iperm (n) its inverse; predicted_nnz (may be NULL): the fill the
 * elimination predicts. Host only (no device needed). Returns SOV_OK, or
 * SOV_SINGULAR when fill_limit > 0 and the prediction passes it (the
 * reference's fill_limit_exceeded). */
int sov_min_degree_order(int32_t n, const int32_t* col_ptr, const int32_t* row_idx, int64_t fill_limit,
                         int32_t* perm, int32_t* iperm, int64_t* predicted_nnz);
/* factor(d): 0 ok, 1 breakdown. d_host has ncols entries. */
int sov_ne_factor(sov_ne* h, const double* d_host, int32_t precision);
/* solve(rhs) in place, nrows entries. */
int sov_ne_solve(sov_ne* h, double* rhs_host);
/* Relative residual ||rhs - M x|| / ||rhs|| of the last solve (inf-norms). */
double sov_ne_last_residual(const sov_ne* h);
int32_t sov_ne_replaced_pivots(const sov_ne* h);
int sov_ne_free(sov_ne* h);

/* ----------------------------------------------------------------- IPM ----
 * InteriorPointSolver (solver/lp/include/sovereign/interior_point.hpp:7-36;
 * Mehrotra loop solve_ipm at solver/lp/src/interior_point.cpp:454-740) with
 * every iteration resident on the GPU. */
typedef struct {
  int32_t max_iterations;      /* 250 */
  double feasibility_tol;      /* 1e-9 */
  double optimality_tol;       /* 1e-9 */
  double fraction_to_boundary; /* 0.999 */
  int32_t enable_scaling;      /* 1 (general-form solve only) */
  /* SOV_PREC_FP64: every factorization in FP64 (the plain FP64 GPU path).
   * SOV_PREC_FP32: mixed precision - FP32 factorizations while
   * mu >= mixed_switch_mu and the FP32-preconditioned CG of an iteration
   * takes at most 8 iterations (and has not stalled), FP64 afterwards;
   * FP64 residuals and refinement. */
  int32_t precision;
  double mixed_switch_mu; /* 1e-6 */
  int32_t refine_max;     /* 30: PCG iterations per solve at most */
  double refine_tol;      /* 1e-12: relative residual target of a solve (an FP32-preconditioned
                           * solve stops at max(refine_tol, min(1e-6, 1e-3 mu))) */
  int32_t normal_path;    /* SOV_NE_AUTO (0): the normal equations' factorization for the one-shot
                           * solves (sov_ipm_solve_form, sov_lp_solve); sov_ipm_create_path sets it
                           * for a resident solver */
} sov_ipm_options;

typedef struct {
  int32_t status; /* sovereign::SolverStatus: 0 Optimal, 1 Feasible, 6 NumericalError, 7 Error */
  int32_t iterations;
  double objective;
  double duality_gap;
  double primal_residual;
  double dual_residual;
  int32_t factorizations;
  int32_t fp32_factorizations;
  int32_t refine_iterations; /* PCG iterations summed over all solves */
  double max_solve_residual; /* worst relative residual of a Newton solve */
  double seconds;            /* whole solve, device-synchronised */
  double factor_seconds;     /* assembly + Cholesky, from HIP events */
  double solve_seconds;      /* triangular solves + refinement, from HIP events */
  int32_t sparse_factor;     /* 1: the normal equations were factored by the sparse path */
} sov_ipm_result;

void sov_ipm_default_options(sov_ipm_options* opt);

/* Resident solver for one constraint matrix (the IPM form's A, m x n):
 * create analyses A once (the normal-equation assembly plan, the role of the
 * NormalEquations constructor, interior_point.cpp:131-161), solve runs the
 * Mehrotra loop for a right-hand side b (m) and cost c (n). The matrix handle
 * must outlive the solver. */
typedef struct sov_ipm sov_ipm;
int sov_ipm_create(sov_mat* a, sov_ipm** out);
/* Same with the normal equations' factorization chosen (SOV_NE_*). */
int sov_ipm_create_path(sov_mat* a, int32_t path, sov_ipm** out);
int sov_ipm_solve(sov_ipm* h, const double* b, const double* c, const sov_ipm_options* opt,
                  sov_ipm_result* res, double* x, double* y, double* s);
int sov_ipm_free(sov_ipm* h);

/* LP already in IPM form: min c'x, A x = b, x >= 0 (A is m x n CSC). x, s
 * (n) and y (m) are host outputs and may be NULL. The objective is c'x. */
int sov_ipm_solve_form(int32_t m, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                       const double* values, const double* b, const double* c,
                       const sov_ipm_options* opt, sov_ipm_result* res, double* x, double* y,
                       double* s);

/* General LP (the array image of sovereign::OptimizationModel for LP,
 * solver/core/include/sovereign/types.hpp:32-59): CSC constraint matrix
 * n_rows x n_vars, row senses 'L' / 'G' / 'E', bounds (ub >= 1e29 = none). */
typedef struct {
  int32_t n_vars;
  int32_t n_rows;
  const int32_t* col_ptr;
  const int32_t* row_idx;
  const double* values;
  const double* obj;
  double obj_constant;
  int32_t maximize;
  const double* lb;
  const double* ub;
  const char* row_sense;
  const double* rhs;
} sov_lp_model;

/* InteriorPointSolver::solve (interior_point.cpp:747-757): build_ipm_form
 * (:224-409) on the host, the Mehrotra loop on the GPU, then the result
 * mapped back (:541-562). primal / var_dual: n_vars, row_dual: n_rows; any
 * may be NULL. */
int sov_lp_solve(const sov_lp_model* model, const sov_ipm_options* opt, sov_ipm_result* res,
                 double* primal, double* row_dual, double* var_dual);

/* ------------------------------------------------------ batched child LPs --
 * The data-parallel unit of strong branching: the 2k child LP relaxations of
 * a branch-and-bound node (solve_two_lps_parallel ×k in pick_strong_branch /
 * pick_reliability_branch, solver/milp/src/branch_and_bound.cpp:554-638,
 * 726-899), solved together in one persistent launch. Each LP is
 *     min c'x  s.t.  A x = b,  0 <= x <= u
 * with a dense column-major A of at most 272 rows (u >= 1e29: no upper bound,
 * u == 0: x fixed at 0). Matrices are uploaded once and shared by LPs;
 * per-LP b lives at b_all + b_off[k] (m entries), c / u / x at
 * + x_off[k] (n entries). status: 0 optimal, 5 iteration limit,
 * 6 numerical breakdown. The objective is c'x of the given data. Each LP's
 * normal matrix is factored in FP32 (LDS) as the preconditioner of FP64 CG;
 * an LP whose FP32 factor drops a pivot or whose solves stall continues
 * with FP64 factorizations in global memory. */
typedef struct sov_batch sov_batch;
typedef struct {
  int32_t status;
  int32_t iterations;
  int32_t pcg_iterations;
  int32_t fp64_from; /* iteration from which the LP's factorizations were FP64 (its FP32 factor had
                      * dropped a pivot or its FP32-preconditioned solves had stalled), or -1 */
  double objective;
  double duality_gap;
  double primal_residual;
  double dual_residual;
} sov_batch_result;

int sov_batch_create(sov_batch** out);
int sov_batch_add_matrix(sov_batch* h, int32_t m, int32_t n, const double* a_colmajor, int32_t* id);
int sov_batch_solve(sov_batch* h, int32_t count, const int32_t* matrix_id, const double* b_all,
                    const int64_t* b_off, const double* c_all, const double* u_all, const int64_t* x_off,
                    const sov_ipm_options* opt, sov_batch_result* res, double* x_all);
/* sov_batch_solve with warm starts and final iterates, for child LPs that
 * start from their parent's point (the interior-point counterpart of the
 * dual simplex warm start from the parent's LpBasis, solver/lp/include/
 * sovereign/dual_simplex.hpp:19-23, used by solve_node_lp, branch_and_bound.cpp:
 * 437-516). An iterate is (x, w, z, v: n each, then y: m) in the batched
 * form's units: x the shifted variables, w = u - x, z and v the multipliers
 * of x >= 0 and x <= u, y the row multipliers. state_in (may be NULL): LP k
 * starts from state_in + in_off[k] when in_off[k] >= 0 (components are
 * moved at least 1e-10 u inside their bounds), else from the default point.
 * state_out (may be NULL): when out_off[k] >= 0, LP k's first iterate whose
 * complementarity mu = (x'z + w'v) / #bounds falls below save_mu (a point
 * near the central path, what a warm start wants) is written to state_out +
 * out_off[k]; its final iterate when none does or save_mu <= 0. */
int sov_batch_solve_warm(sov_batch* h, int32_t count, const int32_t* matrix_id, const double* b_all,
                         const int64_t* b_off, const double* c_all, const double* u_all, const int64_t* x_off,
                         const sov_ipm_options* opt, sov_batch_result* res, double* x_all, const double* state_in,
                         const int64_t* in_off, double* state_out, const int64_t* out_off, double save_mu);
/* Duration of the last solve's kernel (HIP events on the library stream). */
double sov_batch_last_kernel_ms(const sov_batch* h);
int sov_batch_free(sov_batch* h);

#ifdef __cplusplus
}
#endif

#endif /* SOVEREIGN_HIP_H */
