/* sovereign_bnb.h - strong branching and branch-and-bound over batched child
 * LPs (C ABI). Builds on sovereign_hip.h. Status codes as there. */
#ifndef SOVEREIGN_BNB_H
#define SOVEREIGN_BNB_H

#include "sovereign_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SOV_CHILD_BATCHED 0 /* all child LPs of a node in one persistent launch */
#define SOV_CHILD_PER_LP 1  /* each child through the plain FP64 single-LP GPU path */

/* One strong-branching candidate: the down (ub = floor x) and up
 * (lb = ceil x) child LPs and the score of pick_strong_branch
 * (solver/milp/src/branch_and_bound.cpp:726-793): degradation
 * |child objective - parent objective|, 1e6 for an infeasible child, score
 * min(down, up) + 1e-4 max(down, up). Objectives are in the model's sense;
 * a child that did not reach optimality reports the parent objective (as
 * the reference does for a child without an objective value). */
typedef struct {
  int32_t var;
  int32_t status_down; /* 0 optimal, 2 infeasible (empty box), 5/6 unresolved */
  int32_t status_up;
  int32_t iters_down;
  int32_t iters_up;
  int32_t pad;
  double obj_down;
  double obj_up;
  double score;
} sov_sb_child;

/* Strong branching at one node: the node LP is `node` with variable bounds
 * node_lb / node_ub, x_node its LP solution and parent_obj its objective.
 * Candidates are the fractional integer variables (is_int) ordered by
 * closeness of their fraction to 1/2, at most max_candidates of them
 * (the FullStrong rule, branch_and_bound.cpp:731-748). out holds one entry
 * per candidate (max_candidates entries), ncand_out their number, best the
 * index of the first best score (-1 when no candidate is fractional).
 * kernel_ms accumulates the batched launch duration (may be NULL). */
int sov_strong_branch(const sov_lp_model* node, const double* node_lb, const double* node_ub,
                      const int8_t* is_int, const double* x_node, double parent_obj,
                      int32_t max_candidates, int32_t child_solver, const sov_ipm_options* opt,
                      sov_sb_child* out, int32_t* ncand_out, int32_t* best, double* kernel_ms);

/* Ranking callback: fills key[j] for the variables of a node (higher =
 * branch first); used to put the learned surrogate's order in front of the
 * strong-branching candidate cut. */
typedef void (*sov_candidate_order_fn)(int32_t n_vars, const double* x, const double* lb,
                                       const double* ub, double* key, void* user);

/* Branching rules of BranchRule (solver/milp/include/sovereign/branch_and_bound.hpp:7-16). */
#define SOV_BRANCH_MOST_FRACTIONAL 0 /* pick_most_fractional (branch_and_bound.cpp:144-159) */
#define SOV_BRANCH_RELIABILITY 1     /* StrongBranching = pick_reliability_branch (:799-899), the default */
#define SOV_BRANCH_PSEUDOCOST 2      /* pick_pseudo_cost (:161-189) */
#define SOV_BRANCH_FULL_STRONG 3     /* pick_strong_branch (:726-793) */

/* BranchAndBoundOptions (branch_and_bound.hpp:18-59) for the parts built
 * here; heuristics and model presolve are not. */
typedef struct {
  int64_t max_nodes;                /* 100000 */
  int32_t strong_branch_candidates; /* FullStrong: k candidates per node (the reference's default is 4) */
  int32_t child_solver;             /* SOV_CHILD_BATCHED or SOV_CHILD_PER_LP */
  double mip_gap;                   /* 1e-6 relative (branch_and_bound.hpp:23) */
  double integer_tol;               /* 1e-6 */
  double time_limit_seconds;        /* 0 = none */
  sov_candidate_order_fn candidate_order; /* FullStrong only. NULL: closest fraction to 1/2 first */
  void* candidate_order_user;
  int32_t branch_rule;           /* SOV_BRANCH_*, default SOV_BRANCH_RELIABILITY */
  int32_t reliability_threshold; /* 4: observations per direction before a pseudocost is trusted */
  int32_t max_strong_per_node;   /* 8: reliability branching's strong-branch budget per node */
  int32_t plunging;              /* 1: continue with the preferred child before the best-bound queue */
  int32_t max_cut_rounds;        /* 5: cut rounds at the root (0: no cuts), as max_cut_rounds */
  int32_t max_cuts_per_node;     /* 20: cuts added per round */
  int32_t cmir_at_root;          /* 0: cMIR separation off, the reference's default (cmir_max_depth = -1) */
  int32_t warm_start;            /* 1: batched child LPs start from their parent node LP's iterate */
  double warm_mu;                /* 0.1: the parent's iterate is its first with mu below this (iterations
                                  * per child LP on the knapsack trees: 0.3 -> 13-21% fewer than cold,
                                  * 0.1 -> 13-28%, 1e-3 -> from 16% more to 28% fewer) */
} sov_bnb_options;

typedef struct {
  int32_t status; /* 0 optimal, 1 feasible (limit hit or a node LP failed), 2 infeasible,
                   * 5 limit without incumbent, 6 numerical (no incumbent and a node LP failed) */
  int32_t pad;
  double objective;
  double root_bound;
  int64_t nodes;
  int64_t lp_solves;           /* node LPs + strong-branching child LPs + incumbent polishes */
  double seconds;
  double strong_branch_seconds;
  double kernel_ms;            /* batched launches' kernel time */
  int64_t node_lp_solves;      /* node LPs actually solved (a node whose LP strong branching already
                                * solved reuses that solution) */
  int64_t strong_lp_solves;    /* child LPs solved for strong branching */
  int64_t strong_lp_used;      /* of those, the ones the reference's sequential rule would have solved */
  double node_lp_seconds;
  int32_t max_depth;
  int32_t node_lp_failures;    /* subtrees dropped because their LP did not solve */
  int32_t cuts_added;          /* root cuts kept in the model (cover + Gomory + cMIR) */
  int32_t cuts_cover;
  int32_t cuts_gomory;
  int32_t cuts_cmir;
  int32_t cut_rounds;          /* root rounds that added cuts */
  int32_t cut_rejections;      /* candidates the validity gate rejected */
  double root_bound_before_cuts; /* root LP objective (model sense) before the first round; NaN without cuts */
  int32_t root_rows;           /* rows of the model the tree searches (base rows + cuts) */
  int32_t pad2;
  int64_t lp_iterations;       /* interior-point iterations of every LP the tree solved */
} sov_bnb_result;

void sov_bnb_default_options(sov_bnb_options* opt);

/* BranchAndBoundSolver::solve (branch_and_bound.cpp:906-1523) without
 * primal heuristics or model presolve (a node presolve infeasibility test
 * from the rows' activity ranges settles child LPs the interior point cannot
 * certify infeasible), with cuts at the root only: up to
 * max_cut_rounds rounds of apply_cuts (:648-694, :1214-1273 at depth 0:
 * covers, Gomory rounding and, when asked, cMIR, each through the validity
 * gate, at most max_cuts_per_node a round, the root LP re-solved after each)
 * whose cuts every node inherits; the reference also separates one round
 * at every deeper node, which is not done here. With the batched child
 * solver the cuts stop at 272 rows in all. Then: best-bound queue with plunging, the node LP
 * solved when the node is popped (reusing a child LP that strong branching
 * already solved), pseudocost learning from node LPs (:1159-1170) and from
 * strong branching, integrality checked before bound pruning (:1317), and
 * the branching rule of bopt->branch_rule. Reliability branching's strong
 * branches run speculatively as one batched launch of the first
 * max_strong_per_node unreliable candidates; the reference's sequential
 * look-ahead rule is then replayed over their results, so the decision and
 * the pseudocost updates are the reference's. is_int marks integer
 * variables; x_best (n_vars, may be NULL) receives the incumbent. Variables
 * with lb <= -1e29 cannot be shifted into the batched child form; for such
 * models every LP goes through the per-LP path. */
int sov_bnb_solve(const sov_lp_model* milp, const int8_t* is_int, const sov_bnb_options* bopt,
                  const sov_ipm_options* opt, sov_bnb_result* res, double* x_best);

#ifdef __cplusplus
}
#endif

#endif /* SOVEREIGN_BNB_H */
