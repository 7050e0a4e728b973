// Cut separation for the tree search: cover, Chvatal-Gomory rounding and
// complemented MIR cuts from single rows, and the independent validity gate
// every cut passes before it enters a model (solver/cuts/include/sovereign/
// cuts.hpp:17-55, apply_cuts in solver/milp/src/branch_and_bound.cpp:648-694).
// Host code over a row-wise copy of the model; nothing here touches the GPU.
#pragma once

#include "sovereign_hip.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace sov {

enum CutFamily { kCutCover = 1, kCutGomory = 2, kCutCmir = 4 };
enum VarKind : int8_t { kContinuous = 0, kInteger = 1, kBinary = 2 };

struct CutRow {
  std::vector<int> idx;     // variables, each at most once
  std::vector<double> val;  // coefficients
  char sense = 'L';         // 'L' <=, 'G' >=, 'E' =
  double rhs = 0.0;
};

struct CutModel {
  std::vector<double> lb, ub;
  std::vector<int8_t> kind;  // VarKind per variable
  std::vector<CutRow> rows;
  int n_vars() const { return static_cast<int>(lb.size()); }
};

struct Cut {
  CutRow row;
  int family = 0;
  double efficacy = 0.0;  // the cMIR separation score; 0 for the others
};

// generate_cover_cuts (cuts.cpp:96-159): per <= row, the binary terms with
// positive coefficients against the row's right-hand side less the least
// activity of every other term; the greedy cover by decreasing coefficient;
// the cut sum_C x <= |C| - 1 when x violates it by more than int_tol.
std::vector<Cut> separate_covers(const CutModel& m, const double* x, double int_tol, int max_cuts);

// generate_mir_cuts (cuts.cpp:161-251): Chvatal-Gomory rounding of a <= row
// whose columns can all be rounded soundly.
std::vector<Cut> separate_gomory(const CutModel& m, const double* x, double int_tol, int max_cuts);

// generate_cmir_cuts (cuts.cpp:253-420): complemented MIR on the first
// base_rows rows (both sides of an equality), most efficacious first.
std::vector<Cut> separate_cmir(const CutModel& m, const double* x, double int_tol, int max_cuts,
                               std::size_t base_rows);

// check_cut_validity (cuts.cpp:25-94): 0 accepted, 1 the cut removes one of
// the known integer-feasible points (each n_vars long), 2 no point of the
// variables' bound box satisfies it.
int cut_rejection(const CutModel& m, const CutRow& cut, const std::vector<std::vector<double>>& points, double tol);

// apply_cuts (branch_and_bound.cpp:648-694): covers, then cMIR (when asked),
// then Gomory candidates, each through the gate, appended to m.rows until
// max_cuts are in. Returns the number added; rejected counts the gate's
// rejections and added_families receives the family of each cut added (each
// may be NULL).
int apply_cuts(CutModel& m, const double* x, double int_tol, int max_cuts,
               const std::vector<std::vector<double>>& points, std::size_t base_rows, bool with_cmir,
               int* rejected, std::vector<int>* added_families);

// The model's rows (each row's terms in increasing variable order) with the
// variable kinds; SOV_ERR_ARG on a malformed model.
int cut_model_from(const sov_lp_model& md, const int8_t* kind, CutModel& out);

// A model in the C ABI's layout owning its rows: the base model's columns,
// bounds and objective with m's rows (base rows and cuts) as CSC.
struct OwnedLpModel {
  std::vector<int32_t> col_ptr, row_idx;
  std::vector<double> values, rhs;
  std::string sense;
  sov_lp_model model{};
  void build(const sov_lp_model& base, const CutModel& m);
};

}  // namespace sov
