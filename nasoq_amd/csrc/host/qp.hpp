// Host side of the NASOQ dual active-set QP solver on top of the device LBL
// engine:  min 1/2 x'Hx + q'x  s.t.  Ax = a,  Bx <= b.
//
// Reference: nasoq::Nasoq (include/nasoq/nasoq.h:61-493, src/nasoq.cpp) with
// the super-KKT of SolverSettings mode 1 (build_super_matrix,
// src/QP/linear_solver_wrapper.cpp:246-350) and its add/drop edits
// (add_del_matrix_qp, :1020-1137).  The KKT factor of each active set is the
// device factor of the super-KKT in which every inactive inequality row is
// decoupled (its couplings are zero, its diagonal slot is regularised), which
// equals the reference's factor with those supernodes hidden from the tree.
#pragma once
#include <cstdint>
#include <vector>

#include "lblg.h"

namespace lblg {

struct SparseCsc {  // compressed sparse columns, rows sorted within a column
  int nrow = 0, ncol = 0;
  std::vector<int> p, i;
  std::vector<double> x;
};

// variants and statuses: the reference's nasoq_mode / nasoq_status values
enum { QP_FIXED = 0, QP_AUTO = 1, QP_TUNED = 2, QP_PREDET = 3 };
enum { QP_INFEASIBLE = 0, QP_OPTIMAL = 1, QP_INACCURATE = 2, QP_NOT_CONVERGED = 3, QP_NOT_FINISHED = 4 };

struct QpSolver {
  // problem
  int n = 0, me = 0, mi = 0;
  SparseCsc H, A, B, AT, BT;  // H lower; A: me x n; B: mi x n; AT, BT their transposes
  std::vector<double> q, a, b;
  // settings (the reference's public fields, Nasoq::default_setting)
  double diag_perturb = 1e-9, zero_thresh = 1e-9, eps_abs = 1e-3, stop_tol = 1e-15;
  int max_iter = 0, max_iter_nas = 4000, variant = QP_FIXED, ordering = LBLG_ORDER_ND;
  int precision = LBLG_F64;  // value type of the KKT factor (refinement is FP64 either way)
  int separate_rows = 1;     // inequality rows as one-column supernodes (the reference's extra_cols)
  int analysed_separate = -1;  // separate_rows of the current analysis of h
  int inner_iter_ref = 0, outer_iter_ref = 0, auto_reg_en = 0;
  // state and results
  std::vector<double> primal, dual, dual_eq;
  std::vector<int> active_set;
  std::vector<unsigned char> used;
  int num_iter = 0, n_active = 0, warm_start = 0, status = QP_NOT_CONVERGED;
  double primal_obj = 0, dual_obj = 0, objective = 0, lag_res = 0, cons_sat_norm = 0, non_negativity_infn = 0,
         complementarity_infn = 0;
  int64_t factorizations = 0, kkt_solves = 0;
  double factor_seconds = 0, solve_seconds = 0;  // host wall time in the KKT factorizations / solves
  int64_t refactored_supernodes = 0;             // supernodes refactored, summed over the factorizations
  // super-KKT [H A' B'; A 0 0; B 0 0] (lower CSC, explicit diagonal slots)
  SparseCsc SM;
  std::vector<int64_t> bpos;                // SM entry of every entry of B (B's CSC order)
  std::vector<int> brow_ptr, brow_ent;      // entries of B grouped by constraint
  std::vector<double> rhs, sol;
  // super-KKT entries edited since the device last received SM.x, and
  // whether it holds SM.x at all (a new super-KKT is uploaded whole)
  std::vector<int64_t> kkt_changed;
  bool kkt_uploaded = false;
  lblg_handle* h = nullptr;
  bool analysed = false;
  lblg_settings analysed_settings{};  // settings of the current analysis of h

  QpSolver() = default;
  QpSolver(const QpSolver&) = delete;
  QpSolver& operator=(const QpSolver&) = delete;
  ~QpSolver();

  void build_super_matrix();
  void symbolic();
  void analyze_lbl(const lblg_settings& st);  // mode-1 LBL: super-KKT with given settings
  void initialize_x();
  int numeric();
  int solve();
  int solve_warm(const double* x0, const double* dual_in, const int* as, int nas, const double* dual_eq_in);
  int solve_predicted(const int* as, int nas);
  int predicted_prunes = 0;  // KKT solves spent making a predicted active set dual feasible
  int check_status(int converged);
  // pieces of the active-set iteration
  void factor_kkt();  // the device factor of SM.x (the SoMod update after edits)
  void solve_kkt(bool refactor);
  void edit_kkt(bool add, int c);
  int primal_feasibility(int& nxt_add);
  double primal_step_length(int c) const;
  double dual_step_length(int& nxt_drop) const;
  void update_rhs(int c);
  void update_primal(double step);
  void update_dual(double step);
  int update_active_set(bool add, int c);
  // norms and objectives (the reference's KKT checks)
  void compute_objective();
  double constraint_sat_norm();
  double lagrangian_residual_norm();
  double complementarity_norm();
  double non_negativity_norm();
  double hessian_frobenius() const;
};

SparseCsc transpose(const SparseCsc& M);

}  // namespace lblg
