// Drop-in nasoq::Nasoq over the MI355X engine (include/lblg.h, lblg_qp_*).
//
// Same constructors, public settings, result fields and solve() methods as
// the reference's Nasoq (include/nasoq/nasoq.h:26-493 in the reference):
//   min 1/2 x'Hx + q'x  s.t.  Ax = a,  Bx <= b
// H is lower CSC; A (eq) and B (ineq) are CSC with the variables as columns.
// The input arrays are copied at construction.  primal_vars, dual_vars and
// dual_vars_eq are owned by the object (as in the reference) and hold the
// results after solve().  Link with nasoq_amd/liblblg.so.
#ifndef NASOQ_AMD_NASOQ_H
#define NASOQ_AMD_NASOQ_H

#include <string>
#include <vector>

#include "lblg.h"
#include "nasoq/QP/linear_solver_wrapper.h"
#include "nasoq/common/def.h"

namespace nasoq {

struct QPSettings {  // nasoq.h:26-42, QPSettings::QPSettings (nasoq.cpp:21-32)
  double eps = 1e-3, eps_primal = 1e-3, eps_dual = 1e-3, eps_slack = 1e-3, eps_nn = 1e-3;
  double diag_perturb = 1e-9;
  int batch_size = 1;
  double eps_rel = 1e-3;
  int scaling = 0;
  double zero_thresh = 1e-9;
  int inner_iter_ref = 0, outer_iter_ref = 0, max_iter = 0;
  double stop_tol = 1e-15;
  int max_iter_nas = 4000;
  std::string nasoq_variant = "Fixed";
};

enum nasoq_mode { Fixed = 0, AUTO, Tuned, PREDET };
enum nasoq_status { Optimal = 1, Inaccurate = 2, NotConverged = 3, Infeasible = 0, NotFinished = 4 };

struct Nasoq {
  std::string sol_name = "NASOQ-Fixed";
  int hessian_size = 0, eq_const_size = 0, ineq_const_size = 0;
  // settings (Nasoq::default_setting, nasoq.cpp:186-199)
  double diag_perturb = 1e-9, zero_thresh = 1e-9, eps_abs = 1e-3, eps_rel = 1e-3;
  int inner_iter_ref = 0, outer_iter_ref = 0, max_iter = 0, max_iter_nas = 4000;
  double stop_tol = 1e-15;
  nasoq_mode variant = Fixed;
  int scaling = 0;
  // results
  double *primal_vars = nullptr, *dual_vars = nullptr, *dual_vars_eq = nullptr;
  double primal_obj = 0, dual_obj = 0, objective = 0;
  double non_negativity_infn = 0, complementarity_infn = 0, lag_res = 0, cons_sat_norm = 0;
  std::vector<int> active_set;
  int n_active = 0, num_iter = 0, is_converged = 0;
  nasoq_status ret_val = NotConverged;

  // Nasoq(H, q, B, b) (nasoq.cpp:36-115)
  Nasoq(size_t H_size, int* Hp, int* Hi, double* Hx, double* q_in, size_t B_row, size_t B_col, int* Bp, int* Bi,
        double* Bx, double* b_ineq)
      : Nasoq(H_size, Hp, Hi, Hx, q_in, 0, H_size, nullptr, nullptr, nullptr, nullptr, B_row, B_col, Bp, Bi, Bx,
              b_ineq) {}
  // Nasoq(H, q, A, a, B, b) (nasoq.cpp:117-135)
  Nasoq(size_t H_size, int* Hp, int* Hi, double* Hx, double* q_in, size_t A_size1, size_t A_size2, int* Ap, int* Ai,
        double* Ax, double* a_eq, size_t B_size1, size_t B_size2, int* Bp, int* Bi, double* Bx, double* b_ineq) {
    (void)A_size2;
    (void)B_size2;
    hessian_size = (int)H_size;
    eq_const_size = (int)A_size1;
    ineq_const_size = (int)B_size1;
    lblg_qp_create(&qp_, (int)H_size, Hp, Hi, Hx, q_in, (int)A_size1, Ap, Ai, Ax, a_eq, (int)B_size1, Bp, Bi, Bx,
                   b_ineq);
    primal_vars = new double[H_size]();
    dual_vars = new double[B_size1 > 0 ? B_size1 : 1]();
    dual_vars_eq = new double[A_size1 > 0 ? A_size1 : 1]();
  }
  ~Nasoq() {
    if (qp_) lblg_qp_destroy(qp_);
    delete[] primal_vars;
    delete[] dual_vars;
    delete[] dual_vars_eq;
  }
  Nasoq(const Nasoq&) = delete;
  Nasoq& operator=(const Nasoq&) = delete;

  // Nasoq::solve() (nasoq.cpp:656-710): returns the nasoq_status
  int solve() {
    if (!qp_ || !push()) return Infeasible;
    int st = Infeasible;
    if (lblg_qp_solve(qp_, &st) != LBLG_OK) return Infeasible;
    pull();
    detect_solver_name();
    return st;
  }
  // Nasoq::solve(primal, dual, active_set) (nasoq.cpp:740-760): warm start;
  // returns 1 when the active-set loop converged
  int solve(double* primal_in, double* dual_in, const std::vector<int>& as, double* dual_in_eq = nullptr) {
    if (!qp_ || !push()) return 0;
    int conv = 0;
    if (lblg_qp_solve_warm(qp_, primal_in, dual_in, as.data(), (int)as.size(), dual_in_eq, &conv) != LBLG_OK)
      return 0;
    pull();
    return conv;
  }
  void detect_solver_name() {  // nasoq.cpp:730-738
    sol_name = variant == Fixed ? "NASOQ-Fixed" : variant == AUTO ? "NASOQ-AUTO" : variant == Tuned ? "NASOQ-TUNED"
                                                                                                     : "NASOQ";
  }

 private:
  lblg_qp* qp_ = nullptr;
  bool push() {
    lblg_qp_settings s;
    lblg_qp_default_settings(&s);
    s.diag_perturb = diag_perturb;
    s.eps_abs = eps_abs;
    s.stop_tol = stop_tol;
    s.max_iter = max_iter;
    s.max_iter_nas = max_iter_nas;
    s.variant = (int)variant;
    return lblg_qp_set_settings(qp_, &s) == LBLG_OK;
  }
  void pull() {
    lblg_qp_info info;
    lblg_qp_get_info(qp_, &info);
    active_set.assign(info.n_active, 0);
    lblg_qp_get(qp_, primal_vars, dual_vars, dual_vars_eq, active_set.data());
    ret_val = (nasoq_status)info.status;
    num_iter = info.num_iter;
    n_active = info.n_active;
    inner_iter_ref = info.inner_iter_ref;
    outer_iter_ref = info.outer_iter_ref;
    diag_perturb = info.diag_perturb;
    zero_thresh = info.diag_perturb;
    stop_tol = info.stop_tol;
    primal_obj = info.primal_obj;
    dual_obj = info.dual_obj;
    objective = info.objective;
    lag_res = info.lag_res;
    cons_sat_norm = info.cons_sat_norm;
    non_negativity_infn = info.non_negativity_infn;
    complementarity_infn = info.complementarity_infn;
  }
};

}  // namespace nasoq
#endif
