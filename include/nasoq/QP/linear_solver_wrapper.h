// Drop-in nasoq::SolverSettings over the MI355X engine (include/lblg.h).
//
// Same public fields and methods as the reference's SolverSettings
// (include/nasoq/QP/linear_solver_wrapper.h:53-306 in the reference, mode 0:
// symmetric KKT solve).  Ownership follows the reference: A and rhs are
// borrowed; the solution buffer x is allocated by symbolic_analysis() and
// freed by the caller (examples/lbl_main.cpp:68); solve_only(n_rhs, rhs)
// copies rhs into the constructor's rhs buffer (linear_solver_wrapper.cpp:1300-1303).
// Link with nasoq_amd/liblblg.so.
#ifndef NASOQ_AMD_LINEAR_SOLVER_WRAPPER_H
#define NASOQ_AMD_LINEAR_SOLVER_WRAPPER_H

#include <cstdint>
#include <cstring>
#include <vector>

#include "lblg.h"
#include "nasoq/common/def.h"

namespace nasoq {

struct SolverSettings {
  int solver_mode = 0;  // 0: normal solve (mode 1, the QP super-KKT, is driven by Nasoq)
  CSC* A = nullptr;
  double* rhs = nullptr;
  double* x = nullptr;
  // public settings (default_setting, linear_solver_wrapper.cpp:205-244)
  int ldl_variant = 4, ldl_update_variant = 2;
  int req_ref_iter = 2, max_iter = 2, max_inner_iter = 2;
  double tol_abs = 1e-15, tol_rel = 1e-15;
  double reg_diag = 1e-8;
  int regularization = 1;
  int num_thread = 1, cost_param = 1, level_param = -3, final_seq_node = 50, chunk = 1;
  int n_relax[3] = {4, 16, 48};
  double z_relax[3] = {0.8, 0.1, 0.05};
  int num_ref_iter = 0;
  double bwd_err = 0;
  // engine extensions
  int precision = LBLG_F64;
  double target_bwd_err = 0.0;

  SolverSettings(CSC* Amat, double* rhs_in) : A(Amat), rhs(rhs_in) {}
  ~SolverSettings() {
    if (h_) lblg_destroy(h_);
  }
  SolverSettings(const SolverSettings&) = delete;
  SolverSettings& operator=(const SolverSettings&) = delete;

  // returns 1 on success, 0 on failure (linear_solver_wrapper.cpp:396-536)
  int symbolic_analysis() {
    const int n = (int)A->ncol;
    if (h_) lblg_destroy(h_);
    h_ = nullptr;
    if (lblg_create(&h_, n, A->p, A->i, A->x) != LBLG_OK) return 0;
    lblg_settings s = settings();
    if (lblg_analyze(h_, &s, nullptr, nullptr) != LBLG_OK) return 0;
    x = new double[n];
    return 1;
  }
  // returns 1 on success (linear_solver_wrapper.cpp:602-768)
  int numerical_factorization() {
    lblg_settings s = settings();
    if (lblg_set_settings(h_, &s) != LBLG_OK) return -1;
    return lblg_factor(h_) == LBLG_OK ? 1 : -1;
  }
  // solve with refinement for the constructor's rhs (linear_solver_wrapper.cpp:1305-1404)
  double* solve_only() {
    lblg_settings s = settings();
    if (lblg_set_settings(h_, &s) != LBLG_OK) return nullptr;
    const int64_t n = (int64_t)A->ncol;
    if (lblg_solve(h_, 1, rhs, n, x, n, &bwd_err) != LBLG_OK) return nullptr;
    num_ref_iter = lblg_num_ref_iter(h_);
    return x;
  }
  double* solve_only(const int n_rhs, double* rhs_in) {
    (void)n_rhs;  // the reference copies one right-hand side (linear_solver_wrapper.cpp:1300-1303)
    std::memcpy(rhs, rhs_in, sizeof(double) * A->ncol);
    return solve_only();
  }
  // backward error of the last solve (linear_solver_wrapper.cpp:1462-1498)
  double backward_error() { return bwd_err; }
  // inertia of the perturbed matrix (not computed by the reference)
  int inertia(int64_t* pos, int64_t* neg, int64_t* zero) { return lblg_inertia(h_, pos, neg, zero); }
  lblg_handle* handle() { return h_; }

 private:
  lblg_handle* h_ = nullptr;
  lblg_settings settings() const {
    lblg_settings s;
    lblg_default_settings(&s);
    s.ldl_variant = ldl_variant;
    s.req_ref_iter = req_ref_iter;
    s.max_inner_iter = max_inner_iter;
    s.tol_abs = tol_abs;
    s.tol_rel = tol_rel;
    s.reg_diag = reg_diag;
    s.regularization = regularization;
    s.cost_param = cost_param;
    s.level_param = level_param;
    s.final_seq_node = final_seq_node;
    for (int k = 0; k < 3; ++k) {
      s.n_relax[k] = n_relax[k];
      s.z_relax[k] = z_relax[k];
    }
    s.precision = precision;
    s.target_bwd_err = target_bwd_err;
    return s;
  }
};

}  // namespace nasoq
#endif
