// Drop-in nasoq::SolverSettings over the MI355X engine (include/lblg.h).
//
// Same public fields and methods as the reference's SolverSettings
// (include/nasoq/QP/linear_solver_wrapper.h:53-306 in the reference).
//   mode 0 (SolverSettings(A, rhs)): the LBL solve of a symmetric matrix.
//     A and rhs are borrowed; x is allocated by symbolic_analysis() and freed
//     by the caller (examples/lbl_main.cpp:68); solve_only(n_rhs, rhs) copies
//     rhs into the constructor's rhs buffer (linear_solver_wrapper.cpp:1300-1303).
//   mode 1 (SolverSettings(A, rhs, B[, b]) / (A, rhs, B, BT[, C, CT])): the
//     super-KKT [A C' B'; C 0 0; B 0 0] whose inequality rows B are added and
//     dropped with update_somod / downdate_somod and refactored with
//     update_factorization (:995-1290).  rhs and x then have
//     n + rows(C) + rows(B) entries and are owned by the object (:185-188).
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
  int solver_mode = 0;  // 0: symmetric solve; 1: super-KKT with added / dropped rows
  CSC* A = nullptr;
  CSC *B = nullptr, *BT = nullptr, *C = nullptr, *CT = nullptr;
  double* rhs = nullptr;
  double* x = nullptr;
  double* extra_rhs = nullptr;
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
  int base = 0;  // first inequality row of the super-KKT (mode 1)
  // engine extensions
  int precision = LBLG_F64;
  double target_bwd_err = 0.0;

  SolverSettings(CSC* Amat, double* rhs_in) : A(Amat), rhs(rhs_in) {}
  SolverSettings(CSC* Amat, double* rhs_in, CSC* Bmat) : A(Amat), B(Bmat), rhs(rhs_in) { solver_mode = 1; }
  SolverSettings(CSC* Amat, double* rhs_in, CSC* Bmat, double* b) : SolverSettings(Amat, rhs_in, Bmat) {
    extra_rhs = b;
  }
  SolverSettings(CSC* Amat, double* rhs_in, CSC* Bmat, CSC* BTmat) : SolverSettings(Amat, rhs_in, Bmat) {
    BT = BTmat;
  }
  SolverSettings(CSC* Amat, double* rhs_in, CSC* Bmat, CSC* BTmat, CSC* Cmat, CSC* CTmat)
      : SolverSettings(Amat, rhs_in, Bmat, BTmat) {
    C = Cmat;
    CT = CTmat;
  }
  ~SolverSettings() {
    if (h_) lblg_destroy(h_);
    if (kkt_) lblg_qp_destroy(kkt_);
  }
  SolverSettings(const SolverSettings&) = delete;
  SolverSettings& operator=(const SolverSettings&) = delete;

  // returns 1 on success, 0 on failure (linear_solver_wrapper.cpp:396-536)
  int symbolic_analysis() {
    const int n = (int)A->ncol;
    lblg_settings s = settings();
    if (solver_mode == 0) {
      if (h_) lblg_destroy(h_);
      h_ = nullptr;
      if (lblg_create(&h_, n, A->p, A->i, A->x) != LBLG_OK) return 0;
      if (lblg_analyze(h_, &s, nullptr, nullptr) != LBLG_OK) return 0;
      x = new double[n];
      return 1;
    }
    const int me = C ? (int)C->nrow : 0, mi = B ? (int)B->nrow : 0;
    if (kkt_) lblg_qp_destroy(kkt_);
    kkt_ = nullptr;
    std::vector<double> zeros((size_t)n + me + mi, 0.0);
    if (lblg_qp_create(&kkt_, n, A->p, A->i, A->x, zeros.data(), me, me ? C->p : nullptr, me ? C->i : nullptr,
                       me ? C->x : nullptr, zeros.data(), mi, mi ? B->p : nullptr, mi ? B->i : nullptr,
                       mi ? B->x : nullptr, zeros.data()) != LBLG_OK)
      return 0;
    if (lblg_kkt_analyze(kkt_, &s) != LBLG_OK) return 0;
    const size_t N = (size_t)lblg_kkt_size(kkt_);
    sm_rhs_.assign(N, 0.0);
    sm_solution_.assign(N, 0.0);
    for (int i = 0; i < n; ++i) sm_rhs_[i] = rhs[i];
    rhs = sm_rhs_.data();  // the reference redirects rhs to its super-KKT rhs (:418-420)
    x = sm_solution_.data();
    base = n + me;
    return 1;
  }
  // returns 1 on success (linear_solver_wrapper.cpp:602-768)
  int numerical_factorization() {
    lblg_settings s = settings();
    if (solver_mode == 1) return lblg_kkt_factor(kkt_) == LBLG_OK ? 1 : -1;
    if (lblg_set_settings(h_, &s) != LBLG_OK) return -1;
    return lblg_factor(h_) == LBLG_OK ? 1 : -1;
  }
  // mode 1: refactor after rows were added or dropped (:1151-1290)
  int update_factorization() { return solver_mode == 1 && lblg_kkt_factor(kkt_) == LBLG_OK ? 1 : -1; }
  // mode 1: activate rows with their right-hand sides / deactivate rows (:995-1018)
  int update_somod(std::vector<int> rows, const int n_rhs, const double* new_rhs) {
    (void)n_rhs;  // one right-hand side, as in the reference
    for (int k : rows) rhs[base + k] = new_rhs[k];
    return add_del_matrix_qp(1, rows);
  }
  int downdate_somod(std::vector<int> rows, int n_rhs) {
    (void)n_rhs;
    for (int k : rows) rhs[base + k] = 0.0;
    return add_del_matrix_qp(0, rows);
  }
  int add_del_matrix_qp(int add_del, std::vector<int> rows) {
    return lblg_kkt_edit(kkt_, add_del, (int)rows.size(), rows.data()) == LBLG_OK ? 0 : -1;
  }
  // solve with refinement for the rhs buffer (linear_solver_wrapper.cpp:1305-1404)
  double* solve_only() {
    if (solver_mode == 1) {
      if (lblg_kkt_solve(kkt_, rhs, x, &bwd_err) != LBLG_OK) return nullptr;
      return x;
    }
    lblg_settings s = settings();
    if (lblg_set_settings(h_, &s) != LBLG_OK) return nullptr;
    const int64_t n = (int64_t)A->ncol;
    if (lblg_solve(h_, 1, rhs, n, x, n, &bwd_err) != LBLG_OK) return nullptr;
    num_ref_iter = lblg_num_ref_iter(h_);
    return x;
  }
  double* solve_only(const int n_rhs, double* rhs_in) {
    (void)n_rhs;  // the reference copies one right-hand side (linear_solver_wrapper.cpp:1300-1303)
    const size_t N = solver_mode == 1 ? sm_rhs_.size() : A->ncol;
    std::memcpy(rhs, rhs_in, sizeof(double) * N);
    return solve_only();
  }
  // backward error of the last solve (linear_solver_wrapper.cpp:1462-1498)
  double backward_error() { return bwd_err; }
  // inertia of the perturbed matrix (not computed by the reference; mode 0)
  int inertia(int64_t* pos, int64_t* neg, int64_t* zero) { return lblg_inertia(h_, pos, neg, zero); }
  lblg_handle* handle() { return h_; }

 private:
  lblg_handle* h_ = nullptr;
  lblg_qp* kkt_ = nullptr;
  std::vector<double> sm_rhs_, sm_solution_;
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
