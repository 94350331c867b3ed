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

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "lblg.h"
#include "nasoq/common/def.h"

namespace nasoq {

// Timing fields of the reference's profiling_solver_info
// (include/nasoq/QP/linear_solver_wrapper.h:20-46 there), seconds of wall
// time of the last call of each phase.
struct profiling_solver_info {
  double fact_time = 0, analysis_time = 0, solve_time = 0, iter_time = 0, ordering_time = 0;
  double update_time = 0, piv_reord = 0;
};

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
  // readable state of the reference (linear_solver_wrapper.h:55-59, 75-96, 112, 140-141).
  // L and the level sets are filled by symbolic_analysis(); after
  // numerical_factorization() L->Perm / L->IPerm / L->s follow the pivoted
  // order and perm_piv, A_ord and AT_ord are rebuilt (reorder_matrix).
  // valL and d_val are host copies of the device factor: filled by
  // numerical_factorization() when L->xsize <= host_factor_max, else by
  // fetch_factor().
  BCSC* L = nullptr;
  double *valL = nullptr, *d_val = nullptr;
  CSC *A_ord = nullptr, *AT_ord = nullptr;
  int* perm_piv = nullptr;
  int n_level = 0, n_par = 0;
  int *level_ptr = nullptr, *par_ptr = nullptr, *par_set = nullptr;
  int max_sup_wid = 0, max_col = 0;
  double A_l1 = 0, x_l1 = 0, rhs_l1 = 0, res_l1 = 0;
  profiling_solver_info* psi = &psi_;
  // engine extensions
  int precision = LBLG_F64;
  double target_bwd_err = 0.0;
  int64_t host_factor_max = int64_t(1) << 24;  // entries of L copied to valL after each factorization

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
    const auto t0 = std::chrono::steady_clock::now();
    if (solver_mode == 0) {
      if (h_) lblg_destroy(h_);
      h_ = nullptr;
      if (lblg_create(&h_, n, A->p, A->i, A->x) != LBLG_OK) return 0;
      if (lblg_analyze(h_, &s, nullptr, nullptr) != LBLG_OK) return 0;
      x = new double[n];
      fill_symbolic();
      psi_.analysis_time = seconds_since(t0);
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
    fill_symbolic();
    psi_.analysis_time = seconds_since(t0);
    return 1;
  }
  // returns 1 on success (linear_solver_wrapper.cpp:602-768)
  int numerical_factorization() {
    lblg_settings s = settings();
    const auto t0 = std::chrono::steady_clock::now();
    if (solver_mode == 1) {
      if (lblg_kkt_factor(kkt_) != LBLG_OK) return -1;
    } else {
      if (lblg_set_settings(h_, &s) != LBLG_OK) return -1;
      if (lblg_factor(h_) != LBLG_OK) return -1;
    }
    psi_.fact_time = seconds_since(t0);
    return after_factor();
  }
  // mode 1: refactor after rows were added or dropped (:1151-1290)
  int update_factorization() {
    if (solver_mode != 1) return -1;
    const auto t0 = std::chrono::steady_clock::now();
    if (lblg_kkt_factor(kkt_) != LBLG_OK) return -1;
    psi_.update_time = seconds_since(t0);
    return after_factor();
  }
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
    const auto t0 = std::chrono::steady_clock::now();
    if (solver_mode == 1) {
      if (lblg_kkt_solve(kkt_, rhs, x, &bwd_err) != LBLG_OK) return nullptr;
      psi_.solve_time = seconds_since(t0);
      return x;
    }
    lblg_settings s = settings();
    if (lblg_set_settings(h_, &s) != LBLG_OK) return nullptr;
    const int64_t n = (int64_t)A->ncol;
    if (lblg_solve(h_, 1, rhs, n, x, n, &bwd_err) != LBLG_OK) return nullptr;
    num_ref_iter = lblg_num_ref_iter(h_);
    psi_.solve_time = seconds_since(t0);
    return x;
  }
  double* solve_only(const int n_rhs, double* rhs_in) {
    (void)n_rhs;  // the reference copies one right-hand side (linear_solver_wrapper.cpp:1300-1303)
    const size_t N = solver_mode == 1 ? sm_rhs_.size() : A->ncol;
    std::memcpy(rhs, rhs_in, sizeof(double) * N);
    return solve_only();
  }
  // A_ord / AT_ord: lower / upper CSC of P K P' with the current L->Perm
  // (linear_solver_wrapper.cpp:1432-1460: after the factor, the Perm of the
  // symbolic analysis composed with the pivot permutation)
  void reorder_matrix() {
    lblg_handle* h = eng();
    if (!h || !L) return;
    const int N = (int)L->n;
    std::vector<int> sp(N + 1);
    lblg_qp_super_kkt_or_a(sp);
    build_ordered(N, sp.data(), kp_i_.data(), kp_x_.data());
  }
  // res_l1, x_l1, rhs_l1 of the current x and rhs and A_l1 = ||K0||_inf,
  // recomputed on the device (linear_solver_wrapper.cpp:1462-1487)
  void compute_norms() {
    lblg_handle* h = eng();
    if (!h) return;
    const int64_t N = (int64_t)L->n;
    double nm[4];
    if (lblg_residual_norms(h, 1, rhs, N, x, N, nm) != LBLG_OK) return;
    int zp = 0;
    double fs = 0;
    lblg_factor_info(h, &zp, &A_l1, &fs);
    res_l1 = nm[0];
    x_l1 = nm[1];
    rhs_l1 = nm[2];
  }
  // recomputed from the current x and rhs (linear_solver_wrapper.cpp:1489-1498)
  double backward_error() {
    compute_norms();
    const double den = A_l1 * x_l1 + rhs_l1;
    bwd_err = den > 0 ? res_l1 / den : 0.0;
    return bwd_err;
  }
  // host copies of the factor values (valL, d_val) whatever their size
  int fetch_factor() {
    lblg_handle* h = eng();
    if (!h || !L) return -1;
    valL_.resize(L->xsize);
    d_val_.resize(2 * L->n);
    if (lblg_get_factor(h, valL_.data(), d_val_.data(), nullptr, nullptr) != LBLG_OK) return -1;
    valL = valL_.data();
    d_val = d_val_.data();
    L->x = valL;
    return 1;
  }
  // inertia of the perturbed matrix (not computed by the reference)
  int inertia(int64_t* pos, int64_t* neg, int64_t* zero) { return lblg_inertia(eng(), pos, neg, zero); }
  lblg_handle* handle() { return h_; }

 private:
  lblg_handle* h_ = nullptr;
  lblg_qp* kkt_ = nullptr;
  std::vector<double> sm_rhs_, sm_solution_;
  profiling_solver_info psi_;
  BCSC L_{};
  std::vector<int> perm_, iperm_, colcount_, super_, col2sup_, s_, sparent_, parent_, perm_piv_;
  std::vector<size_t> p_, pi_, i_ptr_;
  std::vector<int> level_ptr_, par_ptr_, par_set_;
  std::vector<double> valL_, d_val_;
  std::vector<int> kp_i_;     // the matrix being factored (A, or the super-KKT of mode 1)
  std::vector<double> kp_x_;
  std::vector<int> kp_p_;
  CSC A_ord_{}, AT_ord_{};
  std::vector<int> ao_p_, ao_i_, at_p_, at_i_;
  std::vector<double> ao_x_, at_x_;

  lblg_handle* eng() { return solver_mode == 1 ? (kkt_ ? lblg_kkt_handle(kkt_) : nullptr) : h_; }
  static double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  template <typename D, typename S>
  static void copy_vec(lblg_handle* h, int which, bool longs, std::vector<D>& v) {
    const int64_t len = longs ? lblg_symbolic_longs(h, which, nullptr) : lblg_symbolic_ints(h, which, nullptr);
    std::vector<S> tmp(len > 0 ? len : 0);
    if (len > 0) {
      if constexpr (sizeof(S) == 8) lblg_symbolic_longs(h, which, reinterpret_cast<int64_t*>(tmp.data()));
      else lblg_symbolic_ints(h, which, reinterpret_cast<int*>(tmp.data()));
    }
    v.assign(tmp.begin(), tmp.end());
  }
  // the matrix handed to the engine (mode 0: A; mode 1: the super-KKT with its current values)
  void lblg_qp_super_kkt_or_a(std::vector<int>& p) {
    if (solver_mode == 1) {
      const int64_t nnz = lblg_qp_super_kkt(kkt_, nullptr, nullptr, nullptr);
      kp_i_.resize(nnz);
      kp_x_.resize(nnz);
      lblg_qp_super_kkt(kkt_, p.data(), kp_i_.data(), kp_x_.data());
      return;
    }
    const int n = (int)A->ncol;
    std::memcpy(p.data(), A->p, sizeof(int) * (n + 1));
    kp_i_.assign(A->i, A->i + A->p[n]);
    kp_x_.assign(A->x, A->x + A->p[n]);
  }
  void fill_symbolic() {
    lblg_handle* h = eng();
    copy_vec<int, int>(h, 0, false, perm_);
    copy_vec<int, int>(h, 1, false, iperm_);
    copy_vec<int, int>(h, 2, false, parent_);
    copy_vec<int, int>(h, 3, false, colcount_);
    copy_vec<int, int>(h, 4, false, super_);
    copy_vec<int, int>(h, 5, false, col2sup_);
    copy_vec<int, int>(h, 6, false, sparent_);
    copy_vec<int, int>(h, 7, false, s_);
    copy_vec<int, int>(h, 8, false, level_ptr_);
    copy_vec<int, int>(h, 9, false, par_ptr_);
    copy_vec<int, int>(h, 10, false, par_set_);
    copy_vec<size_t, int64_t>(h, 0, true, pi_);
    copy_vec<size_t, int64_t>(h, 1, true, p_);
    copy_vec<size_t, int64_t>(h, 2, true, i_ptr_);
    L_ = BCSC{};
    L_.n = perm_.size();
    L_.minor = L_.n;
    L_.Perm = perm_.data();
    L_.IPerm = iperm_.data();
    L_.ColCount = colcount_.data();
    L_.p = p_.data();
    L_.nsuper = (size_t)lblg_symbolic_scalar(h, 1);
    L_.ssize = (size_t)lblg_symbolic_scalar(h, 7);
    L_.xsize = (size_t)lblg_symbolic_scalar(h, 6);
    L_.super = super_.data();
    L_.col2Sup = col2sup_.data();
    L_.pi = pi_.data();
    L_.i_ptr = i_ptr_.data();
    L_.s = s_.data();
    L_.sParent = sparent_.data();
    L_.Parent = parent_.data();
    L_.is_super = 1;
    L = &L_;
    n_level = (int)lblg_symbolic_scalar(h, 2);
    n_par = (int)lblg_symbolic_scalar(h, 3);
    level_ptr = level_ptr_.data();
    par_ptr = par_ptr_.data();
    par_set = par_set_.data();
    max_sup_wid = (int)lblg_symbolic_scalar(h, 4);
    max_col = (int)lblg_symbolic_scalar(h, 5);
    iperm_sym_ = iperm_;
    perm_piv_.resize(L_.n);
    for (size_t k = 0; k < L_.n; ++k) perm_piv_[k] = (int)k;
    perm_piv = perm_piv_.data();
    valL = d_val = nullptr;
    reorder_matrix();
  }
  // after a factorization: pivoted Perm / IPerm / row labels, perm_piv, the
  // reordered matrix and (small factors) the host copies of the values
  int after_factor() {
    lblg_handle* h = eng();
    const size_t N = L_.n;
    std::vector<int> pf(N);
    const auto t0 = std::chrono::steady_clock::now();
    if (lblg_get_factor(h, nullptr, nullptr, s_.data(), pf.data()) != LBLG_OK) return -1;
    // perm_piv: pivoted position -> position in the symbolic ordering
    for (size_t k = 0; k < N; ++k) perm_piv_[k] = iperm_sym_[pf[k]];
    perm_ = pf;
    for (size_t k = 0; k < N; ++k) iperm_[perm_[k]] = (int)k;
    L_.Perm = perm_.data();
    L_.IPerm = iperm_.data();
    L_.s = s_.data();
    reorder_matrix();
    psi_.piv_reord = seconds_since(t0);
    valL = d_val = nullptr;
    L_.x = nullptr;
    if ((int64_t)L_.xsize <= host_factor_max && fetch_factor() != 1) return -1;
    return 1;
  }
  std::vector<int> iperm_sym_;  // inverse of the symbolic analysis' ordering
  // lower (A_ord) and upper (AT_ord) CSC of P K P' for perm_ (new = iperm_[old])
  void build_ordered(int N, const int* p, const int* i, const double* xv) {
    std::vector<int> cnt(N + 1, 0);
    for (int j = 0; j < N; ++j)
      for (int q = p[j]; q < p[j + 1]; ++q) {
        const int a = iperm_[i[q]], b = iperm_[j];
        cnt[(a < b ? a : b) + 1]++;
      }
    for (int j = 0; j < N; ++j) cnt[j + 1] += cnt[j];
    ao_p_ = cnt;
    ao_i_.resize(cnt[N]);
    ao_x_.resize(cnt[N]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int q = p[j]; q < p[j + 1]; ++q) {
        const int a = iperm_[i[q]], b = iperm_[j];
        const int c = a < b ? a : b, r = a < b ? b : a;
        const int d = fill[c]++;
        ao_i_[d] = r;
        ao_x_[d] = xv[q];
      }
    for (int c = 0; c < N; ++c) {  // sorted rows, the diagonal first
      std::vector<std::pair<int, double>> col;
      for (int q = ao_p_[c]; q < ao_p_[c + 1]; ++q) col.emplace_back(ao_i_[q], ao_x_[q]);
      std::sort(col.begin(), col.end());
      for (size_t k = 0; k < col.size(); ++k) {
        ao_i_[ao_p_[c] + k] = col[k].first;
        ao_x_[ao_p_[c] + k] = col[k].second;
      }
    }
    // upper CSC = transpose of the lower one
    at_p_.assign(N + 1, 0);
    for (int c = 0; c < N; ++c)
      for (int q = ao_p_[c]; q < ao_p_[c + 1]; ++q) at_p_[ao_i_[q] + 1]++;
    for (int j = 0; j < N; ++j) at_p_[j + 1] += at_p_[j];
    at_i_.resize(ao_i_.size());
    at_x_.resize(ao_x_.size());
    std::vector<int> f2(at_p_.begin(), at_p_.end() - 1);
    for (int c = 0; c < N; ++c)
      for (int q = ao_p_[c]; q < ao_p_[c + 1]; ++q) {
        const int d = f2[ao_i_[q]]++;
        at_i_[d] = c;
        at_x_[d] = ao_x_[q];
      }
    auto set = [&](CSC& M, std::vector<int>& pp, std::vector<int>& ii, std::vector<double>& xx, int stype) {
      M = CSC{};
      M.nrow = M.ncol = (size_t)N;
      M.nzmax = ii.size();
      M.p = pp.data();
      M.i = ii.data();
      M.x = xx.data();
      M.stype = stype;
      M.xtype = CHOLMOD_REAL;
      M.sorted = 1;
      M.packed = 1;
    };
    set(A_ord_, ao_p_, ao_i_, ao_x_, -1);
    set(AT_ord_, at_p_, at_i_, at_x_, 1);
    A_ord = &A_ord_;
    AT_ord = &AT_ord_;
  }
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
