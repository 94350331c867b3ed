// NASOQ dual active-set QP solver over the device LBL engine (see qp.hpp).
// The iteration follows Nasoq::numeric_QP (src/nasoq.cpp:478-610); each
// function cites the reference routine whose steps it restates.
#include "host/qp.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

namespace lblg {

namespace {
void check(int rc, const char* what) {
  if (rc != LBLG_OK) throw std::runtime_error(std::string(what) + ": " + lblg_last_error());
}
// y = M x for a general CSC matrix (spmv_csc_small)
void spmv(const SparseCsc& M, const double* x, double* y) {
  std::fill(y, y + M.nrow, 0.0);
  for (int j = 0; j < M.ncol; ++j)
    for (int k = M.p[j]; k < M.p[j + 1]; ++k) y[M.i[k]] += M.x[k] * x[j];
}
// y = H x for H given by its lower triangle (spmv_csc_sym_one_int)
void spmv_sym(const SparseCsc& H, const double* x, double* y) {
  std::fill(y, y + H.ncol, 0.0);
  for (int j = 0; j < H.ncol; ++j)
    for (int k = H.p[j]; k < H.p[j + 1]; ++k) {
      const int i = H.i[k];
      if (i == j) {
        y[i] += H.x[k] * x[j];
      } else if (i > j) {
        y[i] += H.x[k] * x[j];
        y[j] += H.x[k] * x[i];
      }
    }
}
double inf_norm(const std::vector<double>& v) {  // norm_dense(m, 1, v, 0)
  double s = 0.0;
  for (double e : v) {
    const double a = std::abs(e);
    if ((!std::isfinite(a) || a > s) && std::isfinite(s)) s = a;
  }
  return s;
}
}  // namespace

SparseCsc transpose(const SparseCsc& M) {
  SparseCsc T;
  T.nrow = M.ncol;
  T.ncol = M.nrow;
  T.p.assign(T.ncol + 1, 0);
  for (int k = 0; k < M.p[M.ncol]; ++k) ++T.p[M.i[k] + 1];
  for (int c = 0; c < T.ncol; ++c) T.p[c + 1] += T.p[c];
  T.i.resize(M.p[M.ncol]);
  T.x.resize(M.p[M.ncol]);
  std::vector<int> next(T.p.begin(), T.p.end() - 1);
  for (int j = 0; j < M.ncol; ++j)
    for (int k = M.p[j]; k < M.p[j + 1]; ++k) {
      const int d = next[M.i[k]]++;
      T.i[d] = j;
      T.x[d] = M.x[k];
    }
  return T;
}

QpSolver::~QpSolver() {
  if (h) lblg_destroy(h);
}

// SolverSettings::build_super_matrix (linear_solver_wrapper.cpp:246-350):
// column j < n holds H(:, j), then the equality rows n + A(:, j) (values),
// then the inequality rows n + me + B(:, j) (pattern only, values zero until
// the constraint is activated); columns n.. hold one diagonal slot each.
void QpSolver::build_super_matrix() {
  const int N = n + me + mi;
  // a column of H without its diagonal entry gets an explicit zero diagonal
  // (the engine, like find_perturbation, reads the diagonal first)
  std::vector<int> nodiag(n, 0);
  for (int j = 0; j < n; ++j) nodiag[j] = !(H.p[j] < H.p[j + 1] && H.i[H.p[j]] == j);
  SM.nrow = SM.ncol = N;
  SM.p.assign(N + 1, 0);
  for (int j = 0; j < n; ++j)
    SM.p[j + 1] = SM.p[j] + nodiag[j] + (H.p[j + 1] - H.p[j]) + (me ? A.p[j + 1] - A.p[j] : 0) +
                  (mi ? B.p[j + 1] - B.p[j] : 0);
  for (int j = n; j < N; ++j) SM.p[j + 1] = SM.p[j] + 1;
  SM.i.resize(SM.p[N]);
  SM.x.assign(SM.p[N], 0.0);
  bpos.assign(mi ? B.p[n] : 0, 0);
  for (int j = 0; j < n; ++j) {
    int64_t d = SM.p[j];
    if (nodiag[j]) SM.i[d++] = j;
    for (int k = H.p[j]; k < H.p[j + 1]; ++k, ++d) {
      SM.i[d] = H.i[k];
      SM.x[d] = H.x[k];
    }
    if (me)
      for (int k = A.p[j]; k < A.p[j + 1]; ++k, ++d) {
        SM.i[d] = n + A.i[k];
        SM.x[d] = A.x[k];
      }
    if (mi)
      for (int k = B.p[j]; k < B.p[j + 1]; ++k, ++d) {
        SM.i[d] = n + me + B.i[k];
        bpos[k] = d;
      }
  }
  for (int j = n; j < N; ++j) SM.i[SM.p[j]] = j;
  // B entries grouped by constraint (row of B), in column order
  brow_ptr.assign(mi + 1, 0);
  if (mi) {
    for (int k = 0; k < B.p[n]; ++k) ++brow_ptr[B.i[k] + 1];
    for (int c = 0; c < mi; ++c) brow_ptr[c + 1] += brow_ptr[c];
    brow_ent.resize(B.p[n]);
    std::vector<int> next(brow_ptr.begin(), brow_ptr.end() - 1);
    for (int j = 0; j < n; ++j)
      for (int k = B.p[j]; k < B.p[j + 1]; ++k) brow_ent[next[B.i[k]]++] = k;
  }
  rhs.assign(N, 0.0);
  sol.assign(N, 0.0);
  kkt_changed.clear();
  kkt_uploaded = false;
}

double QpSolver::hessian_frobenius() const {  // norm_sparse_int(n, H, -1, 2)
  double frob = 0.0;
  for (int j = 0; j < n; ++j)
    for (int k = H.p[j]; k < H.p[j + 1]; ++k) {
      const double s = std::abs(H.x[k]);
      if (H.i[k] == j) frob += s * s;
      else if (H.i[k] > j) frob += 2 * s * s;
    }
  return std::sqrt(frob);
}

// Nasoq::symbolic_QP (nasoq.cpp:201-278) with determine_iterations (:375-407)
// and auto_perturbation2 (:321-373)
void QpSolver::symbolic() {
  AT = transpose(A);
  BT = transpose(B);
  const int64_t total_nnz = (int64_t)H.p[n] + (mi ? B.p[n] : 0);
  if (auto_reg_en == 2) {
    const double hn = hessian_frobenius();
    const double ineq_ratio = mi / (double)n;
    diag_perturb = 1e-8;
    int low = 2, high = 9, large = 0;
    if (eps_abs >= 1e-3) {
      low /= 2;
      high /= 2;
    } else if (eps_abs < 1e-8) {
      low *= 2;
      large = large * 2 + 1;
    }
    inner_iter_ref = outer_iter_ref = large;
    if (total_nnz < 3e5) {
      inner_iter_ref = outer_iter_ref = low;
      stop_tol = 1e-10;
      if (hn < 1) {
        diag_perturb = 1e-7;
        stop_tol = 1e-15;
      } else if (ineq_ratio > 2.5) {
        diag_perturb = 1e-10;
        stop_tol = 1e-16;
      } else if (n > 1e4 && mi >= 1e4) {
        diag_perturb = 1e-11;
        inner_iter_ref = outer_iter_ref = high;
        stop_tol = 1e-15;
      }
    } else if (hn < 1) {
      diag_perturb = 1e-7;
      inner_iter_ref = outer_iter_ref = 2;
      stop_tol = 1e-15;
    }
  } else if (auto_reg_en == 1) {
    if (eps_abs >= 1e-3) inner_iter_ref = outer_iter_ref = 0;
    else if (eps_abs <= 1e-4 && eps_abs > 1e-7) inner_iter_ref = outer_iter_ref = 1;
    else if (eps_abs <= 1e-7 && eps_abs > 1e-10) inner_iter_ref = outer_iter_ref = 2;
    else inner_iter_ref = outer_iter_ref = 4;
    if (total_nnz <= 20000) inner_iter_ref = outer_iter_ref = 3;
    stop_tol = 1e-15;
    diag_perturb = hessian_frobenius() < 1 ? 1e-7 : 1e-9;
  } else {
    outer_iter_ref = inner_iter_ref = max_iter;
  }
  zero_thresh = diag_perturb;
  build_super_matrix();
  const int N = SM.ncol;
  lblg_settings st;
  lblg_default_settings(&st);
  st.reg_diag = diag_perturb;
  st.req_ref_iter = outer_iter_ref;
  st.max_inner_iter = inner_iter_ref;
  st.tol_abs = st.tol_rel = stop_tol;
  st.ordering = ordering;
  st.precision = precision;
  if (precision != LBLG_F64) {
    // an FP32 / BF16 factor is a preconditioner: the variant's refinement
    // budget fits an FP64 factor (one GMRES(1) cycle for larger QPs), so the
    // KKT solves refine to an FP64-level backward error instead
    st.target_bwd_err = 1e-13;
    st.max_refine_rounds = 40;
  }
  // the super-KKT pattern depends on the problem only: a later solve of the
  // same problem (new q or b, or the next call) keeps the analysis and the
  // device factor, and the KKT values are refreshed at the next factorization
  if (h && analysed && std::memcmp(&st, &analysed_settings, sizeof(st)) == 0 &&
      analysed_separate == separate_rows)
    return;
  if (h) lblg_destroy(h);
  h = nullptr;
  check(lblg_create(&h, N, SM.p.data(), SM.i.data(), SM.x.data()), "qp: create");
  // inequality rows are kept as single-column supernodes (extra_cols,
  // SYM_REMOV), or join supernodes (separate_rows = 0: an inactive row is
  // decoupled, so it needs no supernode of its own here)
  std::vector<int> extra(N, 0);
  for (int j = n + me; j < N; ++j) extra[j] = 1;
  check(lblg_analyze(h, &st, nullptr, separate_rows ? extra.data() : nullptr), "qp: analyze");
  analysed = true;
  analysed_settings = st;
  analysed_separate = separate_rows;
}

// SolverSettings::symbolic_analysis in mode 1 (linear_solver_wrapper.cpp:396-536)
void QpSolver::analyze_lbl(const lblg_settings& st) {
  AT = transpose(A);
  BT = transpose(B);
  diag_perturb = zero_thresh = st.reg_diag;
  outer_iter_ref = st.req_ref_iter;
  inner_iter_ref = st.max_inner_iter;
  stop_tol = st.tol_abs;
  build_super_matrix();
  const int N = SM.ncol;
  if (h) lblg_destroy(h);
  h = nullptr;
  check(lblg_create(&h, N, SM.p.data(), SM.i.data(), SM.x.data()), "kkt: create");
  std::vector<int> extra(N, 0);
  for (int j = n + me; j < N; ++j) extra[j] = 1;
  lblg_settings s2 = st;
  if (s2.ordering == LBLG_ORDER_GIVEN) s2.ordering = LBLG_ORDER_ND;
  check(lblg_analyze(h, &s2, nullptr, extra.data()), "kkt: analyze");
  used.assign(mi, 0);
  active_set.clear();
  n_active = 0;
  analysed = true;
}

// SolverSettings::update_factorization (linear_solver_wrapper.cpp:1151-1290)
// after add_del_matrix_qp edits, numerical_factorization for a new super-KKT:
// the device receives only the edited entries and refactors the supernodes
// on their tree paths (lblg_update_values), or all of SM.x the first time
void QpSolver::factor_kkt() {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  if (!kkt_uploaded) {
    check(lblg_set_values(h, SM.x.data()), "qp: set values");
  } else if (!kkt_changed.empty()) {
    std::vector<double> v(kkt_changed.size());
    for (size_t k = 0; k < v.size(); ++k) v[k] = SM.x[kkt_changed[k]];
    check(lblg_update_values(h, (int64_t)v.size(), kkt_changed.data(), v.data()), "qp: update values");
  }
  check(lblg_factor(h), "qp: factor");
  kkt_uploaded = true;
  kkt_changed.clear();
  factor_seconds += std::chrono::duration<double>(clock::now() - t0).count();
  refactored_supernodes += lblg_refactored_supernodes(h);
  ++factorizations;
}

// Nasoq::solve_kkt (nasoq.cpp:918-949): refactor (UPDATE / REFACTOR) or reuse
// the factor (SOLVE), then the refined solve
void QpSolver::solve_kkt(bool refactor) {
  using clock = std::chrono::steady_clock;
  if (refactor) factor_kkt();
  const int64_t N = SM.ncol;
  double bwd = 0.0;
  const auto t0 = clock::now();
  check(lblg_solve(h, 1, rhs.data(), N, sol.data(), N, &bwd), "qp: solve");
  solve_seconds += std::chrono::duration<double>(clock::now() - t0).count();
  ++kkt_solves;
}

// SolverSettings::add_del_matrix_qp (linear_solver_wrapper.cpp:1020-1137):
// the constraint's couplings take their values (add) or zero (drop)
void QpSolver::edit_kkt(bool add, int c) {
  for (int e = brow_ptr[c]; e < brow_ptr[c + 1]; ++e) {
    const int k = brow_ent[e];
    SM.x[bpos[k]] = add ? B.x[k] : 0.0;
    kkt_changed.push_back(bpos[k]);
  }
}

// Nasoq::initialize_x (nasoq.cpp:438-476): the equality-constrained solution
void QpSolver::initialize_x() {
  std::fill(rhs.begin(), rhs.end(), 0.0);
  for (int i = 0; i < n; ++i) rhs[i] = -q[i];
  for (int i = 0; i < me; ++i) rhs[n + i] = a[i];
  solve_kkt(true);
  for (int i = 0; i < n; ++i) primal[i] = sol[i];
  for (int i = 0; i < me; ++i) dual_eq[i] = sol[n + i];
}

// Nasoq::primal_feasibility (nasoq.cpp:787-840): count the violated
// inequalities and pick the most violated one
int QpSolver::primal_feasibility(int& nxt_add) {
  nxt_add = -1;
  if (mi == 0) return 0;
  std::vector<double> t(mi);
  spmv(B, primal.data(), t.data());
  const double tol = eps_abs;
  double maxv = tol;
  int violated = 0;
  for (int i = 0; i < mi; ++i) {
    t[i] -= b[i];
    if (t[i] > tol) ++violated;
    if (t[i] > maxv) {
      nxt_add = i;
      maxv = t[i];
    }
  }
  return violated;
}

// Nasoq::dual_step_length (nasoq.cpp:852-864)
double QpSolver::dual_step_length(int& nxt_drop) const {
  double step = std::numeric_limits<double>::max();
  const double* lm = sol.data() + n + me;
  for (int i = 0; i < mi; ++i)
    if (lm[i] > 0 && used[i]) {
      const double t = dual[i] / lm[i];
      if (t < step) {
        step = t;
        nxt_drop = i;
      }
    }
  return step;
}

// Nasoq::primal_step_length (nasoq.cpp:866-878)
double QpSolver::primal_step_length(int c) const {
  double bx = 0.0, bd = 0.0;
  for (int k = BT.p[c]; k < BT.p[c + 1]; ++k) {
    bx += BT.x[k] * primal[BT.i[k]];
    bd += BT.x[k] * sol[BT.i[k]];
  }
  if (bd < zero_thresh) return std::numeric_limits<double>::max();
  return -(b[c] - bx) / bd;
}

// Nasoq::update_rhs (nasoq.cpp:880-888): right-hand side = row c of B
void QpSolver::update_rhs(int c) {
  if (c < 0 || c >= mi) return;
  std::fill(rhs.begin(), rhs.end(), 0.0);
  for (int k = BT.p[c]; k < BT.p[c + 1]; ++k) rhs[BT.i[k]] = BT.x[k];
}

// Nasoq::update_primal / update_dual (nasoq.cpp:612-633)
void QpSolver::update_primal(double step) {
  for (int i = 0; i < n; ++i) primal[i] -= step * sol[i];
}
void QpSolver::update_dual(double step) {
  for (int i = 0; i < me; ++i) dual_eq[i] -= step * sol[n + i];
  for (int c : active_set) dual[c] -= step * sol[n + me + c];
}

// Nasoq::update_active_set (nasoq.cpp:635-654)
int QpSolver::update_active_set(bool add, int c) {
  if (add) {
    ++n_active;
    active_set.push_back(c);
    used[c] = 1;
    return 1;
  }
  --n_active;
  if (active_set.empty()) return 0;
  auto it = std::find(active_set.begin(), active_set.end(), c);
  if ((active_set.size() == 1 && active_set[0] == c) || (it != active_set.end() && active_set.size() > 1)) {
    active_set.erase(it);
    used[c] = 0;
    return 1;
  }
  return 0;
}

// Nasoq::numeric_QP (nasoq.cpp:478-610): the dual active-set iteration
int QpSolver::numeric() {
  int nxt_active = -1, nxt_drop = -1;
  bool to_add = true;
  bool refactor = warm_start != 0;  // REFACTOR after a warm start, else SOLVE
  const double inf = std::numeric_limits<double>::max();
  while (true) {
    if (num_iter > max_iter_nas) return 0;
    if (to_add) {
      const int violated = primal_feasibility(nxt_active);
      if (violated == 0) return 1;
      if (nxt_active == -1) return 0;  // every constraint is used
      update_rhs(nxt_active);
    }
    solve_kkt(refactor);
    const double ps = primal_step_length(nxt_active);
    const double ds = dual_step_length(nxt_drop);
    const double step = ps > ds ? ds : ps;
    if (step == inf) return 0;
    if (ps == inf && step == ds) {  // dual step: drop a constraint
      update_dual(step);
      dual[nxt_active] += step;
      to_add = false;
      edit_kkt(false, nxt_drop);
      refactor = true;
      update_active_set(false, nxt_drop);
    } else {  // full step
      update_primal(step);
      update_dual(step);
      dual[nxt_active] += step;
      if (step == ps) {
        to_add = true;
        edit_kkt(true, nxt_active);
        refactor = true;
        update_active_set(true, nxt_active);
      } else if (step == ds) {
        to_add = false;
        edit_kkt(false, nxt_drop);
        refactor = true;
        update_active_set(false, nxt_drop);
      }
    }
    ++num_iter;
  }
}

// Nasoq::constraint_sat_norm (nasoq.cpp:1097-1131)
double QpSolver::constraint_sat_norm() {
  std::vector<double> t(mi);
  if (mi) spmv(B, primal.data(), t.data());
  for (int i = 0; i < mi; ++i) {
    const double d = t[i] - b[i];
    t[i] = d < 0.0 ? 0.0 : d;
  }
  cons_sat_norm = inf_norm(t);
  if (me > 0) {
    std::vector<double> e(me);
    spmv(A, primal.data(), e.data());
    for (int i = 0; i < me; ++i) e[i] -= a[i];
    cons_sat_norm = std::max(cons_sat_norm, inf_norm(e));
  }
  return cons_sat_norm;
}

// Nasoq::lagrangian_residual_norm (nasoq.cpp:1058-1095): ||Hx + q + A'y + B'z||_inf
double QpSolver::lagrangian_residual_norm() {
  std::vector<double> r(n), t(n);
  spmv_sym(H, primal.data(), r.data());
  if (me > 0) {
    spmv(AT, dual_eq.data(), t.data());
    for (int i = 0; i < n; ++i) r[i] += t[i];
  }
  if (mi > 0) {
    spmv(BT, dual.data(), t.data());
    for (int i = 0; i < n; ++i) r[i] += t[i];
  }
  for (int i = 0; i < n; ++i) r[i] += q[i];
  lag_res = inf_norm(r);
  return lag_res;
}

// Nasoq::complementarity_norm (nasoq.cpp:1029-1049)
double QpSolver::complementarity_norm() {
  std::vector<double> t(mi);
  if (mi) spmv(B, primal.data(), t.data());
  for (int i = 0; i < mi; ++i) t[i] = dual[i] * (t[i] - b[i]);
  complementarity_infn = inf_norm(t);
  return complementarity_infn;
}

// Nasoq::non_negativity_norm (nasoq.cpp:1180-1187)
double QpSolver::non_negativity_norm() {
  std::vector<double> t(mi);
  for (int i = 0; i < mi; ++i) t[i] = dual[i] > 0.0 ? 0.0 : dual[i];
  non_negativity_infn = inf_norm(t);
  return non_negativity_infn;
}

// Nasoq::compute_primal_obj + compute_dual_obj (nasoq.cpp:958-1027)
void QpSolver::compute_objective() {
  std::vector<double> t(n);
  spmv_sym(H, primal.data(), t.data());
  double pq = 0.0, pp = 0.0;
  for (int i = 0; i < n; ++i) {
    pp += t[i] * primal[i];
    pq += q[i] * primal[i];
  }
  primal_obj = 0.5 * pp + pq;
  std::vector<double> s(mi);
  if (mi) spmv(B, primal.data(), s.data());
  dual_obj = 0.0;
  for (int i = 0; i < mi; ++i) dual_obj += (s[i] - b[i]) * dual[i];
  objective = primal_obj + dual_obj;
}

// Nasoq::check_solve_status (nasoq.cpp:712-728)
int QpSolver::check_status(int converged) {
  if (converged == 0) return QP_INFEASIBLE;
  constraint_sat_norm();
  lagrangian_residual_norm();
  complementarity_norm();
  non_negativity_norm();
  if (lag_res <= eps_abs && cons_sat_norm <= eps_abs && non_negativity_infn <= eps_abs) return QP_OPTIMAL;
  if (cons_sat_norm <= eps_abs) return QP_INACCURATE;
  return QP_NOT_CONVERGED;
}

// Nasoq::solve() (nasoq.cpp:656-710)
int QpSolver::solve() {
  status = QP_NOT_CONVERGED;
  warm_start = 0;
  primal.assign(n, 0.0);
  dual.assign(mi, 0.0);
  dual_eq.assign(me, 0.0);
  used.assign(mi, 0);
  active_set.clear();
  n_active = 0;
  num_iter = 0;
  int conv = 0;
  if (variant == QP_TUNED) {
    symbolic();
    for (int r = 0; r < 4; ++r) {
      initialize_x();
      conv = numeric();
    }
  } else {
    auto_reg_en = variant == QP_AUTO ? 2 : (variant == QP_FIXED ? 1 : 0);
    symbolic();
    initialize_x();
    conv = numeric();
  }
  compute_objective();
  status = check_status(conv);
  return status;
}

// Nasoq::solve(primal, dual, active_set) (nasoq.cpp:740-760) with
// set_warm_start (:413-436)
int QpSolver::solve_warm(const double* x0, const double* dual_in, const int* as, int nas,
                         const double* dual_eq_in) {
  primal.assign(x0, x0 + n);
  dual.assign(dual_in, dual_in + mi);
  dual_eq.assign(me, 0.0);
  if (me && dual_eq_in) dual_eq.assign(dual_eq_in, dual_eq_in + me);
  used.assign(mi, 0);
  active_set.clear();
  n_active = 0;
  num_iter = 0;
  symbolic();
  for (int k = 0; k < nas; ++k) {
    const int c = as[k];
    if (c < 0 || c >= mi) throw std::invalid_argument("qp: active constraint out of range");
    active_set.push_back(c);
    edit_kkt(true, c);
    ++n_active;
    used[c] = 1;
  }
  warm_start = 1;
  const int conv = numeric();
  compute_objective();
  constraint_sat_norm();
  lagrangian_residual_norm();
  status = conv;
  return conv;
}

// Warm start from a predicted active set (engine extension for the learned
// warm start): the KKT system with the predicted constraints as equalities
// gives x and their multipliers; constraints with negative multipliers are
// dropped and the system solved again until the multipliers are
// non-negative.  That state (x, z >= 0, Hx + q + A'y + B_W'z = 0) is the one
// the dual active-set loop keeps (Nasoq::numeric_QP, nasoq.cpp:478-610), so
// the loop continues from it as after set_warm_start (:413-436) and ends at
// the same optimum as a cold start; only the iteration count changes.
int QpSolver::solve_predicted(const int* as, int nas) {
  status = QP_NOT_CONVERGED;
  num_iter = 0;
  predicted_prunes = 0;
  auto_reg_en = variant == QP_AUTO ? 2 : (variant == QP_FIXED ? 1 : 0);
  symbolic();
  std::vector<unsigned char> in(mi, 0);
  std::vector<int> act;
  for (int k = 0; k < nas; ++k) {
    const int c = as[k];
    if (c < 0 || c >= mi) throw std::invalid_argument("qp: predicted constraint out of range");
    if (in[c]) continue;
    in[c] = 1;
    act.push_back(c);
    edit_kkt(true, c);
  }
  while (true) {
    std::fill(rhs.begin(), rhs.end(), 0.0);
    for (int i = 0; i < n; ++i) rhs[i] = -q[i];
    for (int i = 0; i < me; ++i) rhs[n + i] = a[i];
    for (int c : act) rhs[n + me + c] = b[c];
    solve_kkt(true);
    ++predicted_prunes;
    std::vector<int> keep;
    bool dropped = false;
    for (int c : act) {
      if (sol[n + me + c] < 0.0) {
        edit_kkt(false, c);
        in[c] = 0;
        dropped = true;
      } else {
        keep.push_back(c);
      }
    }
    act.swap(keep);
    if (!dropped) break;
  }
  primal.assign(sol.begin(), sol.begin() + n);
  dual_eq.assign(sol.begin() + n, sol.begin() + n + me);
  dual.assign(mi, 0.0);
  for (int c : act) dual[c] = sol[n + me + c];
  used.assign(in.begin(), in.end());
  active_set = act;
  n_active = (int)act.size();
  warm_start = 1;
  const int conv = numeric();
  compute_objective();
  status = check_status(conv);
  return status;
}

}  // namespace lblg
