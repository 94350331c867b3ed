// C-ABI of the NASOQ QP solver (include/lblg.h, lblg_qp_*).
#include <cstring>
#include <new>
#include <stdexcept>

#include "host/qp.hpp"
#include "lblg.h"

using lblg::QpSolver;

namespace lblg {
void set_last_error(const char* what);  // capi_numeric.cpp (lblg_last_error)
}

namespace {
template <typename F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    lblg::set_last_error("allocation failed");
    return LBLG_ERR_NOMEM;
  } catch (const std::invalid_argument& e) {
    lblg::set_last_error(e.what());
    return LBLG_ERR_ARG;
  } catch (const std::exception& e) {
    lblg::set_last_error(e.what());
    return LBLG_ERR_DEVICE;
  }
}
bool valid_csc(int nrow, int ncol, const int* p, const int* i) {
  if (!p || (p[ncol] > 0 && !i) || p[0] != 0) return false;
  for (int j = 0; j < ncol; ++j) {
    if (p[j + 1] < p[j]) return false;
    for (int k = p[j]; k < p[j + 1]; ++k) {
      if (i[k] < 0 || i[k] >= nrow) return false;
      if (k > p[j] && i[k] <= i[k - 1]) return false;
    }
  }
  return true;
}
void copy_csc(lblg::SparseCsc& M, int nrow, int ncol, const int* p, const int* i, const double* x) {
  M.nrow = nrow;
  M.ncol = ncol;
  M.p.assign(p, p + ncol + 1);
  M.i.assign(i, i + p[ncol]);
  M.x.assign(x, x + p[ncol]);
}
}  // namespace

extern "C" void lblg_qp_default_settings(lblg_qp_settings* s) {
  if (!s) return;
  s->diag_perturb = 1e-9;
  s->eps_abs = 1e-3;
  s->stop_tol = 1e-15;
  s->max_iter = 0;
  s->max_iter_nas = 4000;
  s->variant = LBLG_QP_FIXED;
  s->ordering = LBLG_ORDER_ND;
  s->precision = LBLG_F64;
  s->separate_rows = 1;
}

extern "C" int lblg_qp_create(lblg_qp** out, int n, const int* Hp, const int* Hi, const double* Hx, const double* q,
                              int me, const int* Ap, const int* Ai, const double* Ax, const double* a, int mi,
                              const int* Bp, const int* Bi, const double* Bx, const double* b) {
  if (!out || n <= 0 || me < 0 || mi < 0 || !Hx || !q) return LBLG_ERR_ARG;
  if (!valid_csc(n, n, Hp, Hi)) return LBLG_ERR_ARG;
  for (int j = 0; j < n; ++j)
    for (int k = Hp[j]; k < Hp[j + 1]; ++k)
      if (Hi[k] < j) return LBLG_ERR_ARG;  // lower triangle only
  if (me > 0 && (!valid_csc(me, n, Ap, Ai) || !Ax || !a)) return LBLG_ERR_ARG;
  if (mi > 0 && (!valid_csc(mi, n, Bp, Bi) || !Bx || !b)) return LBLG_ERR_ARG;
  return guarded([&] {
    QpSolver* s = new QpSolver();
    s->n = n;
    s->me = me;
    s->mi = mi;
    copy_csc(s->H, n, n, Hp, Hi, Hx);
    s->q.assign(q, q + n);
    if (me > 0) {
      copy_csc(s->A, me, n, Ap, Ai, Ax);
      s->a.assign(a, a + me);
    } else {
      s->A.nrow = 0;
      s->A.ncol = n;
      s->A.p.assign(n + 1, 0);
    }
    if (mi > 0) {
      copy_csc(s->B, mi, n, Bp, Bi, Bx);
      s->b.assign(b, b + mi);
    } else {
      s->B.nrow = 0;
      s->B.ncol = n;
      s->B.p.assign(n + 1, 0);
    }
    s->primal.assign(n, 0.0);
    s->dual.assign(mi, 0.0);
    s->dual_eq.assign(me, 0.0);
    *out = reinterpret_cast<lblg_qp*>(s);
    return LBLG_OK;
  });
}

extern "C" int lblg_qp_destroy(lblg_qp* qp) {
  delete reinterpret_cast<QpSolver*>(qp);
  return LBLG_OK;
}

extern "C" int lblg_qp_set_settings(lblg_qp* qp, const lblg_qp_settings* st) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || !st) return LBLG_ERR_ARG;
  if (st->variant < LBLG_QP_FIXED || st->variant > LBLG_QP_PREDET || st->max_iter < 0 || st->max_iter_nas < 0 ||
      !(st->diag_perturb >= 0) || !(st->eps_abs > 0) || !(st->stop_tol >= 0) ||
      (st->ordering != LBLG_ORDER_ND && st->ordering != LBLG_ORDER_NATURAL) ||
      (st->precision != LBLG_F64 && st->precision != LBLG_F32 && st->precision != LBLG_BF16) ||
      (st->separate_rows != 0 && st->separate_rows != 1))
    return LBLG_ERR_ARG;
  s->diag_perturb = st->diag_perturb;
  s->zero_thresh = st->diag_perturb;
  s->eps_abs = st->eps_abs;
  s->stop_tol = st->stop_tol;
  s->max_iter = st->max_iter;
  s->max_iter_nas = st->max_iter_nas;
  s->variant = st->variant;
  s->ordering = st->ordering;
  s->precision = st->precision;
  s->separate_rows = st->separate_rows;
  return LBLG_OK;
}

extern "C" int lblg_qp_set_linear(lblg_qp* qp, const double* q, const double* a, const double* b) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s) return LBLG_ERR_ARG;
  if (q) s->q.assign(q, q + s->n);
  if (a && s->me) s->a.assign(a, a + s->me);
  if (b && s->mi) s->b.assign(b, b + s->mi);
  return LBLG_OK;
}

extern "C" int lblg_qp_analyze(lblg_qp* qp) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s) return LBLG_ERR_ARG;
  return guarded([&] {
    s->auto_reg_en = s->variant == LBLG_QP_AUTO ? 2 : (s->variant == LBLG_QP_FIXED ? 1 : 0);
    s->symbolic();
    return LBLG_OK;
  });
}

extern "C" int lblg_qp_autotune(lblg_qp* qp, int reps, lblg_tuning* t) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || reps < 1) return LBLG_ERR_ARG;
  return guarded([&] {
    s->auto_reg_en = s->variant == LBLG_QP_AUTO ? 2 : (s->variant == LBLG_QP_FIXED ? 1 : 0);
    s->symbolic();  // the super-KKT and its analysis, kept for the solves that follow
    int rc = lblg_set_values(s->h, s->SM.x.data());
    if (rc == LBLG_OK) rc = lblg_autotune(s->h, 1, reps, t);
    if (rc != LBLG_OK) throw std::runtime_error("qp: autotune of the KKT schedule failed");
    return LBLG_OK;
  });
}

extern "C" int lblg_qp_solve(lblg_qp* qp, int* status) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s) return LBLG_ERR_ARG;
  return guarded([&] {
    const int st = s->solve();
    if (status) *status = st;
    return LBLG_OK;
  });
}

extern "C" int lblg_qp_solve_warm(lblg_qp* qp, const double* x0, const double* z0, const int* active, int n_active,
                                  const double* y0, int* converged) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || !x0 || (s->mi > 0 && !z0) || n_active < 0 || (n_active > 0 && !active)) return LBLG_ERR_ARG;
  return guarded([&] {
    std::vector<double> z(s->mi, 0.0);
    if (s->mi) z.assign(z0, z0 + s->mi);
    const int c = s->solve_warm(x0, z.data(), active, n_active, y0);
    if (converged) *converged = c;
    return LBLG_OK;
  });
}

extern "C" int lblg_qp_solve_predicted(lblg_qp* qp, const int* active, int n_active, int* status, int* prune_solves) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || n_active < 0 || (n_active > 0 && !active)) return LBLG_ERR_ARG;
  return guarded([&] {
    const int st = s->solve_predicted(active, n_active);
    if (status) *status = st;
    if (prune_solves) *prune_solves = s->predicted_prunes;
    return LBLG_OK;
  });
}

extern "C" int lblg_qp_get(const lblg_qp* qp, double* primal, double* dual_ineq, double* dual_eq, int* active) {
  const QpSolver* s = reinterpret_cast<const QpSolver*>(qp);
  if (!s) return LBLG_ERR_ARG;
  if (primal) std::memcpy(primal, s->primal.data(), sizeof(double) * s->n);
  if (dual_ineq && s->mi) std::memcpy(dual_ineq, s->dual.data(), sizeof(double) * s->mi);
  if (dual_eq && s->me) std::memcpy(dual_eq, s->dual_eq.data(), sizeof(double) * s->me);
  if (active && !s->active_set.empty())
    std::memcpy(active, s->active_set.data(), sizeof(int) * s->active_set.size());
  return LBLG_OK;
}

extern "C" int lblg_qp_get_info(const lblg_qp* qp, lblg_qp_info* info) {
  const QpSolver* s = reinterpret_cast<const QpSolver*>(qp);
  if (!s || !info) return LBLG_ERR_ARG;
  info->status = s->status;
  info->num_iter = s->num_iter;
  info->n_active = (int)s->active_set.size();
  info->inner_iter_ref = s->inner_iter_ref;
  info->outer_iter_ref = s->outer_iter_ref;
  info->diag_perturb = s->diag_perturb;
  info->stop_tol = s->stop_tol;
  info->primal_obj = s->primal_obj;
  info->dual_obj = s->dual_obj;
  info->objective = s->objective;
  info->lag_res = s->lag_res;
  info->cons_sat_norm = s->cons_sat_norm;
  info->non_negativity_infn = s->non_negativity_infn;
  info->complementarity_infn = s->complementarity_infn;
  info->factorizations = s->factorizations;
  info->kkt_solves = s->kkt_solves;
  info->factor_seconds = s->factor_seconds;
  info->solve_seconds = s->solve_seconds;
  info->refactored_supernodes = s->refactored_supernodes;
  info->supernodes = s->h ? lblg_symbolic_scalar(s->h, 1) : 0;
  return LBLG_OK;
}

extern "C" int64_t lblg_qp_super_kkt(const lblg_qp* qp, int* p, int* i, double* x) {
  const QpSolver* s = reinterpret_cast<const QpSolver*>(qp);
  if (!s || s->SM.p.empty()) return LBLG_ERR_STATE;
  const int64_t nnz = s->SM.p.back();
  if (p) std::memcpy(p, s->SM.p.data(), sizeof(int) * s->SM.p.size());
  if (i) std::memcpy(i, s->SM.i.data(), sizeof(int) * nnz);
  if (x) std::memcpy(x, s->SM.x.data(), sizeof(double) * nnz);
  return nnz;
}

extern "C" int lblg_kkt_analyze(lblg_qp* qp, const lblg_settings* st) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || !st) return LBLG_ERR_ARG;
  if (st->req_ref_iter < 0 || st->max_inner_iter < 0 || !(st->reg_diag >= 0)) return LBLG_ERR_ARG;
  return guarded([&] {
    s->analyze_lbl(*st);
    return LBLG_OK;
  });
}

extern "C" int64_t lblg_kkt_size(const lblg_qp* qp) {
  const QpSolver* s = reinterpret_cast<const QpSolver*>(qp);
  if (!s || !s->analysed) return LBLG_ERR_STATE;
  return s->SM.ncol;
}

// the engine handle of the super-KKT (symbolic introspection, factor copies,
// residual norms of the mode-1 drop-in)
extern "C" lblg_handle* lblg_kkt_handle(lblg_qp* qp) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  return s && s->analysed ? s->h : nullptr;
}

extern "C" int lblg_kkt_edit(lblg_qp* qp, int add, int count, const int* rows) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || count < 0 || (count > 0 && !rows)) return LBLG_ERR_ARG;
  if (!s->analysed) return LBLG_ERR_STATE;
  for (int k = 0; k < count; ++k)
    if (rows[k] < 0 || rows[k] >= s->mi) return LBLG_ERR_ARG;
  for (int k = 0; k < count; ++k) {
    const int c = rows[k];
    s->edit_kkt(add != 0, c);
    if (add && !s->used[c]) s->update_active_set(true, c);
    else if (!add && s->used[c]) s->update_active_set(false, c);
  }
  return LBLG_OK;
}

extern "C" int lblg_kkt_factor(lblg_qp* qp) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s) return LBLG_ERR_ARG;
  if (!s->analysed) return LBLG_ERR_STATE;
  return guarded([&] {
    s->factor_kkt();
    return LBLG_OK;
  });
}

extern "C" int lblg_kkt_solve(lblg_qp* qp, const double* rhs, double* x, double* bwd_err) {
  QpSolver* s = reinterpret_cast<QpSolver*>(qp);
  if (!s || !rhs || !x) return LBLG_ERR_ARG;
  if (!s->analysed) return LBLG_ERR_STATE;
  const int64_t N = s->SM.ncol;
  double bwd = 0.0;
  const int rc = lblg_solve(s->h, 1, rhs, N, x, N, &bwd);
  if (bwd_err) *bwd_err = bwd;
  if (rc == LBLG_OK) ++s->kkt_solves;
  return rc;
}
