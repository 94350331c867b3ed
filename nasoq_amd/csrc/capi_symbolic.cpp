// C-ABI entry points of the host-side phase: settings, handle lifetime and
// symbolic analysis.  See include/lblg.h for the reference interface each
// function replaces.
#include <cstring>
#include <new>
#include <stdexcept>

#include "engine/device.hpp"
#include "engine/handle.hpp"
#include "lblg.h"

using lblg::Handle;

extern "C" void lblg_default_settings(lblg_settings* s) {
  if (!s) return;
  std::memset(s, 0, sizeof(*s));
  s->ldl_variant = 4;
  s->req_ref_iter = 2;
  s->max_inner_iter = 2;
  s->tol_abs = 1e-15;
  s->tol_rel = 1e-15;
  s->reg_diag = 1e-8;
  s->regularization = 1;
  s->cost_param = 1;
  s->level_param = -3;
  s->final_seq_node = 50;
  s->n_relax[0] = 4;
  s->n_relax[1] = 16;
  s->n_relax[2] = 48;
  s->z_relax[0] = 0.8;
  s->z_relax[1] = 0.1;
  s->z_relax[2] = 0.05;
  s->ordering = LBLG_ORDER_ND;
  s->precision = LBLG_F64;
  s->target_bwd_err = 0.0;
  s->max_refine_rounds = 20;
}

static bool valid_lower_csc(int n, const int* Ap, const int* Ai) {
  if (n <= 0 || !Ap || !Ai || Ap[0] != 0) return false;
  for (int j = 0; j < n; ++j) {
    if (Ap[j + 1] <= Ap[j]) return false;        // every column holds its diagonal
    if (Ai[Ap[j]] != j) return false;            // diagonal first
    for (int q = Ap[j] + 1; q < Ap[j + 1]; ++q)
      if (Ai[q] <= Ai[q - 1] || Ai[q] >= n) return false;  // sorted, strictly lower
  }
  return true;
}

extern "C" int lblg_create(lblg_handle** out, int n, const int* Ap, const int* Ai, const double* Ax) {
  if (!out) return LBLG_ERR_ARG;
  *out = nullptr;
  if (!valid_lower_csc(n, Ap, Ai) || !Ax) return LBLG_ERR_ARG;
  try {
    Handle* h = new Handle;
    h->n = n;
    h->Ap.assign(Ap, Ap + n + 1);
    h->Ai.assign(Ai, Ai + Ap[n]);
    h->Ax.assign(Ax, Ax + Ap[n]);
    lblg_default_settings(&h->set);
    *out = reinterpret_cast<lblg_handle*>(h);
    return LBLG_OK;
  } catch (const std::bad_alloc&) {
    return LBLG_ERR_NOMEM;
  }
}

extern "C" int lblg_destroy(lblg_handle* hh) {
  delete reinterpret_cast<Handle*>(hh);
  return LBLG_OK;
}

extern "C" int lblg_order_nd(int n, const int* Ap, const int* Ai, int* perm) {
  if (!valid_lower_csc(n, Ap, Ai) || !perm) return LBLG_ERR_ARG;
  try {
    std::vector<int> p;
    lblg::nested_dissection(lblg::CscView{n, Ap, Ai}, p);
    std::memcpy(perm, p.data(), sizeof(int) * n);
    return LBLG_OK;
  } catch (const std::bad_alloc&) {
    return LBLG_ERR_NOMEM;
  }
}

extern "C" int lblg_analyze(lblg_handle* hh, const lblg_settings* s, const int* perm, const int* extra) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !s) return LBLG_ERR_ARG;
  if (s->ordering == LBLG_ORDER_GIVEN && !perm) return LBLG_ERR_ARG;
  if (s->cost_param < 1 || s->final_seq_node < 1) return LBLG_ERR_ARG;
  try {
    h->set = *s;
    h->dev.reset();  // a new analysis invalidates the device factor
    h->analyzed = false;
    std::vector<int> order(h->n);
    const lblg::CscView A{h->n, h->Ap.data(), h->Ai.data()};
    if (s->ordering == LBLG_ORDER_GIVEN) {
      order.assign(perm, perm + h->n);
    } else if (s->ordering == LBLG_ORDER_NATURAL) {
      for (int k = 0; k < h->n; ++k) order[k] = k;
    } else {
      lblg::nested_dissection(A, order);
    }
    lblg::SymbolicOptions opt;
    for (int k = 0; k < 3; ++k) {
      opt.nrelax[k] = s->n_relax[k];
      opt.zrelax[k] = s->z_relax[k];
    }
    opt.cost_param = s->cost_param;
    opt.level_param = s->level_param;
    opt.final_seq_node = s->final_seq_node;
    h->analyzed = lblg::analyze(A, order.data(), extra, opt, h->sym);
    h->factored = false;
    return h->analyzed ? LBLG_OK : LBLG_ERR_ARG;
  } catch (const std::invalid_argument&) {
    return LBLG_ERR_ARG;
  } catch (const std::bad_alloc&) {
    return LBLG_ERR_NOMEM;
  }
}

extern "C" int64_t lblg_symbolic_scalar(const lblg_handle* hh, int which) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h || !h->analyzed) return LBLG_ERR_STATE;
  const lblg::Symbolic& S = h->sym;
  switch (which) {
    case 0: return S.n;
    case 1: return S.nsuper;
    case 2: return S.lbc.n_level;
    case 3: return (int64_t)S.lbc.par_ptr.size() - 1;
    case 4: return S.max_sup_wid;
    case 5: return S.max_col;
    case 6: return S.xsize;
    case 7: return S.ssize;
  }
  return LBLG_ERR_ARG;
}

extern "C" int64_t lblg_symbolic_ints(const lblg_handle* hh, int which, int* dst) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h || !h->analyzed) return LBLG_ERR_STATE;
  const lblg::Symbolic& S = h->sym;
  const std::vector<int>* v = nullptr;
  std::vector<int> tmp;
  switch (which) {
    case 0: v = &S.perm; break;
    case 1: v = &S.iperm; break;
    case 2: v = &S.parent; break;
    case 3: v = &S.colcount; break;
    case 4: v = &S.super; break;
    case 5: v = &S.col2sup; break;
    case 6: v = &S.sparent; break;
    case 7: tmp.assign(S.s.begin(), S.s.begin() + S.pi[S.nsuper]); v = &tmp; break;
    case 8: v = &S.lbc.level_ptr; break;
    case 9: v = &S.lbc.par_ptr; break;
    case 10: tmp.assign(S.lbc.par_set.begin(), S.lbc.par_set.begin() + S.lbc.par_ptr.back()); v = &tmp; break;
    case 11: v = &S.extra_cols; break;
    default: return LBLG_ERR_ARG;
  }
  if (dst) std::memcpy(dst, v->data(), v->size() * sizeof(int));
  return (int64_t)v->size();
}

extern "C" int64_t lblg_symbolic_longs(const lblg_handle* hh, int which, int64_t* dst) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h || !h->analyzed) return LBLG_ERR_STATE;
  const lblg::Symbolic& S = h->sym;
  const std::vector<int64_t>* v = which == 0 ? &S.pi : which == 1 ? &S.p : which == 2 ? &S.i_ptr : nullptr;
  if (!v) return LBLG_ERR_ARG;
  if (dst) std::memcpy(dst, v->data(), v->size() * sizeof(int64_t));
  return (int64_t)v->size();
}

extern "C" int lblg_symbolic_flops(const lblg_handle* hh, double* fl, double* lnz) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h || !h->analyzed) return LBLG_ERR_STATE;
  if (fl) *fl = h->sym.fl;
  if (lnz) *lnz = h->sym.lnz;
  return LBLG_OK;
}

extern "C" int lblg_set_settings(lblg_handle* hh, const lblg_settings* s) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !s) return LBLG_ERR_ARG;
  if (!h->analyzed) return LBLG_ERR_STATE;
  // the structural fields must not change after the analysis
  const lblg_settings& o = h->set;
  if (s->cost_param != o.cost_param || s->level_param != o.level_param ||
      s->final_seq_node != o.final_seq_node || s->ordering != o.ordering)
    return LBLG_ERR_ARG;
  for (int k = 0; k < 3; ++k)
    if (s->n_relax[k] != o.n_relax[k] || s->z_relax[k] != o.z_relax[k]) return LBLG_ERR_ARG;
  if (s->max_inner_iter < 0 || s->req_ref_iter < 0 || s->max_inner_iter > 64) return LBLG_ERR_ARG;
  const bool refactor = s->reg_diag != o.reg_diag || s->regularization != o.regularization ||
                        s->precision != o.precision;
  h->set = *s;
  if (refactor) h->factored = false;
  return LBLG_OK;
}
