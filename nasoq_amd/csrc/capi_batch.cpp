// Batched systems: many independent symmetric matrices of uneven size held
// by one handle as a block-diagonal matrix.  Each block is ordered by its own
// nested dissection, so the elimination forest is the union of the blocks'
// trees and the symbolic result of every block is the one it would have
// alone; the device factor then processes the supernodes of all blocks of
// one height in the same launches (one launch per level and kernel family
// for the whole batch), and the solve / refinement run on all blocks at once.
// Engine extension: the reference factors one KKT per SolverSettings.
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "engine/handle.hpp"
#include "lblg.h"

using lblg::Handle;

namespace {
bool valid_block(int n, const int* Ap, const int* Ai) {
  if (n <= 0 || !Ap || !Ai || Ap[0] != 0) return false;
  for (int j = 0; j < n; ++j) {
    if (Ap[j + 1] <= Ap[j] || Ai[Ap[j]] != j) return false;
    for (int q = Ap[j] + 1; q < Ap[j + 1]; ++q)
      if (Ai[q] <= Ai[q - 1] || Ai[q] >= n) return false;
  }
  return true;
}
}  // namespace

extern "C" int lblg_batch_create(lblg_handle** out, int count, const int* n, const int* const* Ap,
                                 const int* const* Ai, const double* const* Ax) {
  if (!out) return LBLG_ERR_ARG;
  *out = nullptr;
  if (count < 1 || !n || !Ap || !Ai || !Ax) return LBLG_ERR_ARG;
  int64_t N = 0, nnz = 0;
  for (int k = 0; k < count; ++k) {
    if (!Ax[k] || !valid_block(n[k], Ap[k], Ai[k])) return LBLG_ERR_ARG;
    N += n[k];
    nnz += Ap[k][n[k]];
  }
  if (N > 0x7fffffff || nnz > 0x7fffffff) return LBLG_ERR_ARG;  // the CSC index type is int
  try {
    std::vector<int> P(N + 1), I(nnz);
    std::vector<double> X(nnz);
    std::vector<int64_t> off(count + 1, 0);
    int64_t c = 0, e = 0;
    for (int k = 0; k < count; ++k) {
      off[k] = c;
      for (int j = 0; j < n[k]; ++j) {
        P[c + j] = (int)(e + Ap[k][j]);
      }
      for (int q = 0; q < Ap[k][n[k]]; ++q) {
        I[e + q] = (int)(c + Ai[k][q]);
        X[e + q] = Ax[k][q];
      }
      c += n[k];
      e += Ap[k][n[k]];
    }
    P[N] = (int)e;
    off[count] = N;
    const int rc = lblg_create(out, (int)N, P.data(), I.data(), X.data());
    if (rc != LBLG_OK) return rc;
    reinterpret_cast<Handle*>(*out)->batch_off = std::move(off);
    return LBLG_OK;
  } catch (const std::bad_alloc&) {
    return LBLG_ERR_NOMEM;
  }
}

extern "C" int lblg_batch_analyze(lblg_handle* hh, const lblg_settings* s) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !s || h->batch_off.size() < 2) return LBLG_ERR_ARG;
  try {
    const int count = (int)h->batch_off.size() - 1;
    std::vector<int> perm(h->n);
    std::vector<int> bp, bi, order;
    for (int k = 0; k < count; ++k) {
      const int c0 = (int)h->batch_off[k], nb = (int)(h->batch_off[k + 1] - c0);
      bp.assign(nb + 1, 0);
      const int e0 = h->Ap[c0];
      for (int j = 0; j <= nb; ++j) bp[j] = h->Ap[c0 + j] - e0;
      bi.resize(bp[nb]);
      for (int q = 0; q < bp[nb]; ++q) bi[q] = h->Ai[e0 + q] - c0;
      if (s->ordering == LBLG_ORDER_NATURAL) {
        order.resize(nb);
        for (int i = 0; i < nb; ++i) order[i] = i;
      } else {
        lblg::nested_dissection(lblg::CscView{nb, bp.data(), bi.data()}, order);
      }
      for (int i = 0; i < nb; ++i) perm[c0 + i] = c0 + order[i];
    }
    lblg_settings st = *s;
    st.ordering = LBLG_ORDER_GIVEN;
    return lblg_analyze(hh, &st, perm.data(), nullptr);
  } catch (const std::bad_alloc&) {
    return LBLG_ERR_NOMEM;
  }
}

extern "C" int lblg_batch_info(const lblg_handle* hh, int* count, int64_t* offsets) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h || !count || h->batch_off.size() < 2) return LBLG_ERR_ARG;
  *count = (int)h->batch_off.size() - 1;
  if (offsets) std::memcpy(offsets, h->batch_off.data(), sizeof(int64_t) * h->batch_off.size());
  return LBLG_OK;
}

// backward error of every system and column with the reference's definition
// (linear_solver_wrapper.cpp:1462-1498): ||b - K0 x||_1 / (||K0||_inf ||x||_1 +
// ||b||_1) on the block, K0 = K with the perturbed diagonals set to zero
extern "C" int lblg_batch_backward_errors(const lblg_handle* hh, int nrhs, const double* B, int64_t ldb,
                                          const double* X, int64_t ldx, double* err) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h || !B || !X || !err || nrhs < 1 || ldb < h->n || ldx < h->n || h->batch_off.size() < 2)
    return LBLG_ERR_ARG;
  try {
    const int count = (int)h->batch_off.size() - 1;
    const double reg = h->set.regularization == 1 ? h->set.reg_diag : -1.0;
    auto val = [&](int q, int j) { return (h->Ai[q] == j && std::fabs(h->Ax[q]) < reg) ? 0.0 : h->Ax[q]; };
    std::vector<double> rowabs(h->n, 0.0), r(h->n);
    for (int j = 0; j < h->n; ++j)
      for (int q = h->Ap[j]; q < h->Ap[j + 1]; ++q) {
        const double a = std::fabs(val(q, j));
        rowabs[j] += a;
        if (h->Ai[q] != j) rowabs[h->Ai[q]] += a;
      }
    for (int c = 0; c < nrhs; ++c) {
      const double* b = B + (int64_t)c * ldb;
      const double* x = X + (int64_t)c * ldx;
      for (int i = 0; i < h->n; ++i) r[i] = b[i];
      for (int j = 0; j < h->n; ++j)
        for (int q = h->Ap[j]; q < h->Ap[j + 1]; ++q) {
          const int i = h->Ai[q];
          const double a = val(q, j);
          r[i] -= a * x[j];
          if (i != j) r[j] -= a * x[i];
        }
      for (int k = 0; k < count; ++k) {
        double rn = 0, xn = 0, bn = 0, kn = 0;
        for (int64_t i = h->batch_off[k]; i < h->batch_off[k + 1]; ++i) {
          rn += std::fabs(r[i]);
          xn += std::fabs(x[i]);
          bn += std::fabs(b[i]);
          kn = std::max(kn, rowabs[i]);
        }
        const double den = kn * xn + bn;
        err[(int64_t)c * count + k] = den > 0 ? rn / den : 0.0;
      }
    }
    return LBLG_OK;
  } catch (const std::bad_alloc&) {
    return LBLG_ERR_NOMEM;
  }
}
