// C-ABI entry points of the numeric phase (device factorization, solves,
// refinement, inertia and residuals).  See include/lblg.h for the reference
// interface each function replaces.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "engine/device.hpp"
#include "engine/engine.hpp"
#include "engine/handle.hpp"
#include "engine/kernels.hpp"
#include "lblg.h"

using lblg::Handle;

namespace {
thread_local char g_last_error[512] = {0};
void set_error(const char* what) {
  std::strncpy(g_last_error, what, sizeof(g_last_error) - 1);
  g_last_error[sizeof(g_last_error) - 1] = 0;
}
}  // namespace

namespace lblg {
void set_last_error(const char* what) { set_error(what); }
}  // namespace lblg

namespace {
template <typename F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    set_error("allocation failed");
    return LBLG_ERR_NOMEM;
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return LBLG_ERR_ARG;
  } catch (const std::exception& e) {
    set_error(e.what());
    return LBLG_ERR_DEVICE;
  }
}
}  // namespace

extern "C" const char* lblg_last_error(void) { return g_last_error; }

extern "C" int lblg_device_count(int* count) {
  if (!count) return LBLG_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return LBLG_OK;
}

extern "C" int lblg_factor(lblg_handle* hh) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (!h->analyzed) return LBLG_ERR_STATE;
  if (h->set.precision != LBLG_F64 && h->set.precision != LBLG_F32 && h->set.precision != LBLG_BF16)
    return LBLG_ERR_ARG;
  return guarded([&] {
    auto t0 = std::chrono::steady_clock::now();
    lblg::factor(*h);
    h->factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return LBLG_OK;
  });
}

extern "C" int lblg_refactor_resident(lblg_handle* hh) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (!h->analyzed || !h->dev) return LBLG_ERR_STATE;
  return guarded([&] {
    lblg::factor_device(*h);
    return LBLG_OK;
  });
}

extern "C" int lblg_set_values(lblg_handle* hh, const double* Ax) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !Ax) return LBLG_ERR_ARG;
  std::memcpy(h->Ax.data(), Ax, sizeof(double) * h->Ax.size());
  h->factored = false;
  h->values_dirty = true;
  return LBLG_OK;
}

extern "C" int lblg_update_values(lblg_handle* hh, int64_t count, const int64_t* idx, const double* vals) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || count < 0 || (count > 0 && (!idx || !vals))) return LBLG_ERR_ARG;
  const int64_t nnz = (int64_t)h->Ax.size();
  for (int64_t k = 0; k < count; ++k)
    if (idx[k] < 0 || idx[k] >= nnz) return LBLG_ERR_ARG;
  if (h->changed_flag.size() != h->Ax.size()) h->changed_flag.assign(h->Ax.size(), 0);
  for (int64_t k = 0; k < count; ++k) {
    const int64_t q = idx[k];
    h->Ax[q] = vals[k];
    if (!h->changed_flag[q]) {
      h->changed_flag[q] = 1;
      h->changed.push_back(q);
    }
  }
  if (count > 0) h->factored = false;
  return LBLG_OK;
}

extern "C" int64_t lblg_refactored_supernodes(const lblg_handle* hh) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  return h->refactored;
}

static int do_solve(lblg_handle* hh, int nrhs, const double* B, int64_t ldb, double* X, int64_t ldx, bool dev,
                    double* bwd) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !B || !X || nrhs < 1 || nrhs > lblg::MAX_RHS || ldb < h->n || ldx < h->n) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  if (h->set.max_inner_iter < 0 || h->set.req_ref_iter < 0) return LBLG_ERR_ARG;
  return guarded([&] {
    lblg::solve(*h, nrhs, B, ldb, X, ldx, dev, bwd);
    return LBLG_OK;
  });
}

extern "C" int lblg_solve(lblg_handle* h, int nrhs, const double* B, int64_t ldb, double* X, int64_t ldx,
                          double* bwd_err) {
  return do_solve(h, nrhs, B, ldb, X, ldx, false, bwd_err);
}

extern "C" int lblg_solve_device(lblg_handle* h, int nrhs, const double* dB, int64_t ldb, double* dX,
                                 int64_t ldx, double* bwd_err) {
  return do_solve(h, nrhs, dB, ldb, dX, ldx, true, bwd_err);
}

extern "C" int lblg_num_ref_iter(const lblg_handle* hh) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  return h ? h->num_ref_iter : LBLG_ERR_ARG;
}

extern "C" int lblg_get_tuning(lblg_handle* hh, lblg_tuning* t) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !t) return LBLG_ERR_ARG;
  lblg::resolve_tuning(*h);
  t->group_w = h->tune.group_w;
  t->chain_w = h->tune.chain_w;
  t->small_solve_nnz = h->tune.small_nnz;
  t->schur_big_m = h->tune.big_m;
  t->wide_sub = h->tune.wide_sub;
  return LBLG_OK;
}

extern "C" int lblg_set_tuning(lblg_handle* hh, const lblg_tuning* t) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !t || t->group_w < 64 || t->chain_w < 1 || t->small_solve_nnz < 0 || t->schur_big_m < 1 ||
      t->wide_sub < 1 || t->wide_sub > 16)
    return LBLG_ERR_ARG;
  h->tune.group_w = t->group_w;
  h->tune.chain_w = t->chain_w;
  h->tune.small_nnz = t->small_solve_nnz;
  h->tune.big_m = t->schur_big_m;
  h->tune.wide_sub = t->wide_sub;
  h->dev.reset();  // the device schedule is rebuilt at the next factorization
  h->factored = false;
  return LBLG_OK;
}

extern "C" int lblg_autotune(lblg_handle* hh, int nrhs, int reps, lblg_tuning* t) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || nrhs < 1 || nrhs > lblg::MAX_RHS || reps < 1) return LBLG_ERR_ARG;
  if (!h->analyzed) return LBLG_ERR_STATE;
  return guarded([&] {
    double fms = 0.0, sms = 0.0;
    int tried = 0;
    lblg::autotune(*h, nrhs, reps, fms, sms, tried);
    if (t) {
      lblg_get_tuning(hh, t);
      t->factor_ms = fms;
      t->solve_ms = sms;
      t->schedules_tried = tried;
    }
    return LBLG_OK;
  });
}

extern "C" int lblg_graph_stats(const lblg_handle* hh, int64_t* captures, int64_t* launches) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (captures) *captures = h->dev ? h->dev->graph_captures : 0;
  if (launches) *launches = h->dev ? h->dev->graph_launches : 0;
  return LBLG_OK;
}

extern "C" int lblg_synchronize(lblg_handle* hh) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !h->dev) return LBLG_ERR_STATE;
  return hipStreamSynchronize(h->dev->stream) == hipSuccess ? LBLG_OK : LBLG_ERR_DEVICE;
}

extern "C" void* lblg_stream(lblg_handle* hh) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  return (h && h->dev) ? (void*)h->dev->stream : nullptr;
}

extern "C" int lblg_get_factor(lblg_handle* hh, double* valL, double* d_val, int* rows, int* perm_final) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  return guarded([&] {
    lblg::DeviceFactor& D = *h->dev;
    hipStream_t st = D.stream;
    const lblg::Symbolic& S = h->sym;
    // values of a float factor are widened on the host
    auto fetch = [&](double* dst, const void* src, int64_t count) {
      if (D.prec == 0) {
        if (hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToHost, st) != hipSuccess)
          throw std::runtime_error("copy failed");
        return;
      }
      std::vector<float> tmp(count);
      if (hipMemcpyAsync(tmp.data(), src, sizeof(float) * count, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        throw std::runtime_error("copy failed");
      for (int64_t i = 0; i < count; ++i) dst[i] = tmp[i];
    };
    if (valL) fetch(valL, D.valL, S.p[h->n]);
    if (d_val) fetch(d_val, D.dval, 2 * (int64_t)h->n);
    if (rows) hipMemcpyAsync(rows, D.rows, sizeof(int) * S.pi[S.nsuper], hipMemcpyDeviceToHost, st);
    if (perm_final) hipMemcpyAsync(perm_final, D.perm_f, sizeof(int) * h->n, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("copy failed");
    return LBLG_OK;
  });
}

extern "C" int lblg_get_scaling(lblg_handle* hh, double* s) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !s) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  return guarded([&] {
    lblg::DeviceFactor& D = *h->dev;
    if (D.prec == 0) {
      for (int i = 0; i < h->n; ++i) s[i] = 1.0;
      return LBLG_OK;
    }
    if (hipMemcpyAsync(s, D.scale, sizeof(double) * h->n, hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
        hipStreamSynchronize(D.stream) != hipSuccess)
      throw std::runtime_error("copy failed");
    return LBLG_OK;
  });
}

extern "C" int lblg_inertia(lblg_handle* hh, int64_t* pos, int64_t* neg, int64_t* zero) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !pos || !neg || !zero) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  std::vector<double> d(2 * (size_t)h->n);
  int rc = lblg_get_factor(hh, nullptr, d.data(), nullptr, nullptr);
  if (rc) return rc;
  const int n = h->n;
  int64_t a = 0, b = 0, z = 0;
  for (int i = 0; i < n; ++i) {
    if (d[n + i] == 0.0) {
      if (d[i] > 0) ++a;
      else if (d[i] < 0) ++b;
      else ++z;
    } else {
      const double det = d[i] * d[i + 1] - d[n + i] * d[n + i];
      const double tr = d[i] + d[i + 1];
      if (det < 0) {
        ++a;
        ++b;
      } else if (det > 0) {
        if (tr > 0) a += 2; else b += 2;
      } else {
        ++z;
        if (tr > 0) ++a; else if (tr < 0) ++b; else ++z;
      }
      ++i;
    }
  }
  *pos = a;
  *neg = b;
  *zero = z;
  return LBLG_OK;
}

extern "C" int lblg_factor_info(const lblg_handle* hh, int* zero_pivot, double* k0_norm_inf,
                                double* factor_seconds) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  if (zero_pivot) *zero_pivot = h->zero_pivot;
  if (k0_norm_inf) *k0_norm_inf = h->k0_norm_inf;
  if (factor_seconds) *factor_seconds = h->factor_seconds;
  return LBLG_OK;
}

extern "C" int lblg_pivot_stats(const lblg_handle* hh, int64_t* second_column_steps, int64_t* interchanges) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  if (second_column_steps) *second_column_steps = h->second_col_steps;
  if (interchanges) *interchanges = h->interchanges;
  return LBLG_OK;
}

extern "C" int lblg_panel_stats(const lblg_handle* hh, int64_t* panels, int64_t* panels_pivoted) {
  const Handle* h = reinterpret_cast<const Handle*>(hh);
  if (!h) return LBLG_ERR_ARG;
  if (!h->factored) return LBLG_ERR_STATE;
  if (panels) *panels = h->wide_panels;
  if (panels_pivoted) *panels_pivoted = h->wide_panels_pivoted;
  return LBLG_OK;
}

extern "C" int lblg_selftest_mfma(const double* A, const double* B, double* C, int precision) {
  if (!A || !B || !C || (precision != LBLG_F64 && precision != LBLG_F32)) return LBLG_ERR_ARG;
  return guarded([&] {
    double *dA, *dB, *dC;
    if (hipMalloc(&dA, 64 * 8) != hipSuccess || hipMalloc(&dB, 64 * 8) != hipSuccess ||
        hipMalloc(&dC, 256 * 8) != hipSuccess)
      throw std::runtime_error("alloc");
    hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice);
    hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice);
    lblg::launch_mfma_selftest(dA, dB, dC, precision == LBLG_F32, nullptr);
    hipMemcpy(C, dC, 256 * 8, hipMemcpyDeviceToHost);
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    return LBLG_OK;
  });
}

extern "C" int lblg_set_kernel_timing(lblg_handle* hh, int on) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !h->dev) return LBLG_ERR_STATE;
  h->dev->resolve_timing();
  h->dev->timing = on != 0;
  h->dev->timing_only.clear();
  if (on) h->dev->ktime.clear();
  return LBLG_OK;
}

extern "C" int lblg_set_kernel_timing_only(lblg_handle* hh, const char* name) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !h->dev) return LBLG_ERR_STATE;
  if (!name || !*name) return LBLG_ERR_ARG;
  h->dev->resolve_timing();
  h->dev->timing_only.clear();
  for (const char* p = name; *p;) {  // comma-separated family names
    const char* e = p;
    while (*e && *e != ',') ++e;
    if (e > p) h->dev->timing_only.insert(std::string(p, e));
    p = *e ? e + 1 : e;
  }
  if (h->dev->timing_only.empty()) return LBLG_ERR_ARG;
  h->dev->timing = true;
  h->dev->ktime.clear();
  return LBLG_OK;
}

extern "C" int lblg_kernel_time(lblg_handle* hh, const char* name, double* total_ms, int64_t* launches) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h || !h->dev || !name) return LBLG_ERR_ARG;
  h->dev->resolve_timing();
  auto it = h->dev->ktime.find(name);
  if (total_ms) *total_ms = it == h->dev->ktime.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == h->dev->ktime.end() ? 0 : it->second.second;
  return LBLG_OK;
}
