// Residual norms and backward errors of caller-given solutions, on the device.
//
// The reference's SolverSettings::compute_norms / backward_error
// (src/QP/linear_solver_wrapper.cpp:1462-1498) form r = rhs - K0 x with the
// unperturbed matrix K0 (set_diags(0): perturbed diagonals are zero), then
// res_l1 = ||r||_1, x_l1 = ||x||_1, rhs_l1 = ||rhs||_1 (norm_dense with
// norm_type 0, src/common/Norm.cpp:15-54) and
//   bwd_err = res_l1 / (A_l1 x_l1 + rhs_l1),
// A_l1 the max absolute row sum of K0 (norm_sparse_int, Norm.cpp:115-185).
// Here the product runs over the full CSR of K0 in the caller's ordering
// (one row per thread, no atomics), each workgroup writes its partial sums,
// and one workgroup adds the partials in a fixed order: repeated calls agree
// bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <stdexcept>

#include "engine/device.hpp"
#include "engine/device_util.hpp"
#include "engine/handle.hpp"
#include "engine/kernels.hpp"
#include "lblg.h"

namespace lblg {

void set_last_error(const char* what);  // capi_numeric.cpp

constexpr int RES_BLOCKS = 1024;                      // workgroups of the partial-sum launch, at most
constexpr int RES_PART = RES_BLOCKS * 3 * MAX_RHS;    // partial sums: [block][|r|, |x|, |b|][column]
constexpr int RES_OUT = 4 * MAX_RHS;                  // final sums (3 per column) and errors (1 per column)

static int res_blocks(int n) {
  const int b = (n + 255) / 256;
  return b < 1 ? 1 : (b > RES_BLOCKS ? RES_BLOCKS : b);
}

// partial sums of |b - K0 x|, |x| and |b| per column over the rows of a
// grid-stride loop; B and X column-major with leading dimensions ldb / ldx
template <int NR>
__global__ __launch_bounds__(256) void k_resid_part(int n, const int* rp, const int* ci, const double* Kv,
                                                    const double* B, int64_t ldb, const double* X, int64_t ldx,
                                                    double* part) {
  __shared__ double sh[4][3 * NR];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double acc[3 * NR];
#pragma unroll
  for (int c = 0; c < 3 * NR; ++c) acc[c] = 0.0;
  for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256) {
    double kx[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) kx[c] = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) {
      const double a = Kv[q];
      const int j = ci[q];
#pragma unroll
      for (int c = 0; c < NR; ++c) kx[c] += a * X[(int64_t)c * ldx + j];
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const double b = B[(int64_t)c * ldb + i];
      acc[c] += fabs(b - kx[c]);
      acc[NR + c] += fabs(X[(int64_t)c * ldx + i]);
      acc[2 * NR + c] += fabs(b);
    }
  }
#pragma unroll
  for (int c = 0; c < 3 * NR; ++c) {
    const double v = wave_sum(acc[c]);
    if (ln == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (t < 3 * NR) part[(int64_t)blockIdx.x * 3 * NR + t] = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
}

// one workgroup: the partials of nblk blocks added in block order; sums
// (3 nrhs: res_l1, x_l1, rhs_l1 per column, may be null) and the backward
// errors (nrhs, may be null) with A_l1 = *anorm
__global__ __launch_bounds__(64) void k_resid_final(const double* part, int nblk, int nrhs, const double* anorm,
                                                    double* sums, double* err) {
  __shared__ double s[3 * MAX_RHS];
  const int t = threadIdx.x;
  if (t < 3 * nrhs) {
    double v = 0.0;
    for (int b = 0; b < nblk; ++b) v += part[(int64_t)b * 3 * nrhs + t];
    s[t] = v;
    if (sums) sums[t] = v;
  }
  __syncthreads();
  if (err && t < nrhs) {
    const double den = anorm[0] * s[nrhs + t] + s[2 * nrhs + t];
    err[t] = den > 0.0 ? s[t] / den : 0.0;
  }
}

// the two launches on stream st; part holds RES_PART doubles
static void launch_residual(const DeviceFactor& D, int nrhs, const double* B, int64_t ldb, const double* X,
                            int64_t ldx, double* part, double* sums, double* err, hipStream_t st) {
  const int nb = res_blocks(D.n);
  switch (nrhs) {
#define RES_CASE(k)                                                                                \
  case k:                                                                                          \
    k_resid_part<k><<<nb, 256, 0, st>>>(D.n, D.Krp, D.Kci, D.Kv, B, ldb, X, ldx, part);            \
    break;
    RES_CASE(1) RES_CASE(2) RES_CASE(3) RES_CASE(4) RES_CASE(5) RES_CASE(6) RES_CASE(7) RES_CASE(8)
#undef RES_CASE
    default:
      throw std::invalid_argument("residual: 1 <= nrhs <= 8");
  }
  HIP_LAUNCH_CHECK("resid_part");
  k_resid_final<<<1, 64, 0, st>>>(part, nb, nrhs, D.k0norm, sums, err);
  HIP_LAUNCH_CHECK("resid_final");
}

// a factored single-system handle and arguments in range, else a status
static int residual_args(const Handle* h, int nrhs, const void* B, int64_t ldb, const void* X, int64_t ldx,
                         const void* out) {
  if (!h || !B || !X || !out || nrhs < 1 || nrhs > MAX_RHS || ldb < h->n || ldx < h->n) return LBLG_ERR_ARG;
  if (h->batch_off.size() > 2) {
    set_last_error("residual: a batch handle (use lblg_batch_backward_errors)");
    return LBLG_ERR_ARG;
  }
  if (!h->factored || !h->dev) return LBLG_ERR_STATE;
  return LBLG_OK;
}

static double* owned_alloc(DeviceFactor& D, int64_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * (size_t)count) != hipSuccess) throw std::bad_alloc();
  D.owned.push_back(p);
  return static_cast<double*>(p);
}

template <typename F>
static int guard(F f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    set_last_error("allocation failed");
    return LBLG_ERR_NOMEM;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return LBLG_ERR_ARG;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return LBLG_ERR_DEVICE;
  }
}

}  // namespace lblg

using lblg::DeviceFactor;
using lblg::Handle;

extern "C" int lblg_backward_error_device(lblg_handle* hh, int nrhs, const double* dB, int64_t ldb,
                                          const double* dX, int64_t ldx, double* d_err) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  const int rc = lblg::residual_args(h, nrhs, dB, ldb, dX, ldx, d_err);
  if (rc != LBLG_OK) return rc;
  return lblg::guard([&] {
    DeviceFactor& D = *h->dev;
    if (!D.res_part) D.res_part = lblg::owned_alloc(D, lblg::RES_PART);
    lblg::launch_residual(D, nrhs, dB, ldb, dX, ldx, D.res_part, nullptr, d_err, D.stream);
    return LBLG_OK;
  });
}

extern "C" int lblg_residual_norms(lblg_handle* hh, int nrhs, const double* B, int64_t ldb, const double* X,
                                   int64_t ldx, double* norms) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  const int rc = lblg::residual_args(h, nrhs, B, ldb, X, ldx, norms);
  if (rc != LBLG_OK) return rc;
  return lblg::guard([&] {
    DeviceFactor& D = *h->dev;
    const int n = h->n;
    hipStream_t st = D.stream;
    if (!D.res_part) D.res_part = lblg::owned_alloc(D, lblg::RES_PART);
    if (!D.res_out) D.res_out = lblg::owned_alloc(D, lblg::RES_OUT);
    if (!D.res_bx) D.res_bx = lblg::owned_alloc(D, 2 * (int64_t)n * lblg::MAX_RHS);
    double* dB = D.res_bx;
    double* dX = D.res_bx + (int64_t)n * lblg::MAX_RHS;
    for (int c = 0; c < nrhs; ++c) {
      if (hipMemcpyAsync(dB + (int64_t)c * n, B + (int64_t)c * ldb, sizeof(double) * n, hipMemcpyHostToDevice,
                         st) != hipSuccess ||
          hipMemcpyAsync(dX + (int64_t)c * n, X + (int64_t)c * ldx, sizeof(double) * n, hipMemcpyHostToDevice,
                         st) != hipSuccess)
        throw std::runtime_error("residual: host to device copy failed");
    }
    lblg::launch_residual(D, nrhs, dB, n, dX, n, D.res_part, D.res_out, D.res_out + 3 * nrhs, st);
    double out[lblg::RES_OUT];
    if (hipMemcpyAsync(out, D.res_out, sizeof(double) * 4 * nrhs, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      throw std::runtime_error("residual: device to host copy failed");
    for (int c = 0; c < nrhs; ++c) {
      norms[4 * c + 0] = out[c];
      norms[4 * c + 1] = out[nrhs + c];
      norms[4 * c + 2] = out[2 * nrhs + c];
      norms[4 * c + 3] = out[3 * nrhs + c];
    }
    return LBLG_OK;
  });
}

// ---------------------------------------------------------------------------
// The same backward errors with no allocation at all: the partial sums go to
// a file-scope device array (one copy per GPU, held until the process ends)
// instead of a buffer the handle allocates.  hipGetSymbolAddress gives that
// array's address on a device; the host never passes the array's own name to
// a kernel (on the host it is a stand-in, not a device address).
// ---------------------------------------------------------------------------
#include <mutex>

namespace lblg {
static __device__ __align__(256) double scratch[RES_PART];  // 192 KiB per GPU
constexpr int SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[SCRATCH_DEVICES] = {nullptr};

// the scratch array's address on device dev, looked up once per device with
// that device current
static double* scratch_on(int dev) {
  if (dev < 0 || dev >= SCRATCH_DEVICES) throw std::runtime_error("no-alloc residual: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) throw std::runtime_error("no-alloc residual: hipGetDevice failed");
    if (cur != dev && hipSetDevice(dev) != hipSuccess) throw std::runtime_error("no-alloc residual: hipSetDevice failed");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("no-alloc residual: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}
}  // namespace lblg

extern "C" int lblg_backward_error_device_noalloc(lblg_handle* hh, int nrhs, const double* dB, int64_t ldb,
                                                  const double* dX, int64_t ldx, double* d_err) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (h && (nrhs < 1 || nrhs > lblg::MAX_RHS)) {
    lblg::set_last_error("no-alloc residual: 1 <= nrhs <= 8 (the scratch array holds 1024 x 3 x 8 partial sums)");
    return LBLG_ERR_ARG;
  }
  const int rc = lblg::residual_args(h, nrhs, dB, ldb, dX, ldx, d_err);
  if (rc != LBLG_OK) return rc;
  return lblg::guard([&] {
    DeviceFactor& D = *h->dev;
    int dev = -1;
    if (hipStreamGetDevice(D.stream, &dev) != hipSuccess) throw std::runtime_error("no-alloc residual: stream device");
    double* part = lblg::scratch_on(dev);
    lblg::launch_residual(D, nrhs, dB, ldb, dX, ldx, part, nullptr, d_err, D.stream);
    return LBLG_OK;
  });
}
