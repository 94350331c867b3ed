// gfx950 kernels of the LBL engine: assembly, Bunch-Kaufman factorization of
// diagonal blocks (LDS-resident for narrow supernodes, panel-blocked with a
// workgroup-group barrier for wide ones), FP64 MFMA trailing and
// Schur-complement updates with atomic scatter into ancestor panels,
// off-diagonal TRSM, the blocked L / D / L^T sweeps, the symmetric SpMV and
// the GMRES vector kernels.
//
// Reference counterparts: ldl_left_sn_parallel_02
// (src/ldl/parallel_blocked_ldlt_02.cpp:17-366), LAPACK dsytf2/dlasyf as
// reached through LAPACKE_dsytrf at :155, reorder_after_sytrf,
// blocked_2by2_solver, blocked_2by2_mult (src/common/Sym_BLAS.cpp:128-196,
// 492-529, 573-599), H2LeveledBlockedLsolve / H2LeveledBlockedLTsolve
// (src/triangularSolve/Triangular_BCSC.cpp:378-446, 534-600),
// spmv_csc_sym_one_int (src/matrixVector/spmv_CSC.cpp:26-103) and the vector
// steps of pmgmres_ldlt_auto (src/gmres/mgmres.cpp:1015-1211).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "engine/kernels.hpp"
#include "engine/device_util.hpp"

namespace lblg {

static int grid_for(int64_t count, int block) {
  int64_t g = (count + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

// --------------------------------------------------------------------------
// small helpers
// --------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* p, int64_t count, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_int(int* p, int64_t count, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_iota(int* p, int count) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = i;
}
__global__ void k_copy(double* dst, const double* src, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
template <typename T>
void launch_fill(T* p, int64_t count, double v, hipStream_t st) {
  if (count <= 0) return;
  k_fill<T><<<grid_for(count, 256), 256, 0, st>>>(p, count, (T)v);
  HIP_LAUNCH_CHECK("fill");
}
void launch_fill_int(int* p, int64_t count, int v, hipStream_t st) {
  if (count <= 0) return;
  k_fill_int<<<grid_for(count, 256), 256, 0, st>>>(p, count, v);
  HIP_LAUNCH_CHECK("fill_int");
}
// GMRES state of nrhs columns (stride ss): zero, except word flag_off of each
// column, which is set to 1 (the column's "running" flag)
__global__ void k_gmres_reset(double* gm, int64_t ss, int nrhs, int64_t flag_off) {
  const int64_t count = ss * nrhs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    gm[i] = (i % ss == flag_off) ? 1.0 : 0.0;
}
void launch_gmres_reset(double* gm, int64_t ss, int nrhs, int64_t flag_off, hipStream_t st) {
  k_gmres_reset<<<grid_for(ss * nrhs, 256), 256, 0, st>>>(gm, ss, nrhs, flag_off);
  HIP_LAUNCH_CHECK("gmres_reset");
}
void launch_iota(int* p, int count, hipStream_t st) {
  if (count <= 0) return;
  k_iota<<<grid_for(count, 256), 256, 0, st>>>(p, count);
  HIP_LAUNCH_CHECK("iota");
}
void launch_copy(double* dst, const double* src, int64_t count, hipStream_t st) {
  if (count <= 0) return;
  k_copy<<<grid_for(count, 256), 256, 0, st>>>(dst, src, count);
  HIP_LAUNCH_CHECK("copy");
}
// value conversion between the FP64 refinement vectors and the factor's type
template <typename TO, typename TI>
__global__ void k_convert(TO* dst, const TI* src, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (TO)src[i];
}
template <typename TO, typename TI>
void launch_convert(TO* dst, const TI* src, int64_t count, hipStream_t st) {
  if (count <= 0) return;
  k_convert<TO, TI><<<grid_for(count, 256), 256, 0, st>>>(dst, src, count);
  HIP_LAUNCH_CHECK("convert");
}

// --------------------------------------------------------------------------
// assembly (parallel_blocked_ldlt_02.cpp:81-91) with the diagonal
// perturbation of find_perturbation (linear_solver_wrapper.cpp:352-363)
// --------------------------------------------------------------------------
template <typename T>
__global__ void k_assemble(const int64_t* amap, const double* Kx, const unsigned char* pert, int nnz,
                           double reg, T* valL, const int* asup, const unsigned char* tmask) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
    if (tmask && !tmask[asup[q]]) continue;
    const double v = (pert[q] & 2) ? reg : Kx[q];
    valL[amap[q]] = (T)v;
  }
}
template <typename T>
void launch_assemble(const int64_t* amap, const double* Kx, const unsigned char* pert, int nnz, double reg,
                     T* valL, const int* asup, const unsigned char* tmask, hipStream_t st) {
  if (nnz <= 0) return;
  k_assemble<T><<<grid_for(nnz, 256), 256, 0, st>>>(amap, Kx, pert, nnz, reg, valL, asup, tmask);
  HIP_LAUNCH_CHECK("assemble");
}

// one workgroup per listed supernode: zero its panel (r x w at px[s]) and its
// D entries, identity local pivots over its columns
template <typename T>
__global__ void k_reset_supernodes(const int* list, const int* super, const int64_t* pi, const int64_t* px,
                                   T* valL, T* dval, int* perm_piv, int n) {
  const int s = list[blockIdx.x];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t len = (pi[s + 1] - pi[s]) * (int64_t)w;
  T* P = valL + px[s];
  for (int64_t e = threadIdx.x; e < len; e += blockDim.x) P[e] = (T)0;
  for (int c = threadIdx.x; c < w; c += blockDim.x) {
    dval[c0 + c] = (T)0;
    dval[n + c0 + c] = (T)0;
    perm_piv[c0 + c] = c0 + c;
  }
}
template <typename T>
void launch_reset_supernodes(const int* list, int count, const SnArgs& a, T* valL, T* dval, int* perm_piv,
                             hipStream_t st) {
  if (count <= 0) return;
  k_reset_supernodes<T><<<count, 256, 0, st>>>(list, a.super, a.pi, a.px, valL, dval, perm_piv, a.n);
  HIP_LAUNCH_CHECK("reset_supernodes");
}

__global__ void k_scatter_values(int64_t count, const int64_t* idx, const double* vals, double* Kx) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
    Kx[idx[k]] = vals[k];
}
void launch_scatter_values(int64_t count, const int64_t* idx, const double* vals, double* Kx, hipStream_t st) {
  if (count <= 0) return;
  k_scatter_values<<<grid_for(count, 256), 256, 0, st>>>(count, idx, vals, Kx);
  HIP_LAUNCH_CHECK("scatter_values");
}

// --------------------------------------------------------------------------
// Bunch-Kaufman of a narrow diagonal block (w <= 64) held in LDS.
// One workgroup (256 threads) per supernode.  Pivot decisions follow LAPACK
// dsytf2 (uplo = 'L').  Each interchange is also applied to the earlier
// columns of L, giving the standard form that reorder_after_sytrf
// (Sym_BLAS.cpp:128-196) produces; the 2x2 sub-diagonals go to
// dval[n + c] and the diagonal of the block becomes 1.
// --------------------------------------------------------------------------
#define LDA_S 65
template <typename T>
__global__ __launch_bounds__(256) void k_bk_small(const int* list, const int* super, const int64_t* pi,
                                                  const int64_t* px, T* valL, T* dval,
                                                  int* perm_piv, int n, int* status) {
  // the block, a[i + j * LDA_S] for j < the list's widest supernode (the
  // launch sizes the LDS, so lists of narrow supernodes keep many
  // workgroups per CU)
  extern __shared__ double smem_d[];
  T* const a = reinterpret_cast<T*>(smem_d);
  __shared__ T wk[64], wk1[64];
  __shared__ int lperm[64], kind[64];
  __shared__ int sh_kp, sh_kstep, sh_zero;
  const int s = list[blockIdx.x];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  T* P = valL + px[s];
  const int t = threadIdx.x;
  const int NT = blockDim.x;  // 256, or 64 for lists of supernodes at most 16 wide
  for (int idx0 = 0; idx0 < w * w; idx0 += NT * 8) {  // a thread's loads of a batch issue before its stores
    T v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = idx0 + t + NT * q, i = idx % w, j = idx / w;
      v[q] = (idx < w * w && i >= j) ? P[i + j * r] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = idx0 + t + NT * q, i = idx % w, j = idx / w;
      if (idx < w * w) a[i + j * LDA_S] = v[q];
    }
  }
  if (t < 64) {
    lperm[t] = t;
    kind[t] = 1;
  }
  __syncthreads();
  int k = 0;
  while (k < w) {
    if (t < 64) {  // wave 0 takes the pivot decision
      const double absakk = fabs(a[k + k * LDA_S]);
      const int i = k + 1 + t;
      double v = (i < w) ? fabs(a[i + k * LDA_S]) : -1.0;
      if (!(v == v)) v = -1.0;
      int iv = (i < w) ? i : 0x7fffffff;
      wave_argmax(v, iv);
      const double colmax = (k < w - 1) ? v : 0.0;
      const int imax = (k < w - 1) ? iv : k;
      int kp = k, kstep = 1, zero = 0;
      if (fmax(absakk, colmax) == 0.0 || !(absakk == absakk)) {
        zero = 1;
      } else if (!(absakk >= BK_ALPHA * colmax)) {
        const int j = k + t;
        double rv = -1.0;
        if (j < imax) rv = fabs(a[imax + j * LDA_S]);
        else if (j > imax && j < w) rv = fabs(a[j + imax * LDA_S]);
        int dummy = 0;
        wave_argmax(rv, dummy);
        const double rowmax = rv;
        if (absakk >= BK_ALPHA * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (fabs(a[imax + imax * LDA_S]) >= BK_ALPHA * rowmax) {
          kp = imax;
        } else {
          kp = imax;
          kstep = 2;
        }
      }
      if (t == 0) {
        sh_kp = kp;
        sh_kstep = kstep;
        sh_zero = zero;
      }
    }
    __syncthreads();
    const int kp = sh_kp, kstep = sh_kstep, kk = k + kstep - 1;
    const int zero = sh_zero;
    if (zero && t == 0) atomicOr(status, 1);
    if (kp != kk) {
      for (int q = t; q < 64 * 3; q += NT) {
        const int part = q / 64, x = q % 64;
        if (part == 0) {
          const int i = kp + 1 + x;
          if (i < w) {
            const T tmp = a[i + kk * LDA_S];
            a[i + kk * LDA_S] = a[i + kp * LDA_S];
            a[i + kp * LDA_S] = tmp;
          }
        } else if (part == 1) {
          const int j = kk + 1 + x;
          if (j < kp) {
            const T tmp = a[j + kk * LDA_S];
            a[j + kk * LDA_S] = a[kp + j * LDA_S];
            a[kp + j * LDA_S] = tmp;
          }
        } else {
          const int j = x;
          if (j < k) {
            const T tmp = a[kk + j * LDA_S];
            a[kk + j * LDA_S] = a[kp + j * LDA_S];
            a[kp + j * LDA_S] = tmp;
          }
        }
      }
      if (t == 0) {
        T tmp = a[kk + kk * LDA_S];
        a[kk + kk * LDA_S] = a[kp + kp * LDA_S];
        a[kp + kp * LDA_S] = tmp;
        if (kstep == 2) {
          tmp = a[k + 1 + k * LDA_S];
          a[k + 1 + k * LDA_S] = a[kp + k * LDA_S];
          a[kp + k * LDA_S] = tmp;
        }
        const int tp = lperm[kk];
        lperm[kk] = lperm[kp];
        lperm[kp] = tp;
      }
      __syncthreads();
    }
    if (t == 0 && kstep == 2) {
      kind[k] = 2;
      kind[k + 1] = 0;
    }
    if (!zero) {
      if (kstep == 1) {
        if (k < w - 1) {
          const T akk = a[k + k * LDA_S];
          const bool big = fabs(akk) >= Mfma<T>::sfmin();
          const T d11 = (T)1 / akk;
          const int m = w - k - 1;
          if (!big) {
            for (int i = k + 1 + t; i < w; i += NT) a[i + k * LDA_S] /= akk;
            __syncthreads();
          }
          for (int idx = t; idx < m * m; idx += NT) {
            const int i = k + 1 + idx % m, j = k + 1 + idx / m;
            if (i >= j) {
              const T f = big ? -d11 * a[j + k * LDA_S] : -akk * a[j + k * LDA_S];
              a[i + j * LDA_S] += a[i + k * LDA_S] * f;
            }
          }
          __syncthreads();
          if (big)
            for (int i = k + 1 + t; i < w; i += NT) a[i + k * LDA_S] *= d11;
        }
      } else if (k < w - 2) {
        T d21 = a[k + 1 + k * LDA_S];
        const T d11 = a[k + 1 + (k + 1) * LDA_S] / d21;
        const T d22 = a[k + k * LDA_S] / d21;
        const T tt = (T)1 / (d11 * d22 - (T)1);
        d21 = tt / d21;
        for (int j = k + 2 + t; j < w; j += NT) {
          wk[j] = d21 * (d11 * a[j + k * LDA_S] - a[j + (k + 1) * LDA_S]);
          wk1[j] = d21 * (d22 * a[j + (k + 1) * LDA_S] - a[j + k * LDA_S]);
        }
        __syncthreads();
        const int m = w - k - 2;
        for (int idx = t; idx < m * m; idx += NT) {
          const int i = k + 2 + idx % m, j = k + 2 + idx / m;
          if (i >= j) a[i + j * LDA_S] -= a[i + k * LDA_S] * wk[j] + a[i + (k + 1) * LDA_S] * wk1[j];
        }
        __syncthreads();
        for (int j = k + 2 + t; j < w; j += NT) {
          a[j + k * LDA_S] = wk[j];
          a[j + (k + 1) * LDA_S] = wk1[j];
        }
      }
    }
    __syncthreads();
    k += kstep;
  }
  // D extraction and unit diagonal
  if (t < w) {
    const int l = t;
    dval[c0 + l] = a[l + l * LDA_S];
    dval[n + c0 + l] = 0.0;
    if (kind[l] == 2) dval[n + c0 + l] = a[l + 1 + l * LDA_S];
    perm_piv[c0 + l] = c0 + lperm[l];
  }
  __syncthreads();
  if (t < w) {
    const int l = t;
    if (kind[l] == 2) a[l + 1 + l * LDA_S] = 0.0;
    a[l + l * LDA_S] = 1.0;
  }
  __syncthreads();
  for (int idx = t; idx < w * w; idx += NT) {
    const int i = idx % w, j = idx / w;
    if (i >= j) P[i + j * r] = a[i + j * LDA_S];
  }
}
template <typename T>
void launch_bk_small(const int* list, int count, int wmax, const SnArgs& a, T* valL, T* dval,
                     int* perm_piv, int* status, hipStream_t st) {
  if (count <= 0) return;
  if (wmax < 1 || wmax > 64) throw std::runtime_error("bk_small: supernode width out of range");
  // lists of supernodes at most 16 wide: one-wave workgroups (the block's
  // work fits one wave; more of them per CU)
  k_bk_small<T><<<count, wmax <= 16 ? 64 : 256, sizeof(T) * LDA_S * wmax, st>>>(list, a.super, a.pi, a.px, valL,
                                                                                 dval, perm_piv, a.n, status);
  HIP_LAUNCH_CHECK("bk_small");
}

// --------------------------------------------------------------------------
// Panel Bunch-Kaufman for a wide diagonal block (w > SMALL_W), LAPACK dlasyf
// style: columns of the current panel are formed left-looking from the
// panel's L and W = L*D, pivots are chosen by the dsytf2 rules on the fully
// updated columns, and the trailing block is updated afterwards
// (k_trailing).  Interchanges are applied to all earlier columns of L and to
// the columns of the rows below the diagonal block (the column permutation
// of LAPACKE_dlapmt at parallel_blocked_ldlt_02.cpp:164-165), so no separate
// permutation pass is needed.  G workgroups cooperate on one supernode.
//
// Per-supernode workspace at W + woff[s]:  Wp (w x NB, ld w) | t0 (w) | t1 (w)
//                                         | part (2 G doubles) | parti (2 G ints)
//
// LDSP (one workgroup per supernode, diagonal block of at most ldsm rows):
// the panel's columns [k0, k0 + NB) of the diagonal block's rows [k0, w),
// its W columns and t0 / t1 live in LDS for the whole panel, so each column
// step reads and writes LDS only; the panel and W (rows the trailing update
// reads) are written back at the end.  Entries outside the staged block
// (earlier panels, later columns, the rows below the diagonal block) stay
// in memory.
// --------------------------------------------------------------------------
constexpr int PANEL_UF = 8;  // left-looking terms whose loads are issued together

template <typename T, bool LDSP>
__global__ __launch_bounds__(256) void k_bk_panel(const int* list, int G, const int* super,
                                                  const int64_t* pi, const int64_t* px, T* valL,
                                                  T* dval, int* perm_piv, T* Wbase,
                                                  const int64_t* woff, int* kpos, int* kstart,
                                                  int* sync, int n, int* status, int ldsm) {
  __shared__ double sv[8];
  __shared__ int si[8];
  extern __shared__ double panel_smem[];
  const int sidx = blockIdx.x / G, g = blockIdx.x % G;
  const int s = list[sidx];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  T* A = valL + px[s];
  T* Wp = Wbase + woff[s];
  T* part = Wp + (int64_t)w * PANEL_NB + 2 * (int64_t)w;  // column maxima (exact in T: magnitudes of T values)
  int* parti = reinterpret_cast<int*>(part + 2 * G);
  int* ctr = sync + sidx;
  const int gt = g * 256 + threadIdx.x, GT = G * 256;
  const int k0 = kpos[s];
  if (k0 >= w) return;  // whole group returns together (k0 is uniform)
  const bool last = (w - k0) <= PANEL_NB;
  const int pc1 = min(w, k0 + PANEL_NB);  // end of the panel's columns
  // LDS blocks (LDSP): As(i, j) = L(k0 + i, k0 + j), Ws likewise, ld ldsm
  T* As = reinterpret_cast<T*>(panel_smem);
  T* Ws = As + (int64_t)PANEL_NB * ldsm;
  T* t0 = LDSP ? Ws + (int64_t)PANEL_NB * ldsm - k0 : Wp + (int64_t)w * PANEL_NB;  // indexed by row i >= k0
  T* t1 = LDSP ? Ws + (int64_t)PANEL_NB * ldsm + ldsm - k0 : Wp + (int64_t)w * PANEL_NB + w;
  // element (i, j) of the diagonal block / rows below, wherever it lives
  auto AA = [&](int64_t i, int64_t j) -> T& {
    if (LDSP && j >= k0 && j < pc1 && i >= k0 && i < w) return As[(j - k0) * ldsm + (i - k0)];
    return A[j * r + i];
  };
  auto WW = [&](int64_t i, int64_t j) -> T& {
    if (LDSP) return Ws[(j - k0) * ldsm + (i - k0)];
    return Wp[(j - k0) * w + i];
  };
  if (LDSP) {
    for (int64_t e = threadIdx.x; e < (int64_t)(pc1 - k0) * (w - k0); e += 256) {
      const int64_t j = e / (w - k0), i = e % (w - k0);
      As[j * ldsm + i] = A[(k0 + j) * r + (k0 + i)];
    }
    __syncthreads();
  }
  // v - sum_{j in [k0, k)} L(i, j) W(q, j) for a row i >= k and a row q of W
  auto left_update = [&](T v, int i, int q, int k) -> T {
    if (LDSP) {
      const T* ar = As + (i - k0);
      const T* wr = Ws + (q - k0);
      const int nk = k - k0;
      int j = 0;
      for (; j + PANEL_UF <= nk; j += PANEL_UF) {
        T av[PANEL_UF], wv[PANEL_UF];
#pragma unroll
        for (int u = 0; u < PANEL_UF; ++u) {
          av[u] = ar[(int64_t)(j + u) * ldsm];
          wv[u] = wr[(int64_t)(j + u) * ldsm];
        }
#pragma unroll
        for (int u = 0; u < PANEL_UF; ++u) v -= av[u] * wv[u];
      }
      for (; j < nk; ++j) v -= ar[(int64_t)j * ldsm] * wr[(int64_t)j * ldsm];
    } else {
      int j = k0;
      for (; j + PANEL_UF <= k; j += PANEL_UF) {
        T av[PANEL_UF], wv[PANEL_UF];
#pragma unroll
        for (int u = 0; u < PANEL_UF; ++u) {
          av[u] = A[(int64_t)(j + u) * r + i];
          wv[u] = Wp[(int64_t)(j + u - k0) * w + q];
        }
#pragma unroll
        for (int u = 0; u < PANEL_UF; ++u) v -= av[u] * wv[u];
      }
      for (; j < k; ++j) v -= A[(int64_t)j * r + i] * Wp[(int64_t)(j - k0) * w + q];
    }
    return v;
  };
  int epoch = 0;
  int k = k0;
  while (k < w && (last || k - k0 < PANEL_NB - 1)) {
    // ---- phase A: updated column k for own rows, local argmax
    double bv = -1.0;
    int bi = 0x7fffffff;
    for (int i = k + gt; i < w; i += GT) {
      const T v = left_update(AA(i, k), i, k, k);
      t0[i] = v;
      if (i > k) {
        double av = fabs(v);
        if (!(av == av)) av = -1.0;
        better(bv, bi, av, i);
      }
    }
    block_argmax(bv, bi, sv, si);
    if (G > 1 && threadIdx.x == 0) {
      part[g] = (T)bv;
      parti[g] = bi;
    }
    group_barrier(ctr, G, epoch, status);
    double colmax = -1.0;
    int imax = 0x7fffffff;
    if (G > 1) {
      for (int q = 0; q < G; ++q) better(colmax, imax, part[q], parti[q]);
    } else {  // the block maximum in T, as the cross-group exchange stores it
      colmax = (double)(T)bv;
      imax = bi;
    }
    if (k == w - 1 || colmax < 0) {
      colmax = 0.0;
      imax = k;
    }
    const T akk_v = t0[k];
    const double absakk = fabs(akk_v);
    int kp = k, kstep = 1, zero = 0;
    bool use_t1_as_k = false;
    if (fmax(absakk, colmax) == 0.0 || !(absakk == absakk)) {
      zero = 1;
    } else if (!(absakk >= BK_ALPHA * colmax)) {
      // ---- phase B: updated column imax, row maximum
      double rv = -1.0;
      int ri = 0x7fffffff;
      for (int i = k + gt; i < w; i += GT) {
        const T v = left_update((i < imax) ? AA(imax, i) : AA(i, imax), i, imax, k);
        t1[i] = v;
        if (i != imax) {
          double av = fabs(v);
          if (!(av == av)) av = -1.0;
          better(rv, ri, av, i);
        }
      }
      block_argmax(rv, ri, sv, si);
      if (G > 1 && threadIdx.x == 0) {
        part[G + g] = (T)rv;
        parti[G + g] = ri;
      }
      group_barrier(ctr, G, epoch, status);
      double rowmax = -1.0;
      int rj = 0x7fffffff;
      if (G > 1) {
        for (int q = 0; q < G; ++q) better(rowmax, rj, part[G + q], parti[G + q]);
      } else {
        rowmax = (double)(T)rv;
        rj = ri;
      }
      if (rowmax < 0) rowmax = 0.0;
      if (absakk >= BK_ALPHA * colmax * (colmax / rowmax)) {
        kp = k;
      } else if (fabs(t1[imax]) >= BK_ALPHA * rowmax) {
        kp = imax;
        use_t1_as_k = true;
      } else {
        kp = imax;
        kstep = 2;
      }
    }
    const int kk = k + kstep - 1;
    const bool swap = kp != kk;
    // ---- phase C: interchange, W and L columns
    if (zero && gt == 0) atomicOr(status, 1);
    if (swap) {
      // rows kk/kp of the earlier L columns and of the panel's W columns
      for (int j = gt; j < k; j += GT) {
        const T x = AA(kk, j);
        AA(kk, j) = AA(kp, j);
        AA(kp, j) = x;
      }
      for (int j = k0 + gt; j < k; j += GT) {
        const T x = WW(kk, j);
        WW(kk, j) = WW(kp, j);
        WW(kp, j) = x;
      }
      // columns kk/kp of the rows below the diagonal block
      for (int64_t q = w + gt; q < r; q += GT) {
        const T x = AA(q, kk);
        AA(q, kk) = AA(q, kp);
        AA(q, kp) = x;
      }
      if (gt == 0) {
        const int tp = perm_piv[c0 + kk];
        perm_piv[c0 + kk] = perm_piv[c0 + kp];
        perm_piv[c0 + kp] = tp;
      }
    }
    // own rows: non-updated copy of column kk into column kp, then W and L
    const T* colk = use_t1_as_k ? t1 : t0;
    T D1 = 0, D2 = 0, OD = 0, d21 = 0, d11 = 0, d22 = 0, rk = 0;
    if (kstep == 1) {
      const int sk = swap ? (k == kk ? kp : k) : k;
      D1 = colk[sk];
      rk = (T)1 / D1;
    } else {
      // post-interchange values of W(k,k), W(k+1,k), W(k+1,k+1)
      D1 = t0[k];
      OD = t0[kp];       // row k+1 after the interchange holds row kp
      D2 = t1[kp];
      d21 = OD;
      d11 = D2 / d21;
      d22 = D1 / d21;
      const T tt = (T)1 / (d11 * d22 - (T)1);
      d21 = tt / d21;
    }
    for (int i = k + gt; i < w; i += GT) {
      const int si2 = !swap ? i : (i == kk ? kp : (i == kp ? kk : i));
      if (swap) {
        if (i == kk) AA(kp, kp) = AA(kk, kk);
        else if (i > kk && i < kp) AA(kp, i) = AA(i, kk);
        else if (i > kp) AA(i, kp) = AA(i, kk);
      }
      if (kstep == 1) {
        const T wv = colk[si2];
        WW(i, k) = wv;
        if (i == k) {
          AA(k, k) = 1.0;
          dval[c0 + k] = wv;
          dval[n + c0 + k] = 0.0;
        } else {
          AA(i, k) = zero ? wv : wv * rk;
        }
      } else {
        const T w0 = t0[si2], w1 = t1[si2];
        WW(i, k) = w0;
        WW(i, k + 1) = w1;
        if (i == k) {
          AA(k, k) = 1.0;
          dval[c0 + k] = D1;
          dval[n + c0 + k] = OD;
        } else if (i == k + 1) {
          AA(k + 1, k) = 0.0;
          AA(k + 1, k + 1) = 1.0;
          dval[c0 + k + 1] = D2;
          dval[n + c0 + k + 1] = 0.0;
        } else {
          AA(i, k) = d21 * (d11 * w0 - w1);
          AA(i, k + 1) = d21 * (d22 * w1 - w0);
        }
      }
    }
    group_barrier(ctr, G, epoch, status);
    k += kstep;
  }
  if (LDSP) {
    // the staged columns back to the factor, and the W rows the trailing
    // update reads (rows >= k of the eliminated columns)
    for (int64_t e = threadIdx.x; e < (int64_t)(pc1 - k0) * (w - k0); e += 256) {
      const int64_t j = e / (w - k0), i = e % (w - k0);
      A[(k0 + j) * r + (k0 + i)] = As[j * ldsm + i];
    }
    for (int64_t e = threadIdx.x; e < (int64_t)(k - k0) * (w - k); e += 256) {
      const int64_t j = e / (w - k), i = k + e % (w - k);
      Wp[j * w + i] = Ws[j * ldsm + (i - k0)];
    }
  }
  if (gt == 0) {
    kstart[s] = k0;
    kpos[s] = k;
  }
}

// LDS of the staged panel of a diagonal block with up to `rows` rows
template <typename T>
size_t panel_lds(int rows) {
  return sizeof(T) * (2 * (size_t)PANEL_NB + 2) * (size_t)rows;
}
template <typename T>
int panel_lds_rows() {
  return (int)(PANEL_LDS_BYTES / (sizeof(T) * (2 * (size_t)PANEL_NB + 2)));
}
template int panel_lds_rows<double>();
template int panel_lds_rows<float>();

template <typename T>
void launch_bk_panel(const int* list, int count, int G, const SnArgs& a, T* valL, T* dval,
                     int* perm_piv, T* W, const int64_t* woff, int* kpos, int* kstart, int* sync,
                     int* status, bool cooperative, int lds_rows, hipStream_t st) {
  if (count <= 0) return;
  if (lds_rows > 0) {
    if (G != 1 || cooperative || lds_rows > panel_lds_rows<T>())
      throw std::invalid_argument("bk_panel: LDS panel needs one workgroup per supernode and fitting rows");
    static bool attr = false;
    if (!attr) {
      if (hipFuncSetAttribute((const void*)k_bk_panel<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)PANEL_LDS_BYTES) != hipSuccess)
        throw std::runtime_error("bk_panel: LDS attribute");
      attr = true;
    }
    k_bk_panel<T, true><<<count, 256, panel_lds<T>(lds_rows), st>>>(list, 1, a.super, a.pi, a.px, valL, dval,
                                                                     perm_piv, W, woff, kpos, kstart, sync, a.n,
                                                                     status, lds_rows);
  } else if (!cooperative || G == 1) {
    k_bk_panel<T, false><<<count * G, 256, 0, st>>>(list, G, a.super, a.pi, a.px, valL, dval, perm_piv, W, woff,
                                                    kpos, kstart, sync, a.n, status, 0);
  } else {
    int n = a.n, z = 0;
    void* args[] = {(void*)&list, (void*)&G, (void*)&a.super, (void*)&a.pi, (void*)&a.px,
                    (void*)&valL, (void*)&dval, (void*)&perm_piv, (void*)&W, (void*)&woff,
                    (void*)&kpos, (void*)&kstart, (void*)&sync, (void*)&n, (void*)&status, (void*)&z};
    hipError_t e = hipLaunchCooperativeKernel((const void*)k_bk_panel<T, false>, dim3(count * G), dim3(256), args,
                                              0, st);
    if (e != hipSuccess) throw std::runtime_error(std::string("bk_panel coop: ") + hipGetErrorString(e));
  }
  HIP_LAUNCH_CHECK("bk_panel");
}

// --------------------------------------------------------------------------
// MFMA tile engine: a 256-thread workgroup computes a 64x64 block
// C(i,j) = sum_k A(i,k) B(j,k) with A and B staged in LDS as [k][64]
// (k-major).  Each wave owns a 32x32 quadrant = 2x2 blocks of
// Mfma<T>::mma (16x16x4; lane layout in device_util.hpp).  The tile engine
// passes B first so that the
// accumulator holds C^T: lanes 0..15 then cover 16 consecutive rows of C and
// the epilogues touch contiguous memory.
// --------------------------------------------------------------------------
#define KC 32          // k-chunk staged in LDS
#define LDT 66         // padded row length of a staged tile
template <typename T>
struct TileAcc {
  typename Mfma<T>::acc c[2][2];
};
template <typename T>
__device__ inline void tile_zero(TileAcc<T>& acc) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc.c[a][b] = typename Mfma<T>::acc{0, 0, 0, 0};
}
// As, Bs: [KC][LDT]; kc = number of valid k (multiple of 4 after zero padding)
template <typename T>
__device__ inline void tile_mma(TileAcc<T>& acc, const T* As, const T* Bs, int kc) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 32, q0 = (wv & 1) * 32;
  const int li = ln & 15, lk = ln >> 4;
  for (int k = 0; k < kc; k += 4) {
    const T a0 = As[(k + lk) * LDT + r0 + li];
    const T a1 = As[(k + lk) * LDT + r0 + 16 + li];
    const T b0 = Bs[(k + lk) * LDT + q0 + li];
    const T b1 = Bs[(k + lk) * LDT + q0 + 16 + li];
    // D = B_blk A_blk^T, i.e. the transpose of the C block: lane (ln & 15) walks rows
    acc.c[0][0] = Mfma<T>::mma(b0, a0, acc.c[0][0]);
    acc.c[0][1] = Mfma<T>::mma(b1, a0, acc.c[0][1]);
    acc.c[1][0] = Mfma<T>::mma(b0, a1, acc.c[1][0]);
    acc.c[1][1] = Mfma<T>::mma(b1, a1, acc.c[1][1]);
  }
}
// visit every accumulator element: f(row_in_tile, col_in_tile, value)
template <typename T, typename F>
__device__ inline void tile_foreach(const TileAcc<T>& acc, F f) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 32, q0 = (wv & 1) * 32;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 16 * a + (ln & 15);
        const int col = q0 + 16 * b + Mfma<T>::orow(e, ln);
        f(row, col, acc.c[a][b][e]);
      }
}

// one wave: C(16x16) = A(16x4) * B(4x16) (row-major) through Mfma<T>, each
// accumulator element stored at the (row, column) the engine assumes
template <typename T>
__global__ void k_mfma_selftest(const double* A, const double* B, double* C) {
  const int ln = threadIdx.x;
  typename Mfma<T>::acc acc = {0, 0, 0, 0};
  const T a = (T)A[(ln & 15) * 4 + (ln >> 4)];
  const T b = (T)B[(ln >> 4) * 16 + (ln & 15)];
  acc = Mfma<T>::mma(a, b, acc);
  for (int e = 0; e < 4; ++e) C[Mfma<T>::orow(e, ln) * 16 + (ln & 15)] = (double)acc[e];
}
void launch_mfma_selftest(const double* A, const double* B, double* C, int f32, hipStream_t st) {
  if (f32)
    k_mfma_selftest<float><<<1, 64, 0, st>>>(A, B, C);
  else
    k_mfma_selftest<double><<<1, 64, 0, st>>>(A, B, C);
  HIP_LAUNCH_CHECK("mfma_selftest");
}

// --------------------------------------------------------------------------
// trailing update of a wide diagonal block after a panel:
// A(i,j) -= sum_{l in [k0,k1)} L(i,l) W(j,l)  for k1 <= j <= i < w
// (dlasyf's dgemm/dgemv update of A22).  Tiles cover [lo, w)^2, lo <= k1.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_trailing(const int* list, int tiles_per_sn, int lo,
                                                  const int* super, const int64_t* pi, const int64_t* px,
                                                  T* valL, const T* Wbase, const int64_t* woff,
                                                  const int* kpos, const int* kstart) {
  __shared__ T As[KC * LDT], Bs[KC * LDT];
  const int sidx = blockIdx.x / tiles_per_sn, tix = blockIdx.x % tiles_per_sn;
  const int s = list[sidx];
  const int w = super[s + 1] - super[s];
  const int64_t r = pi[s + 1] - pi[s];
  const int k0 = kstart[s], k1 = kpos[s];
  if (k1 >= w || k1 <= k0) return;
  // lower-triangular tile index -> (ti, tj), tj <= ti
  int ti = (int)((sqrt(8.0 * tix + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= tix) ++ti;
  while (ti * (ti + 1) / 2 > tix) --ti;
  const int tj = tix - ti * (ti + 1) / 2;
  const int i0 = lo + 64 * ti, j0 = lo + 64 * tj;
  if (i0 >= w || i0 + 63 < k1 || j0 >= w || j0 + 63 < k1) return;
  T* A = valL + px[s];
  const T* Wp = Wbase + woff[s];
  const int kb = k1 - k0;
  TileAcc<T> acc;
  tile_zero(acc);
  for (int kc0 = 0; kc0 < kb; kc0 += KC) {
    const int kc = min(KC, kb - kc0);
    const int kcp = (kc + 3) & ~3;
    for (int idx = threadIdx.x; idx < kcp * 64; idx += 256) {
      const int kk = idx / 64, x = idx % 64;
      const int i = i0 + x, j = j0 + x;
      T va = 0, vb = 0;
      if (kk < kc) {
        const int l = k0 + kc0 + kk;
        if (i < w) va = A[(int64_t)l * r + i];
        if (j < w) vb = Wp[(int64_t)(l - k0) * w + j];
      }
      As[kk * LDT + x] = va;
      Bs[kk * LDT + x] = vb;
    }
    __syncthreads();
    tile_mma(acc, As, Bs, kcp);
    __syncthreads();
  }
  // batched read-modify-write: masked elements load from a safe address so
  // the 16 loads of a lane issue back to back
  {
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int r0 = (wv >> 1) * 32, q0 = (wv & 1) * 32;
    T cur[16];
    int64_t off[16];
    bool ok[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int a = q >> 3, b = (q >> 2) & 1, e = q & 3;
      const int i = i0 + r0 + 16 * a + (ln & 15), j = j0 + q0 + 16 * b + Mfma<T>::orow(e, ln);
      ok[q] = i < w && j < w && j >= k1 && i >= j;
      off[q] = ok[q] ? (int64_t)j * r + i : 0;
      cur[q] = A[off[q]];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int a = q >> 3, b = (q >> 2) & 1, e = q & 3;
      if (ok[q]) A[off[q]] = cur[q] - acc.c[a][b][e];
    }
  }
}
template <typename T>
void launch_trailing(const int* list, int count, int tiles_per_sn, int lo, const SnArgs& a, T* valL,
                     const T* W, const int64_t* woff, const int* kpos, const int* kstart,
                     hipStream_t st) {
  if (count <= 0 || tiles_per_sn <= 0) return;
  k_trailing<T><<<(int64_t)count * tiles_per_sn, 256, 0, st>>>(list, tiles_per_sn, lo, a.super, a.pi, a.px,
                                                            valL, W, woff, kpos, kstart);
  HIP_LAUNCH_CHECK("trailing");
}

// --------------------------------------------------------------------------
// off-diagonal rows: B := B P L^{-T} D^{-1}
// (LAPACKE_dlapmt + DTRSM(R,L,T,U) + blocked_2by2_solver,
//  parallel_blocked_ldlt_02.cpp:164-188).  One workgroup per 64-row block.
// gather != 0: columns are permuted here (narrow supernodes); otherwise the
// panel kernel has already permuted them.
// --------------------------------------------------------------------------
// The LDS is sized by the launch for the list's widest supernode wmax: X
// holds 64 rows (pitch 65) x min(64, wmax) columns; the second region holds
// the GEMM's operand chunks (only when wmax > 64), then the column block's
// triangle.  Lists of narrow supernodes thus keep many workgroups per CU.
__host__ __device__ inline int offdiag_lds_elems(int wmax) {
  const int xw = wmax < 64 ? wmax : 64;
  const int sh = wmax > 64 ? (2 * KC * LDT > 64 * 65 ? 2 * KC * LDT : 64 * 65) : 65 * xw;
  return 65 * xw + sh;
}
template <typename T>
__global__ __launch_bounds__(256) void k_offdiag(const int* list, const int* blk_prefix, int count,
                                                 const int* super, const int64_t* pi, const int64_t* px,
                                                 T* valL, const T* dval, const int* perm_piv,
                                                 int n, int wmax, const int* blkpos) {
  extern __shared__ double smem_d[];
  T* const X = reinterpret_cast<T*>(smem_d);  // X[row + col*65], col < min(64, wmax)
  // the GEMM's operand chunks, then the column block's triangle L(cb rows, cb
  // cols) for the substitution: never live together, so one region
  T* const Sh = X + 65 * (wmax < 64 ? wmax : 64);
  T* const As = Sh;
  T* const Bs = Sh + KC * LDT;
  T* const Lt = Sh;
  // (supernode, row block): the block's position in the list from blkpos
  const int lo = blkpos[blockIdx.x];
  const int s = list[lo];
  const int rb = blockIdx.x - blk_prefix[lo];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  T* P = valL + px[s];
  const int64_t q0 = w + 64 * (int64_t)rb;
  const int nr = (int)min((int64_t)64, r - q0);
  const int t = threadIdx.x;
  const int NT = blockDim.x;  // 256; 64 for lists of supernodes at most 16 wide (no GEMM part)
  const bool gather = w <= SMALL_W;
  for (int cb0 = 0; cb0 < w; cb0 += 64) {
    const int cw = min(64, w - cb0);
    // load B block (with column gather for narrow supernodes); a thread's
    // loads of a batch issue before its LDS stores
    for (int idx0 = 0; idx0 < 64 * cw; idx0 += NT * 8) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int idx = idx0 + t + NT * k, i = idx & 63, j = idx >> 6;
        v[k] = 0;
        if (idx < 64 * cw && i < nr) {
          const int srcc = gather ? (perm_piv[c0 + cb0 + j] - c0) : (cb0 + j);
          v[k] = P[q0 + i + (int64_t)srcc * r];
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int idx = idx0 + t + NT * k, i = idx & 63, j = idx >> 6;
        if (idx < 64 * cw) X[i + j * 65] = v[k];
      }
    }
    __syncthreads();
    // X(:,cb) -= X(:, <cb) * L(cb, <cb)^T with MFMA (only for wide supernodes)
    if (cb0 > 0) {
      TileAcc<T> acc;
      tile_zero(acc);
      for (int kc0 = 0; kc0 < cb0; kc0 += KC) {
        const int kc = min(KC, cb0 - kc0);
        constexpr int GQ = KC * 64 / 256;  // a thread's loads all issue before its LDS stores
        T va[GQ], vb[GQ];
#pragma unroll
        for (int q = 0; q < GQ; ++q) {
          const int idx = t + 256 * q, kk = idx / 64, x = idx % 64;
          const int64_t l = kc0 + kk;
          va[q] = (kk < kc && x < nr) ? P[q0 + x + l * r] : (T)0;
          vb[q] = (kk < kc && x < cw) ? P[(cb0 + x) + l * r] : (T)0;
        }
#pragma unroll
        for (int q = 0; q < GQ; ++q) {
          const int idx = t + 256 * q, kk = idx / 64, x = idx % 64;
          As[kk * LDT + x] = va[q];
          Bs[kk * LDT + x] = vb[q];
        }
        __syncthreads();
        tile_mma(acc, As, Bs, (kc + 3) & ~3);
        __syncthreads();
      }
      tile_foreach(acc, [&](int ri, int ci, T v) {
        if (ci < cw) X[ri + ci * 65] -= v;
      });
    }
    // the column block's unit-lower triangle (after the GEMM: shares its buffers)
    for (int idx0 = 0; idx0 < cw * cw; idx0 += NT * 8) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int idx = idx0 + t + NT * k, i = idx % cw, j = idx / cw;
        v[k] = (idx < cw * cw && i > j) ? P[(cb0 + i) + (int64_t)(cb0 + j) * r] : (T)0;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int idx = idx0 + t + NT * k, i = idx % cw, j = idx / cw;
        if (idx < cw * cw) Lt[i + j * 65] = v[k];
      }
    }
    __syncthreads();
    // substitution inside the column block: four adjacent lanes per row (one
    // in one-wave workgroups) split each column's sum (l = sub, sub + 4, ...)
    // and meet by shuffles; the
    // solved entry is written by the first lane of the row before any lane
    // of the row reads it (same wave, LDS in program order)
    {
      const int lpr = NT / 64, row = t / lpr, sub = t % lpr;  // lanes per row: 4, or 1
      if (row < nr) {
        for (int j = 0; j < cw; ++j) {
          T v = 0;
          for (int l = sub; l < j; l += lpr) v += X[row + l * 65] * Lt[j + l * 65];
          if (lpr == 4) {
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
          }
          const T xj = X[row + j * 65] - v;
          if (sub == 0) X[row + j * 65] = xj;
        }
      }
    }
    __syncthreads();
    for (int idx = t; idx < 64 * cw; idx += NT) {
      const int i = idx % 64, j = idx / 64;
      if (i < nr) P[q0 + i + (int64_t)(cb0 + j) * r] = X[i + j * 65];
    }
    __syncthreads();
  }
  // D^{-1}: 1x1 scaling or 2x2 Cramer (blocked_2by2_solver, Sym_BLAS.cpp:492-529),
  // 16 columns per batch so that a row's loads issue together; a 2x2 pair
  // may straddle two batches (look-ahead column / previous raw value)
  if (t < nr) {
    T* row = P + q0 + t;
    const T* dg = dval + c0;
    const T* od = dval + n + c0;
    T xprev = 0;
    for (int l0 = 0; l0 < w; l0 += 16) {
      T xv[17], yv[16];
#pragma unroll
      for (int q = 0; q < 17; ++q) xv[q] = (l0 + q < w) ? row[(int64_t)(l0 + q) * r] : 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int l = l0 + q;
        yv[q] = 0.0;
        if (l < w) {
          const T o = od[l], op = l > 0 ? od[l - 1] : (T)0;
          if (o != 0.0) {
            const T d1 = dg[l], d2 = dg[l + 1];
            yv[q] = (xv[q] * d2 - xv[q + 1] * o) / (d1 * d2 - o * o);
          } else if (op != 0.0) {
            const T d1 = dg[l - 1], d2 = dg[l];
            const T x1 = q > 0 ? xv[q - 1] : xprev;
            yv[q] = (xv[q] * d1 - x1 * op) / (d1 * d2 - op * op);
          } else {
            yv[q] = xv[q] / dg[l];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (l0 + q < w) row[(int64_t)(l0 + q) * r] = yv[q];
      xprev = xv[15];
    }
  }
}
template <typename T>
void launch_offdiag(const int* list, int count, int wmax, const int* blk_prefix, const int* blkpos, int nblocks,
                    const SnArgs& a, T* valL, const T* dval, const int* perm_piv, hipStream_t st) {
  if (count <= 0 || nblocks <= 0) return;
  if (wmax < 1) throw std::runtime_error("offdiag: empty supernode width");
  static bool attr = false;
  if (!attr) {  // lists with supernodes wider than 64 need more than the default dynamic LDS
    if (hipFuncSetAttribute((const void*)k_offdiag<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(T) * offdiag_lds_elems(65))) != hipSuccess)
      throw std::runtime_error("offdiag: cannot raise the dynamic LDS limit");
    attr = true;
  }
  // lists of supernodes at most 16 wide: one-wave workgroups, one lane per row
  k_offdiag<T><<<nblocks, wmax <= 16 ? 64 : 256, sizeof(T) * offdiag_lds_elems(wmax), st>>>(list, blk_prefix, count,
                                                                                            a.super, a.pi,
                                                                          a.px, valL, dval, perm_piv, a.n, wmax,
                                                                          blkpos);
  HIP_LAUNCH_CHECK("offdiag");
}

// --------------------------------------------------------------------------
// Schur update of a factored supernode s onto its ancestors:
// U = L_off D L_off^T (lower tiles, MFMA FP64; the D scaling of
// blocked_2by2_mult is applied while staging the B operand), scattered with
// atomics into the ancestor panels (the DGEMM + scatter-subtract of
// parallel_blocked_ldlt_02.cpp:128-153, applied right-looking).
// --------------------------------------------------------------------------
__device__ inline int64_t row_pos(const int* rows, int64_t b, int64_t e, int key) {
  // position of key in the sorted rows[b, e); key is present
  int64_t lo = b, hi = e - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (rows[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo - b;
}
template <typename T>
__global__ __launch_bounds__(256) void k_schur(const int* list, const int64_t* tile_prefix, int count,
                                               const int* super, const int64_t* pi, const int64_t* px,
                                               const int* rows, const int* col2sup, T* valL,
                                               const T* dval, int n, const int* tilepos,
                                               const unsigned char* tmask) {
  __shared__ T As[KC * LDT], Bs[KC * LDT];
  __shared__ T halo[2][64];
  __shared__ T dd[KC + 2], oo[KC + 2];
  __shared__ int rlab[64], clab[64], ctg[64], dtcol[64];
  __shared__ int tpos[4][64];
  __shared__ int64_t cbase[64];
  __shared__ int dtl[4], ntg, reach;
  const int64_t bid = blockIdx.x;
  const int lo = tilepos[bid];  // the tile's supernode: its position in the list
  const int s = list[lo];
  const int64_t tix = bid - tile_prefix[lo];
  int ti = (int)((sqrt(8.0 * (double)tix + 1.0) - 1.0) * 0.5);
  while ((int64_t)(ti + 1) * (ti + 2) / 2 <= tix) ++ti;
  while ((int64_t)ti * (ti + 1) / 2 > tix) --ti;
  const int tj = (int)(tix - (int64_t)ti * (ti + 1) / 2);
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  const int m = (int)(r - w);
  const T* P = valL + px[s];
  const int i0 = 64 * ti, j0 = 64 * tj;  // within the m off-diagonal rows
  const int t = threadIdx.x;
  if (t < 64) {
    rlab[t] = (i0 + t < m) ? rows[pi[s] + w + i0 + t] : -1;
    clab[t] = (j0 + t < m) ? rows[pi[s] + w + j0 + t] : -1;
    ctg[t] = clab[t] >= 0 ? col2sup[clab[t]] : -1;
  }
  if (tmask) {  // a restricted refactorization: tiles reaching no flagged supernode are skipped
    if (t == 0) reach = 0;
    __syncthreads();
    if (t < 64 && ctg[t] >= 0 && tmask[ctg[t]]) reach = 1;
    __syncthreads();
    if (!reach) return;
  }
  __syncthreads();
  // distinct target supernodes of the tile's columns: labels are sorted, so
  // targets come in runs; run starts are found with one wave ballot
  if (t < 64) {
    const int tg = ctg[t];
    const bool start = tg >= 0 && (t == 0 || ctg[t - 1] != tg);
    const unsigned long long mask = __ballot(start);
    const int run = __popcll(mask & (~0ull >> (63 - t))) - 1;  // run index of column t
    dtcol[t] = (tg >= 0 && run < 4) ? run : -1;
    if (start && run < 4) dtl[run] = tg;
    if (t == 0) ntg = min(4, __popcll(mask));
  }
  __syncthreads();
  for (int idx = t; idx < ntg * 64; idx += 256) {
    const int dt = idx / 64, i = idx % 64;
    const int tg = dtl[dt];
    const int R = rlab[i];
    tpos[dt][i] = (R >= super[tg]) ? (int)row_pos(rows, pi[tg], pi[tg + 1], R) : -1;
  }
  if (t < 64) {
    const int tg = ctg[t];
    cbase[t] = tg >= 0 ? px[tg] + (int64_t)(clab[t] - super[tg]) * (pi[tg + 1] - pi[tg]) : 0;
  }
  TileAcc<T> acc;
  tile_zero(acc);
  for (int kc0 = 0; kc0 < w; kc0 += KC) {
    const int kc = min(KC, w - kc0);
    // raw L columns of the chunk, plus one halo column on each side (2x2
    // blocks) and the chunk's D entries; a thread's loads all issue before
    // its LDS stores
    constexpr int SQ = KC * 64 / 256;
    T av[SQ], bw[SQ], hv = 0, dv = 0, ov = 0;
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
      const int idx = t + 256 * q, kk = idx / 64, x = idx % 64;
      const int64_t la = kc0 + kk;
      av[q] = (kk < kc && i0 + x < m) ? P[w + i0 + x + la * r] : (T)0;
      bw[q] = (kk < kc && j0 + x < m) ? P[w + j0 + x + la * r] : (T)0;
    }
    if (t < 128) {
      const int side = t >> 6, x = t & 63;
      const int l = side == 0 ? kc0 - 1 : kc0 + kc;
      const int j = j0 + x;
      hv = (l >= 0 && l < w && j < m) ? P[w + j + (int64_t)l * r] : (T)0;
    }
    if (t < KC + 2) {
      const int l = kc0 - 1 + t;
      dv = (l >= 0 && l < w) ? dval[c0 + l] : (T)0;
      ov = (l >= 0 && l < w) ? dval[n + c0 + l] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
      const int idx = t + 256 * q, kk = idx / 64, x = idx % 64;
      As[kk * LDT + x] = av[q];
      Bs[kk * LDT + x] = bw[q];
    }
    if (t < 128) halo[t >> 6][t & 63] = hv;
    if (t < KC + 2) {
      dd[t] = dv;
      oo[t] = ov;
    }
    __syncthreads();
    // B = (L D) columns: 1x1 scaling or the matching row of a 2x2 block,
    // computed in registers, then written back in place
    T bv[KC * 64 / 256];
    for (int q = 0; q < KC * 64 / 256; ++q) {
      const int idx = t + 256 * q;
      const int kk = idx / 64, x = idx % 64;
      T vb = 0;
      if (kk < kc) {
        const int h = kk + 1;  // index into dd / oo
        const T x0 = Bs[kk * LDT + x];
        const T xn = (kk + 1 < kc) ? Bs[(kk + 1) * LDT + x] : halo[1][x];
        const T xp = (kk > 0) ? Bs[(kk - 1) * LDT + x] : halo[0][x];
        if (oo[h] != 0.0) vb = dd[h] * x0 + oo[h] * xn;
        else if (oo[h - 1] != 0.0) vb = oo[h - 1] * xp + dd[h] * x0;
        else vb = dd[h] * x0;
      }
      bv[q] = vb;
    }
    __syncthreads();
    for (int q = 0; q < KC * 64 / 256; ++q) {
      const int idx = t + 256 * q;
      Bs[(idx / 64) * LDT + idx % 64] = bv[q];
    }
    __syncthreads();
    tile_mma(acc, As, Bs, (kc + 3) & ~3);
    __syncthreads();
  }
  // target bases and row positions come from LDS on the common path
  tile_foreach(acc, [&](int ri, int ci, T v) {
    const int R = rlab[ri], C = clab[ci];
    if (R < 0 || C < 0 || R < C) return;
    if (tmask && !tmask[ctg[ci]]) return;
    const int dt = dtcol[ci];
    int64_t pos;
    if (dt >= 0) {
      pos = tpos[dt][ri];
    } else {
      const int tg = ctg[ci];
      pos = row_pos(rows, pi[tg], pi[tg + 1], R);
    }
    atomicAdd(valL + cbase[ci] + pos, -v);
  });
}
template <typename T>
void launch_schur(const int* list, int count, const int64_t* tile_prefix, const int* tilepos, int64_t ntiles,
                  const SnArgs& a, T* valL, const T* dval, hipStream_t st) {
  if (count <= 0 || ntiles <= 0) return;
  if (ntiles > 0x7fffffffLL) throw std::runtime_error("schur: too many tiles");
  k_schur<T><<<(unsigned)ntiles, 256, 0, st>>>(list, tile_prefix, count, a.super, a.pi, a.px, a.rows,
                                            a.col2sup, valL, dval, a.n, tilepos, a.tmask);
  HIP_LAUNCH_CHECK("schur");
}

// --------------------------------------------------------------------------
// after all supernodes are factored: relabel off-diagonal row indices into
// the pivoted ordering (row_reordering, Sym_BLAS.cpp:327-490, done as index
// relabelling) and compose the final ordering (reorder_matrix,
// linear_solver_wrapper.cpp:1432-1460)
// --------------------------------------------------------------------------
__global__ void k_newpos(const int* perm_piv, int* newpos, const int* perm0, int* perm_f, int* iperm_f,
                         int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    newpos[perm_piv[k]] = k;
    const int o = perm0[perm_piv[k]];
    perm_f[k] = o;
    iperm_f[o] = k;
  }
}
__global__ void k_relabel(int nsuper, const int* super, const int64_t* pi, const int* rows0, int* rows,
                          const int* newpos) {
  const int s = blockIdx.x;
  if (s >= nsuper) return;
  const int c0 = super[s], w = super[s + 1] - c0;
  for (int64_t q = pi[s] + threadIdx.x; q < pi[s + 1]; q += blockDim.x) {
    const int64_t loc = q - pi[s];
    rows[q] = loc < w ? c0 + (int)loc : newpos[rows0[q]];
  }
}
void launch_relabel(const SnArgs& a, int nsuper, const int* rows0, int* rows, const int* perm_piv,
                    int* newpos, const int* perm0, int* perm_f, int* iperm_f, hipStream_t st) {
  k_newpos<<<grid_for(a.n, 256), 256, 0, st>>>(perm_piv, newpos, perm0, perm_f, iperm_f, a.n);
  HIP_LAUNCH_CHECK("newpos");
  k_relabel<<<nsuper, 256, 0, st>>>(nsuper, a.super, a.pi, rows0, rows, newpos);
  HIP_LAUNCH_CHECK("relabel");
}

// --------------------------------------------------------------------------
// sweeps over solve pieces (column blocks of at most SOLVE_W of a supernode)
// X, Y, Z are n x nrhs, row-interleaved (element (i,c) at i*nrhs + c).
// Rows of a piece below its triangle: local rows q in [c0l+pw, r) with label
// super+q for q < w and rows[pi+q] otherwise.
// --------------------------------------------------------------------------
#define MAXRHS 8
__device__ inline int row_label(const int* rows, int64_t pis, int c0, int w, int64_t q) {
  return q < w ? c0 + (int)q : rows[pis + q];
}

// Sweeps over pieces of <= SOLVE_W columns of a supernode.  Triangles: the
// piece's unit lower triangle is staged in LDS (pitch lp >= pw + 1, dynamic)
// and one wave substitutes with its 64 lanes holding rows (lane, lane + 64);
// the solved entry of each step is broadcast with a lane shuffle.  Row
// blocks below a piece: GROWS rows per workgroup, the four waves split the
// piece's columns, partial sums meet in LDS.
#define GROWS SOLVE_ROWS
#define TSTEP 16  // substitution steps per batch of LDS reads
template <typename T>
__device__ inline T bcast(T v, int lane) { return __shfl(v, lane, 64); }

// Stage the strict lower triangle of a piece's pw x pw diagonal block (pw <=
// SOLVE_W) into LDS: entry L(d0 + i, d0 + j), i > j, at Lt[j * lp + i]
// (colmajor) or Lt[i * lp + j] (the transpose).  Thread (i, h) (i = row, h =
// which half of the workgroup; pw <= blockDim.x / 2) loads the entries
// j = h, h + 2, ... < i of its row,
// 16 loads in flight before the first LDS store; a wave reads 64 consecutive
// rows of one column per load.  Narrow pieces cost a few loads per thread.
template <typename T>
__device__ __forceinline__ void stage_piece_triangle(T* Lt, int lp, const T* P, int64_t r, int d0, int pw,
                                                     bool colmajor) {
  constexpr int B = 16;
  const int half = blockDim.x / 2;  // rows per half: 128 (256 threads) or 32 (64 threads, pw <= 32)
  const int i = threadIdx.x % half, h = threadIdx.x / half;
  if (i >= pw) return;
  const int nj = (i - h + 1) / 2;  // entries j = h + 2 k < i
  const T* Pi = P + (int64_t)(d0 + i) + (int64_t)(d0 + h) * r;
  for (int k0 = 0; k0 < nj; k0 += B) {
    T v[B];
#pragma unroll
    for (int k = 0; k < B; ++k) v[k] = (k0 + k < nj) ? Pi[(int64_t)(2 * (k0 + k)) * r] : (T)0;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int j = h + 2 * (k0 + k);
      if (k0 + k < nj) Lt[colmajor ? j * lp + i : i * lp + j] = v[k];
    }
  }
}
// value of lane `lane` (wave-uniform) of a VGPR, through v_readlane
__device__ inline double readlane_t(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline float readlane_t(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// forward triangle of a piece: Y(p) = L_pp^{-1} X(p)
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_tri(const int* pieces, const int* piece_sn, const int* piece_c0,
                                                 const int* super, const int64_t* pi, const int64_t* px,
                                                 const T* valL, const T* X, T* Y, int nrhs, int lp) {
  extern __shared__ double smem_d[];
  T* Lt = reinterpret_cast<T*>(smem_d);  // Lt[j * lp + i] = L(i, j), i > j
  const int p = pieces[blockIdx.x];
  const int s = piece_sn[p], c0l = piece_c0[p];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int pw = min(SOLVE_W, w - c0l);
  const int64_t r = pi[s + 1] - pi[s];
  const T* P = valL + px[s];
  const int t = threadIdx.x;
  stage_piece_triangle(Lt, lp, P, r, c0l, pw, true);
  __syncthreads();
  // wave wv substitutes right-hand sides wv and wv + 4 (lane ln holds rows
  // ln and ln + 64), so up to four SIMDs run the sequential steps at once
  const int wv = t >> 6, ln = t & 63;
  if (wv >= min(nrhs, 4)) return;
  const int64_t base = (int64_t)(c0 + c0l) * nrhs;
  T x0[2], x1[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = wv + 4 * q;
    x0[q] = (c < nrhs && ln < pw) ? X[base + (int64_t)ln * nrhs + c] : (T)0;
    x1[q] = (c < nrhs && ln + 64 < pw) ? X[base + (int64_t)(ln + 64) * nrhs + c] : (T)0;
  }
  const bool two = wv + 4 < nrhs;
  // column j holds rows j + 1 .. pw - 1.  The steps go in chunks of TSTEP:
  // a lane reads its entries of the chunk's columns together (addresses
  // clamped into the columns, masked afterwards), then substitutes from
  // registers, so an LDS round trip is paid per chunk rather than per step.
  const int n0 = min(64, pw - 1);
  for (int jc = 0; jc < n0; jc += TSTEP) {  // x_j in lane j of x0
    T e0[TSTEP], e1[TSTEP];
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = min(jc + k, n0 - 1);
      e0[k] = Lt[j * lp + min(max(ln, j + 1), pw - 1)];
      e1[k] = Lt[j * lp + min(ln + 64, pw - 1)];
    }
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = jc + k;
      if (j >= n0) break;
      const T l0 = (ln > j && ln < pw) ? e0[k] : (T)0;
      const T l1 = (ln + 64 < pw) ? e1[k] : (T)0;
      const T xa = readlane_t(x0[0], j);
      x0[0] -= l0 * xa;
      x1[0] -= l1 * xa;
      if (two) {
        const T xb = readlane_t(x0[1], j);
        x0[1] -= l0 * xb;
        x1[1] -= l1 * xb;
      }
    }
  }
  for (int jc = 64; jc < pw - 1; jc += TSTEP) {  // x_j in lane j - 64 of x1
    T e1[TSTEP];
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = min(jc + k, pw - 2);
      e1[k] = Lt[j * lp + min(max(ln + 64, j + 1), pw - 1)];
    }
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = jc + k;
      if (j >= pw - 1) break;
      const T l1 = (ln + 64 > j && ln + 64 < pw) ? e1[k] : (T)0;
      x1[0] -= l1 * readlane_t(x1[0], j - 64);
      if (two) x1[1] -= l1 * readlane_t(x1[1], j - 64);
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = wv + 4 * q;
    if (c >= nrhs) break;
    if (ln < pw) Y[base + (int64_t)ln * nrhs + c] = x0[q];
    if (ln + 64 < pw) Y[base + (int64_t)(ln + 64) * nrhs + c] = x1[q];
  }
}
// forward update of the rows below a piece: X(label q) -= L(q, p) Y(p)
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gemv(const int* pieces, const int* blk_prefix, const int* blkpos,
                                                  int blk_base, const int* piece_sn, const int* piece_c0,
                                                  const int* super, const int64_t* pi, const int64_t* px,
                                                  const int* rows, const T* valL, const T* Y, T* X, int nrhs) {
  __shared__ T ys[SOLVE_W * MAXRHS];
  __shared__ T part[4][GROWS * MAXRHS];
  // row block b of the level: piece at position blkpos[b] of the level's list
  const int b = (int)blockIdx.x + blk_base, lo = blkpos[b];
  const int p = pieces[lo], rb = b - blk_prefix[lo];
  const int s = piece_sn[p], c0l = piece_c0[p];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int pw = min(SOLVE_W, w - c0l);
  const int64_t r = pi[s + 1] - pi[s];
  const T* P = valL + px[s];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  for (int idx = t; idx < pw * nrhs; idx += 256) ys[idx] = Y[(int64_t)(c0 + c0l) * nrhs + idx];
  __syncthreads();
  const int64_t q = c0l + pw + (int64_t)GROWS * rb + ln;
  const bool ok = q < r;
  T acc[MAXRHS];
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c) acc[c] = 0.0;
  const int jb = wv * (SOLVE_W / 4), je = min(pw, jb + SOLVE_W / 4);
  if (ok) {
    for (int j = jb; j < je; j += 8) {
      T l[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) l[u] = (j + u < je) ? P[q + (int64_t)(c0l + j + u) * r] : (T)0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int c = 0; c < MAXRHS; ++c)
          if (c < nrhs && j + u < je) acc[c] += l[u] * ys[(j + u) * nrhs + c];
    }
  }
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c)
    if (c < nrhs) part[wv][ln * nrhs + c] = acc[c];
  __syncthreads();
  if (wv == 0 && ok) {
    const int lab = row_label(rows, pi[s], c0, w, q);
    for (int c = 0; c < nrhs; ++c) {
      const T v = part[0][ln * nrhs + c] + part[1][ln * nrhs + c] + part[2][ln * nrhs + c] +
                       part[3][ln * nrhs + c];
      atomicAdd(X + (int64_t)lab * nrhs + c, -v);
    }
  }
}
// D solve: Z = D^{-1} Y (blocked_2by2_solver with n_rhs = 1)
template <typename T>
__global__ void k_dsolve(int n, const T* dval, const T* Y, T* Z, int nrhs) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T od = dval[n + i];
    if (od != 0.0) {
      const T d1 = dval[i], d2 = dval[i + 1];
      const T inv = (T)1 / (d1 * d2 - od * od);
      for (int c = 0; c < nrhs; ++c) {
        const T x1 = Y[(int64_t)i * nrhs + c], x2 = Y[(int64_t)(i + 1) * nrhs + c];
        Z[(int64_t)i * nrhs + c] = (x1 * d2 - x2 * od) * inv;
        Z[(int64_t)(i + 1) * nrhs + c] = (x2 * d1 - x1 * od) * inv;
      }
    } else if (i == 0 || dval[n + i - 1] == 0.0) {
      const T inv = (T)1 / dval[i];
      for (int c = 0; c < nrhs; ++c) Z[(int64_t)i * nrhs + c] = Y[(int64_t)i * nrhs + c] * inv;
    }
  }
}
// backward gather: acc(p) += L(q, p)^T X(label q) over the rows below a piece.
// The GROWS x pw block is staged in LDS (coalesced along the rows), then
// each thread sums a column over half of the rows.  The block's LDS is sized
// by the launch for the level's widest piece, so levels of narrow pieces
// keep many workgroups per CU.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_gemv(const int* pieces, const int* blk_prefix, const int* blkpos,
                                                  int blk_base, const int* piece_sn, const int* piece_c0,
                                                  const int* super, const int64_t* pi, const int64_t* px,
                                                  const int* rows, const T* valL, const T* Xf, T* acc,
                                                  int nrhs) {
  extern __shared__ double smem_d[];
  T* Lb = reinterpret_cast<T*>(smem_d);  // Lb[j * (GROWS + 1) + qq], j < the launch's widest piece
  __shared__ T xq[GROWS * MAXRHS];
  __shared__ T part[SOLVE_W * MAXRHS];
  // row block b of the level (the launch covers blocks from blk_base): piece
  // at position blkpos[b] of the level's list
  const int b = (int)blockIdx.x + blk_base, lo = blkpos[b];
  const int p = pieces[lo], rb = b - blk_prefix[lo];
  const int s = piece_sn[p], c0l = piece_c0[p];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int pw = min(SOLVE_W, w - c0l);
  const int64_t r = pi[s + 1] - pi[s];
  const T* P = valL + px[s];
  const int t = threadIdx.x;
  const int64_t qb = c0l + pw + (int64_t)GROWS * rb;
  const int nq = (int)min((int64_t)GROWS, r - qb);
  // the GROWS x pw block: thread (qq, g) loads columns g, g + 4, ... of row
  // qq, 16 loads in flight before its first LDS store
  const int qq0 = t % GROWS, g0 = t / GROWS;
  const int njb = (pw - g0 + 3) / 4;
  const T* Pq = P + qb + qq0 + (int64_t)(c0l + g0) * r;
  if (t < nq) {
    const int lab = row_label(rows, pi[s], c0, w, qb + t);
    for (int c = 0; c < nrhs; ++c) xq[t * nrhs + c] = Xf[(int64_t)lab * nrhs + c];
  } else if (t < GROWS) {
    for (int c = 0; c < nrhs; ++c) xq[t * nrhs + c] = 0.0;  // rows past the panel: 0 * stale LDS could be NaN
  }
  for (int k0 = 0; k0 < njb; k0 += 16) {
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = (qq0 < nq && k0 + k < njb) ? Pq[(int64_t)(4 * (k0 + k)) * r] : (T)0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k0 + k < njb) Lb[(g0 + 4 * (k0 + k)) * (GROWS + 1) + qq0] = v[k];
  }
  __syncthreads();
  const int j = t & (SOLVE_W - 1), half = t >> 7;  // 2 x 128 threads: column j, rows of one half
  T a[MAXRHS];
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c) a[c] = 0.0;
  if (j < pw) {
    const int q0 = half * (GROWS / 2);
    for (int qq = q0; qq < min(q0 + GROWS / 2, nq); ++qq) {  // the block's rows only
      const T l = Lb[j * (GROWS + 1) + qq];
#pragma unroll
      for (int c = 0; c < MAXRHS; ++c)
        if (c < nrhs) a[c] += l * xq[qq * nrhs + c];
    }
  }
  if (half == 1)
#pragma unroll
    for (int c = 0; c < MAXRHS; ++c)
      if (c < nrhs) part[j * nrhs + c] = a[c];
  __syncthreads();
  if (half == 0 && j < pw)
    for (int c = 0; c < nrhs; ++c) atomicAdd(acc + (int64_t)(c0 + c0l + j) * nrhs + c, a[c] + part[j * nrhs + c]);
}
// backward triangle: X(p) = L_pp^{-T} (Z(p) - acc(p))
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_tri(const int* pieces, const int* piece_sn, const int* piece_c0,
                                                 const int* super, const int64_t* pi, const int64_t* px,
                                                 const T* valL, const T* Z, const T* acc,
                                                 T* Xf, int nrhs, int lp) {
  extern __shared__ double smem_d[];
  T* Lt = reinterpret_cast<T*>(smem_d);  // Lt[i * lp + j] = L(i, j), i > j (row i of L)
  const int p = pieces[blockIdx.x];
  const int s = piece_sn[p], c0l = piece_c0[p];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int pw = min(SOLVE_W, w - c0l);
  const int64_t r = pi[s + 1] - pi[s];
  const T* P = valL + px[s];
  const int t = threadIdx.x;
  stage_piece_triangle(Lt, lp, P, r, c0l, pw, false);
  __syncthreads();
  // wave wv substitutes right-hand sides wv and wv + 4 (lane ln holds rows
  // ln and ln + 64)
  const int wv = t >> 6, ln = t & 63;
  if (wv >= min(nrhs, 4)) return;
  const int64_t base = (int64_t)(c0 + c0l) * nrhs;
  T x0[2], x1[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = wv + 4 * q;
    x0[q] = (c < nrhs && ln < pw) ? Z[base + (int64_t)ln * nrhs + c] - acc[base + (int64_t)ln * nrhs + c] : (T)0;
    x1[q] = (c < nrhs && ln + 64 < pw)
                ? Z[base + (int64_t)(ln + 64) * nrhs + c] - acc[base + (int64_t)(ln + 64) * nrhs + c]
                : (T)0;
  }
  const bool two = wv + 4 < nrhs;
  // rows i < j: x_i -= L(j, i) x_j, from the last column back; row j holds
  // L(j, 0 .. j - 1).  The steps go in chunks of TSTEP rows read together
  // (addresses clamped into the rows, masked afterwards), then substituted
  // from registers.
  for (int jc = pw - 1; jc >= 64; jc -= TSTEP) {  // x_j in lane j - 64 of x1
    T e0[TSTEP], e1[TSTEP];
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = max(jc - k, 64);
      e0[k] = Lt[j * lp + ln];
      e1[k] = Lt[j * lp + min(ln + 64, j - 1)];
    }
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = jc - k;
      if (j < 64) break;
      const T l0 = e0[k];
      const T l1 = (ln + 64 < j) ? e1[k] : (T)0;
      const T xa = readlane_t(x1[0], j - 64);
      x0[0] -= l0 * xa;
      x1[0] -= l1 * xa;
      if (two) {
        const T xb = readlane_t(x1[1], j - 64);
        x0[1] -= l0 * xb;
        x1[1] -= l1 * xb;
      }
    }
  }
  for (int jc = min(pw, 64) - 1; jc > 0; jc -= TSTEP) {  // x_j in lane j of x0
    T e0[TSTEP];
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = max(jc - k, 1);
      e0[k] = Lt[j * lp + min(ln, j - 1)];
    }
#pragma unroll
    for (int k = 0; k < TSTEP; ++k) {
      const int j = jc - k;
      if (j <= 0) break;
      const T l0 = (ln < j) ? e0[k] : (T)0;
      x0[0] -= l0 * readlane_t(x0[0], j);
      if (two) x0[1] -= l0 * readlane_t(x0[1], j);
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = wv + 4 * q;
    if (c >= nrhs) break;
    if (ln < pw) Xf[base + (int64_t)ln * nrhs + c] = x0[q];
    if (ln + 64 < pw) Xf[base + (int64_t)(ln + 64) * nrhs + c] = x1[q];
  }
}

template <typename T>
void launch_fwd_tri(const int* pieces, int np, int maxpw, const int* piece_sn, const int* piece_c0, const SnArgs& a,
                    const T* valL, const T* X, T* Y, int nrhs, hipStream_t st) {
  if (np <= 0) return;
  if (nrhs > MAXRHS) throw std::runtime_error("too many right-hand sides");
  if (maxpw < 1 || maxpw > SOLVE_W) throw std::runtime_error("fwd_tri: piece width out of range");
  const int lp = maxpw + 1;
  static bool attr = false;
  if (!attr) {  // the widest triangle (128 x 129 doubles) needs more than the default dynamic LDS
    if (hipFuncSetAttribute((const void*)k_fwd_tri<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(T) * SOLVE_W * (SOLVE_W + 1))) != hipSuccess)
      throw std::runtime_error("fwd_tri: cannot raise the dynamic LDS limit");
    attr = true;
  }
  // one right-hand side and pieces of at most 32 columns: one wave does all
  // the work, so the launch uses one-wave workgroups (more of them per CU)
  const int nt = (nrhs == 1 && maxpw <= 32) ? 64 : 256;
  k_fwd_tri<T><<<np, nt, sizeof(T) * maxpw * lp, st>>>(pieces, piece_sn, piece_c0, a.super, a.pi, a.px, valL, X,
                                                         Y, nrhs, lp);
  HIP_LAUNCH_CHECK("fwd_tri");
}
template <typename T>
void launch_fwd_gemv(const int* pieces, const int* blk_prefix, const int* blkpos, int nblk, const int* piece_sn,
                     const int* piece_c0, const SnArgs& a, const T* valL, const T* Y, T* X,
                     int nrhs, hipStream_t st) {
  if (nblk <= 0) return;
  k_fwd_gemv<T><<<nblk, 256, 0, st>>>(pieces, blk_prefix, blkpos, 0, piece_sn, piece_c0, a.super, a.pi, a.px,
                                   a.rows, valL, Y, X, nrhs);
  HIP_LAUNCH_CHECK("fwd_gemv");
}
template <typename T>
void launch_dsolve(int n, const T* dval, const T* Y, T* Z, int nrhs, hipStream_t st) {
  k_dsolve<T><<<grid_for(n, 256), 256, 0, st>>>(n, dval, Y, Z, nrhs);
  HIP_LAUNCH_CHECK("dsolve");
}
template <typename T>
void launch_bwd_gemv(const int* pieces, int maxpw, const int* blk_prefix, const int* blkpos, int nblk, int blk_base,
                     const int* piece_sn, const int* piece_c0, const SnArgs& a, const T* valL, const T* Xf, T* acc,
                     int nrhs, hipStream_t st) {
  if (nblk <= 0) return;
  if (maxpw < 1 || maxpw > SOLVE_W) throw std::runtime_error("bwd_gemv: piece width out of range");
  static bool attr = false;
  if (!attr) {  // a level of 128-column pieces needs more than the default dynamic LDS
    if (hipFuncSetAttribute((const void*)k_bwd_gemv<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(T) * SOLVE_W * (GROWS + 1))) != hipSuccess)
      throw std::runtime_error("bwd_gemv: cannot raise the dynamic LDS limit");
    attr = true;
  }
  k_bwd_gemv<T><<<nblk, 256, sizeof(T) * maxpw * (GROWS + 1), st>>>(pieces, blk_prefix, blkpos, blk_base,
                                                                    piece_sn, piece_c0, a.super, a.pi, a.px,
                                                                    a.rows, valL, Xf, acc, nrhs);
  HIP_LAUNCH_CHECK("bwd_gemv");
}
template <typename T>
void launch_bwd_tri(const int* pieces, int np, int maxpw, const int* piece_sn, const int* piece_c0, const SnArgs& a,
                    const T* valL, const T* Z, const T* acc, T* Xf, int nrhs,
                    hipStream_t st) {
  if (np <= 0) return;
  if (maxpw < 1 || maxpw > SOLVE_W) throw std::runtime_error("bwd_tri: piece width out of range");
  const int lp = maxpw + 1;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void*)k_bwd_tri<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(T) * SOLVE_W * (SOLVE_W + 1))) != hipSuccess)
      throw std::runtime_error("bwd_tri: cannot raise the dynamic LDS limit");
    attr = true;
  }
  const int nt = (nrhs == 1 && maxpw <= 32) ? 64 : 256;  // as in launch_fwd_tri
  k_bwd_tri<T><<<np, nt, sizeof(T) * maxpw * lp, st>>>(pieces, piece_sn, piece_c0, a.super, a.pi, a.px, valL, Z,
                                                          acc, Xf, nrhs, lp);
  HIP_LAUNCH_CHECK("bwd_tri");
}

// --------------------------------------------------------------------------
// y_ord = P K0 P^T x_ord with the full CSR of K0 in the original ordering
// --------------------------------------------------------------------------
__global__ void k_spmv(int n, const int* rp, const int* ci, const double* v, const int* perm_f,
                       const int* iperm_f, const double* x, double* y, int nrhs) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int row = perm_f[i];
    double acc[MAXRHS];
    for (int c = 0; c < nrhs; ++c) acc[c] = 0.0;
    for (int q = rp[row]; q < rp[row + 1]; ++q) {
      const double a = v[q];
      const int j = iperm_f[ci[q]];
      for (int c = 0; c < nrhs; ++c) acc[c] += a * x[(int64_t)j * nrhs + c];
    }
    for (int c = 0; c < nrhs; ++c) y[(int64_t)i * nrhs + c] = acc[c];
  }
}
void launch_spmv(int n, const int* rp, const int* ci, const double* v, const int* perm_f, const int* iperm_f,
                 const double* x, double* y, int nrhs, hipStream_t st) {
  k_spmv<<<grid_for(n, 256), 256, 0, st>>>(n, rp, ci, v, perm_f, iperm_f, x, y, nrhs);
  HIP_LAUNCH_CHECK("spmv");
}

// ordered <-> original vector layouts
__global__ void k_gather_rhs(int n, int nrhs, const double* B, int64_t ldb, const int* perm_f, double* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < nrhs; ++c) out[(int64_t)i * nrhs + c] = B[(int64_t)c * ldb + perm_f[i]];
}
__global__ void k_scatter_sol(int n, int nrhs, const double* Xo, const int* perm_f, double* X, int64_t ldx) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < nrhs; ++c) X[(int64_t)c * ldx + perm_f[i]] = Xo[(int64_t)i * nrhs + c];
}
void launch_gather_rhs(int n, int nrhs, const double* B, int64_t ldb, const int* perm_f, double* out,
                       hipStream_t st) {
  k_gather_rhs<<<grid_for(n, 256), 256, 0, st>>>(n, nrhs, B, ldb, perm_f, out);
  HIP_LAUNCH_CHECK("gather_rhs");
}
void launch_scatter_sol(int n, int nrhs, const double* Xo, const int* perm_f, double* X, int64_t ldx,
                        hipStream_t st) {
  k_scatter_sol<<<grid_for(n, 256), 256, 0, st>>>(n, nrhs, Xo, perm_f, X, ldx);
  HIP_LAUNCH_CHECK("scatter_sol");
}

// --------------------------------------------------------------------------
// GMRES(mr) on the device (pmgmres_ldlt_auto, mgmres.cpp:1015-1211), one
// independent iteration per right-hand-side column.  Small state per column:
// --------------------------------------------------------------------------
// state layout per column (doubles), MR = max restart length:
//   h[(MR+1)*MR] | cs[MR+1] | sn[MR+1] | g[MR+1] | y[MR+1] | dots[MR+2] |
//   rho, rho_tol, av, flags(active, inner_done, k_last), used
__device__ inline double* gs_col(double* gm, int c, int MR) {
  return gm + (int64_t)c * ((MR + 1) * MR + 4 * (MR + 1) + (MR + 2) + 8);
}
#define GS_H(st, i, j, MR) st[(i) * (MR) + (j)]
#define GS_CS(st, MR) (st + (MR + 1) * (MR))
#define GS_SN(st, MR) (GS_CS(st, MR) + (MR + 1))
#define GS_G(st, MR) (GS_SN(st, MR) + (MR + 1))
#define GS_Y(st, MR) (GS_G(st, MR) + (MR + 1))
#define GS_D(st, MR) (GS_Y(st, MR) + (MR + 1))
#define GS_S(st, MR) (GS_D(st, MR) + (MR + 2))  // scalars: [0]rho [1]rho_tol [2]av [3]active [4]inner_done [5]k_last [6]used [7]norm_tmp

// partial dot products of nd pairs (a_k, b_k) for every column; part[blk][k][c]
// dot products per column of a (n x nrhs) and b: one partial per workgroup
__global__ void k_dot_partial(int n, int nrhs, const double* a, const double* b, double* part) {
  __shared__ double sh[4][MAXRHS];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double acc[MAXRHS];
  for (int c = 0; c < nrhs; ++c) acc[c] = 0.0;
  for (int i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < nrhs; ++c) acc[c] += a[(int64_t)i * nrhs + c] * b[(int64_t)i * nrhs + c];
  for (int c = 0; c < nrhs; ++c) {
    const double v = wave_sum(acc[c]);
    if (ln == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (t < nrhs) part[(int64_t)blockIdx.x * nrhs + t] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
}
// out[v] = sum over blocks of part[b * nv + v], in a fixed order (strided
// per-thread sums, then a fixed tree), so repeated calls agree bit for bit
__global__ void k_red_final(const double* part, int nblk, int nv, double* out) {
  __shared__ double sh[4];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  for (int v = 0; v < nv; ++v) {
    double s = 0.0;
    for (int b = t; b < nblk; b += 256) s += part[(int64_t)b * nv + v];
    s = wave_sum(s);
    if (ln == 0) sh[wv] = s;
    __syncthreads();
    if (t == 0) out[v] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
  }
}
// workgroups for a reduction over n rows: enough to cover n, at most cap
static int red_blocks(int n, int cap) { return (int)std::max<int64_t>(1, std::min<int64_t>(cap, (n + 255) / 256)); }
// short vectors: one workgroup of 1024 threads, strided per-thread sums and
// a fixed tree (one launch instead of two; repeated calls agree bit for bit)
__global__ __launch_bounds__(1024) void k_dot_one(int n, int nrhs, const double* a, const double* b, double* out) {
  __shared__ double sh[16][MAXRHS];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double acc[MAXRHS];
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c) acc[c] = 0.0;
  for (int i = t; i < n; i += 1024)
#pragma unroll
    for (int c = 0; c < MAXRHS; ++c)
      if (c < nrhs) acc[c] += a[(int64_t)i * nrhs + c] * b[(int64_t)i * nrhs + c];
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c) {
    if (c >= nrhs) break;
    const double v = wave_sum(acc[c]);
    if (ln == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (t < nrhs) {
    double s = 0.0;
    for (int q = 0; q < 16; ++q) s += sh[q][t];
    out[t] = s;
  }
}
#define DOT_ONE_MAX (1 << 16)  // rows x columns up to which one workgroup takes the dot products

void launch_dot(int n, int nrhs, const double* a, const double* b, double* part, int nblk, double* out,
                hipStream_t st) {
  if ((int64_t)n * nrhs <= DOT_ONE_MAX) {
    k_dot_one<<<1, 1024, 0, st>>>(n, nrhs, a, b, out);
    HIP_LAUNCH_CHECK("dot_one");
    return;
  }
  nblk = red_blocks(n, nblk);
  k_dot_partial<<<nblk, 256, 0, st>>>(n, nrhs, a, b, part);
  HIP_LAUNCH_CHECK("dot_partial");
  k_red_final<<<1, 256, 0, st>>>(part, nblk, nrhs, out);
  HIP_LAUNCH_CHECK("dot_final");
}

// elementwise: r = b - y
__global__ void k_sub(int64_t count, const double* b, const double* y, double* r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - y[i];
}
void launch_sub(int64_t count, const double* b, const double* y, double* r, hipStream_t st) {
  k_sub<<<grid_for(count, 256), 256, 0, st>>>(count, b, y, r);
  HIP_LAUNCH_CHECK("sub");
}

// y(:,c) = alpha_c * x(:,c) + beta_c * y(:,c) with per-column coefficients
// read from the device (coef[c*2], coef[c*2+1]); columns whose mask entry is
// zero are left untouched
__global__ void k_axpby(int n, int nrhs, const double* coefs, int cstride, const double* x, double* y,
                        const double* mask, int mstride, int mask_off, int mode) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < nrhs; ++c) {
      if (mask && mask[(int64_t)c * mstride + mask_off] == 0.0) continue;
      const double a = coefs[(int64_t)c * cstride];
      const int64_t e = (int64_t)i * nrhs + c;
      if (mode == 0) y[e] -= a * x[e];          // y -= a x
      else if (mode == 1) {                     // y = x / a (skipped when a == 0)
        if (a != 0.0) y[e] = x[e] / a;
      }
      else y[e] += a * x[e];                    // y += a x
    }
}
void launch_axpby(int n, int nrhs, const double* coefs, int cstride, const double* x, double* y,
                  const double* mask, int mstride, int mask_off, int mode, hipStream_t st) {
  k_axpby<<<grid_for(n, 256), 256, 0, st>>>(n, nrhs, coefs, cstride, x, y, mask, mstride, mask_off, mode);
  HIP_LAUNCH_CHECK("axpby");
}

// scalar GMRES bookkeeping kernels (one thread per column)
// op 0: after r = M(b - K x): rho = sqrt(d0); stop if rho < tol_abs; set rho_tol at itr 0;
//       g = [rho,0..]; h = 0; coefficient for v0 = r / rho
// op 1: av = sqrt(d0)
// op 2: h(j,k) = d0 (MGS coefficient)  [j,k given]
// op 3: h(k+1,k) = sqrt(d0); reorth flag
// op 4: reorth: h(j,k) += d0
// op 5: h(k+1,k) = sqrt(d0) after reorth
// op 6: Givens for column k, convergence check
// op 7: back substitution -> y
__device__ void gmres_scalar_col(int op, int c, int MR, double* gm, double d0, int j, int k, int itr,
                                 double tol_abs, double tol_rel) {
  double* st = gs_col(gm, c, MR);
  double* S = GS_S(st, MR);
  double* H = st;
  double* cs = GS_CS(st, MR);
  double* sn = GS_SN(st, MR);
  double* g = GS_G(st, MR);
  double* y = GS_Y(st, MR);
  double* D = GS_D(st, MR);  // D[0]: coefficient slot used by vector kernels
  const double delta = 1.0e-3;
  if (op == 0) {
    S[7] = 0.0;
    if (S[3] == 0.0) return;  // column already finished
    const double rho = sqrt(d0);
    S[0] = rho;
    if (rho < tol_abs) {
      S[3] = 0.0;
      S[5] = -1.0;
      return;
    }
    S[7] = 1.0;
    if (itr == 0) S[1] = rho * tol_rel;
    g[0] = rho;
    for (int i = 1; i <= MR; ++i) g[i] = 0.0;
    for (int i = 0; i <= MR; ++i)
      for (int q = 0; q < MR; ++q) GS_H(H, i, q, MR) = 0.0;
    S[4] = 0.0;  // inner not done
    S[5] = -1.0;
    D[0] = rho;
  } else if (op == 1) {
    if (S[7] == 0.0) return;
    S[2] = sqrt(d0);
  } else if (op == 2) {
    if (S[7] == 0.0) return;
    GS_H(H, j, k, MR) = d0;
    D[0] = d0;
  } else if (op == 3) {
    D[1] = 0.0;
    if (S[7] == 0.0) return;
    const double hk = sqrt(d0);
    GS_H(H, k + 1, k, MR) = hk;
    D[1] = (S[2] + delta * hk == S[2]) ? 1.0 : 0.0;  // re-orthogonalise?
    D[0] = hk;
  } else if (op == 4) {
    if (D[1] == 0.0) return;
    GS_H(H, j, k, MR) += d0;
    D[0] = d0;
  } else if (op == 5) {
    if (D[1] == 0.0) return;
    const double hk = sqrt(d0);
    GS_H(H, k + 1, k, MR) = hk;
    D[0] = hk;
  } else if (op == 6) {
    if (S[7] == 0.0) return;
    S[5] = (double)k;
    if (k > 0) {
      for (int i = 0; i < k + 2; ++i) y[i] = GS_H(H, i, k, MR);
      for (int q = 0; q < k; ++q) {
        const double g1 = cs[q] * y[q] - sn[q] * y[q + 1];
        const double g2 = sn[q] * y[q] + cs[q] * y[q + 1];
        y[q] = g1;
        y[q + 1] = g2;
      }
      for (int i = 0; i < k + 2; ++i) GS_H(H, i, k, MR) = y[i];
    }
    const double hkk = GS_H(H, k, k, MR), hk1 = GS_H(H, k + 1, k, MR);
    const double mu = sqrt(hkk * hkk + hk1 * hk1);
    if (mu == 0.0) {
      S[3] = 0.0;  // breakdown: stop this column (reference returns)
      S[5] = -2.0;
      S[7] = 0.0;
      return;
    }
    cs[k] = hkk / mu;
    sn[k] = -hk1 / mu;
    GS_H(H, k, k, MR) = cs[k] * hkk - sn[k] * hk1;
    GS_H(H, k + 1, k, MR) = 0.0;
    const double g1 = cs[k] * g[k] - sn[k] * g[k + 1];
    const double g2 = sn[k] * g[k] + cs[k] * g[k + 1];
    g[k] = g1;
    g[k + 1] = g2;
    S[0] = fabs(g[k + 1]);
    S[6] += 1.0;
    if (S[0] <= S[1] && S[0] <= tol_abs) {
      S[4] = 1.0;
      S[7] = 0.0;
    }
  } else if (op == 7) {
    // back substitution for y; coefficients written to y[0..k]
    const int kl = (int)S[5];
    if (S[3] == 0.0 || kl < 0) {
      for (int i = 0; i <= MR; ++i) y[i] = 0.0;
      return;
    }
    if (GS_H(H, kl, kl, MR) == 0.0) {
      for (int i = 0; i <= MR; ++i) y[i] = 0.0;
      S[3] = 0.0;
      return;
    }
    for (int i = kl + 1; i <= MR; ++i) y[i] = 0.0;
    y[kl] = g[kl] / GS_H(H, kl, kl, MR);
    for (int i = kl - 1; i >= 0; --i) {
      y[i] = g[i];
      for (int q = i + 1; q < kl + 1; ++q) y[i] -= GS_H(H, i, q, MR) * y[q];
      y[i] /= GS_H(H, i, i, MR);
    }
    if (S[4] != 0.0) S[3] = 0.0;  // converged: stop outer loop for this column
  }
}
__global__ void k_gmres_scalar(int op, int nrhs, int MR, double* gm, const double* dots, int j, int k,
                               int itr, double tol_abs, double tol_rel) {
  const int c = threadIdx.x;
  if (c >= nrhs) return;
  gmres_scalar_col(op, c, MR, gm, dots[c], j, k, itr, tol_abs, tol_rel);
}
void launch_gmres_scalar(int op, int nrhs, int MR, double* gm, const double* dots, int j, int k, int itr,
                         double tol_abs, double tol_rel, hipStream_t st) {
  k_gmres_scalar<<<1, 64, 0, st>>>(op, nrhs, MR, gm, dots, j, k, itr, tol_abs, tol_rel);
  HIP_LAUNCH_CHECK("gmres_scalar");
}

// dot products of short vectors followed by the GMRES scalar step on them,
// in one workgroup and one launch (launch_dot + launch_gmres_scalar for long
// vectors); the sums are those of k_dot_one
__global__ __launch_bounds__(1024) void k_dot_scalar_one(int n, int nrhs, const double* a, const double* b,
                                                         double* out, int op, int MR, double* gm, int j, int k,
                                                         int itr, double tol_abs, double tol_rel) {
  __shared__ double sh[16][MAXRHS];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double acc[MAXRHS];
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c) acc[c] = 0.0;
  for (int i = t; i < n; i += 1024)
#pragma unroll
    for (int c = 0; c < MAXRHS; ++c)
      if (c < nrhs) acc[c] += a[(int64_t)i * nrhs + c] * b[(int64_t)i * nrhs + c];
#pragma unroll
  for (int c = 0; c < MAXRHS; ++c) {
    if (c >= nrhs) break;
    const double v = wave_sum(acc[c]);
    if (ln == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (t < nrhs) {
    double s = 0.0;
    for (int q = 0; q < 16; ++q) s += sh[q][t];
    out[t] = s;
    gmres_scalar_col(op, t, MR, gm, s, j, k, itr, tol_abs, tol_rel);
  }
}
void launch_dot_scalar(int n, int nrhs, const double* a, const double* b, double* part, int nblk, double* out,
                       int op, int MR, double* gm, int j, int k, int itr, double tol_abs, double tol_rel,
                       hipStream_t st) {
  if ((int64_t)n * nrhs <= DOT_ONE_MAX) {
    k_dot_scalar_one<<<1, 1024, 0, st>>>(n, nrhs, a, b, out, op, MR, gm, j, k, itr, tol_abs, tol_rel);
    HIP_LAUNCH_CHECK("dot_scalar");
    return;
  }
  launch_dot(n, nrhs, a, b, part, nblk, out, st);
  launch_gmres_scalar(op, nrhs, MR, gm, out, j, k, itr, tol_abs, tol_rel, st);
}

// x += V y (columns with active or just-converged state)
__global__ void k_gmres_update(int n, int nrhs, int MR, const double* V, const double* gm, double* x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < nrhs; ++c) {
      const double* st = gm + (int64_t)c * ((MR + 1) * MR + 4 * (MR + 1) + (MR + 2) + 8);
      const double* y = st + (MR + 1) * MR + 3 * (MR + 1);
      double acc = x[(int64_t)i * nrhs + c];
      for (int j = 0; j <= MR; ++j)
        if (y[j] != 0.0) acc += V[((int64_t)j * n + i) * nrhs + c] * y[j];
      x[(int64_t)i * nrhs + c] = acc;
    }
}
void launch_gmres_update(int n, int nrhs, int MR, const double* V, const double* gm, double* x,
                         hipStream_t st) {
  k_gmres_update<<<grid_for(n, 256), 256, 0, st>>>(n, nrhs, MR, V, gm, x);
  HIP_LAUNCH_CHECK("gmres_update");
}

// backward-error ingredients per column: sum|r|, sum|x|, sum|b|
__global__ void k_abs_sums(int n, int nrhs, const double* r, const double* x, const double* b, double* part) {
  __shared__ double sh[4][3 * MAXRHS];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double acc[3 * MAXRHS];
  for (int c = 0; c < 3 * nrhs; ++c) acc[c] = 0.0;
  for (int i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < nrhs; ++c) {
      acc[c] += fabs(r[(int64_t)i * nrhs + c]);
      acc[nrhs + c] += fabs(x[(int64_t)i * nrhs + c]);
      acc[2 * nrhs + c] += fabs(b[(int64_t)i * nrhs + c]);
    }
  for (int c = 0; c < 3 * nrhs; ++c) {
    const double v = wave_sum(acc[c]);
    if (ln == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (t < 3 * nrhs) part[(int64_t)blockIdx.x * 3 * nrhs + t] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
}
void launch_abs_sums(int n, int nrhs, const double* r, const double* x, const double* b, double* part,
                     int nblk, double* out, hipStream_t st) {
  nblk = red_blocks(n, nblk);
  k_abs_sums<<<nblk, 256, 0, st>>>(n, nrhs, r, x, b, part);
  HIP_LAUNCH_CHECK("abs_sums");
  k_red_final<<<1, 256, 0, st>>>(part, nblk, 3 * nrhs, out);
  HIP_LAUNCH_CHECK("abs_final");
}
// the same sums per segment of rows (the systems of a block-diagonal batch):
// out[(3 * k + which) * nrhs + c], one workgroup per segment
__global__ void k_seg_abs_sums(const int64_t* seg, int nrhs, const double* r, const double* x, const double* b,
                               double* out) {
  __shared__ double sh[4][3 * MAXRHS];
  const int k = blockIdx.x, t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double acc[3 * MAXRHS];
  for (int c = 0; c < 3 * MAXRHS; ++c) acc[c] = 0.0;
  for (int64_t i = seg[k] + t; i < seg[k + 1]; i += 256)
    for (int c = 0; c < nrhs; ++c) {
      acc[c] += fabs(r[i * nrhs + c]);
      acc[nrhs + c] += fabs(x[i * nrhs + c]);
      acc[2 * nrhs + c] += fabs(b[i * nrhs + c]);
    }
  for (int c = 0; c < 3 * nrhs; ++c) {
    const double v = wave_sum(acc[c]);
    if (ln == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (t < 3 * nrhs) out[(int64_t)k * 3 * nrhs + t] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
}
void launch_seg_abs_sums(int nseg, const int64_t* seg, int nrhs, const double* r, const double* x, const double* b,
                         double* out, hipStream_t st) {
  if (nseg <= 0) return;
  k_seg_abs_sums<<<nseg, 256, 0, st>>>(seg, nrhs, r, x, b, out);
  HIP_LAUNCH_CHECK("seg_abs_sums");
}
// max row sum of |K0| per segment (norm_sparse_int of each system)
__global__ void k_seg_rowabs_max(const int64_t* seg, const int* rp, const double* Kv, double* out) {
  __shared__ double sh[4];
  const int k = blockIdx.x, t = threadIdx.x, wv = t >> 6, ln = t & 63;
  double m = 0.0;
  for (int64_t i = seg[k] + t; i < seg[k + 1]; i += 256) {
    double s = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) s += fabs(Kv[q]);
    m = fmax(m, s);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if (ln == 0) sh[wv] = m;
  __syncthreads();
  if (t == 0) out[k] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}
void launch_seg_rowabs_max(int nseg, const int64_t* seg, const int* rp, const double* Kv, double* out,
                           hipStream_t st) {
  if (nseg <= 0) return;
  k_seg_rowabs_max<<<nseg, 256, 0, st>>>(seg, rp, Kv, out);
  HIP_LAUNCH_CHECK("seg_rowabs_max");
}

}  // namespace lblg

namespace lblg {
// --------------------------------------------------------------------------
// value-dependent preparation: perturbation flags (find_perturbation), the
// unperturbed refinement matrix K0 (set_diags(0)) and its infinity norm
// (norm_sparse_int, src/common/Norm.cpp:115-185)
// --------------------------------------------------------------------------
// bit 0: |K_jj| < reg (the reference's rule; those diagonals are zero in K0),
// bit 1: |Kf_jj| < reg_factor (replaced by reg_factor in the factored matrix;
// Kf = K, or the diagonally scaled S K S of the BF16 factor)
__global__ void k_pert_flags(int nnz, const unsigned char* isdiag, const double* Kx, const double* Kf, double reg,
                             double reg_factor, unsigned char* pert) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
    const double a = fabs(Kx[q]), af = fabs(Kf[q]);
    pert[q] = isdiag[q] ? (unsigned char)((a < reg ? 1 : 0) | (af < reg_factor ? 2 : 0)) : (unsigned char)0;
  }
}
void launch_pert_flags(int nnz, const unsigned char* isdiag, const double* Kx, const double* Kf, double reg,
                       double reg_factor, int regularization, unsigned char* pert, hipStream_t st) {
  if (nnz <= 0) return;
  const bool on = regularization == 1;
  k_pert_flags<<<grid_for(nnz, 256), 256, 0, st>>>(nnz, isdiag, Kx, Kf, on ? reg : -1.0, on ? reg_factor : -1.0,
                                                   pert);
  HIP_LAUNCH_CHECK("pert_flags");
}
// --------------------------------------------------------------------------
// symmetric diagonal scaling of the BF16 factor: s_i = 1 / sqrt(max_j |K_ij|)
// (one max-norm equilibration step; rows without entries keep 1), the scaled
// entries Kf = s_row K s_col of the lower CSC, and s in the pivoted order of
// the factor (s_ord[k] = s[perm_f[k]])
// --------------------------------------------------------------------------
__global__ void k_row_scale(int n, const int* rp, const int64_t* src, const double* Kx, double* s) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double m = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) m = fmax(m, fabs(Kx[src[q]]));
    s[i] = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
  }
}
__global__ void k_scale_entries(int nnz, const int* row, const int* col, const double* Kx, const double* s,
                                double* Kf) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x)
    Kf[q] = s[row[q]] * Kx[q] * s[col[q]];
}
__global__ void k_gather_scale(int n, const int* perm_f, const double* s, double* s_ord) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) s_ord[k] = s[perm_f[k]];
}
void launch_diag_scaling(int n, int nnz, const int* rp, const int64_t* src, const int* row, const int* col,
                         const double* Kx, double* s, double* Kf, hipStream_t st) {
  k_row_scale<<<grid_for(n, 256), 256, 0, st>>>(n, rp, src, Kx, s);
  HIP_LAUNCH_CHECK("row_scale");
  if (nnz <= 0) return;
  k_scale_entries<<<grid_for(nnz, 256), 256, 0, st>>>(nnz, row, col, Kx, s, Kf);
  HIP_LAUNCH_CHECK("scale_entries");
}
void launch_scale_entries(int nnz, const int* row, const int* col, const double* Kx, const double* s, double* Kf,
                          hipStream_t st) {
  if (nnz <= 0) return;
  k_scale_entries<<<grid_for(nnz, 256), 256, 0, st>>>(nnz, row, col, Kx, s, Kf);
  HIP_LAUNCH_CHECK("scale_entries");
}
void launch_gather_scale(int n, const int* perm_f, const double* s, double* s_ord, hipStream_t st) {
  k_gather_scale<<<grid_for(n, 256), 256, 0, st>>>(n, perm_f, s, s_ord);
  HIP_LAUNCH_CHECK("gather_scale");
}
// dst[i nrhs + c] = s[i] src[i nrhs + c] (row i of a row-interleaved block), converted to TO
template <typename TO, typename TI>
__global__ void k_row_scale_copy(TO* dst, const TI* src, const double* s, int n, int nrhs) {
  const int64_t nn = (int64_t)n * nrhs;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nn; q += (int64_t)gridDim.x * blockDim.x)
    dst[q] = (TO)((double)src[q] * s[q / nrhs]);
}
template <typename TO, typename TI>
void launch_row_scale_copy(TO* dst, const TI* src, const double* s, int n, int nrhs, hipStream_t st) {
  k_row_scale_copy<TO, TI><<<grid_for((int64_t)n * nrhs, 256), 256, 0, st>>>(dst, src, s, n, nrhs);
  HIP_LAUNCH_CHECK("row_scale_copy");
}
template void launch_row_scale_copy<float, double>(float*, const double*, const double*, int, int, hipStream_t);
template void launch_row_scale_copy<double, float>(double*, const float*, const double*, int, int, hipStream_t);
template void launch_row_scale_copy<double, double>(double*, const double*, const double*, int, int, hipStream_t);
__global__ void k_k0_values(int64_t nnzf, const int64_t* src, const double* Kx, const unsigned char* pert,
                            double* Kv) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnzf; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = src[e];
    Kv[e] = (pert[q] & 1) ? 0.0 : Kx[q];
  }
}
__global__ void k_rowabs_max(int n, const int* rp, const double* Kv, double* part) {
  __shared__ double sh[4];
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) s += fabs(Kv[q]);
    m = fmax(m, s);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}
// one wave: lanes over the partial maxima, then a wave maximum
__global__ void k_max_final(const double* part, int nblk, double* out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) m = fmax(m, part[b]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if (threadIdx.x == 0) out[0] = m;
}
void launch_k0(int n, int64_t nnzf, const int* rp, const int64_t* src, const double* Kx,
               const unsigned char* pert, double* Kv, double* part, int nblk, double* norm_out, hipStream_t st) {
  k_k0_values<<<grid_for(nnzf, 256), 256, 0, st>>>(nnzf, src, Kx, pert, Kv);
  HIP_LAUNCH_CHECK("k0_values");
  k_rowabs_max<<<nblk, 256, 0, st>>>(n, rp, Kv, part);
  HIP_LAUNCH_CHECK("rowabs");
  k_max_final<<<1, 64, 0, st>>>(part, nblk, norm_out);
  HIP_LAUNCH_CHECK("max_final");
}
// --------------------------------------------------------------------------
// instantiations for the two factor value types
// --------------------------------------------------------------------------
#define LBLG_INSTANTIATE(T)                                                                                      \
  template void launch_fill<T>(T*, int64_t, double, hipStream_t);                                              \
  template void launch_assemble<T>(const int64_t*, const double*, const unsigned char*, int, double, T*,       \
                                   const int*, const unsigned char*, hipStream_t);                             \
  template void launch_reset_supernodes<T>(const int*, int, const SnArgs&, T*, T*, int*, hipStream_t);          \
  template void launch_bk_small<T>(const int*, int, int, const SnArgs&, T*, T*, int*, int*, hipStream_t);      \
  template void launch_bk_panel<T>(const int*, int, int, const SnArgs&, T*, T*, int*, T*, const int64_t*, int*, \
                                   int*, int*, int*, bool, int, hipStream_t);                                       \
  template void launch_trailing<T>(const int*, int, int, int, const SnArgs&, T*, const T*, const int64_t*,      \
                                   const int*, const int*, hipStream_t);                                       \
  template void launch_offdiag<T>(const int*, int, int, const int*, const int*, int, const SnArgs&, T*,      \
                                  const T*,                                                                    \
                                  const int*, hipStream_t);                                                                \
  template void launch_schur<T>(const int*, int, const int64_t*, const int*, int64_t, const SnArgs&, T*,     \
                                const T*, hipStream_t);                                                                  \
  template void launch_fwd_tri<T>(const int*, int, int, const int*, const int*, const SnArgs&, const T*,       \
                                  const T*, T*, int, hipStream_t);                                             \
  template void launch_fwd_gemv<T>(const int*, const int*, const int*, int, const int*, const int*,            \
                                   const SnArgs&,                                                              \
                                   const T*, const T*, T*, int, hipStream_t);                                  \
  template void launch_dsolve<T>(int, const T*, const T*, T*, int, hipStream_t);                               \
  template void launch_bwd_gemv<T>(const int*, int, const int*, const int*, int, int, const int*, const int*, \
                                   const SnArgs&,                                                              \
                                   const T*, const T*, T*, int, hipStream_t);                                  \
  template void launch_bwd_tri<T>(const int*, int, int, const int*, const int*, const SnArgs&, const T*,       \
                                  const T*, const T*, T*, int, hipStream_t);
LBLG_INSTANTIATE(double)
LBLG_INSTANTIATE(float)
#undef LBLG_INSTANTIATE
template void launch_convert<float, double>(float*, const double*, int64_t, hipStream_t);
template void launch_convert<double, float>(double*, const float*, int64_t, hipStream_t);

}  // namespace lblg


