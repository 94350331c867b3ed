// Blocked L and L^T sweeps of one wide supernode in a single cooperative
// launch (the H2LeveledBlockedLsolve / H2LeveledBlockedLTsolve work of
// src/triangularSolve/Triangular_BCSC.cpp:378-446, 534-600 for that
// supernode).  Processing a wide supernode piece by piece (128 columns per
// piece) costs two launches per piece on a dependent chain; here the
// supernode's rows are split into units of 128 rows, each workgroup owns
// units (u = g, g + G, ...), and the pieces are chained through an exchange
// buffer in global memory:
//
//   forward  (rows below each piece updated right-looking):
//     the owner of diagonal unit p, once the pieces < p are applied to it,
//     solves y_p = L_pp^{-1} x_p (triangle staged in LDS, one wave
//     substitutes) and publishes y_p; every workgroup that owns units below
//     reads y_p and applies x_u -= L(u, p) y_p.  Units of the off-diagonal
//     rows accumulate -L_off y and add it to the ancestors' entries at the end.
//   backward (left-looking):
//     the owner of diagonal unit u gathers a = L_off(:, u)^T x_off and, as
//     the x_v of units v > u are published, a += L(v, u)^T x_v; then
//     x_u = L_uu^{-T} (z_u - a) and publishes x_u.
//
// The exchange is flag-in-word: every published value travels as 32-bit
// halves in 64-bit words whose upper half is the launch's epoch, stored and
// polled with agent-scope atomics, so a value is its own flag and no cache
// maintenance is needed.  Every wait is bounded; on time-out status bit 16
// is set and the workgroup goes on, so the grid always drains.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "engine/device_util.hpp"
#include "engine/kernels.hpp"

namespace lblg {

#define CU 128                // rows per unit = columns per piece
#define CMAXRHS 8
#define CUPW 4                // units per workgroup (at most)
#define CTRI (CU * (CU + 1) / 2)
#define CSPIN (1L << 24)
#define CSTEP 16              // substitution steps per batch of LDS reads

// exchange words: value idx of a published block occupies WPV words
template <typename T>
struct XWords;
template <>
struct XWords<double> {
  static constexpr int WPV = 2;
  static __device__ inline void put(uint64_t* b, int64_t idx, double v, uint32_t ep) {
    const uint64_t x = (uint64_t)__double_as_longlong(v);
    __hip_atomic_store(b + 2 * idx, ((uint64_t)ep << 32) | (x & 0xffffffffull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(b + 2 * idx + 1, ((uint64_t)ep << 32) | (x >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};
template <>
struct XWords<float> {
  static constexpr int WPV = 1;
  static __device__ inline void put(uint64_t* b, int64_t idx, float v, uint32_t ep) {
    __hip_atomic_store(b + idx, ((uint64_t)ep << 32) | (uint64_t)__float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
};
// the low half of word b[k] once it carries epoch ep (bounded wait)
__device__ inline uint32_t xword(const uint64_t* b, int64_t k, uint32_t ep, int* status, int& abort) {
  long spins = 0;
  while (!abort) {
    const uint64_t x = __hip_atomic_load(b + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(x >> 32) == ep) return (uint32_t)x;
    if (++spins > CSPIN) {
      abort = 1;
      atomicOr(status, 16);
    }
  }
  return 0u;
}
// one thread waits (with sleeps) for the last word of a published block, so
// the workgroup's threads do not all poll; the words are then read with
// their own epoch check
__device__ inline void xwait_block(const uint64_t* b, int64_t last_word, uint32_t ep, int* status, int& abort) {
  if (threadIdx.x == 0 && !abort) {
    long spins = 0;
    while ((uint32_t)(__hip_atomic_load(b + last_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) != ep) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > CSPIN) {
        abort = 1;
        atomicOr(status, 16);
        break;
      }
    }
  }
  __syncthreads();
}
// Read the published values idx0 .. idx0 + n - 1 into dst[0 .. n) and zero
// dst[n .. lim), lim <= 256 * NV (thread t handles k = t, t + 256, ...).
// Every word of a thread is loaded before any is checked, so the round trips
// overlap; a word that does not carry the epoch yet is polled again
// (bounded, as in xword).
template <typename T>
__device__ inline T xassemble(const uint64_t* x);
template <>
__device__ inline double xassemble<double>(const uint64_t* x) {
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)x[1] << 32) | (uint32_t)x[0]));
}
template <>
__device__ inline float xassemble<float>(const uint64_t* x) {
  return __uint_as_float((uint32_t)x[0]);
}
template <typename T, int NV>
__device__ inline void xget_block(const uint64_t* b, int64_t idx0, int n, int lim, T* dst, uint32_t ep, int* status,
                                  int& abort) {
  constexpr int W = XWords<T>::WPV;
  uint64_t x[NV][W];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int k = threadIdx.x + 256 * v;
#pragma unroll
    for (int h = 0; h < W; ++h)
      x[v][h] = k < n ? __hip_atomic_load(b + (idx0 + k) * W + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int k = threadIdx.x + 256 * v;
    if (k >= lim) continue;
    if (k < n) {
#pragma unroll
      for (int h = 0; h < W; ++h)
        if ((uint32_t)(x[v][h] >> 32) != ep) x[v][h] = xword(b, (idx0 + k) * W + h, ep, status, abort);
    }
    dst[k] = k < n ? xassemble<T>(x[v]) : (T)0;
  }
}

// stage the strict lower triangle of the pw x pw diagonal block at (d0, d0)
// into LDS, packed by columns (colpack) or by rows; all loads of a thread
// are issued before the first LDS store
// value of lane `lane` (wave-uniform) of a VGPR, through v_readlane
__device__ inline double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <typename T>
__device__ inline void stage_triangle(T* tri, const T* L, int64_t r, int d0, int pw, bool colpack) {
  constexpr int PER = CU * CU / 256 / 2;  // two batches of loads in flight
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    T v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = threadIdx.x + 256 * (k + PER * b), i = idx % CU, j = idx / CU;
      v[k] = (i > j && i < pw && j < pw) ? L[(int64_t)(d0 + i) + (int64_t)(d0 + j) * r] : (T)0;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = threadIdx.x + 256 * (k + PER * b), i = idx % CU, j = idx / CU;
      if (i > j && i < pw && j < pw) {
        if (colpack) tri[j * CU - j * (j + 1) / 2 + (i - j - 1)] = v[k];
        else tri[i * (i - 1) / 2 + j] = v[k];
      }
    }
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_fwd_chain(int s, int G, const int* super, const int64_t* pi,
                                                   const int64_t* px, const int* rows, const T* valL, T* X, T* Y,
                                                   uint64_t* xb, uint32_t epoch, int* status) {
  constexpr int nrhs = NR;
  __shared__ T xs[CUPW][CU * NR];  // diagonal units: x; off-diagonal units: accumulated -L y
  __shared__ T ys[CU * NR];        // y_p, zero beyond the piece's width
  __shared__ T part[2][CU * NR];
  __shared__ T tri[CTRI];               // packed strict lower triangle, column j at j * CU - j (j + 1) / 2
  const int g = blockIdx.x, t = threadIdx.x;
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  const int P = (w + CU - 1) / CU, V = (int)((r - w + CU - 1) / CU), U = P + V;
  const T* L = valL + px[s];
  int abort = 0;
  auto urow0 = [&](int u) -> int64_t { return u < P ? (int64_t)CU * u : w + (int64_t)CU * (u - P); };
  auto ulen = [&](int u) -> int {
    return u < P ? min(CU, w - CU * u) : (int)min((int64_t)CU, r - w - (int64_t)CU * (u - P));
  };
  for (int k = 0; k < CUPW; ++k) {
    const int u = g + k * G;
    if (u >= U) break;
    const int len = ulen(u);
    for (int idx = t; idx < CU * nrhs; idx += 256) {
      const int i = idx / nrhs;
      xs[k][idx] = (u < P && i < len) ? X[(int64_t)(c0 + CU * u) * nrhs + idx] : (T)0;
    }
  }
  int last_own = -1;  // largest unit this workgroup owns
  for (int k = 0; k < CUPW; ++k)
    if (g + k * G < U) last_own = g + k * G;
  // the triangle of the first diagonal unit is staged up front (it depends on nothing)
  int staged = -1;
  if (g < P) {
    stage_triangle(tri, L, r, CU * g, min(CU, w - CU * g), true);
    staged = g;
  }
  __syncthreads();
  const int i = t & (CU - 1), h = t >> 7;
  for (int p = 0; p < P; ++p) {
    const int pw = min(CU, w - CU * p);
    if (p % G == g) {
      // all pieces < p are applied to unit p: solve its triangle
      const int k = p / G;
      if (staged != p) {
        stage_triangle(tri, L, r, CU * p, pw, true);
        staged = p;
        __syncthreads();
      }
      // wave wv substitutes right-hand sides wv and wv + 4 (lane ln holds rows
      // ln and ln + 64), so up to four SIMDs run the sequential steps at once
      constexpr int NW = NR < 4 ? NR : 4, RPW = (NR + 3) / 4;
      const int wv = t >> 6, ln = t & 63;
      if (wv < NW) {
        T x0[RPW], x1[RPW];
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
          const int c = wv + 4 * q;
          x0[q] = (c < NR && ln < pw) ? xs[k][ln * nrhs + c] : (T)0;
          x1[q] = (c < NR && ln + 64 < pw) ? xs[k][(ln + 64) * nrhs + c] : (T)0;
        }
        // column j of the packed triangle holds rows j + 1 .. CU - 1 at
        // tri + cofs(j) + row.  The steps go in chunks of CSTEP: a lane reads
        // its entries of the chunk's columns together (addresses clamped into
        // the columns, masked afterwards), then substitutes from registers,
        // so an LDS round trip is paid per chunk rather than per step.
        auto cofs = [](int j) { return j * CU - j * (j + 1) / 2 - j - 1; };
        const int n0 = min(64, pw - 1);
        for (int jc = 0; jc < n0; jc += CSTEP) {  // x_j in lane j of x0
          T e0[CSTEP], e1[CSTEP];
#pragma unroll
          for (int k = 0; k < CSTEP; ++k) {
            const int j = min(jc + k, n0 - 1);
            e0[k] = tri[cofs(j) + max(ln, j + 1)];
            e1[k] = tri[cofs(j) + ln + 64];
          }
#pragma unroll
          for (int k = 0; k < CSTEP; ++k) {
            const int j = jc + k;
            if (j >= n0) break;
            const T l0 = (ln > j && ln < pw) ? e0[k] : (T)0;
            const T l1 = (ln + 64 < pw) ? e1[k] : (T)0;
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
              const T xj = lane_value(x0[q], j);
              x0[q] -= l0 * xj;
              x1[q] -= l1 * xj;
            }
          }
        }
        for (int jc = 64; jc < pw - 1; jc += CSTEP) {  // x_j in lane j - 64 of x1
          T e1[CSTEP];
#pragma unroll
          for (int k = 0; k < CSTEP; ++k) {
            const int j = min(jc + k, pw - 2);
            e1[k] = tri[cofs(j) + max(ln + 64, j + 1)];
          }
#pragma unroll
          for (int k = 0; k < CSTEP; ++k) {
            const int j = jc + k;
            if (j >= pw - 1) break;
            const T l1 = (ln + 64 > j && ln + 64 < pw) ? e1[k] : (T)0;
#pragma unroll
            for (int q = 0; q < RPW; ++q) x1[q] -= l1 * lane_value(x1[q], j - 64);
          }
        }
        // publish y_p first (the chain waits on it), then the copy for the D solve
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
          const int c = wv + 4 * q;
          if (c >= NR) continue;
          if (ln < pw) XWords<T>::put(xb, ((int64_t)CU * p + ln) * nrhs + c, x0[q], epoch);
          if (ln + 64 < pw) XWords<T>::put(xb, ((int64_t)CU * p + ln + 64) * nrhs + c, x1[q], epoch);
        }
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
          const int c = wv + 4 * q;
          if (c >= NR) continue;
          if (ln < pw) Y[(int64_t)(c0 + CU * p + ln) * nrhs + c] = x0[q];
          if (ln + 64 < pw) Y[(int64_t)(c0 + CU * p + ln + 64) * nrhs + c] = x1[q];
        }
      }
      __syncthreads();
    }
    if (last_own <= p) continue;  // nothing of this workgroup lies below piece p
    // the first unit below p: its block of L is loaded before y_p is awaited
    int k = 0;
    while (k < CUPW && g + k * G <= p) ++k;
    const int jb = 64 * h, je = min(pw, jb + 64);
    T l[64];
#define CHAIN_LOAD_BLOCK(u)                                                          \
  {                                                                                  \
    const bool ok_ = i < ulen(u);                                                    \
    const T* Lc_ = L + urow0(u) + i + (int64_t)(CU * p) * r;                         \
    _Pragma("unroll") for (int e = 0; e < 64; ++e) l[e] =                            \
        (ok_ && jb + e < je) ? Lc_[(int64_t)(jb + e) * r] : (T)0;                    \
  }
    CHAIN_LOAD_BLOCK(g + k * G)
    xwait_block(xb, (((int64_t)CU * p + pw - 1) * nrhs + nrhs - 1) * XWords<T>::WPV + XWords<T>::WPV - 1, epoch,
                status, abort);
    xget_block<T, (CU * NR + 255) / 256>(xb, (int64_t)CU * p * nrhs, pw * nrhs, CU * nrhs, ys, epoch, status, abort);
    __syncthreads();
    for (; k < CUPW; ++k) {
      const int u = g + k * G;
      if (u >= U) break;
      if (k > 0 && g + (k - 1) * G > p) CHAIN_LOAD_BLOCK(u)  // later units of this workgroup
      T a[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) a[c] = 0;
#pragma unroll
      for (int e = 0; e < 64; ++e)  // l[e] and ys are zero past the piece
#pragma unroll
        for (int c = 0; c < NR; ++c) a[c] += l[e] * ys[(jb + e) * nrhs + c];
#pragma unroll
      for (int c = 0; c < NR; ++c) part[h][i * nrhs + c] = a[c];
      __syncthreads();
      const int len = ulen(u);
      for (int idx = t; idx < len * nrhs; idx += 256) xs[k][idx] -= part[0][idx] + part[1][idx];
      __syncthreads();
    }
  }
#undef CHAIN_LOAD_BLOCK
  // off-diagonal units: add the accumulated -L_off y into the ancestors' entries
  for (int k = 0; k < CUPW; ++k) {
    const int u = g + k * G;
    if (u >= U) break;
    if (u < P) continue;
    const int len = ulen(u);
    const int64_t q0 = urow0(u);
    for (int idx = t; idx < len * nrhs; idx += 256) {
      const int ii = idx / nrhs, c = idx % nrhs;
      const int lab = rows[pi[s] + q0 + ii];
      atomicAdd(X + (int64_t)lab * nrhs + c, xs[k][idx]);
    }
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_bwd_chain(int s, int G, const int* super, const int64_t* pi,
                                                   const int64_t* px, const int* rows, const T* valL, const T* Z,
                                                   const T* offacc, T* Xf, uint64_t* xb, uint32_t epoch,
                                                   int* status) {
  constexpr int nrhs = NR;
  __shared__ T stage[64 * (CU + 1)];  // 64 rows x 128 columns of L (pitch CU + 1)
  __shared__ T tri[CTRI];             // row-packed strict lower triangle of the unit being solved
  __shared__ T xv[64 * NR];
  __shared__ T acc[CU * NR];
  __shared__ T part[2][CU * NR];
  const int g = blockIdx.x, t = threadIdx.x;
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  const int P = (w + CU - 1) / CU;
  const T* L = valL + px[s];
  int abort = 0;
  constexpr int PER = 64 * CU / 256;
  T lv[2][PER];  // a 128-row block of L as two 64-row halves, loaded ahead of use
  // issue the loads of rows [q0, q0 + nq) (nq <= 128) x columns of u into lv
  auto load_block = [&](int64_t q0, int nq, int cu0, int pw) {
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = t + 256 * k, qq = idx & 63, j = idx >> 6;
        const int row = 64 * hh + qq;
        lv[hh][k] = (row < nq && j < pw) ? L[q0 + row + (int64_t)(cu0 + j) * r] : (T)0;
      }
  };
  // acc += (loaded block)^T x_v: the block of L in rows of diagonal unit v,
  // x_v from the exchange buffer
  auto gather_loaded = [&](int nq, int pw, int v) {
    for (int hh = 0; hh < 2; ++hh) {
      if (64 * hh >= nq) break;
      const int nh = min(64, nq - 64 * hh);
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = t + 256 * k, qq = idx & 63, j = idx >> 6;
        stage[qq * (CU + 1) + j] = lv[hh][k];
      }
      xget_block<T, (64 * NR + 255) / 256>(xb, ((int64_t)CU * v + 64 * hh) * nrhs, nh * nrhs, 64 * nrhs, xv, epoch,
                                           status, abort);
      __syncthreads();
      const int j = t & (CU - 1), half = t >> 7;
      T a[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) a[c] = 0;
#pragma unroll 8
      for (int qq = 32 * half; qq < 32 * half + 32; ++qq) {  // staged entries past the block are zero
        const T l = stage[qq * (CU + 1) + j];
#pragma unroll
        for (int c = 0; c < NR; ++c) a[c] += l * xv[qq * nrhs + c];
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) part[half][j * nrhs + c] = a[c];
      __syncthreads();
      for (int idx = t; idx < pw * nrhs; idx += 256) acc[idx] += part[0][idx] + part[1][idx];
      __syncthreads();
    }
  };
  int nown = 0;
  for (int k = 0; k < CUPW; ++k)
    if (g + k * G < P) nown = k + 1;
  // the first unit solved here (the largest owned) has its triangle staged up front
  int staged = -1;
  if (nown > 0) {
    staged = g + (nown - 1) * G;
    stage_triangle(tri, L, r, CU * staged, min(CU, w - CU * staged), false);
  }
  for (int k = nown - 1; k >= 0; --k) {
    const int u = g + k * G;
    const int pw = min(CU, w - CU * u);
    // the rows below the diagonal block were gathered into offacc by k_bwd_chain_off
    for (int idx = t; idx < CU * nrhs; idx += 256)
      acc[idx] = idx < pw * nrhs ? offacc[(int64_t)(c0 + CU * u) * nrhs + idx] : (T)0;
    __syncthreads();
    // diagonal units below u, as they are solved (each block loaded before its x is awaited)
    for (int v = P - 1; v > u; --v) {
      const int pv = min(CU, w - CU * v);
      load_block((int64_t)CU * v, pv, CU * u, pw);
      xwait_block(xb, (((int64_t)CU * v + pv - 1) * nrhs + nrhs - 1) * XWords<T>::WPV + XWords<T>::WPV - 1, epoch,
                  status, abort);
      gather_loaded(pv, pw, v);
    }
    // x_u = L_uu^{-T} (z_u - acc): row-packed strict lower triangle, row i at i (i - 1) / 2
    if (staged != u) {
      stage_triangle(tri, L, r, CU * u, pw, false);
      staged = u;
    }
    __syncthreads();
    // wave wv substitutes right-hand sides wv and wv + 4 (lane ln holds rows
    // ln and ln + 64)
    constexpr int NW = NR < 4 ? NR : 4, RPW = (NR + 3) / 4;
    const int wv = t >> 6, ln = t & 63;
    if (wv < NW) {
      const int64_t base = (int64_t)(c0 + CU * u) * nrhs;
      T x0[RPW], x1[RPW];
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        const int c = wv + 4 * q;
        x0[q] = (c < NR && ln < pw) ? Z[base + (int64_t)ln * nrhs + c] - acc[ln * nrhs + c] : (T)0;
        x1[q] = (c < NR && ln + 64 < pw) ? Z[base + (int64_t)(ln + 64) * nrhs + c] - acc[(ln + 64) * nrhs + c]
                                         : (T)0;
      }
      // row j of the packed triangle holds L(j, 0 .. j - 1) at tri + j (j - 1) / 2;
      // the steps go in chunks of CSTEP rows read together (addresses clamped
      // into the rows, masked afterwards), then substituted from registers
      for (int jc = pw - 1; jc >= 64; jc -= CSTEP) {  // x_j in lane j - 64 of x1
        T e0[CSTEP], e1[CSTEP];
#pragma unroll
        for (int k = 0; k < CSTEP; ++k) {
          const int j = max(jc - k, 64);
          e0[k] = tri[j * (j - 1) / 2 + ln];
          e1[k] = tri[j * (j - 1) / 2 + min(ln + 64, j - 1)];
        }
#pragma unroll
        for (int k = 0; k < CSTEP; ++k) {
          const int j = jc - k;
          if (j < 64) break;
          const T l0 = e0[k];
          const T l1 = (ln + 64 < j) ? e1[k] : (T)0;
#pragma unroll
          for (int q = 0; q < RPW; ++q) {
            const T xj = lane_value(x1[q], j - 64);
            x0[q] -= l0 * xj;
            x1[q] -= l1 * xj;
          }
        }
      }
      for (int jc = min(pw, 64) - 1; jc > 0; jc -= CSTEP) {  // x_j in lane j of x0
        T e0[CSTEP];
#pragma unroll
        for (int k = 0; k < CSTEP; ++k) {
          const int j = max(jc - k, 1);
          e0[k] = tri[j * (j - 1) / 2 + min(ln, j - 1)];
        }
#pragma unroll
        for (int k = 0; k < CSTEP; ++k) {
          const int j = jc - k;
          if (j <= 0) break;
          const T l0 = (ln < j) ? e0[k] : (T)0;
#pragma unroll
          for (int q = 0; q < RPW; ++q) x0[q] -= l0 * lane_value(x0[q], j);
        }
      }
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        const int c = wv + 4 * q;
        if (c >= NR) continue;
        if (ln < pw) XWords<T>::put(xb, ((int64_t)CU * u + ln) * nrhs + c, x0[q], epoch);
        if (ln + 64 < pw) XWords<T>::put(xb, ((int64_t)CU * u + ln + 64) * nrhs + c, x1[q], epoch);
      }
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        const int c = wv + 4 * q;
        if (c >= NR) continue;
        if (ln < pw) Xf[base + (int64_t)ln * nrhs + c] = x0[q];
        if (ln + 64 < pw) Xf[base + (int64_t)(ln + 64) * nrhs + c] = x1[q];
      }
    }
    __syncthreads();
  }
}

// Backward sweep, rows below the diagonal block of a chain supernode:
// offacc(cols of piece u) += L(q, cols of u)^T Xf(label q) for the 64-row
// block rb of those rows; one workgroup per (piece, row block), atomics into
// offacc (zeroed with the sweep accumulators)
template <typename T, int NR>
__global__ __launch_bounds__(256) void k_bwd_chain_off(int s, int nrb, const int* super, const int64_t* pi,
                                                       const int64_t* px, const int* rows, const T* valL,
                                                       const T* Xf, T* offacc) {
  __shared__ T stage[64 * (CU + 1)];
  __shared__ T xv[64 * NR];
  __shared__ T part[2][CU * NR];
  const int u = blockIdx.x / nrb, rb = blockIdx.x % nrb, t = threadIdx.x;
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  const int pw = min(CU, w - CU * u);
  const int64_t q0 = w + 64 * (int64_t)rb;
  const int nh = (int)min((int64_t)64, r - q0);
  const T* L = valL + px[s];
  constexpr int PER = 64 * CU / 256;
  T v[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int idx = t + 256 * k, qq = idx & 63, j = idx >> 6;
    v[k] = (qq < nh && j < pw) ? L[q0 + qq + (int64_t)(CU * u + j) * r] : (T)0;
  }
  for (int idx = t; idx < 64 * NR; idx += 256) {
    const int qq = idx / NR, c = idx % NR;
    xv[idx] = qq < nh ? Xf[(int64_t)rows[pi[s] + q0 + qq] * NR + c] : (T)0;
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int idx = t + 256 * k, qq = idx & 63, j = idx >> 6;
    stage[qq * (CU + 1) + j] = v[k];
  }
  __syncthreads();
  const int j = t & (CU - 1), half = t >> 7;
  T a[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) a[c] = 0;
#pragma unroll 8
  for (int qq = 32 * half; qq < 32 * half + 32; ++qq) {
    const T l = stage[qq * (CU + 1) + j];
#pragma unroll
    for (int c = 0; c < NR; ++c) a[c] += l * xv[qq * NR + c];
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) part[half][j * NR + c] = a[c];
  __syncthreads();
  for (int idx = t; idx < pw * NR; idx += 256)
    atomicAdd(offacc + (int64_t)(c0 + CU * u) * NR + idx, part[0][idx] + part[1][idx]);
}

// Workgroups for a chain: the forward sweep needs one unit per workgroup up
// to CUPW units each (units = diagonal pieces + 128-row blocks below the
// diagonal block), the backward sweep the diagonal units only.
int chain_groups(int w, int64_t r, bool forward) {
  const int P = (w + CU - 1) / CU;
  const int U = forward ? P + (int)((r - w + CU - 1) / CU) : P;
  const int G = U < 256 ? U : 256;
  return (U + G - 1) / G <= CUPW ? G : -1;
}

template <typename T, int NR>
static void fwd_chain_nr(int s, int G, const SnArgs& a, const T* valL, T* X, T* Y, uint64_t* xb, uint32_t epoch,
                         int* status, hipStream_t st) {
  void* args[] = {(void*)&s, (void*)&G, (void*)&a.super, (void*)&a.pi, (void*)&a.px, (void*)&a.rows,
                  (void*)&valL, (void*)&X, (void*)&Y, (void*)&xb, (void*)&epoch, (void*)&status};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_fwd_chain<T, NR>, dim3(G), dim3(256), args, 0, st);
  if (e != hipSuccess) throw std::runtime_error(std::string("fwd_chain: ") + hipGetErrorString(e));
}
template <typename T, int NR>
static void bwd_chain_nr(int s, int G, int w, int64_t r, const SnArgs& a, const T* valL, const T* Z, T* offacc,
                         T* Xf, uint64_t* xb, uint32_t epoch, int* status, hipStream_t st) {
  if (r > w) {
    const int P = (w + CU - 1) / CU, nrb = (int)((r - w + 63) / 64);
    k_bwd_chain_off<T, NR><<<P * nrb, 256, 0, st>>>(s, nrb, a.super, a.pi, a.px, a.rows, valL, Xf, offacc);
    HIP_LAUNCH_CHECK("bwd_chain_off");
  }
  const T* oa = offacc;
  void* args[] = {(void*)&s, (void*)&G, (void*)&a.super, (void*)&a.pi, (void*)&a.px, (void*)&a.rows,
                  (void*)&valL, (void*)&Z, (void*)&oa, (void*)&Xf, (void*)&xb, (void*)&epoch, (void*)&status};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_bwd_chain<T, NR>, dim3(G), dim3(256), args, 0, st);
  if (e != hipSuccess) throw std::runtime_error(std::string("bwd_chain: ") + hipGetErrorString(e));
}

// the kernels are specialised on the exact number of right-hand sides
#define CHAIN_DISPATCH(F, ...)                                          \
  switch (nrhs) {                                                       \
    case 1: F<T, 1>(__VA_ARGS__); break;                                \
    case 2: F<T, 2>(__VA_ARGS__); break;                                \
    case 3: F<T, 3>(__VA_ARGS__); break;                                \
    case 4: F<T, 4>(__VA_ARGS__); break;                                \
    case 5: F<T, 5>(__VA_ARGS__); break;                                \
    case 6: F<T, 6>(__VA_ARGS__); break;                                \
    case 7: F<T, 7>(__VA_ARGS__); break;                                \
    case 8: F<T, 8>(__VA_ARGS__); break;                                \
    default: throw std::runtime_error("chain sweep: 1 to 8 right-hand sides"); \
  }

template <typename T>
void launch_fwd_chain(int s, int w, int64_t r, const SnArgs& a, const T* valL, T* X, T* Y, int nrhs,
                      uint64_t* xb, uint32_t epoch, int* status, hipStream_t st) {
  const int G = chain_groups(w, r, true);
  if (G < 1) throw std::runtime_error("fwd_chain: supernode too large for a chain launch");
  CHAIN_DISPATCH(fwd_chain_nr, s, G, a, valL, X, Y, xb, epoch, status, st)
}
template <typename T>
void launch_bwd_chain(int s, int w, int64_t r, const SnArgs& a, const T* valL, const T* Z, T* offacc, T* Xf,
                      int nrhs, uint64_t* xb, uint32_t epoch, int* status, hipStream_t st) {
  const int G = chain_groups(w, r, false);
  if (G < 1) throw std::runtime_error("bwd_chain: supernode too large for a chain launch");
  CHAIN_DISPATCH(bwd_chain_nr, s, G, w, r, a, valL, Z, offacc, Xf, xb, epoch, status, st)
}
#undef CHAIN_DISPATCH

#define LBLG_INSTANTIATE(T)                                                                                    \
  template void launch_fwd_chain<T>(int, int, int64_t, const SnArgs&, const T*, T*, T*, int, uint64_t*, uint32_t, \
                                    int*, hipStream_t);                                                        \
  template void launch_bwd_chain<T>(int, int, int64_t, const SnArgs&, const T*, const T*, T*, T*, int,        \
                                    uint64_t*, uint32_t, int*, hipStream_t);
LBLG_INSTANTIATE(double)
LBLG_INSTANTIATE(float)
#undef LBLG_INSTANTIATE

}  // namespace lblg
