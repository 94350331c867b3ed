// L, D and L^T sweeps of a small factor in ONE workgroup.
//
// The level-scheduled sweeps (k_fwd_tri / k_fwd_gemv / k_bwd_gemv /
// k_bwd_tri) take four launches per level of the supernodal tree; for the
// KKT systems of an active-set iteration (a few hundred to a few thousand
// rows, tens of levels) the launches, not the arithmetic, set the time, and
// GMRES refinement repeats the sweeps several times per solve.  Here one
// workgroup of 16 waves keeps the whole ordered right-hand-side block in
// LDS and walks the same level schedule with a barrier between levels:
// wave k of a level takes supernodes k, k + 16, ...; within a supernode the
// 64 lanes hold 64 columns of the diagonal block (unit-lower triangle by
// lane broadcasts, column by column), and the rows below are updated with
// lanes over rows (forward: LDS atomics into the ancestors' entries, which
// supernodes of one level may share) or gathered (backward: lanes over rows
// with a wave sum per column for narrow chunks, lanes over columns
// otherwise).  Levels holding few wide supernodes (near the root) are
// swept by the whole workgroup, one supernode at a time: each 64-column
// chunk's triangle is staged in LDS and substituted by one wave, and the
// rows below are split over all 16 waves (the host picks the mode of each
// level with a cost model of both).  The arithmetic is the reference's blocked solve
// (src/triangularSolve/BLAS.cpp, blocked L / D / L^T solves over the
// supernodal factor) in the factor's type T.
#include <hip/hip_runtime.h>

#include "engine/device_util.hpp"
#include "engine/kernels.hpp"

namespace lblg {

namespace {

constexpr int SS_THREADS = 1024;
constexpr int SS_WAVES = SS_THREADS / 64;

// value of lane j (uniform j) in every lane
__device__ inline double ss_lane(double v, int j) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), j);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline float ss_lane(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
// The ordered block x: in LDS (XG = false), or in memory when it does not
// fit (XG = true).  In memory it is read with agent-scope loads, which go to
// the L2, so a wave never reads a stale vector-L1 line after another wave's
// atomic update (the atomics are performed in the L2); the workgroup
// barriers between levels order the writes.
// Element i of the block (row i / NR of the ordered matrix) is p[i - off]:
// off = 0 for the whole block, the first element of a system when the LDS
// holds one system of a batch.
template <typename T, bool XG>
struct XBlock {
  T* p;
  int64_t off;
  __device__ T ld(int64_t i) const {
    if constexpr (XG) return __hip_atomic_load(p + (i - off), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return p[i - off];
  }
  __device__ void st(int64_t i, T v) const {
    if constexpr (XG) __hip_atomic_store(p + (i - off), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else p[i - off] = v;
  }
  __device__ void add(int64_t i, T v) const { atomicAdd(p + (i - off), v); }
};

template <typename T>
__device__ inline T ss_wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Loads of the factor are issued several at a time before their values are
// used (the loops below would otherwise wait a full memory latency per
// column), and a lane's gathered right-hand-side rows are read once per
// 64-column chunk rather than once per column.  The number in flight
// shrinks with the right-hand sides a lane carries, so that the kernel keeps
// within the 128 registers of a 1024-thread workgroup.
template <int NR>
constexpr int ss_uf() { return NR <= 2 ? 4 : 2; }

// forward sweep of supernode s by one wave: x(s) = L_ss^{-1} x(s), then
// x(label q) -= L(q, s) x(s) for the rows below
template <typename T, int NR, typename XA>
__device__ void ss_forward(const int4 I, int64_t px, int ln, const SnArgs& a, const T* valL, XA x) {
  constexpr int SS_UF = ss_uf<NR>();
  const int c0 = I.x, w = I.y;
  const int64_t r = I.z, pis = I.w;
  const T* P = valL + px;
  for (int cb = 0; cb < w; cb += 64) {
    const int cw = min(64, w - cb);
    T v[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) v[c] = ln < cw ? x.ld((c0 + cb + ln) * NR + c) : (T)0;
    const T* Pl = P + (cb + ln) + (int64_t)cb * r;  // L(cb + ln, cb + j) = Pl[j r]
    for (int j0 = 0; j0 + 1 < cw; j0 += SS_UF) {
      T l[SS_UF];
#pragma unroll
      for (int u = 0; u < SS_UF; ++u) {
        const int j = j0 + u;
        l[u] = (j + 1 < cw && ln > j && ln < cw) ? Pl[(int64_t)j * r] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < SS_UF; ++u) {
        const int j = j0 + u;
        if (j + 1 < cw) {
#pragma unroll
          for (int c = 0; c < NR; ++c) v[c] -= l[u] * ss_lane(v[c], j);
        }
      }
    }
    if (ln < cw) {
#pragma unroll
      for (int c = 0; c < NR; ++c) x.st((c0 + cb + ln) * NR + c, v[c]);
    }
    for (int64_t q0 = cb + cw; q0 < r; q0 += 64) {
      const int64_t q = q0 + ln;
      const bool ok = q < r;
      const int lab = (ok && q >= w) ? a.rows[pis + q] : 0;
      const T* Pq = P + q + (int64_t)cb * r;  // L(q, cb + j) = Pq[j r]
      T acc[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = (T)0;
      for (int j0 = 0; j0 < cw; j0 += SS_UF) {
        T l[SS_UF];
#pragma unroll
        for (int u = 0; u < SS_UF; ++u) l[u] = (ok && j0 + u < cw) ? Pq[(int64_t)(j0 + u) * r] : (T)0;
#pragma unroll
        for (int u = 0; u < SS_UF; ++u) {
          if (j0 + u < cw) {
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[c] += l[u] * ss_lane(v[c], j0 + u);
          }
        }
      }
      if (ok) {
        if (q < w) {  // a later chunk of this supernode: its rows are this wave's alone
#pragma unroll
          for (int c = 0; c < NR; ++c) x.st((c0 + (int)q) * NR + c, x.ld((c0 + (int)q) * NR + c) - acc[c]);
        } else {      // an ancestor's row, shared with other supernodes of the level
#pragma unroll
          for (int c = 0; c < NR; ++c) x.add(lab * NR + c, -acc[c]);
        }
      }
    }
  }
}

// backward sweep of supernode s by one wave: x(s) = L_ss^{-T} (x(s) - L(q, s)^T x(label q))
template <typename T, int NR, typename XA>
__device__ void ss_backward(const int4 I, int64_t px, int ln, const SnArgs& a, const T* valL, XA x) {
  constexpr int SS_UF = ss_uf<NR>();
  const int c0 = I.x, w = I.y;
  const int64_t r = I.z, pis = I.w;
  const T* P = valL + px;
  // row blocks of 64 whose gathered x values a lane holds at once
  constexpr int RB = NR == 1 ? 4 : 2;
  for (int cb = ((w - 1) / 64) * 64; cb >= 0; cb -= 64) {
    const int cw = min(64, w - cb);
    const int64_t qb = cb + cw, rr = r - qb;  // rows below the chunk
    T v[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) v[c] = ln < cw ? x.ld((c0 + cb + ln) * NR + c) : (T)0;
    if (rr > 0) {
      if ((int64_t)cw * ((rr + 63) / 64 + 6 * NR) < rr) {
        // narrow chunk, many rows: lanes over rows, one wave sum per column;
        // the lane's rows and their x values are loaded once for all columns
        for (int64_t qs = qb; qs < r; qs += 64 * RB) {
          T xs[RB][NR];
          bool okb[RB];
#pragma unroll
          for (int b = 0; b < RB; ++b) {
            const int64_t q = qs + 64 * b + ln;
            okb[b] = q < r;
            const int lab = !okb[b] ? 0 : (q < w ? c0 + (int)q : a.rows[pis + q]);
#pragma unroll
            for (int c = 0; c < NR; ++c) xs[b][c] = okb[b] ? x.ld(lab * NR + c) : (T)0;
          }
          const T* Pq = P + qs + ln + (int64_t)cb * r;  // L(qs + ln + 64 b, cb + j) = Pq[64 b + j r]
          for (int j0 = 0; j0 < cw; j0 += SS_UF / 2) {
            T l[SS_UF / 2][RB];
#pragma unroll
            for (int u = 0; u < SS_UF / 2; ++u)
#pragma unroll
              for (int b = 0; b < RB; ++b)
                l[u][b] = (okb[b] && j0 + u < cw) ? Pq[64 * b + (int64_t)(j0 + u) * r] : (T)0;
#pragma unroll
            for (int u = 0; u < SS_UF / 2; ++u) {
              if (j0 + u < cw) {
#pragma unroll
                for (int c = 0; c < NR; ++c) {
                  T part = (T)0;
#pragma unroll
                  for (int b = 0; b < RB; ++b) part += l[u][b] * xs[b][c];
                  const T tot = ss_wave_sum(part);
                  if (ln == j0 + u) v[c] -= tot;
                }
              }
            }
          }
        }
      } else if (ln < cw) {
        // lanes over columns, rows in order (SS_UF rows per round of loads)
        const T* col = P + (int64_t)(cb + ln) * r;
        for (int64_t q0 = qb; q0 < r; q0 += SS_UF) {
          int lab[SS_UF];
          T l[SS_UF];
#pragma unroll
          for (int u = 0; u < SS_UF; ++u) {
            const int64_t q = q0 + u;
            lab[u] = q < r ? (q < w ? c0 + (int)q : a.rows[pis + q]) : 0;
            l[u] = q < r ? col[q] : (T)0;
          }
#pragma unroll
          for (int u = 0; u < SS_UF; ++u) {
#pragma unroll
            for (int c = 0; c < NR; ++c) v[c] -= l[u] * x.ld(lab[u] * NR + c);
          }
        }
      }
    }
    const T* Pt = P + (int64_t)(cb + ln) * r + cb;  // L(cb + i, cb + ln) = Pt[i]
    for (int i0 = cw - 1; i0 > 0; i0 -= SS_UF) {
      T l[SS_UF];
#pragma unroll
      for (int u = 0; u < SS_UF; ++u) {
        const int i = i0 - u;
        l[u] = (i > 0 && ln < i) ? Pt[i] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < SS_UF; ++u) {
        const int i = i0 - u;
        if (i > 0) {
#pragma unroll
          for (int c = 0; c < NR; ++c) v[c] -= l[u] * ss_lane(v[c], i);
        }
      }
    }
    if (ln < cw) {
#pragma unroll
      for (int c = 0; c < NR; ++c) x.st((c0 + cb + ln) * NR + c, v[c]);
    }
  }
}

// thin supernode (one column, a few rows) by one lane: forward
// x(label q) -= L(q) x(c0) for its rows below the diagonal
template <typename T, int NR, typename XA>
__device__ void ss_thin_forward(const int4 I, int64_t px, const SnArgs& a, const T* valL, XA x) {
  const int c0 = I.x, r = I.z;
  const int64_t pis = I.w;
  const T* P = valL + px;
  T y[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) y[c] = x.ld(c0 * NR + c);
  for (int q = 1; q < r; ++q) {
    const int lab = a.rows[pis + q];
    const T l = P[q];
#pragma unroll
    for (int c = 0; c < NR; ++c) x.add(lab * NR + c, -l * y[c]);
  }
}
// and backward: x(c0) -= sum_q L(q) x(label q)
template <typename T, int NR, typename XA>
__device__ void ss_thin_backward(const int4 I, int64_t px, const SnArgs& a, const T* valL, XA x) {
  const int c0 = I.x, r = I.z;
  const int64_t pis = I.w;
  const T* P = valL + px;
  T v[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) v[c] = x.ld(c0 * NR + c);
  for (int q = 1; q < r; ++q) {
    const int lab = a.rows[pis + q];
    const T l = P[q];
#pragma unroll
    for (int c = 0; c < NR; ++c) v[c] -= l * x.ld(lab * NR + c);
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) x.st(c0 * NR + c, v[c]);
}

constexpr int SS_TP = 65;  // pitch of the staged 64 x 64 triangle

// stage the unit-lower triangle of chunk cb of a supernode: tile[j * SS_TP + i] = L(cb + i, cb + j), i > j
template <typename T>
__device__ void ss_stage(const T* P, int64_t r, int cb, int cw, T* tile) {
  for (int idx = threadIdx.x; idx < cw * cw; idx += SS_THREADS) {
    const int i = idx % cw, j = idx / cw;
    if (i > j) tile[j * SS_TP + i] = P[(cb + i) + (int64_t)(cb + j) * r];
  }
}

// forward sweep of supernode s by the whole workgroup (uniform control flow)
template <typename T, int NR, typename XA>
__device__ void ss_forward_wg(const int4 I, int64_t px, const SnArgs& a, const T* valL, XA x, T* tile) {
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  const int c0 = I.x, w = I.y;
  const int64_t r = I.z, pis = I.w;
  const T* P = valL + px;
  for (int cb = 0; cb < w; cb += 64) {
    const int cw = min(64, w - cb);
    ss_stage(P, r, cb, cw, tile);
    __syncthreads();
    if (wv == 0) {
      T v[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) v[c] = ln < cw ? x.ld((c0 + cb + ln) * NR + c) : (T)0;
      for (int j = 0; j + 1 < cw; ++j) {
        const T l = (ln > j && ln < cw) ? tile[j * SS_TP + ln] : (T)0;
#pragma unroll
        for (int c = 0; c < NR; ++c) v[c] -= l * ss_lane(v[c], j);
      }
      if (ln < cw) {
#pragma unroll
        for (int c = 0; c < NR; ++c) x.st((c0 + cb + ln) * NR + c, v[c]);
      }
    }
    __syncthreads();
    // rows below the chunk in 64-row blocks over the waves; each row has one
    // owner (the level's supernodes are taken one at a time in this mode)
    for (int64_t q0 = cb + cw + (int64_t)wv * 64; q0 < r; q0 += SS_THREADS) {
      const int64_t q = q0 + ln;
      if (q < r) {
        T acc[NR];
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = (T)0;
#pragma unroll 8
        for (int j = 0; j < cw; ++j) {
          const T l = P[q + (int64_t)(cb + j) * r];
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] += l * x.ld((c0 + cb + j) * NR + c);
        }
        const int lab = q < w ? c0 + (int)q : a.rows[pis + q];
#pragma unroll
        for (int c = 0; c < NR; ++c) x.st(lab * NR + c, x.ld(lab * NR + c) - acc[c]);
      }
    }
    __syncthreads();
  }
}

// backward sweep of supernode s by the whole workgroup (uniform control flow)
template <typename T, int NR, typename XA>
__device__ void ss_backward_wg(const int4 I, int64_t px, const SnArgs& a, const T* valL, XA x, T* tile, T* red) {
  constexpr int SS_UF = ss_uf<NR>();
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  const int c0 = I.x, w = I.y;
  const int64_t r = I.z, pis = I.w;
  const T* P = valL + px;
  for (int cb = ((w - 1) / 64) * 64; cb >= 0; cb -= 64) {
    const int cw = min(64, w - cb);
    const int64_t qb = cb + cw, rr = r - qb;
    ss_stage(P, r, cb, cw, tile);
    for (int i = t; i < 64 * NR; i += SS_THREADS) red[i] = (T)0;
    __syncthreads();
    // red(j) = sum_q L(q, cb + j) x(label q): lanes over columns, each wave a
    // contiguous range of rows, partial sums met by LDS atomics
    if (rr > 0 && ln < cw) {
      const int64_t seg = (rr + SS_WAVES - 1) / SS_WAVES;
      const int64_t qs = qb + wv * seg, qe = min(r, qs + seg);
      const T* col = P + (int64_t)(cb + ln) * r;
      T acc[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = (T)0;
      for (int64_t q0 = qs; q0 < qe; q0 += SS_UF) {
        int lab[SS_UF];
        T l[SS_UF];
#pragma unroll
        for (int u = 0; u < SS_UF; ++u) {
          const int64_t q = q0 + u;
          lab[u] = q < qe ? (q < w ? c0 + (int)q : a.rows[pis + q]) : 0;
          l[u] = q < qe ? col[q] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < SS_UF; ++u) {
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] += l[u] * x.ld(lab[u] * NR + c);
        }
      }
      if (qs < qe) {
#pragma unroll
        for (int c = 0; c < NR; ++c) atomicAdd(&red[ln * NR + c], acc[c]);
      }
    }
    __syncthreads();
    if (wv == 0) {
      T v[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) v[c] = ln < cw ? x.ld((c0 + cb + ln) * NR + c) - red[ln * NR + c] : (T)0;
      for (int i = cw - 1; i > 0; --i) {
        const T l = ln < i ? tile[ln * SS_TP + i] : (T)0;  // L(cb + i, cb + ln)
#pragma unroll
        for (int c = 0; c < NR; ++c) v[c] -= l * ss_lane(v[c], i);
      }
      if (ln < cw) {
#pragma unroll
        for (int c = 0; c < NR; ++c) x.st((c0 + cb + ln) * NR + c, v[c]);
      }
    }
    __syncthreads();
  }
}

// out = (L D L^T)^{-1} in for the rows [col0, col0 + ncol) of the ordered,
// row-interleaved blocks in / out (n x NR; the whole matrix, or one system of
// a block-diagonal batch, whose supernodes touch no other rows).  Supernodes
// are listed level by level: info[k] = (first column, width, rows, first row
// index) and ipx[k] = first value of supernode k of the order, level l
// holding k in [lp[l], lp[l+1]); lm[l] & 1 = 1 for levels swept by the whole
// workgroup, lm[l] >> 1 = the level's thin supernodes.  lp and lm are LDS
// copies; each wave loads the description of its first supernode of the
// next level before working on the current one, so a level costs one
// dependent round of value loads instead of four.
template <typename T, int NR, bool XG>
__device__ inline void ss_sweeps(const int4* info, const int64_t* ipx, const int* lp, const int* lm, int nlev,
                                 int col0, int ncol, const SnArgs& a, const T* valL, const T* dval, const double* in,
                                 double* out, const XBlock<T, XG>& x, T* tile, T* red) {
  const int n = a.n, t = threadIdx.x, wv = t >> 6, ln = t & 63;
  const int64_t e0 = (int64_t)col0 * NR, e1 = (int64_t)(col0 + ncol) * NR;
  for (int64_t i = e0 + t; i < e1; i += SS_THREADS) x.st(i, (T)in[i]);
  __syncthreads();
  // first wave-swept item (after the level's thin supernodes) of this wave
  auto first = [&](int l) { return lp[l] + (lm[l] >> 1) + wv; };
  int4 nI = make_int4(0, 0, 0, 0);
  int64_t npx = 0;
  bool nok = first(0) < lp[1];
  if (nok) {
    nI = info[first(0)];
    npx = ipx[first(0)];
  }
  for (int l = 0; l < nlev; ++l) {
    const int4 I = nI;
    const int64_t px = npx;
    const bool ok = nok;
    if (l + 1 < nlev) {  // the next level's first item of this wave, in flight meanwhile
      const int k2 = first(l + 1);
      nok = k2 < lp[l + 2];
      if (nok) {
        nI = info[k2];
        npx = ipx[k2];
      }
    }
    const int lt = lp[l] + (lm[l] >> 1);  // end of the thin supernodes
    for (int k = lp[l] + t; k < lt; k += SS_THREADS) ss_thin_forward<T, NR>(info[k], ipx[k], a, valL, x);
    if (lm[l] & 1) {
      __syncthreads();  // the thin sweeps' atomics before the workgroup's plain updates
      for (int k = lt; k < lp[l + 1]; ++k) ss_forward_wg<T, NR>(info[k], ipx[k], a, valL, x, tile);
      __syncthreads();  // also where the level holds thin supernodes only
    } else {
      // each supernode's description is loaded while the one before is swept
      int4 cI = I;
      int64_t cpx = px;
      for (int k = first(l); ok && k < lp[l + 1]; k += SS_WAVES) {
        const bool more = k + SS_WAVES < lp[l + 1];
        int4 fI = cI;
        int64_t fpx = cpx;
        if (more) {
          fI = info[k + SS_WAVES];
          fpx = ipx[k + SS_WAVES];
        }
        ss_forward<T, NR>(cI, cpx, ln, a, valL, x);
        cI = fI;
        cpx = fpx;
      }
      __syncthreads();
    }
  }
  // D: 1x1 pivots, and 2x2 pivots (off-diagonal in dval[n + i]) by their first row
  for (int i = col0 + t; i < col0 + ncol; i += SS_THREADS) {
    const T od = dval[n + i];
    if (od != (T)0) {
      const T d1 = dval[i], d2 = dval[i + 1];
      const T inv = (T)1 / (d1 * d2 - od * od);
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const T x1 = x.ld(i * NR + c), x2 = x.ld((i + 1) * NR + c);
        x.st(i * NR + c, (x1 * d2 - x2 * od) * inv);
        x.st((i + 1) * NR + c, (x2 * d1 - x1 * od) * inv);
      }
    } else if (i == 0 || dval[n + i - 1] == (T)0) {
      const T inv = (T)1 / dval[i];
#pragma unroll
      for (int c = 0; c < NR; ++c) x.st(i * NR + c, x.ld(i * NR + c) * inv);
    }
  }
  nok = first(nlev - 1) < lp[nlev];
  if (nok) {
    nI = info[first(nlev - 1)];
    npx = ipx[first(nlev - 1)];
  }
  __syncthreads();
  for (int l = nlev - 1; l >= 0; --l) {
    const int4 I = nI;
    const int64_t px = npx;
    const bool ok = nok;
    if (l > 0) {
      const int k2 = first(l - 1);
      nok = k2 < lp[l];
      if (nok) {
        nI = info[k2];
        npx = ipx[k2];
      }
    }
    const int lt = lp[l] + (lm[l] >> 1);
    for (int k = lp[l] + t; k < lt; k += SS_THREADS) ss_thin_backward<T, NR>(info[k], ipx[k], a, valL, x);
    if (lm[l] & 1) {
      for (int k = lp[l + 1] - 1; k >= lt; --k) ss_backward_wg<T, NR>(info[k], ipx[k], a, valL, x, tile, red);
      __syncthreads();
    } else {
      int4 cI = I;
      int64_t cpx = px;
      for (int k = first(l); ok && k < lp[l + 1]; k += SS_WAVES) {
        const bool more = k + SS_WAVES < lp[l + 1];
        int4 fI = cI;
        int64_t fpx = cpx;
        if (more) {
          fI = info[k + SS_WAVES];
          fpx = ipx[k + SS_WAVES];
        }
        ss_backward<T, NR>(cI, cpx, ln, a, valL, x);
        cI = fI;
        cpx = fpx;
      }
      __syncthreads();
    }
  }
  for (int64_t i = e0 + t; i < e1; i += SS_THREADS) out[i] = (double)x.ld(i);
}


template <typename T, int NR, bool XG>
__global__ __launch_bounds__(SS_THREADS) void k_solve_small(const int4* info, const int64_t* ipx, const int* levptr,
                                                            const int* mode, int nlev, SnArgs a, const T* valL,
                                                            const T* dval, const double* in, double* out,
                                                            const double* active, int64_t astride, T* xg) {
  if (active) {  // inside a GMRES step: when no column takes part, no sweep is needed
    bool any = false;
#pragma unroll
    for (int c = 0; c < NR; ++c) any |= active[(int64_t)c * astride] != 0.0;
    if (!any) return;
  }
  extern __shared__ double smem_d[];
  T* lds = reinterpret_cast<T*>(smem_d);
  const int n = a.n, t = threadIdx.x;
  const XBlock<T, XG> x{XG ? xg : lds, 0};
  T* tile = XG ? lds : lds + ((n * NR + 1) & ~1);  // 64 x SS_TP staged triangle
  T* red = tile + 64 * SS_TP;         // 64 x NR partial sums
  int* lp = reinterpret_cast<int*>(red + 64 * NR);  // nlev + 1 level pointers
  int* lm = lp + nlev + 1;                          // nlev modes
  for (int i = t; i <= nlev; i += SS_THREADS) lp[i] = levptr[i];
  for (int i = t; i < nlev; i += SS_THREADS) lm[i] = mode[i];
  ss_sweeps<T, NR, XG>(info, ipx, lp, lm, nlev, 0, n, a, valL, dval, in, out, x, tile, red);
}

// The systems of a block-diagonal batch, one workgroup each: system g has
// sysd[6g ..] = (first level pointer, first mode, levels, first column,
// columns, -); its level pointers index info / ipx directly.  A system whose
// block has at most fit_elems values is kept in LDS, a larger one in xg (the
// same rows of an n x NR block in memory).
template <typename T, int NR>
__global__ __launch_bounds__(SS_THREADS) void k_solve_batch(const int* sysd, const int4* info, const int64_t* ipx,
                                                            const int* levptr, const int* mode, int fit_elems,
                                                            int maxlev, SnArgs a, const T* valL, const T* dval,
                                                            const double* in, double* out, const double* active,
                                                            int64_t astride, T* xg) {
  if (active) {
    bool any = false;
#pragma unroll
    for (int c = 0; c < NR; ++c) any |= active[(int64_t)c * astride] != 0.0;
    if (!any) return;
  }
  const int* d = sysd + 6 * (int64_t)blockIdx.x;
  const int lpo = d[0], mdo = d[1], nlev = d[2], col0 = d[3], ncol = d[4];
  extern __shared__ double smem_d[];
  T* lds = reinterpret_cast<T*>(smem_d);
  const int t = threadIdx.x;
  T* tile = lds + ((fit_elems + 1) & ~1);
  T* red = tile + 64 * SS_TP;
  int* lp = reinterpret_cast<int*>(red + 64 * NR);
  int* lm = lp + maxlev + 1;
  for (int i = t; i <= nlev; i += SS_THREADS) lp[i] = levptr[lpo + i];
  for (int i = t; i < nlev; i += SS_THREADS) lm[i] = mode[mdo + i];
  if ((int64_t)ncol * NR <= fit_elems) {
    const XBlock<T, false> x{lds, (int64_t)col0 * NR};
    ss_sweeps<T, NR, false>(info, ipx, lp, lm, nlev, col0, ncol, a, valL, dval, in, out, x, tile, red);
  } else {
    const XBlock<T, true> x{xg, 0};
    ss_sweeps<T, NR, true>(info, ipx, lp, lm, nlev, col0, ncol, a, valL, dval, in, out, x, tile, red);
  }
}

// LDS of the kernel: the block (n x NR, padded to an even count), the staged
// triangle and the partial sums
template <typename T>
size_t ss_lds(int n, int nr, int nlev) {
  return sizeof(T) * ((((size_t)n * nr + 1) & ~(size_t)1) + 64 * SS_TP + 64 * (size_t)nr) +
         sizeof(int) * (2 * (size_t)nlev + 1);
}

template <typename T, int NR>
void solve_small_nr(const int4* info, const int64_t* ipx, const int* levptr, const int* mode, int nlev,
                    const SnArgs& a, const T* valL, const T* dval, const double* in, double* out,
                    const double* active, int64_t astride, T* xg, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void*)k_solve_small<T, NR, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(SOLVE_SMALL_LDS + ss_lds<T>(0, MAX_RHS, SOLVE_SMALL_LEVELS))) != hipSuccess)
      throw std::runtime_error("solve_small: LDS attribute");
    attr = true;
  }
  if (xg) {  // the block in memory
    k_solve_small<T, NR, true><<<1, SS_THREADS, ss_lds<T>(0, NR, nlev), st>>>(info, ipx, levptr, mode, nlev, a, valL,
                                                                              dval, in, out, active, astride, xg);
  } else {
    k_solve_small<T, NR, false><<<1, SS_THREADS, ss_lds<T>(a.n, NR, nlev), st>>>(
        info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, nullptr);
  }
  HIP_LAUNCH_CHECK("solve_small");
}

template <typename T, int NR>
void solve_batch_nr(const int* sysd, int nsys, const int4* info, const int64_t* ipx, const int* levptr,
                    const int* mode, int fit_elems, int maxlev, const SnArgs& a, const T* valL, const T* dval,
                    const double* in, double* out, const double* active, int64_t astride, T* xg, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void*)k_solve_batch<T, NR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(SOLVE_SMALL_LDS + ss_lds<T>(0, MAX_RHS, SOLVE_SMALL_LEVELS))) != hipSuccess)
      throw std::runtime_error("solve_batch: LDS attribute");
    attr = true;
  }
  const size_t lds = sizeof(T) * (((size_t)fit_elems + 1) & ~(size_t)1) + ss_lds<T>(0, NR, maxlev);
  k_solve_batch<T, NR><<<nsys, SS_THREADS, lds, st>>>(sysd, info, ipx, levptr, mode, fit_elems, maxlev, a, valL,
                                                      dval, in, out, active, astride, xg);
  HIP_LAUNCH_CHECK("solve_batch");
}

}  // namespace

template <typename T>
void launch_solve_batch(const int* sysd, int nsys, const int4* info, const int64_t* ipx, const int* levptr,
                        const int* mode, int fit_cols, int maxlev, const SnArgs& a, const T* valL, const T* dval,
                        const double* in, double* out, int nrhs, const double* active, int64_t astride, T* xg,
                        hipStream_t st) {
  const int64_t fit = (int64_t)fit_cols * nrhs;
  if (nsys < 1 || nrhs < 1 || nrhs > MAX_RHS || maxlev < 1 || maxlev > SOLVE_SMALL_LEVELS ||
      sizeof(T) * (size_t)fit > SOLVE_SMALL_LDS || !xg)
    throw std::invalid_argument("solve_batch: schedule or block does not fit the workgroup's LDS");
  const int fe = (int)fit;
#define SB(NR) solve_batch_nr<T, NR>(sysd, nsys, info, ipx, levptr, mode, fe, maxlev, a, valL, dval, in, out, active, \
                                     astride, xg, st)
  switch (nrhs) {
    case 1: SB(1); break;
    case 2: SB(2); break;
    case 3: SB(3); break;
    case 4: SB(4); break;
    case 5: SB(5); break;
    case 6: SB(6); break;
    case 7: SB(7); break;
    default: SB(8); break;
  }
#undef SB
}

template <typename T>
void launch_solve_small(const int4* info, const int64_t* ipx, const int* levptr, const int* mode, int nlev,
                        const SnArgs& a, const T* valL, const T* dval, const double* in, double* out, int nrhs,
                        const double* active, int64_t astride, T* xg, hipStream_t st) {
  if (nrhs < 1 || nrhs > MAX_RHS || (!xg && sizeof(T) * (size_t)a.n * nrhs > SOLVE_SMALL_LDS) || nlev < 1 ||
      nlev > SOLVE_SMALL_LEVELS)
    throw std::invalid_argument("solve_small: block or level schedule does not fit the workgroup's LDS");
  switch (nrhs) {
    case 1: solve_small_nr<T, 1>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    case 2: solve_small_nr<T, 2>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    case 3: solve_small_nr<T, 3>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    case 4: solve_small_nr<T, 4>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    case 5: solve_small_nr<T, 5>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    case 6: solve_small_nr<T, 6>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    case 7: solve_small_nr<T, 7>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
    default: solve_small_nr<T, 8>(info, ipx, levptr, mode, nlev, a, valL, dval, in, out, active, astride, xg, st); break;
  }
}

template void launch_solve_batch<double>(const int*, int, const int4*, const int64_t*, const int*, const int*, int,
                                         int, const SnArgs&, const double*, const double*, const double*, double*, int,
                                         const double*, int64_t, double*, hipStream_t);
template void launch_solve_batch<float>(const int*, int, const int4*, const int64_t*, const int*, const int*, int, int,
                                        const SnArgs&, const float*, const float*, const double*, double*, int,
                                        const double*, int64_t, float*, hipStream_t);
template void launch_solve_small<double>(const int4*, const int64_t*, const int*, const int*, int, const SnArgs&,
                                         const double*, const double*, const double*, double*, int, const double*,
                                         int64_t, double*, hipStream_t);
template void launch_solve_small<float>(const int4*, const int64_t*, const int*, const int*, int, const SnArgs&,
                                        const float*, const float*, const double*, double*, int, const double*,
                                        int64_t, float*, hipStream_t);

}  // namespace lblg
