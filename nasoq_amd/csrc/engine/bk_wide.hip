// Cooperative panel Bunch-Kaufman for very wide diagonal blocks (the root
// separators of KKT matrices), LAPACK dlasyf-style: the columns of a panel
// of nb <= 64 columns are formed left-looking from the panel's L and
// W = L D, pivots follow the dsytf2 rules (alpha = (1+sqrt(17))/8) on the
// fully updated columns, and the rank-nb trailing update runs afterwards in
// k_trailing.  Reference: LAPACKE_dsytrf on the diagonal block at
// src/ldl/parallel_blocked_ldlt_02.cpp:155 and the column permutation of the
// rows below it (LAPACKE_dlapmt, :164-165).
//
// Work split.  The rows [k0, r) of the panel are the nb "diagonal" rows
// [k0, k0+nb) and the "far" rows below them: the rest of the diagonal block
// and the supernode's off-diagonal rows [w, r), which get their L in the same
// steps (so no separate off-diagonal solve is needed) but are never pivot
// candidates.  G workgroups each own up to WCH far rows (one per thread of
// every wave but the last) and keep those rows of the panel (L for
// eliminated columns, A for pending ones) in LDS.  Every
// workgroup also keeps the diagonal rows (nb x nb) and computes their
// updated values itself, so the W row needed by each column step is always
// local.
//
// Exchange.  Per column step the workgroups publish one small record (the
// largest candidate of the column) as flag-in-word 64-bit values written
// and polled with agent-scope atomics; the flag is a per-step epoch, so the
// record is its own barrier.  A step that has to inspect a second column
// (absakk < alpha * colmax) adds the candidate row of the winning group
// (published speculatively with the record) and a second record round.
//
// Interchanges.  Within the panel nothing is written to the matrix: the
// interchange of logical rows/columns kk and kp only updates a
// logical-to-physical map (the diagonal rows' map and a table of moved far
// rows) and the LDS row copies.  At the end of the panel the accumulated
// permutation is applied in place, through a scratch copy between two group
// barriers, to the trailing block, to the earlier columns of L and to the
// columns of the rows below the diagonal block; then L and W of the panel
// are written.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "engine/device_util.hpp"
#include "engine/kernels.hpp"

namespace lblg {

#define WNB 64                         // widest panel
#define WMM (2 * WNB)                  // moved indices per panel
#define OSB 16                         // columns per sub-block of the optimistic pass
#define LL_SPIN (1L << 22)

// exchange buffer layout, in 64-bit words (see kernels.hpp: LL_WORDS)
#define LL_RECA 0                            // [2][256][4]
#define LL_RECB (LL_RECA + 2 * 256 * 4)      // [2][256][8]
#define LL_PAY (LL_RECB + 2 * 256 * 8)       // [2][256][2 * WNB]

static_assert(LL_PAY + 2 * 256 * 2 * WNB == LL_WORDS, "exchange buffer size");

__device__ inline void ll_store(uint64_t* p, uint32_t v, uint32_t flag) {
  __hip_atomic_store(p, ((uint64_t)flag << 32) | (uint64_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The first wait of a factorization that runs out records where: status[8]
// = site + 16 x (word index - ll), status[9] = the waiting workgroup,
// status[10] = the flag it waited for, status[11] = the flag it last saw.
struct LlDiag {
  int* status;
  int site;
  int g;
  const uint64_t* base;
};
__device__ inline void ll_note(const LlDiag& d, const uint64_t* p, uint32_t flag, uint32_t seen) {
  if (atomicCAS(d.status + 8, 0, d.site + 16 * (int)(p - d.base)) == 0) {
    d.status[9] = d.g;
    d.status[10] = (int)flag;
    d.status[11] = (int)seen;
  }
}
// wait (bounded) for the word to carry this epoch's flag; abort != 0 skips the wait
__device__ inline uint32_t ll_load(uint64_t* p, uint32_t flag, int& abort, const LlDiag& d) {
  long spins = 0;
  while (!abort) {
    const uint64_t x = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(x >> 32) == flag) return (uint32_t)x;
    if (++spins > LL_SPIN) {
      abort = 1;
      ll_note(d, p, flag, (uint32_t)(x >> 32));
    }
  }
  return 0u;
}
__device__ inline void ll_store_d(uint64_t* p, double v, uint32_t flag) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  ll_store(p, (uint32_t)b, flag);
  ll_store(p + 1, (uint32_t)(b >> 32), flag);
}
__device__ inline double ll_load_d(uint64_t* p, uint32_t flag, int& abort, const LlDiag& d) {
  const uint64_t lo = ll_load(p, flag, abort, d);
  const uint64_t hi = ll_load(p + 1, flag, abort, d);
  return __longlong_as_double((long long)((hi << 32) | lo));
}

// W(i, j) = sum_c L(i, c) D(c, j) for an eliminated panel column j; l(j) reads L(i, j)
template <typename F, typename T>
__device__ inline double w_entry(F l, int j, const T* dd, const T* dsb) {
  if (dsb[j] != 0.0) return l(j) * dd[j] + l(j + 1) * dsb[j];
  if (j > 0 && dsb[j - 1] != 0.0) return l(j - 1) * dsb[j - 1] + l(j) * dd[j];
  return l(j) * dd[j];
}

__device__ inline int far_phys(int a, int fstart, const unsigned* mbits, const int* mlog, const int* mphys,
                               int nm) {
  const int b = a - fstart;
  if (!(mbits[b >> 5] & (1u << (b & 31)))) return a;
  for (int q = 0; q < nm; ++q)
    if (mlog[q] == a) return mphys[q];
  return a;
}

// NT threads per workgroup: waves 0 .. NT/64 - 2 own one far row per
// thread (WCH = NT - 64 rows per workgroup), the last wave computes the
// diagonal rows; NBM is the widest panel the launch takes.  256 threads with
// 64-column panels keep 192 far rows per workgroup (G <= 256: 49,152 rows);
// 512 threads keep 448 (114,688 rows) with 32-column FP64 or 64-column FP32
// panels, whose LDS rows are the same 256 bytes.
template <typename T, int NT, int NBM>
__global__ __launch_bounds__(NT) void k_bk_wide(const int* list, int G, int nb, int sub, int wcols, const int* super,
                                                 const int64_t* pi, const int64_t* px, T* valL,
                                                 T* dval, int* perm_piv, T* Wbase,
                                                 const int64_t* woff, int* kpos, int* kstart, int* sync,
                                                 int* sync_next, uint64_t* ll, uint32_t ebase, int n,
                                                 int* status) {
  constexpr int NW = NT / 64;                   // waves; wave NW - 1 holds the diagonal rows
  constexpr int WCH = NT - 64;                  // far rows per workgroup
  constexpr int WLDA = NBM + 1;                 // padded row length of the diagonal rows
  constexpr int MBITS_WORDS = 256 * WCH / 32;   // moved-far-row bitmap
  static_assert(NBM <= WNB && NT % 64 == 0 && NT >= 256, "wide panel launch shape");
  __shared__ T LC[NBM * WCH];       // own far rows, column-major [column][row]
  __shared__ T Ad[NBM * WLDA];      // diagonal rows: L (eliminated) | symmetric A (pending)
  __shared__ T wk[NBM], wi[NBM], td0[NBM], td1[NBM], dd[NBM], dsb[NBM], prow[NBM];
  __shared__ int dphys[NBM];
  __shared__ int mlog[WMM], mphys[WMM], Ml[WMM], Mp[WMM];
  __shared__ unsigned mbits[MBITS_WORDS];
  __shared__ double redv[NW];
  __shared__ int redi[NW];
  __shared__ double bcd[4];
  __shared__ int bci[4];
  __shared__ int s_nm, s_abort, s_nM, s_nMa, s_fail;

  const int g = blockIdx.x, t = threadIdx.x, wv = t >> 6, ln = t & 63;
  // consecutive launches alternate between two barrier counters: this one
  // counts from zero, and zeroes the other for the next launch (the launch
  // before used it and has ended; the next has not started)
  if (g == 0 && t == 0) __hip_atomic_store(sync_next, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int s = list[0];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  T* A = valL + px[s];
  T* Wp = Wbase + woff[s];
  const T* Pp = Wp + (int64_t)w * wcols;  // this panel's columns with the block's earlier panels applied
  T* scr = Wp + (int64_t)w * wcols + (int64_t)r * nb;
  const int k0 = kpos[s];
  if (k0 >= w) return;
  // sub > 0: this panel continues the block of panels that started at K0.
  // The matrix holds no contribution of the block's earlier panels (the
  // trailing block is updated once per block); this panel's columns with
  // those contributions applied are in Pp, and W holds the earlier panels'
  // columns [0, off)
  const int K0 = sub == 0 ? k0 : kstart[s];
  const int off = k0 - K0;
  const int nbe = min(nb, w - k0);
  const bool last = (w - k0) <= nb;
  // far rows: the rest of the diagonal block, then the rows below it (those
  // take part in every update and get their L here, but are never pivots)
  const int fstart = k0 + nbe, nfar = (int)(r - fstart);
  const int chunk = nfar > 0 ? (nfar + G - 1) / G : 0;
  if (chunk > WCH || nb > NBM || off < 0 || off + nb > wcols) {  // uniform over the grid: every workgroup returns
    if (g == 0 && t == 0) atomicOr(status, 8);
    return;
  }
  const int fbeg = fstart + g * chunk;
  const int fcnt = max(0, min(chunk, (int)r - fbeg));
  const bool hasrow = t < fcnt;
  const int fr = fbeg + t;
  int myphys = fr;
#define AG(i, j) A[(int64_t)(j) * r + (i)]
#define LCR(j) LC[(j) * WCH + t]
#define ADR(i, j) Ad[(i) * WLDA + (j)]

  // ---- load the panel rows
  // panel column j, row i >= j (with the block's earlier panels applied)
  auto pcol = [&](int i, int j) -> T { return sub ? Pp[(int64_t)j * r + i] : AG(i, k0 + j); };
  // a thread's loads of a batch issue before its LDS stores (one round trip
  // per batch rather than per entry)
  if (hasrow) {
#pragma unroll
    for (int j0 = 0; j0 < NBM; j0 += 16) {
      T v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = (j0 + q < nbe) ? pcol(fr, j0 + q) : (T)0;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (j0 + q < nbe) LCR(j0 + q) = v[q];
    }
  }
  {
    constexpr int Q = NBM * NBM / NT;
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int idx = t + NT * q, i = idx % nbe, j = idx / nbe;
      v[q] = idx < nbe * nbe ? ((i >= j) ? pcol(k0 + i, j) : pcol(k0 + j, i)) : (T)0;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int idx = t + NT * q, i = idx % nbe, j = idx / nbe;
      if (idx < nbe * nbe) ADR(i, j) = v[q];
    }
  }
  if (t < NBM) {
    dphys[t] = k0 + t;
    dd[t] = 0.0;
    dsb[t] = 0.0;
  }
  for (int q = t; q < MBITS_WORDS; q += NT) mbits[q] = 0u;
  if (t == 0) {
    s_nm = 0;
    s_abort = 0;
    s_fail = nbe;
  }
  __syncthreads();

  // ---- optimistic pass.  Bunch-Kaufman keeps the diagonal pivot of column
  // k without interchange when |a_kk| >= alpha * max_{i>k, i<w} |a_ik| on the
  // updated column.  Every workgroup runs all steps of the panel on that
  // assumption with no exchange, and each thread notes the first step at
  // which one of its candidate rows breaks the test.  One exchange then
  // gives the first failing step J: steps [0, J) are the steps the pivoting
  // loop below would take on the same values, and that loop resumes at J
  // from the panel's original columns.
  // The steps run in sub-blocks of OSB columns.  A sub-block's columns first
  // receive the earlier sub-blocks' contribution as 16x16 FP64 MFMA tiles
  // (far rows: a(i,j) -= sum_c L(i,c) W(j,c); diagonal rows:
  // a(i,j) -= sum_c W(i,c) L(j,c)); inside the sub-block every wave computes
  // the diagonal rows itself, so the steps need no block barrier.  During
  // the pass the upper triangle of the diagonal rows holds
  // W(i, j) = L(i, j) d_j (row j, column i) and wk holds 1 / d_j.
  {
    const bool drow = ln < nbe;
    const int li = ln & 15, lk = ln >> 4;
    int myfail = nbe;
    for (int cb = 0; cb < nbe; cb += OSB) {
      const int ce = min(nbe, cb + OSB);
      if (cb > 0) {
        __syncthreads();
        const int nft = (fcnt + 15) >> 4;      // far-row tiles (rows < WCH)
        const int ndt = (nbe - cb + 15) >> 4;  // diagonal-row tiles, rows [cb, 64)
        for (int tile = wv; tile < nft + ndt; tile += NW) {
          const bool isfar = tile < nft;
          const int i0 = isfar ? tile * 16 : cb + (tile - nft) * 16;
          typename Mfma<T>::acc acc = {0, 0, 0, 0};
          for (int c = 0; c < cb; c += 4) {
            const int cc = c + lk;
            // first operand row = column j of the sub-block, second operand col = row i
            const T x = isfar ? ADR(cc, cb + li) : ADR(cc, cb + li) * wk[cc];
            const T y = isfar ? LC[cc * WCH + i0 + li] : ADR(cc, i0 + li);
            acc = Mfma<T>::mma(x, y, acc);
          }
          // acc[e] = C(j = cb + orow(e), i = i0 + li)
          const int i = i0 + li;
          for (int e = 0; e < 4; ++e) {
            const int j = cb + Mfma<T>::orow(e, ln);
            if (j < ce) {
              if (isfar) {
                if (i < fcnt) LC[j * WCH + i] -= acc[e];
              } else if (i < nbe && i >= j) {
                ADR(i, j) -= acc[e];
              }
            }
          }
        }
        __syncthreads();
      }
      for (int kr = cb; kr < ce; ++kr) {
        double vd = 0.0;
        if (drow && ln >= kr) {
          double v = ADR(ln, kr);
          for (int c = cb; c < kr; ++c) v -= (ADR(c, ln) * wk[c]) * ADR(c, kr);
          vd = v;
        }
        double v0 = 0.0;
        if (hasrow) {
          double v = LCR(kr);
          for (int c = cb; c < kr; ++c) v -= LCR(c) * ADR(c, kr);
          v0 = v;
        }
        const double d = __shfl(vd, kr, 64);
        const double absakk = fabs(d);
        if (!(absakk > 0.0)) myfail = min(myfail, kr);
        if (wv == NW - 1 && drow && ln > kr) {
          const double x = fabs(vd);
          if (x == x && !(absakk >= BK_ALPHA * x)) myfail = min(myfail, kr);
        }
        if (hasrow && fr < w) {
          const double x = fabs(v0);
          if (x == x && !(absakk >= BK_ALPHA * x)) myfail = min(myfail, kr);
        }
        const double rk = 1.0 / d;
        if (hasrow) LCR(kr) = v0 * rk;
        if (drow && ln > kr) ADR(kr, ln) = vd;  // every wave stores the same value
        if (ln == 0) {
          wk[kr] = rk;
          dd[kr] = d;
        }
        // LDS accesses of one wave complete in order; keep the compiler from
        // moving the next step's loads above these stores
        __asm__ volatile("" ::: "memory");
      }
    }
    if (myfail < nbe) atomicMin(&s_fail, myfail);
    __syncthreads();
    if (wv == 0) {
      const uint32_t fO = ebase + 3u * (uint32_t)nb + 1u;
      int abort = 0;
      if (ln == 0) ll_store(ll + LL_RECA + (256 + (int64_t)g) * 4, (uint32_t)s_fail, fO);
      int jm = nbe;
      const LlDiag dg{status, 1, g, ll};
      for (int q = ln; q < G; q += 64)
        jm = min(jm, (int)ll_load(ll + LL_RECA + (256 + (int64_t)q) * 4, fO, abort, dg));
      for (int o = 32; o > 0; o >>= 1) jm = min(jm, __shfl_xor(jm, o, 64));
      const bool any_abort = __ballot(abort) != 0;
      if (ln == 0) {
        bci[2] = any_abort ? 0 : jm;
        if (any_abort) s_abort = 1;
      }
    }
    __syncthreads();
  }
  const int J = bci[2];
  // back to the pivoting loop's layout: L below the diagonal for the
  // accepted columns, the original columns from J on
  if (J < nbe && hasrow)
    for (int j = J; j < nbe; ++j) LCR(j) = pcol(fr, j);
  for (int idx = t; idx < nbe * nbe; idx += NT) {
    const int i = idx % nbe, j = idx / nbe;
    if (min(i, j) >= J) ADR(i, j) = i >= j ? pcol(k0 + i, j) : pcol(k0 + j, i);
    else if (i > j) ADR(i, j) = ADR(j, i) * wk[j];
    else if (i == j) ADR(i, i) = 1.0;
  }
  if (g == 0 && t < J) {
    dval[c0 + k0 + t] = dd[t];
    dval[n + c0 + k0 + t] = 0.0;
  }
  if (g == 0 && t == 0) {
    atomicAdd(status + 4, 1);
    if (J < nbe) atomicAdd(status + 5, 1);
  }
  __syncthreads();

  int k = k0 + J, step = 0;
  while (k < w && (last || k - k0 < nb - 1)) {
    const int kr = k - k0, par = step & 1;
    const uint32_t fA = ebase + 3u * (uint32_t)step + 1u, fB = fA + 1u;
    int abort = s_abort;
    // ---- W row k over the eliminated panel columns
    if (t < kr) wk[t] = w_entry([&](int j) { return ADR(kr, j); }, t, dd, dsb);
    __syncthreads();
    // ---- column k updated: own far rows and (wave 3) the diagonal rows
    double v0 = 0.0, v1 = 0.0;
    if (hasrow) {
      double v = LCR(kr);
      for (int j = 0; j < kr; ++j) v -= LCR(j) * wk[j];
      v0 = v;
    }
    if (wv == NW - 1 && ln >= kr && ln < nbe) {
      double v = ADR(ln, kr);
      for (int j = 0; j < kr; ++j) v -= ADR(ln, j) * wk[j];
      td0[ln] = v;
    }
    double av = -1.0;
    int ai = 0x7fffffff;
    if (hasrow && fr < w) {
      av = fabs(v0);
      if (!(av == av)) av = -1.0;
      ai = fr;
    }
    __syncthreads();
    if (wv == NW - 1 && ln > kr && ln < nbe) {
      av = fabs(td0[ln]);
      if (!(av == av)) av = -1.0;
      ai = k0 + ln;
    }
    block_argmax(av, ai, redv, redi);
    const double absakk = fabs(td0[kr]);
    // ---- record A: the group's largest candidate; candidate row for a second column
    {
      uint64_t* rec = ll + LL_RECA + ((int64_t)par * 256 + g) * 4;
      const bool cand = av >= 0.0;
      const bool own_far = cand && hasrow && ai == fr;
      // lanes of the last wave from nbe on hold no diagonal row (the far rows
      // k0 + nbe .. k0 + 63 must not match them)
      const bool own_diag = cand && wv == NW - 1 && ln < nbe && ai == k0 + ln;
      if (own_far || own_diag) {
        ll_store_d(rec, own_far ? v0 : td0[ln], fA);
        ll_store(rec + 2, (uint32_t)ai, fA);
      } else if (!cand && t == 0) {
        ll_store_d(rec, 0.0, fA);
        ll_store(rec + 2, 0x7fffffffu, fA);
      }
      if (cand && ai >= fbeg && ai < fbeg + fcnt && absakk < BK_ALPHA * av && t < nbe) {
        uint64_t* pay = ll + LL_PAY + ((int64_t)par * 256 + g) * 2 * WNB;
        ll_store_d(pay + 2 * t, LC[t * WCH + (ai - fbeg)], fA);
      }
    }
    if (wv == 0) {
      double ba = -1.0, bv = 0.0;
      int bi = 0x7fffffff, bg = 0;
      for (int q = ln; q < G; q += 64) {
        uint64_t* rec = ll + LL_RECA + ((int64_t)par * 256 + q) * 4;
        const double v = ll_load_d(rec, fA, abort, LlDiag{status, 2, g, ll});
        const int idx = (int)ll_load(rec + 2, fA, abort, LlDiag{status, 3, g, ll});
        double x = fabs(v);
        if (!(x == x) || idx == 0x7fffffff) x = -1.0;
        if (x > ba || (x == ba && idx < bi)) {
          ba = x;
          bi = idx;
          bv = v;
          bg = q;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double oa = __shfl_xor(ba, off, 64), ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64), og = __shfl_xor(bg, off, 64);
        if (oa > ba || (oa == ba && oi < bi)) {
          ba = oa;
          bi = oi;
          bv = ov;
          bg = og;
        }
      }
      const bool any_abort = __ballot(abort) != 0;
      if (ln == 0) {
        bcd[0] = ba;
        bcd[1] = bv;
        bci[0] = bi;
        bci[1] = bg;
        if (any_abort) s_abort = 1;
      }
    }
    __syncthreads();
    double colmax = bcd[0];
    int imax = bci[0];
    const int gwin = bci[1];
    if (k == w - 1 || colmax < 0.0) {
      colmax = 0.0;
      imax = k;
    }
    const int imr = imax - k0;
    const bool im_diag = imr < nbe;
    const double t0imax = im_diag ? td0[imr] : bcd[1];
    int kp = k, kstep = 1, zero = 0;
    bool use_t1 = false;
    double t1imax = 0.0;
    abort = s_abort;
    if (fmax(absakk, colmax) == 0.0 || !(absakk == absakk)) {
      zero = 1;
    } else if (!(absakk >= BK_ALPHA * colmax)) {
      // ---- second column: row imax of L / A, then column imax updated
      if (im_diag) {
        if (t < nbe) prow[t] = ADR(imr, t);
      } else if (wv == 0) {
        uint64_t* pay = ll + LL_PAY + ((int64_t)par * 256 + gwin) * 2 * WNB;
        for (int j = ln; j < nbe; j += 64) prow[j] = ll_load_d(pay + 2 * j, fA, abort, LlDiag{status, 4, g, ll});
      }
      __syncthreads();
      if (t < kr) wi[t] = w_entry([&](int j) { return prow[j]; }, t, dd, dsb);
      const int pim = im_diag ? dphys[imr] : far_phys(imax, fstart, mbits, mlog, mphys, s_nm);
      __syncthreads();
      if (hasrow) {
        const int pa = myphys;
        double v = (pa >= pim) ? AG(pa, pim) : AG(pim, pa);
        // the matrix lacks the block's earlier panels
        for (int c = 0; c < off; ++c) v -= AG(pa, K0 + c) * Wp[(int64_t)c * w + pim];
        for (int j = 0; j < kr; ++j) v -= LCR(j) * wi[j];
        v1 = v;
      }
      if (wv == NW - 1 && ln >= kr && ln < nbe) {
        double v = im_diag ? ADR(ln, imr) : prow[ln];
        for (int j = 0; j < kr; ++j) v -= ADR(ln, j) * wi[j];
        td1[ln] = v;
      }
      double bv2 = -1.0;
      int bi2 = 0x7fffffff;
      if (hasrow && fr < w && fr != imax) {
        bv2 = fabs(v1);
        if (!(bv2 == bv2)) bv2 = -1.0;
        bi2 = fr;
      }
      __syncthreads();
      if (wv == NW - 1 && ln >= kr && ln < nbe && k0 + ln != imax) {
        bv2 = fabs(td1[ln]);
        if (!(bv2 == bv2)) bv2 = -1.0;
        bi2 = k0 + ln;
      }
      block_argmax(bv2, bi2, redv, redi);
      {
        uint64_t* rec = ll + LL_RECB + ((int64_t)par * 256 + g) * 8;
        if (t == 0) ll_store_d(rec, bv2, fB);
        if (hasrow && fr == imax) ll_store_d(rec + 4, v1, fB);
      }
      if (wv == 0) {
        double rm = -1.0;
        for (int q = ln; q < G; q += 64) {
          const double x = ll_load_d(ll + LL_RECB + ((int64_t)par * 256 + q) * 8, fB, abort, LlDiag{status, 5, g, ll});
          rm = fmax(rm, x);
        }
        for (int off = 32; off > 0; off >>= 1) rm = fmax(rm, __shfl_xor(rm, off, 64));
        double ti = 0.0;
        if (!im_diag && ln == 0)
          ti = ll_load_d(ll + LL_RECB + ((int64_t)par * 256 + gwin) * 8 + 4, fB, abort, LlDiag{status, 6, g, ll});
        const bool any_abort = __ballot(abort) != 0;
        if (ln == 0) {
          bcd[2] = rm;
          bcd[3] = ti;
          if (any_abort) s_abort = 1;
        }
      }
      __syncthreads();
      double rowmax = bcd[2];
      if (rowmax < 0.0) rowmax = 0.0;
      t1imax = im_diag ? td1[imr] : bcd[3];
      if (absakk >= BK_ALPHA * colmax * (colmax / rowmax)) {
        kp = k;
      } else if (fabs(t1imax) >= BK_ALPHA * rowmax) {
        kp = imax;
        use_t1 = true;
      } else {
        kp = imax;
        kstep = 2;
      }
    }
    const int kk = k + kstep - 1, kkr = kk - k0, kpr = kp - k0;
    const bool swap = kp != kk;
    const bool kp_diag = kpr < nbe;
    double D1 = 0, D2 = 0, OD = 0, d11 = 0, d22 = 0, d21 = 0, rk = 0;
    if (kstep == 1) {
      D1 = use_t1 ? t1imax : td0[kr];
      rk = 1.0 / D1;
    } else {
      D1 = td0[kr];
      OD = t0imax;
      D2 = t1imax;
      d21 = OD;
      d11 = D2 / d21;
      d22 = D1 / d21;
      const double tt = 1.0 / (d11 * d22 - 1.0);
      d21 = tt / d21;
    }
    // ---- interchange kk <-> kp: row histories and the logical-to-physical map
    bool recv = false;  // this thread's far row becomes the old diagonal row kk
    if (swap) {
      if (kp_diag) {
        if (hasrow) LCR(kpr) = LCR(kkr);
      } else if (hasrow && fr == kp) {
        recv = true;
        myphys = dphys[kkr];
        for (int j = 0; j < nbe; ++j)
          if (j < kr || j > kkr) LCR(j) = ADR(kkr, j);
      }
      __syncthreads();
      if (kp_diag) {
        if (t < kr) {
          const double x = ADR(kkr, t);
          ADR(kkr, t) = ADR(kpr, t);
          ADR(kpr, t) = x;
        } else if (t > kk - k0 && t < nbe && t != kpr) {
          const double x = ADR(kkr, t);
          ADR(kpr, t) = x;
          ADR(t, kpr) = x;
        }
        if (t == 0) {
          ADR(kpr, kpr) = ADR(kkr, kkr);
          const int x = dphys[kkr];
          dphys[kkr] = dphys[kpr];
          dphys[kpr] = x;
        }
      } else {
        if (t < kr) ADR(kkr, t) = prow[t];
        if (t == 0) {
          const int old_kk = dphys[kkr];
          const int b = kp - fstart;
          int q = -1;
          if (mbits[b >> 5] & (1u << (b & 31)))
            for (int x = 0; x < s_nm; ++x)
              if (mlog[x] == kp) q = x;
          int old_kp = kp;
          if (q >= 0) {
            old_kp = mphys[q];
            mphys[q] = old_kk;
          } else {
            mlog[s_nm] = kp;
            mphys[s_nm] = old_kk;
            ++s_nm;
            mbits[b >> 5] |= 1u << (b & 31);
          }
          dphys[kkr] = old_kp;
        }
      }
      __syncthreads();
    }
    // ---- L column(s) of this step
    if (kstep == 1) {
      if (hasrow) {
        const double x = recv ? (use_t1 ? td1[kkr] : td0[kkr]) : (use_t1 ? v1 : v0);
        LCR(kr) = zero ? x : x * rk;
      }
      if (t > kr && t < nbe) {
        const int sr = (swap && kp_diag && t == kpr) ? kkr : t;
        const double x = use_t1 ? td1[sr] : td0[sr];
        ADR(t, kr) = zero ? x : x * rk;
      }
    } else {
      if (hasrow) {
        const double a0 = recv ? td0[kkr] : v0, a1 = recv ? td1[kkr] : v1;
        LCR(kr) = d21 * (d11 * a0 - a1);
        LCR(kr + 1) = d21 * (d22 * a1 - a0);
      }
      if (t > kr + 1 && t < nbe) {
        const int sr = (swap && kp_diag && t == kpr) ? kkr : t;
        ADR(t, kr) = d21 * (d11 * td0[sr] - td1[sr]);
        ADR(t, kr + 1) = d21 * (d22 * td1[sr] - td0[sr]);
      }
    }
    if (t == 0) {
      ADR(kr, kr) = 1.0;
      dd[kr] = D1;
      dsb[kr] = kstep == 2 ? OD : 0.0;
      if (kstep == 2) {
        ADR(kr + 1, kr) = 0.0;
        ADR(kr + 1, kr + 1) = 1.0;
        dd[kr + 1] = D2;
        dsb[kr + 1] = 0.0;
      }
      if (g == 0) {
        dval[c0 + k] = D1;
        dval[n + c0 + k] = kstep == 2 ? OD : 0.0;
        if (kstep == 2) {
          dval[c0 + k + 1] = D2;
          dval[n + c0 + k + 1] = 0.0;
        }
        if (swap) {
          const int x = perm_piv[c0 + kk];
          perm_piv[c0 + kk] = perm_piv[c0 + kp];
          perm_piv[c0 + kp] = x;
        }
        if (zero) atomicOr(status, 1);
        if (colmax > 0.0 && !(absakk >= BK_ALPHA * colmax)) atomicAdd(status + 2, 1);
        if (swap) atomicAdd(status + 3, 1);
      }
    }
    __syncthreads();
    k += kstep;
    ++step;
  }
  const int kend = k, ker = kend - k0;
  if (t == 0 && s_abort) atomicOr(status, 4);

  // ---- moved indices of this panel: (logical, physical)
  if (t == 0) {
    int c = 0, ca = 0;
    for (int i = 0; i < nbe; ++i)
      if (dphys[i] != k0 + i && k0 + i >= kend) {
        Ml[c] = k0 + i;
        Mp[c++] = dphys[i];
      }
    for (int q = 0; q < s_nm; ++q)
      if (mphys[q] != mlog[q]) {
        Ml[c] = mlog[q];
        Mp[c++] = mphys[q];
      }
    ca = c;  // the first ca entries lie in the trailing block
    for (int i = 0; i < nbe; ++i)
      if (dphys[i] != k0 + i && k0 + i < kend) {
        Ml[c] = k0 + i;
        Mp[c++] = dphys[i];
      }
    s_nM = c;
    s_nMa = ca;
  }
  __syncthreads();
  const int nM = s_nM, nMa = s_nMa;
  const int GT = G * NT, gt = g * NT + t;
  const int nt = w - kend;                 // trailing block size
  const int64_t no = r - w;                // rows below the diagonal block
  T* S1 = scr;                                  // [nMa][nt]
  T* S2 = scr + (int64_t)WMM * w;               // [nM][no]
  T* S3 = S2 + (int64_t)WMM * (no > 0 ? no : 0);  // [nM][k0]
  T* S4 = S3 + (int64_t)WMM * w;                 // [nM][off] rows of W of the block's earlier panels
  auto phys = [&](int j) -> int {
    if (j < k0 + nbe) return dphys[j - k0];
    return far_phys(j, fstart, mbits, mlog, mphys, s_nm);
  };
  int epoch = 0;
  if (nM > 0) {
    // gather the permuted values from the untouched panel-start storage
    for (int64_t idx = gt; idx < (int64_t)nMa * nt; idx += GT) {
      const int m = (int)(idx / nt), j = kend + (int)(idx % nt);
      const int pa = Mp[m], pj = phys(j);
      S1[idx] = pa >= pj ? AG(pa, pj) : AG(pj, pa);
    }
    // rows below the diagonal block: only the trailing columns move (their
    // panel columns hold L in logical order already)
    for (int64_t idx = gt; idx < (int64_t)nMa * no; idx += GT) {
      const int m = (int)(idx / no);
      S2[idx] = AG(w + idx % no, Mp[m]);
    }
    for (int64_t idx = gt; idx < (int64_t)nM * k0; idx += GT) {
      const int m = (int)(idx / k0), j = (int)(idx % k0);
      S3[idx] = AG(Mp[m], j);
    }
    for (int64_t idx = gt; idx < (int64_t)nM * off; idx += GT) {
      const int m = (int)(idx / off), c = (int)(idx % off);
      S4[idx] = Wp[(int64_t)c * w + Mp[m]];
    }
    group_barrier(sync, G, epoch, status);
    for (int64_t idx = gt; idx < (int64_t)nMa * nt; idx += GT) {
      const int m = (int)(idx / nt), j = kend + (int)(idx % nt);
      const int a = Ml[m];
      if (a >= j) AG(a, j) = S1[idx];
      else AG(j, a) = S1[idx];
    }
    for (int64_t idx = gt; idx < (int64_t)nMa * no; idx += GT) {
      const int m = (int)(idx / no);
      AG(w + idx % no, Ml[m]) = S2[idx];
    }
    for (int64_t idx = gt; idx < (int64_t)nM * k0; idx += GT) {
      const int m = (int)(idx / k0), j = (int)(idx % k0);
      AG(Ml[m], j) = S3[idx];
    }
    for (int64_t idx = gt; idx < (int64_t)nM * off; idx += GT) {
      const int m = (int)(idx / off), c = (int)(idx % off);
      Wp[(int64_t)c * w + Ml[m]] = S4[idx];
    }
  }
  // ---- L of the panel and W = L D for the rows of the trailing block
  if (hasrow) {
    for (int j = 0; j < ker; ++j) {
      AG(fr, k0 + j) = LCR(j);
      if (fr < w) Wp[(int64_t)(off + j) * w + fr] = w_entry([&](int c) { return LCR(c); }, j, dd, dsb);
    }
  }
  for (int idx = gt; idx < nbe * ker; idx += GT) {
    const int i = idx % nbe, j = idx / nbe;
    if (j <= i) AG(k0 + i, k0 + j) = ADR(i, j);
    if (k0 + i >= kend) Wp[(int64_t)(off + j) * w + k0 + i] = w_entry([&](int c) { return ADR(i, c); }, j, dd, dsb);
  }
  if (g == 0 && t == 0) {
    kstart[s] = K0;
    kpos[s] = kend;
  }
#undef AG
#undef LCR
#undef ADR
}

template <typename T>
void launch_bk_wide(const int* list, int G, int nt, int nb, int sub, int wcols, const SnArgs& a, T* valL, T* dval,
                    int* perm_piv, T* W, const int64_t* woff, int* kpos, int* kstart, int* sync,
                    int* sync_next, uint64_t* ll, uint32_t ebase, int* status, hipStream_t st) {
  if (nb > wide_panel_nb(nt, sizeof(T)) || nb < 4 || G < 1 || G > 256 || (nt != 256 && nt != WIDE_NT_TALL))
    throw std::runtime_error("bk_wide: launch shape out of range");
  int n = a.n;
  void* args[] = {(void*)&list, (void*)&G, (void*)&nb, (void*)&sub, (void*)&wcols, (void*)&a.super, (void*)&a.pi, (void*)&a.px,
                  (void*)&valL, (void*)&dval, (void*)&perm_piv, (void*)&W, (void*)&woff,
                  (void*)&kpos, (void*)&kstart, (void*)&sync, (void*)&sync_next, (void*)&ll, (void*)&ebase, (void*)&n,
                  (void*)&status};
  // tall supernodes: 448 far rows per workgroup in the same 256-byte LDS rows
  // (32 FP64 or 64 FP32 columns)
  constexpr int NBT = sizeof(T) == 8 ? 32 : 64;
  const void* fn = nt == 256 ? (const void*)k_bk_wide<T, 256, 64> : (const void*)k_bk_wide<T, WIDE_NT_TALL, NBT>;
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(G), dim3(nt), args, 0, st);
  if (e != hipSuccess) throw std::runtime_error(std::string("bk_wide: ") + hipGetErrorString(e));
}
template void launch_bk_wide<double>(const int*, int, int, int, int, int, const SnArgs&, double*, double*, int*,
                                     double*, const int64_t*, int*, int*, int*, int*, uint64_t*, uint32_t, int*,
                                     hipStream_t);
template void launch_bk_wide<float>(const int*, int, int, int, int, int, const SnArgs&, float*, float*, int*,
                                    float*, const int64_t*, int*, int*, int*, int*, uint64_t*, uint32_t, int*,
                                    hipStream_t);

}  // namespace lblg
