// 128 x 128 MFMA tiles (FP64, or FP32 for the FP32 factor) for the large
// dense updates of the factorization:
//
//   k_schur_big     U = L_off D L_off^T of a supernode, scatter-added into
//                   the ancestors' panels (the DGEMM + scatter of
//                   ldl_left_sn_parallel_02, src/ldl/parallel_blocked_ldlt_02.cpp:128-153,
//                   in push form)
//   k_trailing_big  A22 -= L21 W21^T after a wide Bunch-Kaufman panel
//                   (dlasyf's trailing update inside LAPACKE_dsytrf, :155)
//
// A workgroup (4 waves) owns a 128 x 128 tile; each wave a 64 x 64 quadrant
// = 4 x 4 blocks of the 16x16x4 MFMA of the value type.  By default every
// lane loads its own MFMA operands straight from memory (the L2 serves the
// reuse between the waves of a tile), two k-steps ahead, with no LDS staging
// and no barrier in the k loop (big_gemm_direct).  The staged form
// (big_gemm, LBLG_SCHUR_DIRECT=0 / LBLG_TRAILING_DIRECT=0) moves chunks of 16
// columns of k through LDS, k-major with a row pitch of 144 doubles (rows of
// one k land in distinct bank halves for the four k of a step), prefetching
// the next chunk into registers while the current one is multiplied.  As in the 64 x 64 engine, the second operand is passed first
// so the accumulator holds the transposed block and lanes 0..15 cover 16
// consecutive rows in the epilogue.
//
// BF16 forms (the FP32 factor of precision LBLG_BF16): chunks of 32 columns
// of k are loaded as FP32, rounded to BF16 once and staged in LDS row by row
// ([row][k], 80-byte pitch), so the four waves of a tile share their
// operands; each wave multiplies them on v_mfma_f32_16x16x16_bf16 with FP32
// accumulation (big_gemm_bf16).  The next chunk is loaded into registers
// while the current one is multiplied.  With three terms (GEMM_BF16X3) every
// operand x is staged as hi = bf16(x) and lo = bf16(x - hi), and the product
// takes hi hi + hi lo + lo hi: three MFMAs, relative error ~2^-16 per term
// instead of ~2^-8.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "engine/device_util.hpp"
#include "engine/kernels.hpp"

namespace lblg {

#define BT 128
#define BKC 16
#define BLD 144
#define BRUNS 8  // target supernodes per tile with cached row positions

template <typename T>
struct BigAcc {
  typename Mfma<T>::acc c[4][4];
};

template <typename T>
__device__ inline void big_mma(BigAcc<T>& acc, const T* As, const T* Bs) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 64, q0 = (wv & 1) * 64;
  const int li = ln & 15, lk = ln >> 4;
#pragma unroll
  for (int k = 0; k < BKC; k += 4) {
    T a[4], b[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      a[x] = As[(k + lk) * BLD + r0 + 16 * x + li];
      b[x] = Bs[(k + lk) * BLD + q0 + 16 * x + li];
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y)
        acc.c[x][y] = Mfma<T>::mma(b[y], a[x], acc.c[x][y]);
  }
}

// acc = sum_k A(x, k) B(y, k) over k < K; la(k, x) / lb(k, x) return the
// operands (zero outside the matrix).  Thread t stages row t % 128 of k
// entries (t / 128) * 8 .. +8 of each chunk.
template <typename T, typename LA, typename LB>
__device__ inline void big_gemm(BigAcc<T>& acc, int K, T* As, T* Bs, LA la, LB lb) {
  const int t = threadIdx.x, x = t & 127, kq = (t >> 7) * 8;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc.c[a][b] = typename Mfma<T>::acc{0, 0, 0, 0};
  T ra[8], rb[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    ra[q] = la(kq + q, x);
    rb[q] = lb(kq + q, x);
  }
  for (int kc0 = 0; kc0 < K; kc0 += BKC) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      As[(kq + q) * BLD + x] = ra[q];
      Bs[(kq + q) * BLD + x] = rb[q];
    }
    __syncthreads();
    if (kc0 + BKC < K) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        ra[q] = la(kc0 + BKC + kq + q, x);
        rb[q] = lb(kc0 + BKC + kq + q, x);
      }
    }
    big_mma(acc, As, Bs);
  }
}

// acc = sum_k A(x, k) B(y, k) with A(x, k) = Ab[x + k lda], B(y, k) = Bb[y + k ldb],
// k < K, without LDS: every lane loads its own MFMA operands (rows 16 x + lane % 16
// of its quadrant, k = 4 step + lane / 16) straight from memory, two k-steps
// ahead of the products.  Rows beyond xa and columns beyond xb are clamped to
// the last valid one (the caller discards those results); k beyond K is zero.
template <typename T>
__device__ inline void big_gemm_direct(BigAcc<T>& acc, int K, const T* Ab, int64_t lda, int xa, const T* Bb,
                                       int64_t ldb, int xb) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 64, q0 = (wv & 1) * 64;
  const int li = ln & 15, lk = ln >> 4;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc.c[a][b] = typename Mfma<T>::acc{0, 0, 0, 0};
  const T* pa[4];
  const T* pb[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    pa[x] = Ab + min(r0 + 16 * x + li, xa - 1) + (int64_t)lk * lda;
    pb[x] = Bb + min(q0 + 16 * x + li, xb - 1) + (int64_t)lk * ldb;
  }
  const int64_t sa = 4 * lda, sb = 4 * ldb;  // one k-step
  const int nfull = K / 4;
  T a0[4], b0[4], a1[4], b1[4];
  auto load = [&](T* a, T* b, int ks) {
    const int64_t oa = ks * sa, ob = ks * sb;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      a[x] = pa[x][oa];
      b[x] = pb[x][ob];
    }
  };
  auto mul = [&](const T* a, const T* b) {
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) acc.c[x][y] = Mfma<T>::mma(b[y], a[x], acc.c[x][y]);
  };
  if (nfull > 0) load(a0, b0, 0);
  if (nfull > 1) load(a1, b1, 1);
  int ks = 0;
  for (; ks + 1 < nfull; ks += 2) {
    mul(a0, b0);
    if (ks + 2 < nfull) load(a0, b0, ks + 2);
    mul(a1, b1);
    if (ks + 3 < nfull) load(a1, b1, ks + 3);
  }
  if (ks < nfull) mul(a0, b0);
  if (4 * nfull < K) {  // last partial k-step
    const bool kok = 4 * nfull + lk < K;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      a0[x] = kok ? pa[x][nfull * sa] : (T)0;
      b0[x] = kok ? pb[x][nfull * sb] : (T)0;
    }
    mul(a0, b0);
  }
}

typedef short s4v __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_rne(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

#define B16K 32  // k per staged chunk
#define B16P 40  // LDS row pitch in BF16 values
#define B16_LDS (2 * BT * B16P)  // BF16 values of the A and B stages (one term)

// acc = sum_k A(x, k) B(y, k) with A(x, k) = Ab[x + k lda], B(y, k) = Bb[y + k ldb],
// k < K, the operands rounded to BF16 (TERMS = 1) or split into two BF16
// terms (TERMS = 3: hi hi + hi lo + lo hi) and accumulated in FP32.  Thread
// t stages row t % 128 and k = (t / 128) 16 .. + 15 of each 32-wide chunk.
// Rows beyond xa / xb are clamped to the last valid one (their results are
// discarded by the caller); k beyond K is zero.  S holds TERMS == 3 ? 2 : 1
// stages of B16_LDS values.
template <int TERMS>
__device__ inline void big_gemm_bf16(BigAcc<float>& acc, int K, const float* Ab, int64_t lda, int xa,
                                     const float* Bb, int64_t ldb, int xb, unsigned short* S) {
  unsigned short* As = S;
  unsigned short* Bs = S + BT * B16P;
  unsigned short* Al = S + B16_LDS;           // lo terms (TERMS == 3)
  unsigned short* Bl = S + B16_LDS + BT * B16P;
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  const int r0 = (wv >> 1) * 64, q0 = (wv & 1) * 64;
  const int li = ln & 15, lk = ln >> 4;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc.c[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const int x = t & 127, kq = (t >> 7) * 16;
  const float* pa = Ab + min(x, xa - 1);
  const float* pb = Bb + min(x, xb - 1);
  float ra[16], rb[16];
  auto load = [&](int kc) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = kc + kq + q;
      ra[q] = k < K ? pa[(int64_t)k * lda] : 0.f;
      rb[q] = k < K ? pb[(int64_t)k * ldb] : 0.f;
    }
  };
  // 16 FP32 values -> 16 BF16 (hi) and, for three terms, their residuals (lo)
  auto pack = [&](const float* v, uint4* hi, uint4* lo) {
    unsigned* uh = reinterpret_cast<unsigned*>(hi);
    unsigned* ul = reinterpret_cast<unsigned*>(lo);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const unsigned short h0 = bf16_rne(v[2 * q]), h1 = bf16_rne(v[2 * q + 1]);
      uh[q] = (unsigned)h0 | ((unsigned)h1 << 16);
      if constexpr (TERMS == 3) {
        const float r0 = v[2 * q] - __uint_as_float((unsigned)h0 << 16);
        const float r1 = v[2 * q + 1] - __uint_as_float((unsigned)h1 << 16);
        ul[q] = (unsigned)bf16_rne(r0) | ((unsigned)bf16_rne(r1) << 16);
      }
    }
  };
  auto stage = [&]() {
    uint4 ha[2], hb[2], la[2], lb[2];
    pack(ra, ha, la);
    pack(rb, hb, lb);
    uint4* da = reinterpret_cast<uint4*>(As + x * B16P + kq);
    uint4* db = reinterpret_cast<uint4*>(Bs + x * B16P + kq);
    da[0] = ha[0];
    da[1] = ha[1];
    db[0] = hb[0];
    db[1] = hb[1];
    if constexpr (TERMS == 3) {
      uint4* ea = reinterpret_cast<uint4*>(Al + x * B16P + kq);
      uint4* eb = reinterpret_cast<uint4*>(Bl + x * B16P + kq);
      ea[0] = la[0];
      ea[1] = la[1];
      eb[0] = lb[0];
      eb[1] = lb[1];
    }
  };
  load(0);
  for (int kc = 0; kc < K; kc += B16K) {
    __syncthreads();  // the previous chunk's operands have been read
    stage();
    __syncthreads();
    if (kc + B16K < K) load(kc + B16K);
#pragma unroll
    for (int kk = 0; kk < B16K; kk += 16) {
      s4v av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = *reinterpret_cast<const s4v*>(As + (r0 + 16 * q + li) * B16P + kk + 4 * lk);
        bv[q] = *reinterpret_cast<const s4v*>(Bs + (q0 + 16 * q + li) * B16P + kk + 4 * lk);
      }
      if constexpr (TERMS == 3) {
        s4v al[4], bl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          al[q] = *reinterpret_cast<const s4v*>(Al + (r0 + 16 * q + li) * B16P + kk + 4 * lk);
          bl[q] = *reinterpret_cast<const s4v*>(Bl + (q0 + 16 * q + li) * B16P + kk + 4 * lk);
        }
        // the small terms first, then hi hi
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            acc.c[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bl[b], av[a], acc.c[a][b], 0, 0, 0);
            acc.c[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bv[b], al[a], acc.c[a][b], 0, 0, 0);
          }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc.c[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bv[b], av[a], acc.c[a][b], 0, 0, 0);
    }
  }
}

template <typename T, typename F>
__device__ inline void big_foreach(const BigAcc<T>& acc, F f) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 64, q0 = (wv & 1) * 64;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) f(r0 + 16 * a + (ln & 15), q0 + 16 * b + Mfma<T>::orow(e, ln), acc.c[a][b][e]);
}

__device__ inline void tile_of(int64_t tix, int& ti, int& tj) {
  ti = (int)((sqrt(8.0 * (double)tix + 1.0) - 1.0) * 0.5);
  while ((int64_t)(ti + 1) * (ti + 2) / 2 <= tix) ++ti;
  while ((int64_t)ti * (ti + 1) / 2 > tix) --ti;
  tj = (int)(tix - (int64_t)ti * (ti + 1) / 2);
}

__device__ inline int64_t big_row_pos(const int* rows, int64_t b, int64_t e, int key) {
  int64_t lo = b, hi = e - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (rows[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo - b;
}

// MODE: GEMM_STAGED (operands staged through LDS, big_gemm), GEMM_DIRECT
// (loaded by each lane from memory, big_gemm_direct) or GEMM_BF16 (float
// operands rounded to BF16 in LDS, big_gemm_bf16)
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_schur_big(const int* list, const int64_t* tile_prefix, int count,
                                                   const int* super, const int64_t* pi, const int64_t* px,
                                                   const int* rows, const int* col2sup, T* valL,
                                                   const T* LD, const int64_t* ldoff, const unsigned char* tmask) {
  __shared__ T As[MODE == GEMM_STAGED ? BKC * BLD : 1], Bs[MODE == GEMM_STAGED ? BKC * BLD : 1];
  __shared__ __align__(16) unsigned short S16[MODE == GEMM_BF16X3 ? 2 * B16_LDS : MODE == GEMM_BF16 ? B16_LDS : 1];
  __shared__ int rlab[BT], clab[BT], ctg[BT], dtcol[BT];
  __shared__ int tpos[BRUNS][BT];
  __shared__ int64_t cbase[BT];
  __shared__ int dtl[BRUNS], ntg, reach;
  int lo = 0, hi = count - 1;
  const int64_t bid = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_prefix[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  const int s = list[lo];
  int ti, tj;
  tile_of(bid - tile_prefix[lo], ti, tj);
  const int w = super[s + 1] - super[s];
  const int64_t r = pi[s + 1] - pi[s];
  const int m = (int)(r - w);
  const T* P = valL + px[s];
  const int i0 = BT * ti, j0 = BT * tj;  // within the m off-diagonal rows
  const int t = threadIdx.x;
  if (t < BT) {
    rlab[t] = (i0 + t < m) ? rows[pi[s] + w + i0 + t] : -1;
    clab[t] = (j0 + t < m) ? rows[pi[s] + w + j0 + t] : -1;
    ctg[t] = clab[t] >= 0 ? col2sup[clab[t]] : -1;
  }
  if (tmask) {  // a restricted refactorization: tiles reaching no flagged supernode are skipped
    if (t == 0) reach = 0;
    __syncthreads();
    if (t < BT && ctg[t] >= 0 && tmask[ctg[t]]) reach = 1;
    __syncthreads();
    if (!reach) return;
  }
  __syncthreads();
  // distinct target supernodes of the tile's columns come in runs (labels are sorted)
  if (t == 0) {
    int run = -1, prev = -2;
    for (int x = 0; x < BT; ++x) {
      const int tg = ctg[x];
      if (tg >= 0 && tg != prev) {
        ++run;
        if (run < BRUNS) dtl[run] = tg;
        prev = tg;
      }
      dtcol[x] = (tg >= 0 && run < BRUNS) ? run : -1;
    }
    ntg = min(BRUNS, run + 1);
  }
  __syncthreads();
  for (int idx = t; idx < ntg * BT; idx += 256) {
    const int dt = idx / BT, i = idx % BT;
    const int tg = dtl[dt];
    const int R = rlab[i];
    tpos[dt][i] = (R >= super[tg]) ? (int)big_row_pos(rows, pi[tg], pi[tg + 1], R) : -1;
  }
  if (t < BT) {
    const int tg = ctg[t];
    cbase[t] = tg >= 0 ? px[tg] + (int64_t)(clab[t] - super[tg]) * (pi[tg + 1] - pi[tg]) : 0;
  }
  // the epilogue reads every thread's tpos / cbase entries (the direct GEMM has
  // no barrier of its own)
  __syncthreads();
  const T* LDs = LD + ldoff[lo];
  BigAcc<T> acc;
  if constexpr (MODE == GEMM_BF16 || MODE == GEMM_BF16X3) {
    big_gemm_bf16<MODE == GEMM_BF16X3 ? 3 : 1>(acc, w, P + w + i0, r, m - i0, LDs + j0, (int64_t)m, m - j0, S16);
  } else if constexpr (MODE == GEMM_DIRECT) {
    big_gemm_direct(acc, w, P + w + i0, r, m - i0, LDs + j0, (int64_t)m, m - j0);
  } else {
    big_gemm(
        acc, w, As, Bs,
        [&](int k, int x) -> T { return (k < w && i0 + x < m) ? P[w + i0 + x + (int64_t)k * r] : (T)0; },
        [&](int k, int x) -> T { return (k < w && j0 + x < m) ? LDs[(int64_t)k * m + j0 + x] : (T)0; });
  }
  // scatter-add: no global reads on the common path (target bases and row
  // positions come from LDS), so the no-return atomics stream out
  big_foreach(acc, [&](int ri, int ci, T v) {
    const int R = rlab[ri], C = clab[ci];
    if (R < 0 || C < 0 || R < C) return;
    if (tmask && !tmask[ctg[ci]]) return;
    const int dt = dtcol[ci];
    int64_t pos;
    if (dt >= 0) {
      pos = tpos[dt][ri];
    } else {
      const int tg = ctg[ci];
      pos = big_row_pos(rows, pi[tg], pi[tg + 1], R);
    }
    atomicAdd(valL + cbase[ci] + pos, -v);
  });
}

// A(i, j) -= sum_{l in [k0, k1)} L(i, l) W(j, l) for k1 <= j < w, j <= i < r
// (the trailing diagonal block and the rows below it), with W(j, l) at
// Wp[(l - k0) w + j] (k0 = kstart, k1 = kpos).
//   strip = 0: in place; 128 x 128 tiles over rows [lo, r) and columns
//              [lo, w), lo <= k1: the lower triangle of the square part, then
//              the full rows below
//   strip > 0: the next panel's columns j < k1 + strip (strip <= 128) only,
//              written to the panel buffer Pp[(j - k1) r + i] (the matrix itself
//              stays without the block's updates until the block ends); one
//              column of tiles starting at k1
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_trailing_big(const int* list, int lo, int strip, int wcols, const int* super,
                                                      const int64_t* pi, const int64_t* px, T* valL,
                                                      T* Wbase, const int64_t* woff,
                                                      const int* kpos, const int* kstart) {
  __shared__ T As[MODE == GEMM_STAGED ? BKC * BLD : 1], Bs[MODE == GEMM_STAGED ? BKC * BLD : 1];
  __shared__ __align__(16) unsigned short S16[MODE == GEMM_BF16X3 ? 2 * B16_LDS : MODE == GEMM_BF16 ? B16_LDS : 1];
  const int s = list[0];
  const int w = super[s + 1] - super[s];
  const int64_t r = pi[s + 1] - pi[s];
  const int k0 = kstart[s], k1 = kpos[s];
  if (k1 >= w || k1 <= k0) return;
  int i0, j0;
  if (strip) {
    i0 = k1 + BT * (int)blockIdx.x;
    j0 = k1;
  } else {
    const int Tj = (w - lo + BT - 1) / BT;
    const int64_t ntri = (int64_t)Tj * (Tj + 1) / 2;
    int ti, tj;
    if ((int64_t)blockIdx.x < ntri) {
      tile_of(blockIdx.x, ti, tj);
    } else {
      const int64_t b = blockIdx.x - ntri;
      ti = Tj + (int)(b / Tj);
      tj = (int)(b % Tj);
    }
    i0 = lo + BT * ti;
    j0 = lo + BT * tj;
  }
  const int jend = strip ? min(w, k1 + strip) : w;
  if (i0 >= r || i0 + BT - 1 < k1 || j0 >= jend || j0 + BT - 1 < k1) return;
  T* A = valL + px[s];
  const T* Wp = Wbase + woff[s];
  T* Pp = Wbase + woff[s] + (int64_t)w * wcols;  // r x strip, pitch r
  const int kb = k1 - k0;
  BigAcc<T> acc;
  if constexpr (MODE == GEMM_BF16 || MODE == GEMM_BF16X3) {
    big_gemm_bf16<MODE == GEMM_BF16X3 ? 3 : 1>(acc, kb, A + (int64_t)k0 * r + i0, r, (int)min<int64_t>(r - i0, BT),
                                              Wp + j0, (int64_t)w, w - j0, S16);
  } else if constexpr (MODE == GEMM_DIRECT) {
    big_gemm_direct(acc, kb, A + (int64_t)k0 * r + i0, r, (int)min<int64_t>(r - i0, BT), Wp + j0, (int64_t)w,
                    w - j0);
  } else {
    big_gemm(
        acc, kb, As, Bs,
        [&](int k, int x) -> T { return (k < kb && i0 + x < r) ? A[(int64_t)(k0 + k) * r + i0 + x] : (T)0; },
        [&](int k, int x) -> T { return (k < kb && j0 + x < w) ? Wp[(int64_t)k * w + j0 + x] : (T)0; });
  }
  // read-modify-write in batches of 16 values per lane: the loads use a
  // safe address for masked elements so that they issue back to back
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 64, q0 = (wv & 1) * 64;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T cur[16];
    int64_t off[16];
    bool ok[16];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + r0 + 16 * a + (ln & 15), j = j0 + q0 + 16 * b + Mfma<T>::orow(e, ln);
        const int q = 4 * b + e;
        ok[q] = i < r && j < jend && j >= k1 && i >= j;
        off[q] = ok[q] ? (int64_t)j * r + i : 0;
        cur[q] = A[off[q]];
      }
    if (strip) {
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int q = 4 * b + e;
          const int i = i0 + r0 + 16 * a + (ln & 15), j = j0 + q0 + 16 * b + Mfma<T>::orow(e, ln);
          if (ok[q]) Pp[(int64_t)(j - k1) * r + i] = cur[q] - acc.c[a][b][e];
        }
    } else {
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int q = 4 * b + e;
          if (ok[q]) A[off[q]] = cur[q] - acc.c[a][b][e];
        }
    }
  }
}

// The strip pre-update of the next wide panel (the k_trailing_big strip
// form) in 64 x 64 tiles: A(i, j) for the panel's columns k1 <= j < k1 +
// strip and rows j <= i < r, minus sum_{l in [k0, k1)} L(i, l) W(j, l),
// written to the panel buffer Pp[(j - k1) r + i].  A strip is one column of
// tiles and its launches hold few of them, so a launch takes as long as one
// tile: each wave here takes a 32 x 32 quarter (2 x 2 MFMA blocks) instead of
// a 64 x 64 one, with its operands four k-steps ahead.
#define SPF 4
template <typename T>
__global__ __launch_bounds__(256) void k_strip_update(const int* list, int strip, int wcols, const int* super,
                                                      const int64_t* pi, const int64_t* px, const T* valL, T* Wbase,
                                                      const int64_t* woff, const int* kpos, const int* kstart) {
  const int s = list[0];
  const int w = super[s + 1] - super[s];
  const int64_t r = pi[s + 1] - pi[s];
  const int k0 = kstart[s], k1 = kpos[s];
  if (k1 >= w || k1 <= k0) return;
  const int64_t i0 = k1 + 64 * (int64_t)blockIdx.x;
  const int j0 = k1 + 64 * (int)blockIdx.y;
  const int jend = min(w, k1 + strip);
  if (i0 >= r || j0 >= jend) return;
  const T* A = valL + px[s];
  const T* Wp = Wbase + woff[s];
  T* Pp = Wbase + woff[s] + (int64_t)w * wcols;  // r x strip, pitch r
  const int kb = k1 - k0;
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int r0 = (wv >> 1) * 32, q0 = (wv & 1) * 32;
  const int li = ln & 15, lk = ln >> 4;
  const int xa = (int)min<int64_t>(r - i0, 64), xb = w - j0;
  const T* pa[2];
  const T* pb[2];
#pragma unroll
  for (int x = 0; x < 2; ++x) {  // rows beyond the operands read their last row (results discarded)
    pa[x] = A + (int64_t)k0 * r + i0 + min(r0 + 16 * x + li, xa - 1) + (int64_t)lk * r;
    pb[x] = Wp + j0 + min(q0 + 16 * x + li, xb - 1) + (int64_t)lk * w;
  }
  const int64_t sa = 4 * r, sb = 4 * (int64_t)w;  // one k-step
  typename Mfma<T>::acc acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = typename Mfma<T>::acc{0, 0, 0, 0};
  const int nfull = kb / 4;
  T ra[SPF][2], rb[SPF][2];
  auto load = [&](T* a, T* b, int ks) {
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      a[x] = pa[x][ks * sa];
      b[x] = pb[x][ks * sb];
    }
  };
  auto mul = [&](const T* a, const T* b) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) acc[x][y] = Mfma<T>::mma(b[y], a[x], acc[x][y]);
  };
#pragma unroll
  for (int p = 0; p < SPF; ++p)
    if (p < nfull) load(ra[p], rb[p], p);
  int ks = 0;
  for (; ks + SPF <= nfull; ks += SPF) {
#pragma unroll
    for (int p = 0; p < SPF; ++p) {
      mul(ra[p], rb[p]);
      if (ks + SPF + p < nfull) load(ra[p], rb[p], ks + SPF + p);
    }
  }
#pragma unroll
  for (int p = 0; p < SPF; ++p)
    if (ks + p < nfull) mul(ra[p], rb[p]);
  if (4 * nfull < kb) {  // last partial k-step: k beyond kb is zero
    const bool kok = 4 * nfull + lk < kb;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      ra[0][x] = kok ? pa[x][nfull * sa] : (T)0;
      rb[0][x] = kok ? pb[x][nfull * sb] : (T)0;
    }
    mul(ra[0], rb[0]);
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t i = i0 + r0 + 16 * x + (ln & 15);
        const int j = j0 + q0 + 16 * y + Mfma<T>::orow(e, ln);
        if (i < r && j < jend && j >= k1 && i >= j)
          Pp[(int64_t)(j - k1) * r + i] = A[(int64_t)j * r + i] - acc[x][y][e];
      }
}

template <typename T>
void launch_strip_update(const int* list, int row_tiles, int strip, int wcols, const SnArgs& a, const T* valL, T* W,
                         const int64_t* woff, const int* kpos, const int* kstart, hipStream_t st) {
  if (row_tiles <= 0 || strip <= 0) return;
  if (strip > BT) throw std::runtime_error("strip_update: strip wider than a tile");
  const dim3 grid((unsigned)row_tiles, (unsigned)((strip + 63) / 64));
  k_strip_update<T><<<grid, 256, 0, st>>>(list, strip, wcols, a.super, a.pi, a.px, valL, W, woff, kpos, kstart);
  HIP_LAUNCH_CHECK("strip_update");
}

// LD(:, k) = (L_off D)(:, k) of each listed supernode, column-major m x w at
// LD + ldoff[q] (the D scaling of blocked_2by2_mult, Sym_BLAS.cpp:573-599)
template <typename T>
__global__ __launch_bounds__(256) void k_scale_ld(const int* list, const int* col_prefix, int count,
                                                  const int* super, const int64_t* pi, const int64_t* px,
                                                  const T* valL, const T* dval, int n, T* LD,
                                                  const int64_t* ldoff) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (col_prefix[mid] <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const int s = list[lo], k = blockIdx.x - col_prefix[lo];
  const int c0 = super[s], w = super[s + 1] - c0;
  const int64_t r = pi[s + 1] - pi[s];
  const int m = (int)(r - w);
  const T* col = valL + px[s] + w;
  const T dk = dval[c0 + k], ok = dval[n + c0 + k], op = k > 0 ? dval[n + c0 + k - 1] : (T)0;
  T* out = LD + ldoff[lo] + (int64_t)k * m;
  for (int j = threadIdx.x; j < m; j += 256) {
    const T l0 = col[(int64_t)k * r + j];
    T v;
    if (ok != 0.0) v = l0 * dk + col[(int64_t)(k + 1) * r + j] * ok;
    else if (op != 0.0) v = col[(int64_t)(k - 1) * r + j] * op + l0 * dk;
    else v = l0 * dk;
    out[j] = v;
  }
}

template <typename T>
void launch_scale_ld(const int* list, int count, const int* col_prefix, int ncols, const SnArgs& a,
                     const T* valL, const T* dval, T* LD, const int64_t* ldoff, hipStream_t st) {
  if (count <= 0 || ncols <= 0) return;
  k_scale_ld<T><<<ncols, 256, 0, st>>>(list, col_prefix, count, a.super, a.pi, a.px, valL, dval, a.n, LD, ldoff);
  HIP_LAUNCH_CHECK("scale_ld");
}

template <typename T>
void launch_schur_big(const int* list, int count, const int64_t* tile_prefix, int64_t ntiles, const SnArgs& a,
                      T* valL, const T* LD, const int64_t* ldoff, int mode, hipStream_t st) {
  if (count <= 0 || ntiles <= 0) return;
  if (ntiles > 0x7fffffffLL) throw std::runtime_error("schur_big: too many tiles");
  const dim3 grid((unsigned)ntiles);
  if (mode == GEMM_BF16 || mode == GEMM_BF16X3) {
    if constexpr (std::is_same<T, float>::value) {
      if (mode == GEMM_BF16X3)
        k_schur_big<T, GEMM_BF16X3><<<grid, 256, 0, st>>>(list, tile_prefix, count, a.super, a.pi, a.px, a.rows,
                                                          a.col2sup, valL, LD, ldoff, a.tmask);
      else
        k_schur_big<T, GEMM_BF16><<<grid, 256, 0, st>>>(list, tile_prefix, count, a.super, a.pi, a.px, a.rows,
                                                        a.col2sup, valL, LD, ldoff, a.tmask);
    } else {
      throw std::runtime_error("schur_big: BF16 operands need the FP32 factor");
    }
  } else if (mode == GEMM_DIRECT) {
    k_schur_big<T, GEMM_DIRECT><<<grid, 256, 0, st>>>(list, tile_prefix, count, a.super, a.pi, a.px, a.rows,
                                                      a.col2sup, valL, LD, ldoff, a.tmask);
  } else {
    k_schur_big<T, GEMM_STAGED><<<grid, 256, 0, st>>>(list, tile_prefix, count, a.super, a.pi, a.px, a.rows,
                                                      a.col2sup, valL, LD, ldoff, a.tmask);
  }
  HIP_LAUNCH_CHECK("schur_big");
}

template <typename T>
void launch_trailing_big(const int* list, int ntiles, int lo, int strip, int wcols, const SnArgs& a, T* valL,
                         T* W, const int64_t* woff, const int* kpos, const int* kstart, int mode, hipStream_t st) {
  if (ntiles <= 0) return;
  if (strip < 0 || strip > BT) throw std::runtime_error("trailing_big: strip wider than a tile");
  if (mode == GEMM_BF16 || mode == GEMM_BF16X3) {
    if constexpr (std::is_same<T, float>::value) {
      if (mode == GEMM_BF16X3)
        k_trailing_big<T, GEMM_BF16X3><<<ntiles, 256, 0, st>>>(list, lo, strip, wcols, a.super, a.pi, a.px, valL,
                                                               W, woff, kpos, kstart);
      else
        k_trailing_big<T, GEMM_BF16><<<ntiles, 256, 0, st>>>(list, lo, strip, wcols, a.super, a.pi, a.px, valL, W,
                                                             woff, kpos, kstart);
    } else {
      throw std::runtime_error("trailing_big: BF16 operands need the FP32 factor");
    }
  } else if (mode == GEMM_DIRECT) {
    k_trailing_big<T, GEMM_DIRECT><<<ntiles, 256, 0, st>>>(list, lo, strip, wcols, a.super, a.pi, a.px, valL, W,
                                                           woff, kpos, kstart);
  } else {
    k_trailing_big<T, GEMM_STAGED><<<ntiles, 256, 0, st>>>(list, lo, strip, wcols, a.super, a.pi, a.px, valL, W,
                                                           woff, kpos, kstart);
  }
  HIP_LAUNCH_CHECK("trailing_big");
}

#define LBLG_INSTANTIATE(T)                                                                                   \
  template void launch_strip_update<T>(const int*, int, int, int, const SnArgs&, const T*, T*, const int64_t*,  \
                                       const int*, const int*, hipStream_t);                                    \
  template void launch_scale_ld<T>(const int*, int, const int*, int, const SnArgs&, const T*, const T*, T*,  \
                                   const int64_t*, hipStream_t);                                            \
  template void launch_schur_big<T>(const int*, int, const int64_t*, int64_t, const SnArgs&, T*, const T*,   \
                                    const int64_t*, int, hipStream_t);                                      \
  template void launch_trailing_big<T>(const int*, int, int, int, int, const SnArgs&, T*, T*, const int64_t*, \
                                       const int*, const int*, int, hipStream_t);
LBLG_INSTANTIATE(double)
LBLG_INSTANTIATE(float)
#undef LBLG_INSTANTIATE

}  // namespace lblg
