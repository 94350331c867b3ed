// Active-set predictor of the learned warm start: an MLP
//   in -> h1 -> h2 -> out   (ReLU hidden layers, sigmoid outputs)
// with BF16 weights and FP32 accumulation on v_mfma_f32_16x16x16bf16_1k,
// trained with binary cross-entropy and Adam by the kernels below.  The
// parameters are kept in FP32 (the Adam master copy) and rounded to BF16 for
// every product; activations are FP32 in memory and rounded to BF16 when
// staged as MFMA operands.  No reference counterpart: NASOQ itself judges
// every prediction (lblg_qp_solve_predicted), so the returned solution is
// always NASOQ's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine/device_util.hpp"
#include "lblg.h"

namespace lblg {

typedef short s4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ inline float tof(float v) { return v; }
__device__ inline float tof(unsigned short v) { return bf2f(v); }

#define GT 64   // output tile
#define GKC 32  // k chunk staged in LDS
// C(i, j) = alpha sum_k a(i, k) b(k, j) + beta C(i, j), a(i, k) = A[i sa0 + k sa1],
// b(k, j) = B[k sb0 + j sb1]; operands rounded to BF16, FP32 accumulation.
// A wave owns a 32 x 32 quadrant = 2 x 2 MFMA blocks of 16 x 16: operand lane l
// holds row (or column) l % 16 and k = 4 (l / 16) .. + 3, accumulator element e
// of lane l is row 4 (l / 16) + e, column l % 16.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const TA* A, int64_t sa0, int64_t sa1,
                                                   const TB* B, int64_t sb0, int64_t sb1, float* C, int ldc,
                                                   float alpha, float beta) {
  __shared__ unsigned short As[GT][GKC + 4];  // [row][k]
  __shared__ unsigned short Bs[GT][GKC + 4];  // [column][k]
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  f4 acc[2][2];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
  const int srow = t >> 2, sk = (t & 3) * 8;
  for (int k0 = 0; k0 < K; k0 += GKC) {
    {
      const int i = i0 + srow, j = j0 + srow;
      float va[8], vb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + sk + e;
        va[e] = (i < M && k < K) ? tof(A[(int64_t)i * sa0 + (int64_t)k * sa1]) : 0.f;
        vb[e] = (j < N && k < K) ? tof(B[(int64_t)k * sb0 + (int64_t)j * sb1]) : 0.f;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        As[srow][sk + e] = f2bf(va[e]);
        Bs[srow][sk + e] = f2bf(vb[e]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GKC; kk += 16) {
      s4 a[2], b[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const unsigned short* pa = &As[wr + 16 * r + (ln & 15)][kk + 4 * (ln >> 4)];
        const unsigned short* pb = &Bs[wc + 16 * r + (ln & 15)][kk + 4 * (ln >> 4)];
        a[r] = s4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
        b[r] = s4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
      }
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[r], b[c], acc[r][c], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + wr + 16 * r + 4 * (ln >> 4) + e, j = j0 + wc + 16 * c + (ln & 15);
        if (i < M && j < N) {
          float* p = C + (int64_t)i * ldc + j;
          *p = alpha * acc[r][c][e] + (beta != 0.f ? beta * *p : 0.f);
        }
      }
}
template <typename TA, typename TB>
static void gemm(int M, int N, int K, const TA* A, int64_t sa0, int64_t sa1, const TB* B, int64_t sb0, int64_t sb1,
                 float* C, int ldc, hipStream_t st) {
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  k_gemm_bf16<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, sa0, sa1, B, sb0, sb1, C, ldc, 1.f, 0.f);
  HIP_LAUNCH_CHECK("gemm_bf16");
}

// Z(b, j) += bias(j), then ReLU in place when relu != 0
__global__ void k_bias_act(float* Z, const float* bias, int rows, int cols, int relu) {
  const int64_t n = (int64_t)rows * cols;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    float v = Z[q] + bias[q % cols];
    Z[q] = (relu && v < 0.f) ? 0.f : v;
  }
}
// dA *= (A > 0)
__global__ void k_relu_back(float* dA, const float* A, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    if (!(A[q] > 0.f)) dA[q] = 0.f;
}
// logits Z, labels Y (n = batch x out entries): dZ = (sigmoid(Z) - Y) / n and
// loss += the mean of the stable BCE terms max(z, 0) - z y + log(1 + e^-|z|)
__global__ void k_bce(const float* Z, const float* Y, int64_t n, float* dZ, float* loss) {
  __shared__ float sh[4];
  float acc = 0.f;
  const float inv = 1.f / (float)n;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const float z = Z[q], y = Y[q];
    const float p = 1.f / (1.f + expf(-z));
    dZ[q] = (p - y) * inv;
    acc += fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z)));
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (sh[0] + sh[1] + sh[2] + sh[3]) * inv);
}
__global__ void k_sigmoid(const float* Z, float* P, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    P[q] = 1.f / (1.f + expf(-Z[q]));
}
// db(j) = sum_b dZ(b, j)
__global__ void k_colsum(const float* dZ, int rows, int cols, float* db) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < cols; j += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < rows; ++b) s += dZ[(int64_t)b * cols + j];
    db[j] = s;
  }
}
// Adam (Kingma & Ba 2015) on the FP32 master parameters, then their BF16 copy
__global__ void k_adam(float* P, unsigned short* Pb, const float* G, float* Mo, float* Vo, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const float g = G[q];
    const float m = b1 * Mo[q] + (1.f - b1) * g;
    const float v = b2 * Vo[q] + (1.f - b2) * g * g;
    Mo[q] = m;
    Vo[q] = v;
    const float w = P[q] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[q] = w;
    Pb[q] = f2bf(w);
  }
}
__global__ void k_zero(float* p, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0.f;
}
__global__ void k_to_bf16(const float* P, unsigned short* Pb, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    Pb[q] = f2bf(P[q]);
}
// one wave: C(16x16) = A(16x16) B(16x16) through the BF16 MFMA (lane layout check)
__global__ void k_bf16_selftest(const float* A, const float* B, float* C) {
  const int ln = threadIdx.x;
  s4 a, b;
  for (int i = 0; i < 4; ++i) {
    a[i] = (short)f2bf(A[(ln & 15) * 16 + 4 * (ln >> 4) + i]);
    b[i] = (short)f2bf(B[(4 * (ln >> 4) + i) * 16 + (ln & 15)]);
  }
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  for (int e = 0; e < 4; ++e) C[(4 * (ln >> 4) + e) * 16 + (ln & 15)] = acc[e];
}

#define MLP_HIP(call)                                                  \
  do {                                                                 \
    if ((call) != hipSuccess) throw std::runtime_error("mlp: " #call); \
  } while (0)

static int ew_grid(int64_t n) { return (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n + 255) / 256)); }

struct Mlp {
  int in = 0, h1 = 0, h2 = 0, out = 0, bmax = 0;
  int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, np = 0;
  float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
  unsigned short* Pb = nullptr;
  float *X = nullptr, *Y = nullptr, *A1 = nullptr, *A2 = nullptr, *Z3 = nullptr, *dZ3 = nullptr, *dA2 = nullptr,
        *dA1 = nullptr, *loss = nullptr;
  int64_t t = 0;  // Adam steps taken
  hipStream_t st = nullptr;
  std::vector<void*> owned;
  template <typename T>
  T* alloc(int64_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)) != hipSuccess) throw std::bad_alloc();
    owned.push_back(p);
    return static_cast<T*>(p);
  }
  ~Mlp() {
    for (void* p : owned) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  void forward(int B) {
    gemm(B, h1, in, X, in, 1, Pb + oW1, 1, in, A1, h1, st);
    k_bias_act<<<ew_grid((int64_t)B * h1), 256, 0, st>>>(A1, P + ob1, B, h1, 1);
    gemm(B, h2, h1, A1, h1, 1, Pb + oW2, 1, h1, A2, h2, st);
    k_bias_act<<<ew_grid((int64_t)B * h2), 256, 0, st>>>(A2, P + ob2, B, h2, 1);
    gemm(B, out, h2, A2, h2, 1, Pb + oW3, 1, h2, Z3, out, st);
    k_bias_act<<<ew_grid((int64_t)B * out), 256, 0, st>>>(Z3, P + ob3, B, out, 0);
    HIP_LAUNCH_CHECK("mlp forward");
  }
  // gradients of the mean BCE over the batch into G; loss into *loss (device)
  void backward(int B) {
    k_zero<<<1, 64, 0, st>>>(loss, 1);
    const int64_t no = (int64_t)B * out;
    k_bce<<<ew_grid(no), 256, 0, st>>>(Z3, Y, no, dZ3, loss);
    gemm(out, h2, B, dZ3, 1, out, A2, h2, 1, G + oW3, h2, st);  // dW3 = dZ3^T A2
    k_colsum<<<ew_grid(out), 256, 0, st>>>(dZ3, B, out, G + ob3);
    gemm(B, h2, out, dZ3, out, 1, Pb + oW3, h2, 1, dA2, h2, st);  // dA2 = dZ3 W3
    k_relu_back<<<ew_grid((int64_t)B * h2), 256, 0, st>>>(dA2, A2, (int64_t)B * h2);
    gemm(h2, h1, B, dA2, 1, h2, A1, h1, 1, G + oW2, h1, st);  // dW2 = dA2^T A1
    k_colsum<<<ew_grid(h2), 256, 0, st>>>(dA2, B, h2, G + ob2);
    gemm(B, h1, h2, dA2, h2, 1, Pb + oW2, h1, 1, dA1, h1, st);  // dA1 = dA2 W2
    k_relu_back<<<ew_grid((int64_t)B * h1), 256, 0, st>>>(dA1, A1, (int64_t)B * h1);
    gemm(h1, in, B, dA1, 1, h1, X, in, 1, G + oW1, in, st);  // dW1 = dA1^T X
    k_colsum<<<ew_grid(h1), 256, 0, st>>>(dA1, B, h1, G + ob1);
    HIP_LAUNCH_CHECK("mlp backward");
  }
  void adam(float lr) {
    ++t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    k_adam<<<ew_grid(np), 256, 0, st>>>(P, Pb, G, Mo, Vo, np, lr, b1, b2, 1e-8f, c1, c2);
    HIP_LAUNCH_CHECK("adam");
  }
};

}  // namespace lblg

using lblg::Mlp;

namespace lblg {
void set_last_error(const char* what);  // capi_numeric.cpp (lblg_last_error)
}

namespace {
template <typename F>
int mlp_guard(F f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    lblg::set_last_error("mlp: allocation failed");
    return LBLG_ERR_NOMEM;
  } catch (const std::exception& e) {
    lblg::set_last_error(e.what());
    return LBLG_ERR_DEVICE;
  }
}
}  // namespace

extern "C" int lblg_mlp_create(lblg_mlp** out, int in, int h1, int h2, int nout, uint64_t seed, int max_batch) {
  if (!out || in < 1 || h1 < 1 || h2 < 1 || nout < 1 || max_batch < 1) return LBLG_ERR_ARG;
  *out = nullptr;
  return mlp_guard([&] {
    Mlp* m = new Mlp;
    try {
      m->in = in;
      m->h1 = h1;
      m->h2 = h2;
      m->out = nout;
      m->bmax = max_batch;
      m->oW1 = 0;
      m->ob1 = m->oW1 + (int64_t)h1 * in;
      m->oW2 = m->ob1 + h1;
      m->ob2 = m->oW2 + (int64_t)h2 * h1;
      m->oW3 = m->ob2 + h2;
      m->ob3 = m->oW3 + (int64_t)nout * h2;
      m->np = m->ob3 + nout;
      if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) throw std::runtime_error("stream");
      m->P = m->alloc<float>(m->np);
      m->G = m->alloc<float>(m->np);
      m->Mo = m->alloc<float>(m->np);
      m->Vo = m->alloc<float>(m->np);
      m->Pb = m->alloc<unsigned short>(m->np);
      m->X = m->alloc<float>((int64_t)max_batch * in);
      m->Y = m->alloc<float>((int64_t)max_batch * nout);
      m->A1 = m->alloc<float>((int64_t)max_batch * h1);
      m->A2 = m->alloc<float>((int64_t)max_batch * h2);
      m->Z3 = m->alloc<float>((int64_t)max_batch * nout);
      m->dZ3 = m->alloc<float>((int64_t)max_batch * nout);
      m->dA2 = m->alloc<float>((int64_t)max_batch * h2);
      m->dA1 = m->alloc<float>((int64_t)max_batch * h1);
      m->loss = m->alloc<float>(1);
      // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases
      std::vector<float> p(m->np, 0.f);
      std::mt19937_64 rng(seed);
      auto fill = [&](int64_t off, int rows, int cols) {
        std::uniform_real_distribution<float> u(-1.f / std::sqrt((float)cols), 1.f / std::sqrt((float)cols));
        for (int64_t q = 0; q < (int64_t)rows * cols; ++q) p[off + q] = u(rng);
      };
      fill(m->oW1, h1, in);
      fill(m->oW2, h2, h1);
      fill(m->oW3, nout, h2);
      if (hipMemcpy(m->P, p.data(), sizeof(float) * m->np, hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("upload");
      lblg::k_zero<<<lblg::ew_grid(m->np), 256, 0, m->st>>>(m->Mo, m->np);
      lblg::k_zero<<<lblg::ew_grid(m->np), 256, 0, m->st>>>(m->Vo, m->np);
      lblg::k_to_bf16<<<lblg::ew_grid(m->np), 256, 0, m->st>>>(m->P, m->Pb, m->np);
      if (hipStreamSynchronize(m->st) != hipSuccess) throw std::runtime_error("init");
    } catch (...) {
      delete m;
      throw;
    }
    *out = reinterpret_cast<lblg_mlp*>(m);
    return LBLG_OK;
  });
}

extern "C" int lblg_mlp_destroy(lblg_mlp* m) {
  delete reinterpret_cast<Mlp*>(m);
  return LBLG_OK;
}

extern "C" int64_t lblg_mlp_num_params(const lblg_mlp* mm) {
  const Mlp* m = reinterpret_cast<const Mlp*>(mm);
  return m ? m->np : -1;
}

extern "C" int lblg_mlp_get_params(lblg_mlp* mm, float* p) {
  Mlp* m = reinterpret_cast<Mlp*>(mm);
  if (!m || !p) return LBLG_ERR_ARG;
  return hipMemcpy(p, m->P, sizeof(float) * m->np, hipMemcpyDeviceToHost) == hipSuccess ? LBLG_OK : LBLG_ERR_DEVICE;
}

extern "C" int lblg_mlp_set_params(lblg_mlp* mm, const float* p) {
  Mlp* m = reinterpret_cast<Mlp*>(mm);
  if (!m || !p) return LBLG_ERR_ARG;
  return mlp_guard([&] {
    if (hipMemcpy(m->P, p, sizeof(float) * m->np, hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("upload");
    lblg::k_to_bf16<<<lblg::ew_grid(m->np), 256, 0, m->st>>>(m->P, m->Pb, m->np);
    lblg::k_zero<<<lblg::ew_grid(m->np), 256, 0, m->st>>>(m->Mo, m->np);
    lblg::k_zero<<<lblg::ew_grid(m->np), 256, 0, m->st>>>(m->Vo, m->np);
    m->t = 0;
    if (hipStreamSynchronize(m->st) != hipSuccess) throw std::runtime_error("sync");
    return LBLG_OK;
  });
}

extern "C" int lblg_mlp_gradients(lblg_mlp* mm, const float* X, const float* Y, int batch, float* grad, float* loss) {
  Mlp* m = reinterpret_cast<Mlp*>(mm);
  if (!m || !X || !Y || !grad || batch < 1 || batch > m->bmax) return LBLG_ERR_ARG;
  return mlp_guard([&] {
    MLP_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * batch * m->in, hipMemcpyHostToDevice, m->st));
    MLP_HIP(hipMemcpyAsync(m->Y, Y, sizeof(float) * batch * m->out, hipMemcpyHostToDevice, m->st));
    m->forward(batch);
    m->backward(batch);
    MLP_HIP(hipMemcpyAsync(grad, m->G, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->st));
    float l = 0.f;
    MLP_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    if (hipStreamSynchronize(m->st) != hipSuccess) throw std::runtime_error("gradients");
    if (loss) *loss = l;
    return LBLG_OK;
  });
}

extern "C" int lblg_mlp_train(lblg_mlp* mm, const float* X, const float* Y, int nsamples, int batch, int epochs,
                              float lr, uint64_t seed, float* last_loss) {
  Mlp* m = reinterpret_cast<Mlp*>(mm);
  if (!m || !X || !Y || nsamples < 1 || batch < 1 || batch > m->bmax || epochs < 0) return LBLG_ERR_ARG;
  return mlp_guard([&] {
    std::vector<int> order(nsamples);
    for (int i = 0; i < nsamples; ++i) order[i] = i;
    std::mt19937_64 rng(seed);
    std::vector<float> xb((size_t)batch * m->in), yb((size_t)batch * m->out);
    float l = 0.f, lsum = 0.f;
    int nb = 0;
    for (int e = 0; e < epochs; ++e) {
      std::shuffle(order.begin(), order.end(), rng);
      lsum = 0.f;
      nb = 0;
      for (int s0 = 0; s0 + batch <= nsamples; s0 += batch) {  // full minibatches
        for (int b = 0; b < batch; ++b) {
          std::memcpy(&xb[(size_t)b * m->in], X + (size_t)order[s0 + b] * m->in, sizeof(float) * m->in);
          std::memcpy(&yb[(size_t)b * m->out], Y + (size_t)order[s0 + b] * m->out, sizeof(float) * m->out);
        }
        MLP_HIP(hipMemcpyAsync(m->X, xb.data(), sizeof(float) * xb.size(), hipMemcpyHostToDevice, m->st));
        MLP_HIP(hipMemcpyAsync(m->Y, yb.data(), sizeof(float) * yb.size(), hipMemcpyHostToDevice, m->st));
        m->forward(batch);
        m->backward(batch);
        m->adam(lr);
        MLP_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
        if (hipStreamSynchronize(m->st) != hipSuccess) throw std::runtime_error("train");
        lsum += l;
        ++nb;
      }
    }
    if (last_loss) *last_loss = nb ? lsum / nb : 0.f;
    return LBLG_OK;
  });
}

extern "C" int lblg_mlp_predict(lblg_mlp* mm, const float* X, int nsamples, float* P) {
  Mlp* m = reinterpret_cast<Mlp*>(mm);
  if (!m || !X || !P || nsamples < 1) return LBLG_ERR_ARG;
  return mlp_guard([&] {
    for (int s0 = 0; s0 < nsamples; s0 += m->bmax) {
      const int B = std::min(m->bmax, nsamples - s0);
      MLP_HIP(hipMemcpyAsync(m->X, X + (size_t)s0 * m->in, sizeof(float) * B * m->in, hipMemcpyHostToDevice, m->st));
      m->forward(B);
      lblg::k_sigmoid<<<lblg::ew_grid((int64_t)B * m->out), 256, 0, m->st>>>(m->Z3, m->dZ3, (int64_t)B * m->out);
      MLP_HIP(hipMemcpyAsync(P + (size_t)s0 * m->out, m->dZ3, sizeof(float) * B * m->out, hipMemcpyDeviceToHost, m->st));
      if (hipStreamSynchronize(m->st) != hipSuccess) throw std::runtime_error("predict");
    }
    return LBLG_OK;
  });
}

extern "C" int lblg_selftest_mfma_bf16(const float* A, const float* B, float* C) {
  if (!A || !B || !C) return LBLG_ERR_ARG;
  return mlp_guard([&] {
    float* d = nullptr;  // A, B, C: 256 floats each
    MLP_HIP(hipMalloc(&d, 3 * 256 * sizeof(float)));
    struct Free {
      float* p;
      ~Free() { (void)hipFree(p); }
    } guard{d};
    MLP_HIP(hipMemcpy(d, A, 256 * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(d + 256, B, 256 * sizeof(float), hipMemcpyHostToDevice));
    lblg::k_bf16_selftest<<<1, 64>>>(d, d + 256, d + 512);
    HIP_LAUNCH_CHECK("bf16 selftest");
    MLP_HIP(hipMemcpy(C, d + 512, 256 * sizeof(float), hipMemcpyDeviceToHost));
    return LBLG_OK;
  });
}
