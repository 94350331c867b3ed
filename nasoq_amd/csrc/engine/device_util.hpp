// Device helpers shared by the kernel sources: launch checks, LAPACK-style
// argmax reductions and the bounded group barrier.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace lblg {

#define HIP_LAUNCH_CHECK(what)                                                                     \
  do {                                                                                             \
    hipError_t e_ = hipGetLastError();                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define BK_ALPHA 0.6403882032022076  // (1 + sqrt(17)) / 8

// --------------------------------------------------------------------------
// The factor is stored and computed in one value type T (double: the plain
// FP64 factor; float: the FP32 factor that GMRES refines in FP64).  Mfma<T>
// is the 16x16x4 matrix instruction of T: v_mfma_f64_16x16x4f64 or
// v_mfma_f32_16x16x4f32.  Operand lanes are the same for both (first
// operand: row = lane % 16, k = lane / 16; second operand: column =
// lane % 16, k = lane / 16); accumulator element e of a lane holds output
// column lane % 16 and output row orow(e, lane), which differs between the
// two instructions (checked by lblg_selftest_mfma).
// --------------------------------------------------------------------------
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc;
  static __device__ inline d4 mma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int orow(int e, int ln) { return 4 * e + (ln >> 4); }
  static __device__ inline double sfmin() { return 2.2250738585072014e-308; }  // smallest normal
};
template <>
struct Mfma<float> {
  typedef f4 acc;
  static __device__ inline f4 mma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int orow(int e, int ln) { return 4 * (ln >> 4) + e; }
  static __device__ inline float sfmin() { return 1.17549435e-38f; }
};

// --------------------------------------------------------------------------
// reductions with LAPACK idamax tie-breaking (first index of the maximum)
// --------------------------------------------------------------------------
__device__ inline void better(double& v, int& i, double ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
__device__ inline void wave_argmax(double& v, int& i) {
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(i, off, 64);
    better(v, i, ov, oi);
  }
}
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// workgroup argmax over 256 threads; result valid in all threads
__device__ inline void block_argmax(double& v, int& i, double* sv, int* si) {
  wave_argmax(v, i);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) {
    sv[wv] = v;
    si[wv] = i;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) better(v, i, sv[q], si[q]);
  __syncthreads();
}

// --------------------------------------------------------------------------
// group barrier among the G workgroups that share one supernode.  The wait
// is bounded: after the spin limit the barrier flags status bit 2 and
// proceeds, so no wave can wait forever.
// --------------------------------------------------------------------------
__device__ inline void group_barrier(int* ctr, int G, int& epoch, int* status) {
  __syncthreads();
  ++epoch;
  if (G > 1) {
    if (threadIdx.x == 0) {
      __threadfence();  // release: the workgroup's stores (all in L2 after the barrier) become device-visible
      atomicAdd(ctr, 1);
      const int target = epoch * G;
      long spins = 0;
      while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1L << 26)) {
          atomicOr(status, 4);
          break;
        }
      }
      __threadfence();  // acquire: drop stale cache lines before the group reads shared data
    }
    __syncthreads();
  }
}

}  // namespace lblg
