// FP64 reduction helpers and the Pauli term grouping shared by the kernels of
// libqmhip.so (kernels.hip, batch.hip).  Device code only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qm {

template <typename T>
__device__ __forceinline__ double prob_of(const T &a) {
  const double re = (double)a.re, im = (double)a.im;
  return re * re + im * im;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
  return r;  // valid in thread 0
}

__device__ __forceinline__ long long block_max(long long v, long long *sh) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    long long u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  long long r = -1;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = sh[i] > r ? sh[i] : r;
  return r;
}

// Pauli terms sharing one flip mask, evaluated from one read of the state.
constexpr int kTermsPerGroup = 8;
struct PauliGroup {
  uint64_t flip;
  int32_t count;
  int32_t first;  // index of the group's first term in the output order
  uint64_t sign[kTermsPerGroup];
  int32_t y[kTermsPerGroup];
};

}  // namespace qm
