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


// One Pauli group's numerators over amplitudes [j0, j1) (flip 0) or pairs
// [j0, j1) (flip != 0) of one state, C terms (the group's count).  A pair
// (j, j2 = j ^ flip) contributes s(j) f_y(o) + s(j2) f_y(conj o) with o =
// conj(a[j2]) a[j], s the sign parity of the term and f_y the part the Y count
// selects (0: Re, 1: -Im, 2: -Re, 3: Im; StateVector.Observations.cs:107-141).
// With s(j2) = s(j) sigma, sigma the parity of flip & sign, that is s(j) times
// alpha (1 + sigma) Re o + beta (1 - sigma) Im o, (alpha, beta) the
// coefficients of (Re o, Im o) in f_y: one product per term, equal bit for
// bit to the two-part sum (x + x = 2x and x - x = 0 are exact).
template <typename T, int C, int UNROLL = 1>
__device__ __forceinline__ void group_numerators(const T *__restrict__ s, const PauliGroup &g, uint64_t j0,
                                                 uint64_t j1, uint32_t t, uint32_t NT, double (&acc)[C]) {
  uint64_t sign[C];
  double cr[C], ci[C];
#pragma unroll
  for (int k = 0; k < C; k++) {
    sign[k] = g.sign[k];
    const int y = g.y[k] & 3;
    const double alpha = y == 0 ? 1.0 : y == 2 ? -1.0 : 0.0, beta = y == 1 ? -1.0 : y == 3 ? 1.0 : 0.0;
    const bool odd = (__popcll(g.flip & g.sign[k]) & 1) != 0;
    cr[k] = odd ? 0.0 : 2.0 * alpha;
    ci[k] = odd ? 2.0 * beta : 0.0;
    acc[k] = 0;
  }
  if (g.flip == 0) {
    for (uint64_t j = j0 + t; j < j1; j += NT) {
      const double p = prob_of(s[j]);
#pragma unroll
      for (int k = 0; k < C; k++) {
        const int y = g.y[k] & 3;
        double c = y == 0 ? p : y == 2 ? -p : 0.0;
        if (__popcll(j & sign[k]) & 1) c = -c;
        acc[k] += c;
      }
    }
    return;
  }
  const int h = 63 - __clzll((long long)g.flip);
  const uint64_t low = (1ull << h) - 1;
  auto add_pair = [&](uint64_t j, const T &x, const T &y) {
    const double ar = x.re, ai = x.im, pr = y.re, pim = -(double)y.im;
    const double ore = pr * ar - pim * ai;  // Complex.Conjugate(s[j2]) * s[j]
    const double oim = pim * ar + pr * ai;
#pragma unroll
    for (int k = 0; k < C; k++) {
      double v = cr[k] * ore + ci[k] * oim;
      if (__popcll(j & sign[k]) & 1) v = -v;
      acc[k] += v;
    }
  };
  // UNROLL pairs per thread in flight (reads from the last-level cache are
  // latency-bound otherwise); each thread's pairs are added in increasing q.
  uint64_t q = j0 + t;
  for (; q + (UNROLL - 1) * (uint64_t)NT < j1; q += UNROLL * (uint64_t)NT) {
    uint64_t j[UNROLL];
    T x[UNROLL], y[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
      const uint64_t qu = q + u * (uint64_t)NT;
      j[u] = ((qu & ~low) << 1) | (qu & low);
      x[u] = s[j[u]];
      y[u] = s[j[u] ^ g.flip];
    }
#pragma unroll
    for (int u = 0; u < UNROLL; u++) add_pair(j[u], x[u], y[u]);
  }
  for (; q < j1; q += NT) {
    const uint64_t jj = ((q & ~low) << 1) | (q & low);
    add_pair(jj, s[jj], s[jj ^ g.flip]);
  }
}

}  // namespace qm
