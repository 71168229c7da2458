// HIP kernels of libqmhip.so for gfx950 (CDNA4, 64-wide wavefronts).
//
// State layout: 2^n amplitudes, interleaved (re, im), 16 B (complex128) or
// 8 B (complex64) each, qubit 0 the least significant index bit -- the
// layout of System.Numerics.Complex[] in the reference (StateVector.cs:11).
// All index arithmetic is 64-bit.  Built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "dev_types.hpp"
#include "gate_math.hpp"
#include "qm_internal.hpp"
#include "dev_reduce.hpp"

namespace qm {

// ---------------------------------------------------------------------------
// Per-gate kernels: one launch per gate, enumerating only the amplitudes the
// gate can change (pairs with controls satisfied; the one-half of a
// diag(1, D)).  This is the reference's CUDA kernel shape
// (CudaStateVectorKernel.cs:38-77) with 64-bit indices and touched-fraction
// enumeration.
// ---------------------------------------------------------------------------
struct Fixed {
  int32_t count;
  int32_t pos[4];   // ascending bit positions inserted as zeros
  uint64_t value;   // bits set after insertion
};

__device__ __forceinline__ uint64_t insert_bits(uint64_t i, const Fixed &f) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (k < f.count) {
      const int p = f.pos[k];
      i = ((i >> p) << (p + 1)) | (i & ((1ull << p) - 1));
    }
  }
  return i | f.value;
}

template <typename T, int KIND>
__global__ __launch_bounds__(kBlock) void k_gate(T *__restrict__ s, uint64_t items, Fixed f,
                                                 uint64_t tbit, uint64_t obit, DevGate g) {
  using F = typename Real<T>::type;
  F c[8];
#pragma unroll
  for (int i = 0; i < 8; i++) c[i] = (F)g.c[i];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < items; i += stride) {
    const uint64_t first = insert_bits(i, f);
    if constexpr (KIND == G_DIAG) {
      T a = s[first];
      const bool hi = (first & tbit) != 0;
      mul_phase<F>(hi ? c[6] : c[0], hi ? c[7] : c[1], a.re, a.im);
      s[first] = a;
    } else if constexpr (KIND == G_DIAG1) {
      T a = s[first];  // first has the target bit set (inserted as a fixed one)
      mul_phase<F>(c[6], c[7], a.re, a.im);
      s[first] = a;
    } else {
      const uint64_t second = KIND == G_SWAP ? (first ^ tbit ^ obit) : (first | tbit);
      T x = s[first], y = s[second];
      if constexpr (KIND == G_X || KIND == G_SWAP) {
        s[first] = y;
        s[second] = x;
      } else {
        if constexpr (KIND == G_MAT) pair_mat<F>(c, x.re, x.im, y.re, y.im);
        if constexpr (KIND == G_RX) pair_rx<F>(c[0], c[1], x.re, x.im, y.re, y.im);
        if constexpr (KIND == G_RY) pair_ry<F>(c[0], c[1], x.re, x.im, y.re, y.im);
        if constexpr (KIND == G_H) pair_h<F>(c[0], x.re, x.im, y.re, y.im);
        if constexpr (KIND == G_Y) pair_y<F>(x.re, x.im, y.re, y.im);
        s[first] = x;
        s[second] = y;
      }
    }
  }
}

static int launch_blocks(uint64_t items) {
  uint64_t b = (items + kBlock - 1) / kBlock;
  return (int)std::min<uint64_t>(std::max<uint64_t>(b, 1), 65536);
}

template <typename T>
static void launch_gate_t(T *s, int n, const DevGate &g, hipStream_t st) {
  Fixed f{};
  uint64_t tbit = 1ull << g.target, obit = 0;
  std::vector<int> pos;
  uint64_t value = g.ctrl_mask;
  for (int q = 0; q < n; q++)
    if (g.ctrl_mask & (1ull << q)) pos.push_back(q);
  if (g.kind == G_SWAP) {
    obit = 1ull << g.swap_other;
    pos.push_back(g.swap_other);
    value |= obit;  // a = 1, b = 0
  }
  if (g.kind != G_DIAG) pos.push_back(g.target);
  if (g.kind == G_DIAG1) value |= tbit;
  std::sort(pos.begin(), pos.end());
  f.count = (int)pos.size();
  for (int i = 0; i < f.count; i++) f.pos[i] = pos[i];
  f.value = value;
  const uint64_t items = 1ull << (n - f.count);
  const int blocks = launch_blocks(items);
  switch (g.kind) {
#define QM_CASE(K) \
  case K: k_gate<T, K><<<blocks, kBlock, 0, st>>>(s, items, f, tbit, obit, g); break;
    QM_CASE(G_MAT)
    QM_CASE(G_RX)
    QM_CASE(G_RY)
    QM_CASE(G_H)
    QM_CASE(G_Y)
    QM_CASE(G_X)
    QM_CASE(G_DIAG)
    QM_CASE(G_DIAG1)
    QM_CASE(G_SWAP)
#undef QM_CASE
    default:
      fail(QM_ERR_UNSUPPORTED, "unknown device gate kind");
  }
  QM_HIP(hipGetLastError());
}

void launch_gate(State *s, const DevGate &g, hipStream_t st) {
  if (s->precision == QM_C128)
    launch_gate_t<C128>((C128 *)s->data, s->qubits, g, st);
  else
    launch_gate_t<C64>((C64 *)s->data, s->qubits, g, st);
}

// The fused tile-pass kernel lives in fused.hip.

// ---------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_basis(T *__restrict__ s, uint64_t n, uint64_t index) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    T a;
    a.re = i == index ? 1 : 0;
    a.im = 0;
    s[i] = a;
  }
}

void launch_set_basis(State *s, int64_t index, hipStream_t st) {
  const uint64_t n = (uint64_t)s->length;
  if (s->precision == QM_C128)
    k_fill_basis<C128><<<launch_blocks(n), kBlock, 0, st>>>((C128 *)s->data, n, (uint64_t)index);
  else
    k_fill_basis<C64><<<launch_blocks(n), kBlock, 0, st>>>((C64 *)s->data, n, (uint64_t)index);
  QM_HIP(hipGetLastError());
}

// ApplyDamping commit (CpuNoiseChannelBackend.cs:118-142).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_damp(T *__restrict__ s, uint64_t pairs, uint64_t bit,
                                                 int jump, double scale, double jump_scale,
                                                 double stay_scale) {
  using F = typename Real<T>::type;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < pairs; p += stride) {
    const uint64_t low = p & (bit - 1);
    const uint64_t first = ((p - low) << 1) | low;
    const uint64_t second = first | bit;
    if (jump) {
      T y = s[second];
      T a;
      a.re = (F)(jump_scale * (double)y.re);
      a.im = (F)(jump_scale * (double)y.im);
      s[first] = a;
      T z;
      z.re = 0;
      z.im = 0;
      s[second] = z;
    } else {
      T x = s[first], y = s[second];
      x.re = (F)((double)x.re * scale);
      x.im = (F)((double)x.im * scale);
      y.re = (F)((double)y.re * stay_scale);
      y.im = (F)((double)y.im * stay_scale);
      s[first] = x;
      s[second] = y;
    }
  }
}

void launch_scale_pairs(State *s, int target, bool jump, double scale, double jump_scale,
                        double stay_scale, hipStream_t st) {
  const uint64_t pairs = (uint64_t)s->length / 2, bit = 1ull << target;
  if (s->precision == QM_C128)
    k_damp<C128><<<launch_blocks(pairs), kBlock, 0, st>>>((C128 *)s->data, pairs, bit, jump, scale,
                                                          jump_scale, stay_scale);
  else
    k_damp<C64><<<launch_blocks(pairs), kBlock, 0, st>>>((C64 *)s->data, pairs, bit, jump, scale,
                                                         jump_scale, stay_scale);
  QM_HIP(hipGetLastError());
}

// Measure commit (StateVector.Observations.cs:241-247): keep the branch divided
// component-wise by the norm, zero the rest.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_collapse(T *__restrict__ s, uint64_t n, uint64_t mask,
                                                     uint64_t pattern, double norm) {
  using F = typename Real<T>::type;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    T a = s[i];
    if ((i & mask) == pattern) {
      a.re = (F)((double)a.re / norm);
      a.im = (F)((double)a.im / norm);
    } else {
      a.re = 0;
      a.im = 0;
    }
    s[i] = a;
  }
}

void launch_collapse_subset(State *s, uint64_t mask, uint64_t pattern, double norm, hipStream_t st) {
  const uint64_t n = (uint64_t)s->length;
  if (s->precision == QM_C128)
    k_collapse<C128><<<launch_blocks(n), kBlock, 0, st>>>((C128 *)s->data, n, mask, pattern, norm);
  else
    k_collapse<C64><<<launch_blocks(n), kBlock, 0, st>>>((C64 *)s->data, n, mask, pattern, norm);
  QM_HIP(hipGetLastError());
}

template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_convert(const A *__restrict__ a, B *__restrict__ b, uint64_t n) {
  using F = typename Real<B>::type;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    A x = a[i];
    B y;
    y.re = (F)x.re;
    y.im = (F)x.im;
    b[i] = y;
  }
}

void launch_convert(const State *src, State *dst, hipStream_t st) {
  const uint64_t n = (uint64_t)src->length;
  const int blocks = launch_blocks(n);
  if (src->precision == QM_C128 && dst->precision == QM_C64)
    k_convert<C128, C64><<<blocks, kBlock, 0, st>>>((const C128 *)src->data, (C64 *)dst->data, n);
  else if (src->precision == QM_C64 && dst->precision == QM_C128)
    k_convert<C64, C128><<<blocks, kBlock, 0, st>>>((const C64 *)src->data, (C128 *)dst->data, n);
  else
    QM_HIP(hipMemcpyAsync(dst->data, src->data, (size_t)src->bytes, hipMemcpyDeviceToDevice, st));
  QM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// FP64 reductions.  Block partials in a fixed order, then one block folds the
// partials in a fixed order: results are deterministic run to run.
// ---------------------------------------------------------------------------

// mode 0: total mass + last supported index; mode 1: masked mass; mode 2:
// masses of bit 0 / bit 1 of `mask` (stored as partial pairs).
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_mass(const T *__restrict__ s, uint64_t n, uint64_t per,
                                                 uint64_t mask, uint64_t pattern, double *partial,
                                                 long long *ipartial) {
  __shared__ double sh[kBlock / 64];
  __shared__ long long shl[kBlock / 64];
  const uint64_t begin = (uint64_t)blockIdx.x * per;
  const uint64_t end = begin + per < n ? begin + per : n;
  double a0 = 0, a1 = 0;
  long long last = -1;
  for (uint64_t i = begin + threadIdx.x; i < end; i += kBlock) {
    const double p = prob_of(s[i]);
    if constexpr (MODE == 0) {
      a0 += p;
      if (p > 0) last = (long long)i;
    } else if constexpr (MODE == 1) {
      if ((i & mask) == pattern) a0 += p;
    } else {
      if (i & mask)
        a1 += p;
      else
        a0 += p;
    }
  }
  const double r0 = block_sum(a0, sh);
  if constexpr (MODE == 2) {
    const double r1 = block_sum(a1, sh);
    if (threadIdx.x == 0) {
      partial[2 * blockIdx.x] = r0;
      partial[2 * blockIdx.x + 1] = r1;
    }
  } else {
    if (threadIdx.x == 0) partial[blockIdx.x] = r0;
  }
  if constexpr (MODE == 0) {
    const long long m = block_max(last, shl);
    if (threadIdx.x == 0) ipartial[blockIdx.x] = m;
  }
}

// Folds partial columns: out[c] = sum_b partial[b*bstride + c*cstride], b < count.
__global__ __launch_bounds__(kBlock) void k_fold(const double *__restrict__ partial, int64_t count,
                                                 int64_t bstride, int64_t cstride, double *out,
                                                 const long long *__restrict__ ipartial,
                                                 long long *iout) {
  __shared__ double sh[kBlock / 64];
  __shared__ long long shl[kBlock / 64];
  const int64_t c = blockIdx.x;
  double a = 0;
  for (int64_t b = threadIdx.x; b < count; b += kBlock) a += partial[b * bstride + c * cstride];
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) out[c] = r;
  if (ipartial) {
    long long m = -1;
    for (int64_t b = threadIdx.x; b < count; b += kBlock) m = ipartial[b] > m ? ipartial[b] : m;
    const long long mm = block_max(m, shl);
    if (threadIdx.x == 0) *iout = mm;
  }
}

static int64_t reduction_blocks(uint64_t n, uint64_t *per) {
  uint64_t blocks = std::min<uint64_t>(std::max<uint64_t>((n + 4095) / 4096, 1), 4096);
  *per = (n + blocks - 1) / blocks;
  blocks = (n + *per - 1) / *per;
  return (int64_t)blocks;
}

template <int MODE>
static void run_mass(State *s, uint64_t mask, uint64_t pattern, double *out_host,
                     long long *last_host, hipStream_t st) {
  uint64_t per;
  const int64_t blocks = reduction_blocks((uint64_t)s->length, &per);
  const int width = MODE == 2 ? 2 : 1;
  double *part = ctx_partials(s->ctx, blocks * width + 8);
  int64_t *ipart = ctx_ipartials(s->ctx, blocks + 8);
  double *res = part + blocks * width;  // width doubles of results
  long long *ires = (long long *)(ipart + blocks);
  if (s->precision == QM_C128)
    k_mass<C128, MODE><<<(int)blocks, kBlock, 0, st>>>((const C128 *)s->data, (uint64_t)s->length, per,
                                                       mask, pattern, part, (long long *)ipart);
  else
    k_mass<C64, MODE><<<(int)blocks, kBlock, 0, st>>>((const C64 *)s->data, (uint64_t)s->length, per,
                                                      mask, pattern, part, (long long *)ipart);
  QM_HIP(hipGetLastError());
  k_fold<<<width, kBlock, 0, st>>>(part, blocks, width, 1, res,
                                    MODE == 0 ? (long long *)ipart : nullptr, ires);
  QM_HIP(hipGetLastError());
  double *pin = ctx_pinned(s->ctx, 4);
  QM_HIP(hipMemcpyAsync(pin, res, sizeof(double) * width, hipMemcpyDeviceToHost, st));
  if (MODE == 0) QM_HIP(hipMemcpyAsync(pin + 2, ires, sizeof(long long), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < width; i++) out_host[i] = pin[i];
  if (MODE == 0) std::memcpy(last_host, pin + 2, sizeof(long long));
}

MassResult reduce_mass(State *s, hipStream_t st) {
  double total;
  long long last;
  run_mass<0>(s, 0, 0, &total, &last, st);
  return MassResult{total, (int64_t)last};
}

double reduce_masked_mass(State *s, uint64_t mask, uint64_t pattern, hipStream_t st) {
  double v;
  run_mass<1>(s, mask, pattern, &v, nullptr, st);
  return v;
}

void reduce_bit_masses(State *s, int target, double *m0, double *m1, hipStream_t st) {
  double v[2];
  run_mass<2>(s, 1ull << target, 0, v, nullptr, st);
  *m0 = v[0];
  *m1 = v[1];
}

// Kahan masses in the reference's sequential order (CpuNoiseChannelBackend.cs:88-106).
template <typename T>
__global__ void k_serial_bit_masses(const T *__restrict__ s, uint64_t n, uint64_t bit, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double z = 0, o = 0, zc = 0, oc = 0;
  for (uint64_t i = 0; i < n; i++) {
    const double m = prob_of(s[i]);
    double &sum = (i & bit) ? o : z;
    double &comp = (i & bit) ? oc : zc;
    const double corrected = m - comp;
    const double next = sum + corrected;
    comp = (next - sum) - corrected;
    sum = next;
  }
  out[0] = z;
  out[1] = o;
}

void serial_bit_masses(State *s, int target, double *m0, double *m1, hipStream_t st) {
  double *res = ctx_partials(s->ctx, 2);
  if (s->precision == QM_C128)
    k_serial_bit_masses<C128><<<1, 64, 0, st>>>((const C128 *)s->data, (uint64_t)s->length, 1ull << target, res);
  else
    k_serial_bit_masses<C64><<<1, 64, 0, st>>>((const C64 *)s->data, (uint64_t)s->length, 1ull << target, res);
  QM_HIP(hipGetLastError());
  double *pin = ctx_pinned(s->ctx, 2);
  QM_HIP(hipMemcpyAsync(pin, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  *m0 = pin[0];
  *m1 = pin[1];
}

// Pauli numerators (StateVector.Observations.cs:269-286): for each term of a
// flip-mask group, sum_j (-1)^popcount(j & sign) f_y(conj(a[j ^ flip]) * a[j]).

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pauli(const T *__restrict__ s, uint64_t n, uint64_t per,
                                                  const PauliGroup *__restrict__ groups,
                                                  double *partial, int64_t blocks) {
  __shared__ double sh[kBlock / 64];
  const PauliGroup &g = groups[blockIdx.y];
  const uint64_t begin = (uint64_t)blockIdx.x * per;
  const uint64_t end = begin + per < n ? begin + per : n;
  double acc[kTermsPerGroup];
  switch (g.count) {
#define QM_GROUP_CASE(C)                                                                              \
  case C: {                                                                                           \
    double part[C];                                                                                   \
    group_numerators<T, C>(s, g, g.flip == 0 ? begin : begin / 2, g.flip == 0 ? end : end / 2,        \
                           threadIdx.x, kBlock, part);                                                \
    for (int k = 0; k < C; k++) acc[k] = part[k];                                                     \
  } break;
    QM_GROUP_CASE(1) QM_GROUP_CASE(2) QM_GROUP_CASE(3) QM_GROUP_CASE(4)
    QM_GROUP_CASE(5) QM_GROUP_CASE(6) QM_GROUP_CASE(7) QM_GROUP_CASE(8)
#undef QM_GROUP_CASE
    default: break;
  }
  for (int k = 0; k < kTermsPerGroup; k++) {
    if (k >= g.count) break;
    const double r = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partial[(int64_t)(g.first + k) * blocks + blockIdx.x] = r;
  }
}

void reduce_pauli(State *s, const qm_pauli *terms, int64_t n_terms, double *numerators,
                  hipStream_t st) {
  if (n_terms == 0) return;
  // Group terms by flip mask; groups hold up to kTermsPerGroup terms.
  std::vector<int64_t> order(n_terms);
  for (int64_t i = 0; i < n_terms; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return terms[a].flip_mask < terms[b].flip_mask; });
  std::vector<PauliGroup> groups;
  for (int64_t k = 0; k < n_terms;) {
    PauliGroup g{};
    g.flip = terms[order[k]].flip_mask;
    g.first = (int32_t)k;
    while (k < n_terms && g.count < kTermsPerGroup && terms[order[k]].flip_mask == g.flip) {
      g.sign[g.count] = terms[order[k]].sign_mask;
      g.y[g.count] = terms[order[k]].y_count;
      g.count++;
      k++;
    }
    groups.push_back(g);
  }
  uint64_t per;
  const int64_t blocks = reduction_blocks((uint64_t)s->length, &per);
  const int64_t ng = (int64_t)groups.size();
  const int64_t gbytes = ng * (int64_t)sizeof(PauliGroup);
  double *part = ctx_partials(s->ctx, n_terms * blocks + n_terms + (gbytes + 7) / 8);
  double *res = part + n_terms * blocks;
  PauliGroup *d_groups = (PauliGroup *)(res + n_terms);
  QM_HIP(hipMemcpyAsync(d_groups, groups.data(), (size_t)gbytes, hipMemcpyHostToDevice, st));
  for (int64_t g0 = 0; g0 < ng; g0 += 65535) {
    const int64_t gn = std::min<int64_t>(65535, ng - g0);
    dim3 grid((unsigned)blocks, (unsigned)gn);
    if (s->precision == QM_C128)
      k_pauli<C128><<<grid, kBlock, 0, st>>>((const C128 *)s->data, (uint64_t)s->length, per,
                                             d_groups + g0, part, blocks);
    else
      k_pauli<C64><<<grid, kBlock, 0, st>>>((const C64 *)s->data, (uint64_t)s->length, per,
                                            d_groups + g0, part, blocks);
    QM_HIP(hipGetLastError());
  }
  // Fold: for term k (in sorted order), the partial row k.
  k_fold<<<(unsigned)n_terms, kBlock, 0, st>>>(part, blocks, 1, blocks, res, nullptr, nullptr);
  QM_HIP(hipGetLastError());
  std::vector<double> sorted(n_terms);
  double *pin = ctx_pinned(s->ctx, n_terms);
  QM_HIP(hipMemcpyAsync(pin, res, sizeof(double) * n_terms, hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  for (int64_t k = 0; k < n_terms; k++) numerators[order[k]] = pin[k];
}

// Inner product <a|b> with FP64 accumulation.
template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_inner(const A *__restrict__ a, const B *__restrict__ b,
                                                  uint64_t n, uint64_t per, double *partial) {
  __shared__ double sh[kBlock / 64];
  const uint64_t begin = (uint64_t)blockIdx.x * per;
  const uint64_t end = begin + per < n ? begin + per : n;
  double re = 0, im = 0;
  for (uint64_t i = begin + threadIdx.x; i < end; i += kBlock) {
    const double ar = a[i].re, ai = a[i].im, br = b[i].re, bi = b[i].im;
    re += ar * br + ai * bi;
    im += ar * bi - ai * br;
  }
  const double r0 = block_sum(re, sh);
  const double r1 = block_sum(im, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r0;
    partial[2 * blockIdx.x + 1] = r1;
  }
}

void reduce_inner(State *a, State *b, double *re, double *im, hipStream_t st) {
  uint64_t per;
  const int64_t blocks = reduction_blocks((uint64_t)a->length, &per);
  double *part = ctx_partials(a->ctx, 2 * blocks + 2);
  double *res = part + 2 * blocks;
  const uint64_t n = (uint64_t)a->length;
  if (a->precision == QM_C128 && b->precision == QM_C128)
    k_inner<C128, C128><<<(int)blocks, kBlock, 0, st>>>((const C128 *)a->data, (const C128 *)b->data, n, per, part);
  else if (a->precision == QM_C128)
    k_inner<C128, C64><<<(int)blocks, kBlock, 0, st>>>((const C128 *)a->data, (const C64 *)b->data, n, per, part);
  else if (b->precision == QM_C128)
    k_inner<C64, C128><<<(int)blocks, kBlock, 0, st>>>((const C64 *)a->data, (const C128 *)b->data, n, per, part);
  else
    k_inner<C64, C64><<<(int)blocks, kBlock, 0, st>>>((const C64 *)a->data, (const C64 *)b->data, n, per, part);
  QM_HIP(hipGetLastError());
  k_fold<<<2, kBlock, 0, st>>>(part, blocks, 2, 1, res, nullptr, nullptr);
  QM_HIP(hipGetLastError());
  double *pin = ctx_pinned(a->ctx, 2);
  QM_HIP(hipMemcpyAsync(pin, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  *re = pin[0];
  *im = pin[1];
}

// ---------------------------------------------------------------------------
// CDF search (StateVectorSimulator.cs:44-90, StateVector.cs:131-147), exact.
//
// The reference's cumulative is the left-to-right double sum
// S_i = fl(S_{i-1} + p_i).  This reproduces it bit for bit in parallel:
// while S stays in one binade [2^e, 2^(e+1)), every S is a multiple of
// u = 2^(e-52), so fl(S + p) = S + k*u with k = p/u rounded to nearest and a
// tie (frac = 1/2) rounded to the even multiple, which depends only on the
// parity of S/u.  Per block of 4096 amplitudes, one wave computes the block's
// increment in units of u exactly with integer prefix sums, for both start
// parities, assuming the binade predicted by an approximate (tree) prefix.
// One wave then walks the blocks in order with the exact running S: a block
// whose exact start has the predicted binade and whose end stays below
// 2^(e+1) is one integer add; any other block (a binade crossing, S = 0,
// subnormal sums) is summed sequentially by the wave, exactly as the
// reference does.  All-zero blocks leave S unchanged.  Each draw then
// binary-searches the exact block starts and walks one block sequentially
// from its exact start.  No decision depends on a rounding margin.
// ---------------------------------------------------------------------------
constexpr int64_t kChunk = 4096;
constexpr int kNoBinade = -2147483647 - 1;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const T *__restrict__ s, uint64_t n,
                                                       uint64_t chunk, double *sums) {
  __shared__ double sh[kBlock / 64];
  const uint64_t begin = (uint64_t)blockIdx.x * chunk;
  const uint64_t end = begin + chunk < n ? begin + chunk : n;
  double a = 0;
  for (uint64_t i = begin + threadIdx.x; i < end; i += kBlock) a += prob_of(s[i]);
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) sums[blockIdx.x] = r;
}

// Exclusive scan of `count` chunk sums by one 1024-thread block; total at out[count].
__global__ __launch_bounds__(1024) void k_scan(const double *__restrict__ sums, int64_t count,
                                               double *prefix) {
  __shared__ double sh[1024];
  const int t = threadIdx.x;
  const int64_t seg = (count + 1023) / 1024;
  const int64_t b = t * seg, e = b + seg < count ? b + seg : count;
  double local = 0;
  for (int64_t i = b; i < e; i++) local += sums[i];
  sh[t] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    double v = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  double run = t == 0 ? 0 : sh[t - 1];
  for (int64_t i = b; i < e; i++) {
    prefix[i] = run;
    run += sums[i];
  }
  if (t == 1023) prefix[count] = sh[1023];
}

// Per block (one wave): the increment of S over the block in units of
// u = 2^(e-52) for start parity 0 and 1, e = the predicted binade.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_block_delta(const T *__restrict__ s, uint64_t n,
                                                        uint64_t chunk, int64_t blocks,
                                                        const double *__restrict__ prefix,
                                                        int *epred, long long *delta) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (b >= blocks) return;  // whole wave
  const double a0 = prefix[b], a1 = prefix[b + 1];
  int e = kNoBinade;
  if (a0 >= 0x1p-1022 && isfinite(a1) && ilogb(a0) == ilogb(a1)) e = ilogb(a0);
  if (e == kNoBinade) {
    if (lane == 0) epred[b] = kNoBinade;
    return;
  }
  const int shift = 52 - e;
  const uint64_t begin = (uint64_t)b * chunk;
  const uint64_t len = begin + chunk < n ? chunk : n - begin;
  long long t0 = 0, t1 = 0;
  int ok = 1;
  for (uint64_t g = 0; g < len; g += 64) {
    const double p = g + lane < len ? prob_of(s[begin + g + lane]) : 0.0;
    const double q = ldexp(p, shift);  // exact: a power-of-two scaling
    if (!(q < 0x1p62)) ok = 0;
    const double kf = floor(q);
    const double fr = q - kf;
    const bool tie = fr == 0.5;
    const long long k = ok ? (long long)kf + (fr > 0.5 ? 1 : 0) : 0;  // a tie counts its floor
    long long inc = k;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long v = __shfl_up(inc, o, 64);
      if (lane >= o) inc += v;
    }
    const long long total = __shfl(inc, 63, 64);
    unsigned long long ties = __ballot(tie);
    long long x0 = 0, x1 = 0;  // extra units from ties rounded up, per start parity
    while (ties) {
      const int j = __ffsll((long long)ties) - 1;
      ties &= ties - 1;
      const long long kj = __shfl(k, j, 64);
      const long long before = __shfl(inc, j, 64) - kj;
      x0 += (0 + t0 + before + x0 + kj) & 1;
      x1 += (1 + t1 + before + x1 + kj) & 1;
    }
    t0 += total + x0;
    t1 += total + x1;
  }
  ok = __all(ok);
  if (lane == 0) {
    epred[b] = ok ? e : kNoBinade;
    delta[2 * b] = t0;
    delta[2 * b + 1] = t1;
  }
}

// The reference's sequential sum over one block, one add per amplitude in
// index order, starting from S (every lane holds the same S).
template <typename T>
__device__ double serial_block(const T *__restrict__ s, uint64_t begin, uint64_t len, double S,
                               int lane) {
  for (uint64_t g = 0; g < len; g += 256) {
    double p[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const uint64_t off = g + (uint64_t)r * 64 + lane;
      p[r] = off < len ? prob_of(s[begin + off]) : 0.0;  // S + 0 == S: padding is exact
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
      for (int j = 0; j < 64; j++) S = S + __shfl(p[r], j, 64);
  }
  return S;
}

// Exact block starts S_b (sstart[0..blocks]), one wave walking the blocks.
template <typename T>
__global__ __launch_bounds__(64) void k_chain(const T *__restrict__ s, uint64_t n, uint64_t chunk,
                                              int64_t blocks, const double *__restrict__ sums,
                                              const int *__restrict__ epred,
                                              const long long *__restrict__ delta, double *sstart) {
  const int lane = threadIdx.x;
  double S = 0;
  for (int64_t g0 = 0; g0 < blocks; g0 += 64) {
    const int64_t mine = g0 + lane;
    const bool valid = mine < blocks;
    const double msum = valid ? sums[mine] : 0.0;
    const int me = valid ? epred[mine] : kNoBinade;
    const long long d0 = valid ? delta[2 * mine] : 0, d1 = valid ? delta[2 * mine + 1] : 0;
    double mystart = 0;
    const int cnt = blocks - g0 < 64 ? (int)(blocks - g0) : 64;
    for (int j = 0; j < cnt; j++) {
      if (lane == j) mystart = S;
      if (__shfl(msum, j, 64) == 0) continue;  // every amplitude of the block is zero
      const int e = __shfl(me, j, 64);
      const long long dj0 = __shfl(d0, j, 64), dj1 = __shfl(d1, j, 64);
      bool done = false;
      if (e != kNoBinade && S >= 0x1p-1022 && ilogb(S) == e) {
        const double u = ldexp(1.0, e - 52);
        const long long m = (long long)(S / u);  // exact: S is a multiple of u below 2^(e+1)
        const long long t = m + ((m & 1) ? dj1 : dj0);
        if (t < (1ll << 53)) {  // the block ends inside the binade
          S = (double)t * u;
          done = true;
        }
      }
      if (!done) {
        const uint64_t begin = (uint64_t)(g0 + j) * chunk;
        S = serial_block(s, begin, begin + chunk < n ? chunk : n - begin, S, lane);
      }
    }
    if (valid) sstart[mine] = mystart;
  }
  if (lane == 0) sstart[blocks] = S;
}

// One thread per draw: first index whose sequential cumulative exceeds u * total.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_exact_search(const T *__restrict__ s, uint64_t n,
                                                         uint64_t chunk, int64_t blocks,
                                                         const double *__restrict__ sstart,
                                                         const double *__restrict__ u, int64_t shots,
                                                         long long last, long long *out) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= shots) return;
  const double total = sstart[blocks];
  const double draw = u[k] * total;
  if (!(draw < total)) {  // past the last cumulative: the last supported index
    out[k] = last;
    return;
  }
  int64_t lo = 0, hi = blocks - 1;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo + 1) / 2;
    if (sstart[mid] <= draw)
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint64_t begin = (uint64_t)lo * chunk;
  const uint64_t end = begin + chunk < n ? begin + chunk : n;
  double cum = sstart[lo];
  long long found = last;
  for (uint64_t i = begin; i < end; i++) {
    cum = cum + prob_of(s[i]);
    if (draw < cum) {
      found = (long long)i;
      break;
    }
  }
  out[k] = found;
}

int64_t cdf_search(State *s, const double *uniforms, int64_t shots, int64_t *outcomes, hipStream_t st,
                   double *total_out) {
  const uint64_t n = (uint64_t)s->length;
  const uint64_t chunk = std::min<uint64_t>((uint64_t)kChunk, n);
  const int64_t blocks = (int64_t)((n + chunk - 1) / chunk);
  // Last supported index and the mass check before the scratch below is taken.
  MassResult m = reduce_mass(s, st);
  if (!std::isfinite(m.total) || !(m.total > 0))
    fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
  // Scratch (doubles): sums[B], prefix[B+1], sstart[B+1], delta[2B], epred[B/2+1],
  // draws[shots], out[shots].
  const int64_t need = blocks + (blocks + 1) + (blocks + 1) + 2 * blocks + (blocks / 2 + 1) +
                       shots + shots + 8;
  double *buf = ctx_partials(s->ctx, need);
  double *sums = buf, *prefix = sums + blocks, *sstart = prefix + blocks + 1;
  long long *delta = (long long *)(sstart + blocks + 1);
  int *epred = (int *)(delta + 2 * blocks);
  double *draws = (double *)(epred + 2 * (blocks / 2 + 1));
  long long *out = (long long *)(draws + shots);
  const int dblocks = (int)((blocks + kBlock / 64 - 1) / (kBlock / 64));
  const int sblocks = (int)((shots + kBlock - 1) / kBlock);
  QM_HIP(hipMemcpyAsync(draws, uniforms, sizeof(double) * shots, hipMemcpyHostToDevice, st));
#define QM_CDF(TT)                                                                                   \
  k_chunk_sums<TT><<<(int)blocks, kBlock, 0, st>>>((const TT *)s->data, n, chunk, sums);            \
  k_scan<<<1, 1024, 0, st>>>(sums, blocks, prefix);                                                \
  k_block_delta<TT><<<dblocks, kBlock, 0, st>>>((const TT *)s->data, n, chunk, blocks, prefix,      \
                                                epred, delta);                                       \
  k_chain<TT><<<1, 64, 0, st>>>((const TT *)s->data, n, chunk, blocks, sums, epred, delta, sstart);  \
  k_exact_search<TT><<<sblocks, kBlock, 0, st>>>((const TT *)s->data, n, chunk, blocks, sstart,    \
                                                 draws, shots, (long long)m.last_supported, out);
  if (s->precision == QM_C128) {
    QM_CDF(C128)
  } else {
    QM_CDF(C64)
  }
#undef QM_CDF
  QM_HIP(hipGetLastError());
  double *pin = ctx_pinned(s->ctx, 1);
  QM_HIP(hipMemcpyAsync(pin, sstart + blocks, sizeof(double), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(outcomes, out, sizeof(long long) * shots, hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  if (total_out) *total_out = pin[0];
  return 0;
}


// ---------------------------------------------------------------------------
// Energy reductions without allocation (qm_state_energy_noalloc).
//
// The same kernels, block counts and fold order as reduce_mass /
// reduce_pauli, so the mass and every numerator are bit-identical to theirs,
// but every device buffer they need lives in one file-scope array of this
// library's code object: mass partials [blocks] | term partials
// [terms x blocks] | results [1 + terms] | term groups.  Host tables are on
// the stack, copies go to and from the caller's pageable memory.  Each device
// holds its own copy of the array, at its own address, from the first call of
// noalloc_setup on that device until the process ends.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kNaBlocks = 4096;  // reduction_blocks' maximum
constexpr int64_t kNaTerms = QM_NOALLOC_MAX_TERMS;
constexpr int64_t kNaGroupDoubles = (int64_t)((sizeof(PauliGroup) + 7) / 8);
constexpr int64_t kNaDoubles = kNaBlocks * (1 + kNaTerms) + (1 + kNaTerms) + kNaTerms * kNaGroupDoubles;

static __device__ __align__(256) double scratch[kNaDoubles];

constexpr int kNaDevices = 64;
struct NaSlot {
  std::mutex mu;           // held by a call from set-up to its last copy: calls on one device serialise
  double *ptr = nullptr;   // device address of `scratch` on this device
};
NaSlot g_na[kNaDevices];

}  // namespace

size_t noalloc_scratch_bytes() { return sizeof(double) * (size_t)kNaDoubles; }

void noalloc_setup(int device) {
  if (device < 0 || device >= kNaDevices) fail(QM_ERR_RANGE, "device index beyond the scratch table");
  NaSlot &slot = g_na[device];
  std::lock_guard<std::mutex> lk(slot.mu);
  if (slot.ptr) return;
  void *p = nullptr;
  // `scratch` names a host-side stand-in; the device copy's address comes from the runtime.
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    fail(QM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  slot.ptr = static_cast<double *>(p);
}

void reduce_energy_noalloc(State *s, const qm_pauli *terms, int64_t n_terms, double *total,
                           double *numerators, hipStream_t st) {
  if (n_terms < 0 || n_terms > kNaTerms) fail(QM_ERR_RANGE, "too many terms for the scratch buffer");
  const int dev = s->ctx->device;
  if (dev < 0 || dev >= kNaDevices) fail(QM_ERR_RANGE, "device index beyond the scratch table");
  NaSlot &slot = g_na[dev];
  std::lock_guard<std::mutex> lk(slot.mu);
  if (!slot.ptr) fail(QM_ERR_STATE, "the scratch buffer of this device is not set up");
  uint64_t per;
  const int64_t blocks = reduction_blocks((uint64_t)s->length, &per);
  double *mass_part = slot.ptr;
  double *term_part = mass_part + kNaBlocks;
  double *res = term_part + kNaBlocks * kNaTerms;
  PauliGroup *d_groups = reinterpret_cast<PauliGroup *>(res + 1 + kNaTerms);
  // Mass: k_mass's masked form with an empty mask sums every probability in
  // the same order as the total of reduce_mass.
  if (s->precision == QM_C128)
    k_mass<C128, 1><<<(int)blocks, kBlock, 0, st>>>((const C128 *)s->data, (uint64_t)s->length, per, 0, 0,
                                                    mass_part, nullptr);
  else
    k_mass<C64, 1><<<(int)blocks, kBlock, 0, st>>>((const C64 *)s->data, (uint64_t)s->length, per, 0, 0,
                                                   mass_part, nullptr);
  QM_HIP(hipGetLastError());
  k_fold<<<1, kBlock, 0, st>>>(mass_part, blocks, 1, 1, res, nullptr, nullptr);
  QM_HIP(hipGetLastError());
  // Term groups as reduce_pauli forms them: a stable order by flip mask
  // (std::sort on (flip, position) pairs, which needs no buffer).
  std::pair<uint64_t, int64_t> order[kNaTerms];
  for (int64_t i = 0; i < n_terms; i++) order[i] = {terms[i].flip_mask, i};
  std::sort(order, order + n_terms);
  PauliGroup groups[kNaTerms];
  int64_t ng = 0;
  for (int64_t k = 0; k < n_terms;) {
    PauliGroup g{};
    g.flip = order[k].first;
    g.first = (int32_t)k;
    while (k < n_terms && g.count < kTermsPerGroup && order[k].first == g.flip) {
      g.sign[g.count] = terms[order[k].second].sign_mask;
      g.y[g.count] = terms[order[k].second].y_count;
      g.count++;
      k++;
    }
    groups[ng++] = g;
  }
  if (ng > 0) {
    QM_HIP(hipMemcpyAsync(d_groups, groups, (size_t)ng * sizeof(PauliGroup), hipMemcpyHostToDevice, st));
    dim3 grid((unsigned)blocks, (unsigned)ng);
    if (s->precision == QM_C128)
      k_pauli<C128><<<grid, kBlock, 0, st>>>((const C128 *)s->data, (uint64_t)s->length, per, d_groups,
                                             term_part, blocks);
    else
      k_pauli<C64><<<grid, kBlock, 0, st>>>((const C64 *)s->data, (uint64_t)s->length, per, d_groups,
                                            term_part, blocks);
    QM_HIP(hipGetLastError());
    k_fold<<<(unsigned)n_terms, kBlock, 0, st>>>(term_part, blocks, 1, blocks, res + 1, nullptr, nullptr);
    QM_HIP(hipGetLastError());
  }
  double out[1 + kNaTerms];
  QM_HIP(hipMemcpyAsync(out, res, sizeof(double) * (size_t)(1 + n_terms), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  *total = out[0];
  for (int64_t k = 0; k < n_terms; k++) numerators[order[k].second] = out[1 + k];
}

}  // namespace qm
