// Dense-block passes on the matrix cores (gfx950 MFMA: FP64 for complex128
// states, FP32 for complex64 states).
//
// A block of consecutive gates on k <= 5 qubits is one 2^k x 2^k complex
// matrix U.  Viewing the state as 2^(n-k) columns of 2^k amplitudes (the
// block qubits varying within a column), the pass computes Y = U X.  In real
// form, [Re Y; Im Y] = [[Re U, -Im U], [Im U, Re U]] [Re X; Im X]: an
// M x M by M x (columns) product, M = 2^(k+1), done 16 columns at a time per
// wavefront with v_mfma_f64_16x16x4f64 or v_mfma_f32_16x16x4f32 (A 16x4,
// B 4x16, D 16x16 per issue).  The matrix sits in LDS.  Row r and row r + 2^k
// of the real form are the real and imaginary parts of the same amplitude,
// so for k >= 2 each lane loads whole complex amplitudes (one 16- or 8-byte
// access for two B rows) and, where the output layout puts both parts in the
// same lane, stores whole complex amplitudes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "dense.hpp"
#include "dev_types.hpp"
#include "qm_internal.hpp"

namespace qm {
namespace {

typedef double qm_double4 __attribute__((ext_vector_type(4)));
typedef float qm_float4 __attribute__((ext_vector_type(4)));

struct DenseArgs {
  uint64_t ncols;         // 2^(n - k)
  uint64_t loc_off[32];   // state offset of local index l (block qubits)
  int32_t nruns;          // runs of non-block qubits: column index -> state base
  uint8_t run_src[8], run_dst[8], run_len[8];
};

// Lane layouts (lane l, output register i): both instructions take
// A[l % 16][l / 16] and B[l / 16][l % 16] from lane l.  The FP64 instruction
// leaves D[l / 16 + 4 i][l % 16] in register i (established on the device by a
// probe that matched every output against the product); the FP32 one leaves
// D[4 (l / 16) + i][l % 16] (checked against the host reference by the
// complex64 dense tests).
template <typename T> struct DenseMfma;
template <> struct DenseMfma<C128> {
  typedef double R;
  typedef qm_double4 Acc;
  static __device__ __forceinline__ Acc mfma(R a, R b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static constexpr __host__ __device__ int row(int mt, int grp, int i) { return 16 * mt + grp + 4 * i; }
};
template <> struct DenseMfma<C64> {
  typedef float R;
  typedef qm_float4 Acc;
  static __device__ __forceinline__ Acc mfma(R a, R b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static constexpr __host__ __device__ int row(int mt, int grp, int i) { return 16 * mt + 4 * grp + i; }
};

// Where the imaginary part of output row `row` (< D) sits: the register and
// tile that hold row + D in the same lane, or -1 when it is in another lane.
template <typename T, int K>
__device__ __forceinline__ constexpr int partner_reg(int mt, int i) {
  constexpr int D = 1 << K;
  if (D >= 16) return i;  // tile mt + D / 16, same register
  if (DenseMfma<T>::row(0, 0, 1) == 4) return i + D / 4;  // FP64 layout: rows step by 4 per register
  return -1;
}

// TPI column tiles per loop iteration, their loads issued together so more
// of the state is in flight per wave.  The best count depends on the width
// and the precision (the registers the matrix operand and the loads take set
// the occupancy); dense_tiles holds the measured choice.
template <typename T, int K, int TPI>
__global__ __launch_bounds__(256) void k_dense(T *__restrict__ s, const typename DenseMfma<T>::R *__restrict__ A,
                                               const DenseArgs args) {
  using Mf = DenseMfma<T>;
  using R = typename Mf::R;
  using Acc = typename Mf::Acc;
  constexpr int D = 1 << K, M = 2 * D, MT = (M + 15) / 16, KS = M / 4;
  __shared__ R a_lds[M * M];
  for (int i = threadIdx.x; i < M * M; i += blockDim.x) a_lds[i] = A[i];
  __syncthreads();
  R *const sr = reinterpret_cast<R *>(s);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane & 15, grp = lane >> 4;
  const uint64_t tiles = (args.ncols + 15) / 16;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  // Column base of a lane's column in tile ct (state bits of the non-block
  // qubits), and whether the column exists.
  auto column = [&](uint64_t ct, uint64_t &base) {
    const uint64_t col = ct * 16 + cl;
    base = 0;
    for (int r = 0; r < args.nruns; r++)
      base |= ((col >> args.run_src[r]) & ((1ull << args.run_len[r]) - 1)) << args.run_dst[r];
    return col < args.ncols;
  };
  auto load_b = [&](R (&b)[KS], bool valid, uint64_t base) {
    if constexpr (K >= 2) {
      // B rows 4 kk + grp (kk < KS / 2) are the real parts of local index
      // 4 kk + grp, rows 4 kk + grp + D their imaginary parts.
#pragma unroll
      for (int kk = 0; kk < KS / 2; kk++) {
        T v{0, 0};
        if (valid) v = s[base | args.loc_off[4 * kk + grp]];
        b[kk] = v.re;
        b[kk + KS / 2] = v.im;
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < KS; kk++) {
        const int row = 4 * kk + grp;
        const int local = row & (D - 1), part = row >> K;
        b[kk] = valid ? sr[2 * (base | args.loc_off[local]) + part] : (R)0;
      }
    }
  };
  // Y = U X for one tile: the products, then the stores of a valid column.
  auto product = [&](const R (&b)[KS], bool valid, uint64_t base) {
    Acc acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[mt] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KS; kk++) {
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        const int arow = 16 * mt + cl;
        const R a = arow < M ? a_lds[arow * M + 4 * kk + grp] : (R)0;
        acc[mt] = Mf::mfma(a, b[kk], acc[mt]);
      }
    }
    if (!valid) return;
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = Mf::row(mt, grp, i);
        if (row >= M) continue;
        constexpr bool paired = K >= 2 && partner_reg<T, K>(0, 0) >= 0;
        if constexpr (paired) {
          if (row >= D) continue;  // written with its real part
          const int mt2 = D >= 16 ? mt + D / 16 : mt;
          const int i2 = partner_reg<T, K>(mt, i);
          T v;
          v.re = acc[mt][i];
          v.im = acc[mt2][i2];
          s[base | args.loc_off[row]] = v;
        } else {
          const int local = row & (D - 1), part = row >> K;
          sr[2 * (base | args.loc_off[local]) + part] = acc[mt][i];
        }
      }
    }
  };
  const uint64_t first = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  for (uint64_t ct0 = first; ct0 < tiles; ct0 += waves * TPI) {
    R b[TPI][KS];
    bool valid[TPI];
    uint64_t base[TPI];
#pragma unroll
    for (int u = 0; u < TPI; u++) {
      valid[u] = false;
      base[u] = 0;
      if (ct0 + u * waves < tiles) valid[u] = column(ct0 + u * waves, base[u]);
      load_b(b[u], valid[u], base[u]);
    }
#pragma unroll
    for (int u = 0; u < TPI; u++)
      if (ct0 + u * waves < tiles) product(b[u], valid[u], base[u]);
  }
}

// Column tiles per iteration by precision and width, measured on the
// 30-qubit layer circuit with blocks of k = 2..5 (tuning knob QM_DENSE_TPI:
// 1, 2 or 4 for every width).
int dense_tiles(int precision, int k) {
  static const int knob = [] {
    const char *e = std::getenv("QM_DENSE_TPI");
    return e ? std::atoi(e) : 0;
  }();
  if (knob == 1 || knob == 2 || knob == 4) return knob;
  if (precision == QM_C128) return k <= 2 ? 4 : 2;
  return k <= 3 ? 4 : k == 4 ? 2 : 1;
}

template <typename T, int K, int TPI>
void launch_dense_v(State *st, const DenseArgs &a, const void *d_matrix, hipStream_t stream) {
  const uint64_t tiles = (a.ncols + 15) / 16;
  const uint64_t blocks = std::min<uint64_t>((tiles + 3) / 4, (uint64_t)st->ctx->compute_units * 8);
  k_dense<T, K, TPI><<<(unsigned)std::max<uint64_t>(blocks, 1), 256, 0, stream>>>(
      (T *)st->data, (const typename DenseMfma<T>::R *)d_matrix, a);
  QM_HIP(hipGetLastError());
}

template <typename T, int K>
void launch_dense_t(State *st, const DenseArgs &a, const void *d_matrix, hipStream_t stream) {
  switch (dense_tiles(st->precision, K)) {
    case 4: launch_dense_v<T, K, 4>(st, a, d_matrix, stream); break;
    case 2: launch_dense_v<T, K, 2>(st, a, d_matrix, stream); break;
    default: launch_dense_v<T, K, 1>(st, a, d_matrix, stream);
  }
}

template <typename T>
void launch_dense_k(State *st, int k, const DenseArgs &a, const void *d_matrix, hipStream_t stream) {
  switch (k) {
    case 1: launch_dense_t<T, 1>(st, a, d_matrix, stream); break;
    case 2: launch_dense_t<T, 2>(st, a, d_matrix, stream); break;
    case 3: launch_dense_t<T, 3>(st, a, d_matrix, stream); break;
    case 4: launch_dense_t<T, 4>(st, a, d_matrix, stream); break;
    case 5: launch_dense_t<T, 5>(st, a, d_matrix, stream); break;
    default: fail(QM_ERR_UNSUPPORTED, "dense block width must be 1..5");
  }
}

}  // namespace

void launch_dense(State *st, const DenseBlock &b, const void *d_matrix, hipStream_t stream) {
  DenseArgs a{};
  const int n = st->qubits;
  a.ncols = 1ull << (n - b.k);
  int l = 0;
  std::vector<int> qpos;
  for (int q = 0; q < n; q++)
    if (b.qmask & (1ull << q)) qpos.push_back(q);
  for (int loc = 0; loc < (1 << b.k); loc++) {
    uint64_t off = 0;
    for (int i = 0; i < b.k; i++)
      if (loc & (1 << i)) off |= 1ull << qpos[i];
    a.loc_off[loc] = off;
  }
  // Runs of non-block qubits, low to high: column bit src..src+len -> state bit dst...
  int src = 0;
  for (int q = 0; q < n;) {
    if (b.qmask & (1ull << q)) {
      q++;
      continue;
    }
    const int start = q;
    while (q < n && !(b.qmask & (1ull << q))) q++;
    a.run_src[l] = (uint8_t)src;
    a.run_dst[l] = (uint8_t)start;
    a.run_len[l] = (uint8_t)(q - start);
    src += q - start;
    l++;
  }
  a.nruns = l;
  if (st->precision == QM_C128)
    launch_dense_k<C128>(st, b.k, a, d_matrix, stream);
  else
    launch_dense_k<C64>(st, b.k, a, d_matrix, stream);
}

}  // namespace qm
