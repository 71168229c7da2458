// Dense-block passes on the matrix cores (gfx950 FP64 MFMA).
//
// A block of consecutive gates on k <= 5 qubits is one 2^k x 2^k complex
// matrix U.  Viewing the state as 2^(n-k) columns of 2^k amplitudes (the
// block qubits varying within a column), the pass computes Y = U X.  In real
// form, [Re Y; Im Y] = [[Re U, -Im U], [Im U, Re U]] [Re X; Im X]: an
// M x M by M x (columns) product, M = 2^(k+1), done 16 columns at a time per
// wavefront with v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16 per issue).
// The matrix sits in LDS; each lane loads the four real values of the B
// operand it owns straight from the state and stores the D rows it holds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dense.hpp"
#include "qm_internal.hpp"

namespace qm {
namespace {

typedef double qm_double4 __attribute__((ext_vector_type(4)));

struct DenseArgs {
  uint64_t ncols;         // 2^(n - k)
  uint64_t loc_off[32];   // state offset of local index l (block qubits)
  int32_t nruns;          // runs of non-block qubits: column index -> state base
  uint8_t run_src[8], run_dst[8], run_len[8];
};

// Lane layout of v_mfma_f64_16x16x4f64 (established on the device by a probe
// that matched every output against the product): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[l / 16 + 4 i][l % 16]
// in register i = 0..3.
template <int K>
__global__ __launch_bounds__(256) void k_dense(double *__restrict__ s, const double *__restrict__ A,
                                               const DenseArgs args) {
  constexpr int D = 1 << K, M = 2 * D, MT = (M + 15) / 16, KS = M / 4;
  __shared__ double a_lds[M * M];
  for (int i = threadIdx.x; i < M * M; i += blockDim.x) a_lds[i] = A[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane & 15, grp = lane >> 4;
  const uint64_t tiles = (args.ncols + 15) / 16;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t ct = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wave; ct < tiles; ct += waves) {
    const uint64_t col = ct * 16 + cl;
    const bool valid = col < args.ncols;
    uint64_t base = 0;
    for (int r = 0; r < args.nruns; r++)
      base |= ((col >> args.run_src[r]) & ((1ull << args.run_len[r]) - 1)) << args.run_dst[r];
    double b[KS];
#pragma unroll
    for (int kk = 0; kk < KS; kk++) {
      const int row = 4 * kk + grp;  // < M
      const int local = row & (D - 1), part = row >> K;
      b[kk] = valid ? s[2 * (base | args.loc_off[local]) + part] : 0.0;
    }
    qm_double4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[mt] = qm_double4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KS; kk++) {
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        const int arow = 16 * mt + cl;
        const double a = arow < M ? a_lds[arow * M + 4 * kk + grp] : 0.0;
        acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[kk], acc[mt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 16 * mt + grp + 4 * i;
        if (row < M && valid) {
          const int local = row & (D - 1), part = row >> K;
          s[2 * (base | args.loc_off[local]) + part] = acc[mt][i];
        }
      }
    }
  }
}

template <int K>
void launch_dense_t(State *st, const DenseArgs &a, const double *d_matrix, hipStream_t stream) {
  const uint64_t tiles = (a.ncols + 15) / 16;
  const uint64_t blocks = std::min<uint64_t>((tiles + 3) / 4, (uint64_t)st->ctx->compute_units * 8);
  k_dense<K><<<(unsigned)std::max<uint64_t>(blocks, 1), 256, 0, stream>>>((double *)st->data, d_matrix, a);
  QM_HIP(hipGetLastError());
}

}  // namespace

void launch_dense(State *st, const DenseBlock &b, const double *d_matrix, hipStream_t stream) {
  if (st->precision != QM_C128) fail(QM_ERR_UNSUPPORTED, "dense blocks run on complex128 states");
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
  switch (b.k) {
    case 1: launch_dense_t<1>(st, a, d_matrix, stream); break;
    case 2: launch_dense_t<2>(st, a, d_matrix, stream); break;
    case 3: launch_dense_t<3>(st, a, d_matrix, stream); break;
    case 4: launch_dense_t<4>(st, a, d_matrix, stream); break;
    case 5: launch_dense_t<5>(st, a, d_matrix, stream); break;
    default: fail(QM_ERR_UNSUPPORTED, "dense block width must be 1..5");
  }
}

}  // namespace qm
