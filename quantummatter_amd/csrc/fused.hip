// Fused tile-pass kernel of libqmhip.so (gfx950).
//
// One workgroup owns one tile of 2^K amplitudes at a time (grid-stride over
// tiles).  Each of the 2^L threads keeps 2^R amplitudes in registers; a gate
// whose target is a register bit runs entirely in registers, a diagonal gate
// runs wherever its target lives, and the set of register bits changes
// between phases by an LDS transpose.  The pass reads and writes every
// amplitude of the state exactly once.  The plan (plan.cpp) supplies, per pass,
// the tile qubits, the register set of each phase and the ops of each phase.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "dev_types.hpp"
#include "gate_math.hpp"
#include "qm_internal.hpp"
#include "tile_ops.hpp"

namespace qm {

template <typename T, int L, int R>
__global__ __launch_bounds__(1 << L, 2) void k_fused(T *__restrict__ s, const DevPass *__restrict__ pass,
                                                     const TileOp *__restrict__ prog, uint64_t tiles) {
  using F = typename Real<T>::type;
  constexpr int REGS = 1 << R;
  extern __shared__ char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const DevPass &d = *pass;
  const uint32_t t = threadIdx.x;
  uint64_t load_lane = 0, store_lane = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    const bool on = (t >> k) & 1;
    load_lane |= on ? d.load_lane_off[k] : 0ull;
    store_lane |= on ? d.store_lane_off[k] : 0ull;
  }
  const int nph = d.nphases;
  const bool load_direct = d.load_direct != 0, store_direct = d.store_direct != 0;
  const TileOp *__restrict__ ops = prog + d.op_offset;
  F vr[REGS], vi[REGS];

  for (uint64_t tau = blockIdx.x; tau < tiles; tau += gridDim.x) {
    uint64_t base = 0;
    for (int k = 0; k < d.nruns; k++)
      base |= ((tau >> d.run_src[k]) & ((1ull << d.run_len[k]) - 1)) << d.run_dst[k];
    const T *src = s + (base | load_lane);
#pragma unroll
    for (int r = 0; r < REGS; r++) {
      const T a = src[d.load_reg_off[r]];
      vr[r] = a.re;
      vi[r] = a.im;
    }
    int cur = load_direct ? 0 : kIo;
    for (int ph = 0; ph < nph; ph++) {
      if (cur != ph) lds_transpose<T, F, L, REGS>(lds, d, cur, ph, t, vr, vi);
      cur = ph;
      const TileOp *__restrict__ po = ops + d.phase_op_begin[ph];
      const int cnt = d.phase_op_count[ph];
      for (int o = 0; o < cnt; o++) apply_tile_op<F, REGS>(po + o, vr, vi, t, base);
      const int dcnt = d.phase_diag_count[ph];
      if (dcnt) {
        // Merged lane/outside diagonals: one per-lane product, one multiplication.
        F pr, pi;
        phase_product<F>(ops + d.phase_diag_begin[ph], dcnt, t, base, pr, pi);
#pragma unroll
        for (int r = 0; r < REGS; r++) mul_phase_fma<F>(pr, pi, vr[r], vi[r]);
      }
    }
    if (!store_direct) lds_transpose<T, F, L, REGS>(lds, d, cur, kIo, t, vr, vi);
    T *dst = s + (base | store_lane);
#pragma unroll
    for (int r = 0; r < REGS; r++) {
      T a;
      a.re = vr[r];
      a.im = vi[r];
      dst[d.store_reg_off[r]] = a;
    }
  }
}

template <typename T, int L, int R>
static void launch_fused_t(State *st, const void *d_pass, const void *d_ops, hipStream_t stream) {
  constexpr int K = L + R;
  const uint64_t tiles = 1ull << (st->qubits - K);
  const size_t lds = sizeof(T) << K;
  // Workgroups per CU for this instance, set up once per process (thread-safe
  // initialisation: several lanes may launch concurrently).
  static const int resident = [&] {
    QM_HIP(hipFuncSetAttribute((const void *)k_fused<T, L, R>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    QM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused<T, L, R>, 1 << L, lds));
    return std::max(per_cu, 1);
  }();
  // Persistent grid: every workgroup resident, each walking its share of tiles.
  const uint64_t want = (uint64_t)st->ctx->compute_units * resident;
  const int blocks = (int)std::min<uint64_t>(tiles, std::max<uint64_t>(want, 1));
  k_fused<T, L, R><<<blocks, 1 << L, lds, stream>>>((T *)st->data, (const DevPass *)d_pass,
                                                     (const TileOp *)d_ops, tiles);
  QM_HIP(hipGetLastError());
}

// Compiled tile shapes (lane bits L, register bits R): K = L + R tile qubits.
#define QM_FUSED_SHAPES(X) \
  X(C128, 8, 4) X(C128, 7, 4) X(C128, 9, 3) X(C128, 9, 4) X(C128, 8, 3) \
  X(C64, 8, 5) X(C64, 8, 4) X(C64, 9, 4) X(C64, 7, 5) X(C64, 9, 3)

bool fused_shape_supported(int precision, int L, int R) {
#define QM_SUPPORTED(T, l, r) \
  if ((precision == QM_C128) == std::is_same<T, C128>::value && L == l && R == r) return true;
  QM_FUSED_SHAPES(QM_SUPPORTED)
#undef QM_SUPPORTED
  return false;
}

FusedShape default_fused_shape(int precision) {
  // complex128: 4096-amplitude tiles (64 KiB LDS), 256 threads x 16 registers.
  // complex64: 8192-amplitude tiles (64 KiB LDS), 256 threads x 32 registers.
  if (precision == QM_C128) return FusedShape{12, 8, 4, 3};
  return FusedShape{13, 8, 5, 4};
}

void launch_fused_pass(State *s, const FusedShape &shape, const void *d_pass, const void *d_ops,
                       hipStream_t st) {
#define QM_LAUNCH(T, l, r)                                                           \
  if ((s->precision == QM_C128) == std::is_same<T, C128>::value && shape.L == l && \
      shape.R == r) {                                                                \
    launch_fused_t<T, l, r>(s, d_pass, d_ops, st);                                   \
    return;                                                                          \
  }
  QM_FUSED_SHAPES(QM_LAUNCH)
#undef QM_LAUNCH
  fail(QM_ERR_UNSUPPORTED, "fused tile shape is not compiled");
}

}  // namespace qm
