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
#include <cstdlib>
#include <type_traits>

#include "dev_types.hpp"
#include "gate_math.hpp"
#include "qm_internal.hpp"
#include "tile_ops.hpp"

namespace qm {

// MODE bit 0 (split): transposes through 2^K reals (real, then imaginary
// parts), halving the LDS footprint so more workgroups share a CU.
// MODE bit 1 (prefetch): the next tile's loads are issued before the current
// tile is computed, so their latency overlaps the arithmetic (for R <= 3, where
// a second register set fits).
template <typename T, int L, int R, int MODE>
__global__ __launch_bounds__(1 << L, (MODE & 1) ? 3 : 2) void k_fused(T *__restrict__ s,
                                                                      const DevPass *__restrict__ pass,
                                                                      const TileOp *__restrict__ prog,
                                                                      uint64_t tiles) {
  using F = typename Real<T>::type;
  constexpr int REGS = 1 << R;
  constexpr bool SPLIT = (MODE & 1) != 0;
  constexpr bool PF = (MODE & 2) != 0;
  extern __shared__ char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  F *ldsf = reinterpret_cast<F *>(lds_raw);
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
  F nr[PF ? REGS : 1], ni[PF ? REGS : 1];
  auto tile_base = [&](uint64_t tau) {
    uint64_t b = 0;
    for (int k = 0; k < d.nruns; k++)
      b |= ((tau >> d.run_src[k]) & ((1ull << d.run_len[k]) - 1)) << d.run_dst[k];
    return b;
  };
  uint64_t nbase = tile_base(blockIdx.x);
  if constexpr (PF) {
    if (blockIdx.x < tiles) {
      const T *src = s + (nbase | load_lane);
#pragma unroll
      for (int r = 0; r < REGS; r++) {
        const T a = src[d.load_reg_off[r]];
        nr[r] = a.re;
        ni[r] = a.im;
      }
    }
  }

  for (uint64_t tau = blockIdx.x; tau < tiles; tau += gridDim.x) {
    const uint64_t base = nbase;
    if constexpr (PF) {
#pragma unroll
      for (int r = 0; r < REGS; r++) {
        vr[r] = nr[r];
        vi[r] = ni[r];
      }
      const uint64_t next = tau + gridDim.x;
      if (next < tiles) {
        nbase = tile_base(next);
        const T *src = s + (nbase | load_lane);
#pragma unroll
        for (int r = 0; r < REGS; r++) {
          const T a = src[d.load_reg_off[r]];
          nr[r] = a.re;
          ni[r] = a.im;
        }
      }
    } else {
      const T *src = s + (base | load_lane);
#pragma unroll
      for (int r = 0; r < REGS; r++) {
        const T a = src[d.load_reg_off[r]];
        vr[r] = a.re;
        vi[r] = a.im;
      }
      nbase = tile_base(tau + gridDim.x);
    }
    int cur = load_direct ? 0 : kIo;
    for (int ph = 0; ph < nph; ph++) {
      if (cur != ph) {
        if constexpr (SPLIT)
          lds_transpose_split<F, L, REGS>(ldsf, d, cur, ph, t, vr, vi);
        else
          lds_transpose<T, F, L, REGS>(lds, d, cur, ph, t, vr, vi);
      }
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
    if (!store_direct) {
      if constexpr (SPLIT)
        lds_transpose_split<F, L, REGS>(ldsf, d, cur, kIoStore, t, vr, vi);
      else
        lds_transpose<T, F, L, REGS>(lds, d, cur, kIoStore, t, vr, vi);
    }
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

template <typename T, int L, int R, int MODE>
static void launch_fused_t(State *st, const void *d_pass, const void *d_ops, hipStream_t stream) {
  constexpr int K = L + R;
  const uint64_t tiles = 1ull << (st->qubits - K);
  const size_t lds = ((MODE & 1) ? sizeof(typename Real<T>::type) : sizeof(T)) << K;
  // Workgroups per CU for this instance, set up once per process (thread-safe
  // initialisation: several lanes may launch concurrently).
  static const int resident = [&] {
    QM_HIP(hipFuncSetAttribute((const void *)k_fused<T, L, R, MODE>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    QM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused<T, L, R, MODE>, 1 << L, lds));
    return std::max(per_cu, 1);
  }();
  // Persistent grid: every workgroup resident, each walking its share of tiles.
  const uint64_t want = (uint64_t)st->ctx->compute_units * resident;
  const int blocks = (int)std::min<uint64_t>(tiles, std::max<uint64_t>(want, 1));
  k_fused<T, L, R, MODE><<<blocks, 1 << L, lds, stream>>>((T *)st->data, (const DevPass *)d_pass,
                                                            (const TileOp *)d_ops, tiles);
  QM_HIP(hipGetLastError());
}

// Compiled tile shapes (lane bits L, register bits R, MODE): K = L + R tile
// qubits.  The first entry of a shape is its default mode.
#define QM_FUSED_SHAPES(X)                                                                  \
  X(C128, 8, 4, 0) X(C128, 7, 4, 0) X(C128, 9, 3, 3) X(C128, 9, 3, 1) X(C128, 9, 3, 0)        \
  X(C128, 8, 3, 3) X(C128, 8, 3, 1) X(C128, 10, 3, 3) X(C128, 10, 3, 1)                      \
  X(C64, 8, 5, 0) X(C64, 8, 4, 3) X(C64, 8, 4, 0) X(C64, 9, 4, 3) X(C64, 9, 3, 3) X(C64, 7, 5, 0)

bool fused_shape_supported(int precision, int L, int R) {
#define QM_SUPPORTED(T, l, r, m) \
  if ((precision == QM_C128) == std::is_same<T, C128>::value && L == l && R == r) return true;
  QM_FUSED_SHAPES(QM_SUPPORTED)
#undef QM_SUPPORTED
  return false;
}

std::vector<FusedShape> fused_shapes(int precision) {
  std::vector<FusedShape> out;
  const int contiguous = precision == QM_C128 ? 3 : 4;
#define QM_LIST(T, l, r, m)                                                   \
  if ((precision == QM_C128) == std::is_same<T, C128>::value)                \
    out.push_back(FusedShape{l + r, l, r, contiguous, m});
  QM_FUSED_SHAPES(QM_LIST)
#undef QM_LIST
  return out;
}

FusedShape default_fused_shape(int precision) {
  // complex128: 4096-amplitude tiles (64 KiB LDS), 256 threads x 16 registers.
  // complex64: 8192-amplitude tiles (64 KiB LDS), 256 threads x 32 registers.
  if (precision == QM_C128) return FusedShape{12, 8, 4, 3};
  return FusedShape{13, 8, 5, 4};
}

static int mode_override() {
  static const int mode = [] {
    const char *e = std::getenv("QM_FUSED_MODE");  // tuning knob: -1 = per-shape default
    return e ? std::atoi(e) : -1;
  }();
  return mode;
}

void launch_fused_pass(State *s, const FusedShape &shape, const void *d_pass, const void *d_ops,
                       hipStream_t st) {
  const int want = shape.variant >= 0 ? shape.variant : mode_override();
  bool seen = false;
  // First pass: the requested mode; second: the shape's default (first listed).
  for (int round = 0; round < 2; round++) {
#define QM_LAUNCH(T, l, r, m)                                                              \
    if ((s->precision == QM_C128) == std::is_same<T, C128>::value && shape.L == l &&     \
        shape.R == r && (round == 1 || m == want)) {                                     \
      launch_fused_t<T, l, r, m>(s, d_pass, d_ops, st);                                  \
      return;                                                                            \
    }
    QM_FUSED_SHAPES(QM_LAUNCH)
#undef QM_LAUNCH
    (void)seen;
  }
  fail(QM_ERR_UNSUPPORTED, "fused tile shape is not compiled");
}

}  // namespace qm
