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

namespace qm {

// ---- register-resident gate application -------------------------------------
// Every op is straight-line code over the registers: the register bit of the
// target is a template parameter (dispatched once per op), and control bits
// held in registers select, per pair, either the gate or the identity by a
// scalar select of coefficients -- both exact, so there are no per-register
// branches and no register copies around them.

template <typename F, int REGS, int B, int KIND>
__device__ __forceinline__ void reg_pair(F (&vr)[REGS], F (&vi)[REGS], F c0, F c1) {
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    if constexpr (true) {
      if (r & (1 << B)) continue;
      const int r1 = r | (1 << B);
      if constexpr (KIND == G_RX) pair_rx_fma<F>(c0, c1, vr[r], vi[r], vr[r1], vi[r1]);
      if constexpr (KIND == G_RY) pair_ry_fma<F>(c0, c1, vr[r], vi[r], vr[r1], vi[r1]);
      if constexpr (KIND == G_H) pair_h_fma<F>(c0, vr[r], vi[r], vr[r1], vi[r1]);
      if constexpr (KIND == G_Y) pair_y<F>(vr[r], vi[r], vr[r1], vi[r1]);
      if constexpr (KIND == G_X) {
        const F tr = vr[r], ti = vi[r];
        vr[r] = vr[r1];
        vi[r] = vi[r1];
        vr[r1] = tr;
        vi[r1] = ti;
      }
    }
  }
}

// General 2x2 (the reference's form) with a register control mask m.
template <typename F, int REGS, int B>
__device__ __forceinline__ void reg_mat(F (&vr)[REGS], F (&vi)[REGS], uint32_t m, const F (&c)[8]) {
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    if (r & (1 << B)) continue;
    const int r1 = r | (1 << B);
    const bool on = ((uint32_t)r & m) == m;
    const F k[8] = {on ? c[0] : (F)1, on ? c[1] : (F)0, on ? c[2] : (F)0, on ? c[3] : (F)0,
                    on ? c[4] : (F)0, on ? c[5] : (F)0, on ? c[6] : (F)1, on ? c[7] : (F)0};
    pair_mat_fma<F>(k, vr[r], vi[r], vr[r1], vi[r1]);
  }
}

// Controlled exchange with a register control mask m.
template <typename F, int REGS, int B>
__device__ __forceinline__ void reg_cx(F (&vr)[REGS], F (&vi)[REGS], uint32_t m) {
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    if (r & (1 << B)) continue;
    const int r1 = r | (1 << B);
    const bool on = ((uint32_t)r & m) == m;
    const F ar = vr[r], ai = vi[r], br = vr[r1], bi = vi[r1];
    vr[r] = on ? br : ar;
    vi[r] = on ? bi : ai;
    vr[r1] = on ? ar : br;
    vi[r1] = on ? ai : bi;
  }
}

// Diagonal on register bit B: A on the zero half (skipped for diag(1, D)), D on
// the one half; registers failing the control mask m get the identity.
template <typename F, int REGS, int B, bool FULL>
__device__ __forceinline__ void reg_diag(F (&vr)[REGS], F (&vi)[REGS], uint32_t m, F ar, F ai, F dr,
                                         F di) {
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    const bool on = ((uint32_t)r & m) == m;
    if (r & (1 << B)) {
      mul_phase_fma<F>(on ? dr : (F)1, on ? di : (F)0, vr[r], vi[r]);
    } else if constexpr (FULL) {
      mul_phase_fma<F>(on ? ar : (F)1, on ? ai : (F)0, vr[r], vi[r]);
    }
  }
}

// One phase for every register (target on a lane bit or outside the tile).
template <typename F, int REGS>
__device__ __forceinline__ void reg_phase_all(F (&vr)[REGS], F (&vi)[REGS], uint32_t m, F pr, F pi) {
  if (m == 0) {
#pragma unroll
    for (int r = 0; r < REGS; r++) mul_phase_fma<F>(pr, pi, vr[r], vi[r]);
    return;
  }
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    const bool on = ((uint32_t)r & m) == m;
    mul_phase_fma<F>(on ? pr : (F)1, on ? pi : (F)0, vr[r], vi[r]);
  }
}

#define QM_BIT_SWITCH(b, CALL)             \
  switch (b) {                             \
    case 0: { constexpr int BB = 0; CALL; } break; \
    case 1: { constexpr int BB = 1; CALL; } break; \
    case 2: { constexpr int BB = 2; CALL; } break; \
    case 3: { constexpr int BB = 3; CALL; } break; \
    default: if constexpr (REGS > 16) { constexpr int BB = 4; CALL; } break; \
  }

template <typename F, int REGS>
__device__ __forceinline__ int apply_tile_op(const TileOp *__restrict__ op, F (&vr)[REGS],
                                             F (&vi)[REGS], uint32_t t, uint64_t base) {
  const int kind = op->kind;
  const int rbit = op->rbit;
  const uint32_t m = op->reg_cmask;
  if ((base & op->out_cmask) != op->out_cmask) return 0;
  if ((t & op->lane_cmask) != op->lane_cmask) return 0;
  if (rbit >= 0) {
    if (kind == G_X) {
      if (m == 0) {
        QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_X>(vr, vi, 0, 0)))
      } else {
        QM_BIT_SWITCH(rbit, (reg_cx<F, REGS, BB>(vr, vi, m)))
      }
      return 0;
    }
    if (kind == G_MAT) {
      F c[8];
#pragma unroll
      for (int i = 0; i < 8; i++) c[i] = (F)op->c[i];
      QM_BIT_SWITCH(rbit, (reg_mat<F, REGS, BB>(vr, vi, m, c)))
      return 0;
    }
    const F c0 = (F)op->c[0], c1 = (F)op->c[1];
    switch (kind) {
      case G_RX: QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_RX>(vr, vi, c0, c1))) break;
      case G_RY: QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_RY>(vr, vi, c0, c1))) break;
      case G_H: QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_H>(vr, vi, c0, c1))) break;
      default: QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_Y>(vr, vi, c0, c1))) break;
    }
    return 0;
  }
  const bool full = kind == G_DIAG;
  const F ar = (F)op->c[0], ai = (F)op->c[1], dr = (F)op->c[6], di = (F)op->c[7];
  const uint32_t dreg = op->dreg;
  if (dreg != 0) {
    const int b = __builtin_ctz(dreg);
    if (full) {
      QM_BIT_SWITCH(b, (reg_diag<F, REGS, BB, true>(vr, vi, m, ar, ai, dr, di)))
    } else {
      QM_BIT_SWITCH(b, (reg_diag<F, REGS, BB, false>(vr, vi, m, ar, ai, dr, di)))
    }
    return 0;
  }
  // Target on a lane bit (per-lane choice) or outside the tile (uniform choice).
  const bool hi = ((base & op->dout) != 0) || ((t & op->dlane) != 0);
  if (!full && op->dlane == 0 && !hi) return 0;  // diag(1, D) on a zero half: nothing to do
  const F pr = hi ? dr : (full ? ar : (F)1);
  const F pi = hi ? di : (full ? ai : (F)0);
  reg_phase_all<F, REGS>(vr, vi, m, pr, pi);
  return 0;
}

template <int L>
__device__ __forceinline__ uint32_t lane_swizzled(const DevPass &d, int slot, uint32_t t) {
  uint32_t j = 0;
#pragma unroll
  for (int k = 0; k < L; k++) j ^= ((t >> k) & 1) ? d.lane_swz[slot][k] : 0u;
  return j;
}

template <typename T, typename F, int L, int REGS>
__device__ __forceinline__ void lds_transpose(T *__restrict__ lds, const DevPass &d, int from, int to,
                                              uint32_t t, F (&vr)[REGS], F (&vi)[REGS]) {
  __syncthreads();
  const uint32_t jf = lane_swizzled<L>(d, from, t);
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    T a;
    a.re = vr[r];
    a.im = vi[r];
    lds[jf ^ d.regswz[from][r]] = a;
  }
  __syncthreads();
  const uint32_t jt = lane_swizzled<L>(d, to, t);
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    const T a = lds[jt ^ d.regswz[to][r]];
    vr[r] = a.re;
    vi[r] = a.im;
  }
}

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
  static int resident = 0;  // workgroups per CU for this instance (one device per process)
  if (resident == 0) {
    QM_HIP(hipFuncSetAttribute((const void *)k_fused<T, L, R>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    QM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused<T, L, R>, 1 << L, lds));
    resident = std::max(per_cu, 1);
  }
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
