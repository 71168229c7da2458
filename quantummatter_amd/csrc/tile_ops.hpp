// Register-resident tile operations shared by the fused and batched tile-pass
// kernels (fused.hip, batch.hip).  Device code only.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_types.hpp"
#include "gate_math.hpp"
#include "plan.hpp"

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

// Dispatches a register bit to a template parameter; bits at or above R are
// never planned and compile to nothing.
#define QM_BIT_SWITCH(b, CALL)                                                  \
  switch (b) {                                                                  \
    case 0: { constexpr int BB = 0; if constexpr ((1 << BB) < REGS) { CALL; } } break; \
    case 1: { constexpr int BB = 1; if constexpr ((1 << BB) < REGS) { CALL; } } break; \
    case 2: { constexpr int BB = 2; if constexpr ((1 << BB) < REGS) { CALL; } } break; \
    case 3: { constexpr int BB = 3; if constexpr ((1 << BB) < REGS) { CALL; } } break; \
    default: { constexpr int BB = 4; if constexpr ((1 << BB) < REGS) { CALL; } } break; \
  }

// A diagonal op (control conditions already met) with coefficients A, D:
// target on a register bit, a lane bit or outside the tile.
template <typename F, int REGS>
__device__ __forceinline__ void apply_diag(const TileOp *__restrict__ op, F (&vr)[REGS], F (&vi)[REGS],
                                           uint32_t t, uint64_t base, bool full, F ar, F ai, F dr,
                                           F di) {
  const uint32_t m = op->reg_cmask;
  const uint32_t dreg = op->dreg;
  if (dreg != 0) {
    const int b = __builtin_ctz(dreg);
    if (full) {
      QM_BIT_SWITCH(b, (reg_diag<F, REGS, BB, true>(vr, vi, m, ar, ai, dr, di)))
    } else {
      QM_BIT_SWITCH(b, (reg_diag<F, REGS, BB, false>(vr, vi, m, ar, ai, dr, di)))
    }
    return;
  }
  // Target on a lane bit (per-lane choice) or outside the tile (uniform choice).
  const bool hi = ((base & op->dout) != 0) || ((t & op->dlane) != 0);
  if (!full && op->dlane == 0 && !hi) return;  // diag(1, D) on a zero half: nothing to do
  const F pr = hi ? dr : (full ? ar : (F)1);
  const F pi = hi ? di : (full ? ai : (F)0);
  reg_phase_all<F, REGS>(vr, vi, m, pr, pi);
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
  apply_diag<F, REGS>(op, vr, vi, t, base, full, (F)op->c[0], (F)op->c[1], (F)op->c[6], (F)op->c[7]);
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


// The same exchange through an LDS buffer of 2^K reals (half the bytes): real
// parts first, then imaginary parts.  Halving the LDS footprint lets more
// workgroups share a CU.
template <typename F, int L, int REGS>
__device__ __forceinline__ void lds_transpose_split(F *__restrict__ lds, const DevPass &d, int from,
                                                    int to, uint32_t t, F (&vr)[REGS],
                                                    F (&vi)[REGS]) {
  const uint32_t jf = lane_swizzled<L>(d, from, t);
  const uint32_t jt = lane_swizzled<L>(d, to, t);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < REGS; r++) lds[jf ^ d.regswz[from][r]] = vr[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < REGS; r++) vr[r] = lds[jt ^ d.regswz[to][r]];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < REGS; r++) lds[jf ^ d.regswz[from][r]] = vi[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < REGS; r++) vi[r] = lds[jt ^ d.regswz[to][r]];
}

}  // namespace qm
