// Batched tile-pass kernel of libqmhip.so (gfx950): one program run on many
// same-width states, one workgroup per state.
//
// The workgroup that owns a state runs every pass of the program over all of
// its tiles, so everything a state decides for itself stays inside the
// workgroup: Pauli-channel branches read the state's own uniform pool, the
// masses behind an amplitude-damping decision are a workgroup reduction, and
// the decision is applied by the first op of the next segment.  After the
// program the same workgroup reduces the state's mass and Pauli numerators and
// samples it (CDF search with the sequential-order margin rule).  No host
// round trip happens between the first gate and the last sample.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "batch.hpp"
#include "dev_reduce.hpp"
#include "dev_types.hpp"
#include "gate_math.hpp"
#include "qm_internal.hpp"
#include "tile_ops.hpp"


namespace qm {

namespace {

struct DampDecision {
  int jump;
  double scale, jump_scale, stay_scale;
};

// CpuNoiseChannelBackend.cs:118-142: the chosen Kraus branch, renormalised.
template <typename F, int REGS, int B>
__device__ __forceinline__ void reg_damp(F (&vr)[REGS], F (&vi)[REGS], const DampDecision &d) {
  if (d.jump) {
    const double js = d.jump_scale;
#pragma unroll
    for (int r = 0; r < REGS; r++) {
      if (r & (1 << B)) continue;
      const int r1 = r | (1 << B);
      vr[r] = (F)(js * (double)vr[r1]);
      vi[r] = (F)(js * (double)vi[r1]);
      vr[r1] = 0;
      vi[r1] = 0;
    }
    return;
  }
  const double sc = d.scale, ss = d.stay_scale;
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    if (r & (1 << B)) continue;
    const int r1 = r | (1 << B);
    vr[r] = (F)((double)vr[r] * sc);
    vi[r] = (F)((double)vi[r] * sc);
    vr[r1] = (F)((double)vr[r1] * ss);
    vi[r1] = (F)((double)vi[r1] * ss);
  }
}

// Z branch of a Pauli channel: the one half negated (CpuNoiseChannelBackend.cs:70-72).
template <typename F, int REGS, int B>
__device__ __forceinline__ void reg_neg_one(F (&vr)[REGS], F (&vi)[REGS]) {
#pragma unroll
  for (int r = 0; r < REGS; r++) {
    if (!(r & (1 << B))) continue;
    vr[r] = -vr[r];
    vi[r] = -vi[r];
  }
}

// Ops whose action depends on the state: noise branches, damping, table rotations.
template <typename F, int REGS>
__device__ __forceinline__ void apply_dynamic(const TileOp *__restrict__ op, F (&vr)[REGS],
                                              F (&vi)[REGS], uint32_t t, uint64_t base,
                                              const double *__restrict__ draws,
                                              const double *__restrict__ angles,
                                              const DampDecision &damp) {
  const int kind = op->kind;
  const int rbit = op->rbit;
  if (kind == G_NOISE) {
    // CpuNoiseChannelBackend.cs:38-52: one draw picks the branch.
    const double p = op->c[0];
    int pauli;
    if (op->aux1 == 2) {
      const double r = draws[op->aux0];
      const double quarter = p / 4;
      pauli = r < quarter ? 1 : r < 2 * quarter ? 2 : r < 3 * quarter ? 3 : 0;
    } else {
      const bool apply = op->aux0 < 0 || draws[op->aux0] < p;
      pauli = !apply ? 0 : op->aux1 == 0 ? 1 : 3;
    }
    if (pauli == 1) {
      QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_X>(vr, vi, 0, 0)))
    } else if (pauli == 2) {
      QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_Y>(vr, vi, 0, 0)))
    } else if (pauli == 3) {
      QM_BIT_SWITCH(rbit, (reg_neg_one<F, REGS, BB>(vr, vi)))
    }
    return;
  }
  if (kind == G_DAMP) {
    QM_BIT_SWITCH(rbit, (reg_damp<F, REGS, BB>(vr, vi, damp)))
    return;
  }
  const F c0 = (F)angles[2 * op->aux0], c1 = (F)angles[2 * op->aux0 + 1];
  if (kind == G_PROT) {
    if (op->aux1 == G_RX) {
      QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_RX>(vr, vi, c0, c1)))
    } else {
      QM_BIT_SWITCH(rbit, (reg_pair<F, REGS, BB, G_RY>(vr, vi, c0, c1)))
    }
    return;
  }
  // G_PROTZ: Rz = diag(conj(D), D) with D = (cos, sin)(theta/2).
  apply_diag<F, REGS>(op, vr, vi, t, base, true, c0, -c1, c0, c1);
}

// One pass of the program over every tile of one state.
template <typename T, typename F, int L, int R>
__device__ __forceinline__ void batch_pass(T *__restrict__ s, const DevPass &d,
                                           const TileOp *__restrict__ prog, T *__restrict__ lds,
                                           uint64_t tiles, bool reset, const double *draws,
                                           const double *angles, const DampDecision &damp,
                                           double &m0, double &m1) {
  constexpr int REGS = 1 << R;
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
  for (uint64_t tau = 0; tau < tiles; tau++) {
    uint64_t base = 0;
    for (int k = 0; k < d.nruns; k++)
      base |= ((tau >> d.run_src[k]) & ((1ull << d.run_len[k]) - 1)) << d.run_dst[k];
    if (reset) {
      // |0...0>: the first pass starts from the basis state instead of loading.
#pragma unroll
      for (int r = 0; r < REGS; r++) {
        vr[r] = ((base | load_lane | d.load_reg_off[r]) == 0) ? (F)1 : (F)0;
        vi[r] = 0;
      }
    } else {
      const T *src = s + (base | load_lane);
#pragma unroll
      for (int r = 0; r < REGS; r++) {
        const T a = src[d.load_reg_off[r]];
        vr[r] = a.re;
        vi[r] = a.im;
      }
    }
    int cur = load_direct ? 0 : kIo;
    for (int ph = 0; ph < nph; ph++) {
      if (cur != ph) lds_transpose<T, F, L, REGS>(lds, d, cur, ph, t, vr, vi);
      cur = ph;
      const TileOp *__restrict__ po = ops + d.phase_op_begin[ph];
      const int cnt = d.phase_op_count[ph];
      for (int o = 0; o < cnt; o++) {
        if (po[o].kind >= G_NOISE)
          apply_dynamic<F, REGS>(po + o, vr, vi, t, base, draws, angles, damp);
        else
          apply_tile_op<F, REGS>(po + o, vr, vi, t, base);
      }
      const int dcnt = d.phase_diag_count[ph];
      if (dcnt) {
        F pr, pi;
        phase_product<F>(ops + d.phase_diag_begin[ph], dcnt, t, base, pr, pi);
#pragma unroll
        for (int r = 0; r < REGS; r++) mul_phase_fma<F>(pr, pi, vr[r], vi[r]);
      }
    }
    if (d.mass_on) {
      // Masses of the damping target's halves after the segment, in FP64.
      const bool outer = (base & d.mass_dout) != 0 || (t & d.mass_dlane) != 0;
#pragma unroll
      for (int r = 0; r < REGS; r++) {
        const double re = (double)vr[r], im = (double)vi[r];
        const double p = re * re + im * im;
        if (outer || (r & d.mass_dreg))
          m1 += p;
        else
          m0 += p;
      }
    }
    if (!store_direct) lds_transpose<T, F, L, REGS>(lds, d, cur, kIoStore, t, vr, vi);
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

// Kahan masses of a bit's halves in index order (CpuNoiseChannelBackend.cs:88-106), one lane.
template <typename T>
__device__ void serial_masses(const T *s, uint64_t n, uint64_t bit, double &z, double &o) {
  double zc = 0, oc = 0;
  z = 0, o = 0;
  for (uint64_t i = 0; i < n; i++) {
    const double m = prob_of(s[i]);
    double &sum = (i & bit) ? o : z;
    double &comp = (i & bit) ? oc : zc;
    const double corrected = m - comp;
    const double next = sum + corrected;
    comp = (next - sum) - corrected;
    sum = next;
  }
}

constexpr int kMaxAmbiguous = 64;

}  // namespace

// Mass and Pauli numerators of one item (StateVector.Observations.cs:107-141),
// FP64, by the item's workgroup after its program.  Out of line: the program
// loop's register tile is dead here, and this code's registers then do not add
// to the program loop's (spills there cost every pass).
template <typename T, int NT>
__device__ __noinline__ void item_numerators(const T *__restrict__ s, uint64_t n, const BatchItem &it,
                                             int64_t item, const PauliGroup *__restrict__ groups,
                                             double *__restrict__ numer, double *__restrict__ mass,
                                             double *sh_red) {
  const uint32_t t = threadIdx.x;
  double ms = 0;
  for (uint64_t j = t; j < n; j += NT) ms += prob_of(s[j]);
  ms = block_sum(ms, sh_red);
  if (t == 0 && it.want_mass) mass[item] = ms;
  for (int gi = it.group_begin; gi < it.group_begin + it.group_count; gi++) {
    const PauliGroup &g = groups[gi];
    double acc[kTermsPerGroup];
    switch (g.count) {
#define QM_GROUP_CASE(C)                                                                      \
  case C: {                                                                                   \
    double part[C];                                                                           \
    group_numerators<T, C>(s, g, 0, g.flip == 0 ? n : n / 2, t, NT, part);                    \
    for (int k = 0; k < C; k++) acc[k] = part[k];                                             \
  } break;
      QM_GROUP_CASE(1) QM_GROUP_CASE(2) QM_GROUP_CASE(3) QM_GROUP_CASE(4)
      QM_GROUP_CASE(5) QM_GROUP_CASE(6) QM_GROUP_CASE(7) QM_GROUP_CASE(8)
#undef QM_GROUP_CASE
      default: break;
    }
    for (int k = 0; k < kTermsPerGroup; k++) {
      if (k >= g.count) break;
      const double r = block_sum(acc[k], sh_red);
      if (t == 0) numer[it.numer_off + g.first + k] = r;
    }
  }
  __syncthreads();
}

template <typename T, int L, int R>
__global__ __launch_bounds__(1 << L, 2) void k_batch(const BatchArgs a) {
  using F = typename Real<T>::type;
  constexpr int NT = 1 << L;
  extern __shared__ char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  double *prefix = reinterpret_cast<double *>(lds_raw);  // final stage only: NT + 1 sums
  __shared__ double sh_red[16];
  __shared__ double sh_dec[4];
  __shared__ int sh_int[4];
  __shared__ int sh_amb[kMaxAmbiguous];
  const uint32_t t = threadIdx.x;

  for (int64_t item = blockIdx.x; item < a.n_items; item += gridDim.x) {
    const BatchItem it = a.items[item];
    const uint64_t n = it.len;
    T *s = reinterpret_cast<T *>(a.states) + (a.keep ? it.state_off : (int64_t)blockIdx.x * a.slot_len);
    const double *pool = a.pools ? a.pools + item * a.stride : nullptr;
    const double *angles = a.angles ? a.angles + item * (int64_t)a.slots * 2 : nullptr;
    int64_t cursor = 0;
    int err = 0;
    DampDecision damp{0, 1.0, 0.0, 1.0};

    for (int sg = it.seg_begin; sg < it.seg_begin + it.seg_count; sg++) {
      const BatchSeg seg = a.segs[sg];
      if (seg.draws > 0 && cursor + seg.draws > a.stride) {
        err = QM_ERR_POOL;  // NoiseChannel draws past the end of the pool
        break;
      }
      double m0 = 0, m1 = 0;
      const double *draws = pool ? pool + cursor : nullptr;
      for (int p = 0; p < seg.pass_count; p++) {
        const int pi = seg.pass_begin + p;
        batch_pass<T, F, L, R>(s, a.passes[pi], a.ops, lds, it.tiles, p == 0 && sg == it.seg_begin,
                               draws, angles, damp, m0, m1);
        __syncthreads();  // the pass's global writes are visible to the whole workgroup
      }
      cursor += seg.draws;
      if (seg.damp_target < 0) continue;
      // Amplitude damping (CpuNoiseChannelBackend.cs:85-142), decided by lane 0.
      const double z = block_sum(m0, sh_red);
      const double o = block_sum(m1, sh_red);
      if (t == 0) {
        const double g = seg.gamma;
        double zero = z, one = o;
        double jump = g * one, stay = zero + (1 - g) * one, total = jump + stay;
        int e = 0, do_jump = 0, drew = 0;
        if (!isfinite(total) || !(total > 0)) {
          e = QM_ERR_MASS;
        } else if (!(jump > 0)) {
          do_jump = 0;
        } else if (stay == 0) {
          do_jump = 1;
        } else if (cursor >= a.stride) {
          e = QM_ERR_POOL;
        } else {
          const double r = pool[cursor];
          drew = 1;
          const double ratio = jump / total;
          if (fabs(r - ratio) <= 1e-12 * ratio) {
            // Within the summation-order margin: the reference's Kahan order decides.
            serial_masses(s, n, 1ull << seg.damp_target, zero, one);
            jump = g * one, stay = zero + (1 - g) * one, total = jump + stay;
            atomicAdd(a.fallbacks, 1ull);
          }
          do_jump = r < jump / total;
        }
        const double branch = do_jump ? jump : stay;
        const double scale = 1 / sqrt(branch);
        sh_dec[0] = scale;
        sh_dec[1] = sqrt(g) * scale;
        sh_dec[2] = sqrt(1 - g) * scale;
        sh_int[0] = do_jump;
        sh_int[1] = e;
        sh_int[2] = drew;
      }
      __syncthreads();
      damp.jump = sh_int[0];
      damp.scale = sh_dec[0];
      damp.jump_scale = sh_dec[1];
      damp.stay_scale = sh_dec[2];
      err = sh_int[1];
      cursor += sh_int[2];
      __syncthreads();
      if (err) break;
    }

    // Mass and Pauli numerators (StateVector.Observations.cs:107-141), FP64.
    if (!err && (it.want_mass || it.group_count > 0))
      item_numerators<T, NT>(s, n, it, item, a.groups, a.numer, a.mass, sh_red);

    // Terminal sampling (StateVectorSimulator.cs:44-90) and MeasureAll (StateVector.cs:106-153).
    if (!err && (it.shots > 0 || it.measure)) {
      const uint64_t chunk = n / NT;
      const int lane = t & 63, w = t >> 6, nw = NT / 64;
      // Chunk sums, one wave per chunk, coalesced.
      for (uint64_t c = w; c < (uint64_t)NT; c += nw) {
        double v = 0;
        for (uint64_t i = c * chunk + lane; i < (c + 1) * chunk; i += 64) v += prob_of(s[i]);
        v = wave_sum(v);
        if (lane == 0) prefix[c + 1] = v;
      }
      __syncthreads();
      if (t == 0) {
        prefix[0] = 0;
        for (int c = 0; c < NT; c++) prefix[c + 1] += prefix[c];
        sh_int[0] = 0;  // ambiguous draws recorded
        sh_int[1] = 0;  // overflow of the record
        const double total = prefix[NT];
        sh_int[2] = (!isfinite(total) || !(total > 0)) ? QM_ERR_MASS : 0;
        if (!sh_int[2] && it.measure && cursor >= a.stride) sh_int[2] = QM_ERR_POOL;
      }
      __syncthreads();
      err = sh_int[2];
      if (!err) {
        const double total = prefix[NT];
        const double delta = (5.0 * (double)n + 8.0) * 0x1p-53 * total;
        const int64_t draws_n = it.shots + (it.measure ? 1 : 0);
        for (int64_t k = t; k < draws_n; k += NT) {
          const double u = k < it.shots ? a.shot_u[it.shot_off + k] : pool[cursor];
          const double draw = u * total;
          long long found = -1;
          bool amb = false;
          if (draw >= total) {
            amb = true;  // only by rounding of u * total
          } else {
            int lo = 0, hi = NT - 1;
            while (lo < hi) {
              const int mid = lo + (hi - lo + 1) / 2;
              if (prefix[mid] <= draw)
                lo = mid;
              else
                hi = mid - 1;
            }
            double cum = prefix[lo], before = cum, after = cum;
            for (uint64_t i = (uint64_t)lo * chunk; i < (uint64_t)(lo + 1) * chunk; i++) {
              before = cum;
              cum += prob_of(s[i]);
              if (draw < cum) {
                found = (long long)i;
                after = cum;
                break;
              }
            }
            amb = found < 0 || (draw - before) <= delta || (after - draw) <= delta;
          }
          if (amb) {
            const int slot = atomicAdd(&sh_int[0], 1);
            if (slot < kMaxAmbiguous)
              sh_amb[slot] = (int)k;
            else
              sh_int[1] = 1;
          } else if (k < it.shots) {
            a.samples[it.shot_off + k] = found;
          } else {
            a.outcomes[item] = found;
          }
        }
        __syncthreads();
        if (t == 0 && sh_int[0] > 0) {
          // Sequential order for the flagged draws (or all of them after an overflow).
          double stotal = 0;
          long long slast = -1;
          for (uint64_t i = 0; i < n; i++) {
            const double p = prob_of(s[i]);
            stotal += p;
            if (p > 0) slast = (long long)i;
          }
          const bool all = sh_int[1] != 0;
          const int64_t count = all ? draws_n : sh_int[0];
          for (int64_t f = 0; f < count; f++) {
            const int64_t k = all ? f : sh_amb[f];
            const double u = k < it.shots ? a.shot_u[it.shot_off + k] : pool[cursor];
            const double draw = u * stotal;
            long long found = slast;
            double cum = 0;
            for (uint64_t i = 0; i < n; i++) {
              cum += prob_of(s[i]);
              if (draw < cum) {
                found = (long long)i;
                break;
              }
            }
            if (k < it.shots)
              a.samples[it.shot_off + k] = found;
            else
              a.outcomes[item] = found;
          }
          atomicAdd(a.fallbacks, (unsigned long long)count);
        }
        if (it.measure) cursor += 1;
      }
      __syncthreads();
    }
    if (t == 0) {
      if (a.consumed) a.consumed[item] = cursor;
      a.err[item] = err;
    }
    __syncthreads();
  }
}

// Mass and Pauli numerators of kept batch states (StateVector.Observations.cs:
// 107-141, FP64), reduced after the batch kernel by many workgroups per state.
// Work unit u = (item, slice) in item order, so the grid works on a few
// consecutive states at a time and they stay in the device's last-level cache
// across their reads (one per flip group), where one workgroup per state would
// stream each state from HBM once per group with every state in flight.  Slice
// sl of an item covers amplitudes [sl A, (sl + 1) A) for the mass and the flip
// 0 group, and pair indices [sl P, (sl + 1) P) of every other group (A =
// len / slices, P = len / 2 / slices), so every amplitude and every pair is
// counted once.  Partial sums go to part[(numer_off + first + k) * slices + sl]
// and part_mass[item * slices + sl]; k_pauli_fold adds them in slice order.
template <typename T>
__global__ __launch_bounds__(256) void k_pauli_slices(const T *__restrict__ states,
                                                      const BatchItem *__restrict__ items, int64_t n_items,
                                                      const PauliGroup *__restrict__ groups, int32_t group_begin,
                                                      int32_t group_count, int32_t slices, double *__restrict__ part,
                                                      double *__restrict__ part_mass) {
  __shared__ double sh_red[16];
  const uint32_t t = threadIdx.x, NT = blockDim.x;
  const int64_t units = n_items * slices;
  for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
    const int64_t item = u / slices;
    const uint64_t sl = (uint64_t)(u % slices);
    const BatchItem it = items[item];
    const T *s = states + it.state_off;
    const uint64_t n = it.len, A = n / (uint64_t)slices, P = A / 2;
    double ms = 0;
    for (uint64_t j = sl * A + t; j < (sl + 1) * A; j += NT) ms += prob_of(s[j]);
    ms = block_sum(ms, sh_red);
    if (t == 0) part_mass[u] = ms;
    for (int gi = group_begin; gi < group_begin + group_count; gi++) {
      const PauliGroup &g = groups[gi];
      const uint64_t j0 = g.flip == 0 ? sl * A : sl * P, j1 = g.flip == 0 ? (sl + 1) * A : (sl + 1) * P;
      double out[kTermsPerGroup];
      switch (g.count) {
#define QM_GROUP_CASE(C)                                                   \
  case C: {                                                                \
    double acc[C];                                                         \
    group_numerators<T, C, 4>(s, g, j0, j1, t, NT, acc);                   \
    for (int k = 0; k < C; k++) out[k] = acc[k];                           \
  } break;
        QM_GROUP_CASE(1) QM_GROUP_CASE(2) QM_GROUP_CASE(3) QM_GROUP_CASE(4)
        QM_GROUP_CASE(5) QM_GROUP_CASE(6) QM_GROUP_CASE(7) QM_GROUP_CASE(8)
#undef QM_GROUP_CASE
        default: break;
      }
      for (int k = 0; k < kTermsPerGroup; k++) {
        if (k >= g.count) break;
        const double r = block_sum(out[k], sh_red);
        if (t == 0) part[(it.numer_off + g.first + k) * slices + (int64_t)sl] = r;
      }
    }
  }
}

// Mass and Pauli numerators of kept batch states by tile pairs.  A state of
// 2^n amplitudes is 2^(n - kPauliTileBits) tiles; a pair (j, j ^ flip) joins
// tile h = j >> kPauliTileBits with tile h ^ v, v = flip >> kPauliTileBits.
// Work unit (item, tile h): the unit holds tile h in registers (lane t holds
// offsets t + 256 r, r < 16) and, for every group with h < h ^ v, or v = 0,
// adds the pairs whose lower tile is h (for v = 0 the pairs inside the tile
// whose top flip bit is 0), reading the partner amplitudes, all 16 of a lane
// at once.  Groups are sorted by flip mask, so the groups of one v are
// consecutive and share the partner tile in cache.  Sign parities split into
// a tile part (uniform), a register part (uniform per r) and a lane part
// applied once per group; exactly one of a term's Re/Im weights is nonzero,
// so a term costs two multiply-adds per pair.  Per-wave sums of each term go
// to LDS slots [term][wave] and are added in wave order at the end of the
// unit into part[(numer_off + term) * tiles + h]; k_pauli_fold adds the tiles
// in order.  Terms of the groups a unit skips get 0 from it.  Units are
// ordered XCD-major when the grid is a multiple of 8 and there are at least
// 8 items: the workgroups of one XCD (block id mod 8) reduce items x, x + 8,
// ... so a state's partner tiles are shared through that XCD's L2.
constexpr int kPauliTileBits = kPauliTileBitsHost;
constexpr int kPauliTileThreads = 256;
constexpr int kPauliTileRegs = (1 << kPauliTileBits) / kPauliTileThreads;
static_assert(kPauliTileThreads == 256 && kPauliTileRegs == 16, "lane/register split below assumes 8 + 4 bits");

struct TileHome {
  double re[kPauliTileRegs], im[kPauliTileRegs];
};

// One group's per-lane numerators over tile h (C = the group's term count).
template <typename T, int C>
__device__ __forceinline__ void tile_group(const T *__restrict__ s, uint64_t h, const PauliGroup &g, uint32_t t,
                                           const TileHome &a, double *__restrict__ lds_terms, uint32_t lane,
                                           uint32_t w) {
  constexpr int RG = kPauliTileRegs;
  const uint64_t hb = h << kPauliTileBits;
  double ar[C], ai[C], cr[C], ci[C], sl[C];
#pragma unroll
  for (int k = 0; k < C; k++) {
    ar[k] = 0;
    ai[k] = 0;
    const int y = g.y[k] & 3;
    const double alpha = y == 0 ? 1.0 : y == 2 ? -1.0 : 0.0, beta = y == 1 ? -1.0 : y == 3 ? 1.0 : 0.0;
    if (g.flip == 0) {
      cr[k] = alpha;  // |a_j|^2 with the term's Y factor
      ci[k] = 0;
    } else {
      const bool odd = (__popcll(g.flip & g.sign[k]) & 1) != 0;
      cr[k] = odd ? 0.0 : 2.0 * alpha;
      ci[k] = odd ? 2.0 * beta : 0.0;
    }
    // Tile and lane parts of the term's sign parity at j = hb | t | r << 8.
    const int par = (__popcll(hb & g.sign[k]) + __popc(t & (uint32_t)g.sign[k] & 0xFFu)) & 1;
    sl[k] = par ? -1.0 : 1.0;
  }
  if (g.flip == 0) {
#pragma unroll
    for (int r = 0; r < RG; r++) {
      const double p = a.re[r] * a.re[r] + a.im[r] * a.im[r];
#pragma unroll
      for (int k = 0; k < C; k++) {
        const double sg = (__popc(((uint32_t)r << 8) & (uint32_t)g.sign[k]) & 1) ? -1.0 : 1.0;
        ar[k] = fma(sg, p, ar[k]);
      }
    }
  } else {
    const uint64_t v = g.flip >> kPauliTileBits;
    const uint32_t fl = (uint32_t)g.flip & ((1u << kPauliTileBits) - 1);
    const uint32_t fr = fl >> 8;
    const T *ps = s + (((h ^ v) << kPauliTileBits) | (t ^ (fl & 0xFFu)));
    // v = 0: pairs inside the tile, counted from the member whose top flip
    // bit is 0 (a register condition when that bit is >= 8, else a lane one).
    const uint32_t top = v == 0 ? 1u << (31 - __clz(fl)) : 0u;
    T b[RG];
#pragma unroll
    for (int r = 0; r < RG; r++) b[r] = ps[(uint32_t)(r ^ fr) << 8];
#pragma unroll
    for (int r = 0; r < RG; r++) {
      if (((uint32_t)r << 8) & top) continue;  // uniform
      const double br = b[r].re, bi = b[r].im;
      const double ore = fma(br, a.re[r], bi * a.im[r]);  // Complex.Conjugate(s[j ^ flip]) * s[j]
      const double oim = fma(br, a.im[r], -(bi * a.re[r]));
#pragma unroll
      for (int k = 0; k < C; k++) {
        const double sg = (__popc(((uint32_t)r << 8) & (uint32_t)g.sign[k]) & 1) ? -1.0 : 1.0;
        ar[k] = fma(sg, ore, ar[k]);
        ai[k] = fma(sg, oim, ai[k]);
      }
    }
    if (t & top) {
#pragma unroll
      for (int k = 0; k < C; k++) sl[k] = 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < C; k++) {
    const double x = wave_sum(sl[k] * (cr[k] * ar[k] + ci[k] * ai[k]));
    if (lane == 0) lds_terms[(g.first + k) * 4 + w] = x;
  }
}

template <typename T>
__global__ __launch_bounds__(kPauliTileThreads, 2) void k_pauli_tiles(
    const T *__restrict__ states, const BatchItem *__restrict__ items, int64_t n_items,
    const PauliGroup *__restrict__ groups, int32_t group_begin, int32_t group_count, int32_t n_terms,
    double *__restrict__ part, double *__restrict__ part_mass) {
  constexpr int NT = kPauliTileThreads, RG = kPauliTileRegs, NX = 8;
  extern __shared__ double lds_terms[];  // [n_terms][4]
  __shared__ double sh_mass[4];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t tiles = items[0].len >> kPauliTileBits;
  const bool xm = gridDim.x % NX == 0 && n_items >= NX;
  const int64_t x = xm ? blockIdx.x % NX : 0, slot = xm ? blockIdx.x / NX : blockIdx.x;
  const int64_t stride = xm ? gridDim.x / NX : gridDim.x;
  const int64_t my_items = xm ? (n_items - x + NX - 1) / NX : n_items;
  for (int64_t k = slot; k < my_items * (int64_t)tiles; k += stride) {
    const int64_t item = xm ? x + NX * (k / (int64_t)tiles) : k / (int64_t)tiles;
    const uint64_t h = (uint64_t)(k % (int64_t)tiles);
    const BatchItem it = items[item];
    const T *s = states + it.state_off;
    const uint64_t hb = h << kPauliTileBits;
    __syncthreads();  // the previous unit's LDS slots are read
    for (int i = (int)t; i < n_terms * 4; i += NT) lds_terms[i] = 0;
    TileHome a;
    double ms = 0;
#pragma unroll
    for (int r = 0; r < RG; r++) {
      const T v = s[hb | (t + (uint64_t)NT * r)];
      a.re[r] = v.re;
      a.im[r] = v.im;
      ms += a.re[r] * a.re[r] + a.im[r] * a.im[r];
    }
    ms = wave_sum(ms);
    if (lane == 0) sh_mass[w] = ms;
    __syncthreads();  // slots cleared
    for (int gi = group_begin; gi < group_begin + group_count; gi++) {
      const PauliGroup &g = groups[gi];
      const uint64_t v = g.flip >> kPauliTileBits;
      if (v != 0 && (h ^ v) < h) continue;  // the pair tile's unit adds these pairs
      switch (g.count) {
#define QM_TILE_CASE(C) \
  case C:               \
    tile_group<T, C>(s, h, g, t, a, lds_terms, lane, w); \
    break;
        QM_TILE_CASE(1) QM_TILE_CASE(2) QM_TILE_CASE(3) QM_TILE_CASE(4)
        QM_TILE_CASE(5) QM_TILE_CASE(6) QM_TILE_CASE(7) QM_TILE_CASE(8)
#undef QM_TILE_CASE
        default: break;
      }
    }
    __syncthreads();
    const int64_t u = item * (int64_t)tiles + (int64_t)h;
    for (int i = (int)t; i < n_terms; i += NT)
      part[(it.numer_off + i) * (int64_t)tiles + (int64_t)h] =
          ((lds_terms[i * 4] + lds_terms[i * 4 + 1]) + lds_terms[i * 4 + 2]) + lds_terms[i * 4 + 3];
    if (t == 0) part_mass[u] = ((sh_mass[0] + sh_mass[1]) + sh_mass[2]) + sh_mass[3];
  }
}

// Sums of k_pauli_slices' partials in slice order: one thread per item (mass)
// and per (item, group) (the group's numerators).
__global__ void k_pauli_fold(const BatchItem *__restrict__ items, int64_t n_items,
                             const PauliGroup *__restrict__ groups, int32_t group_begin, int32_t group_count,
                             int32_t slices, const double *__restrict__ part,
                             const double *__restrict__ part_mass, double *__restrict__ numer,
                             double *__restrict__ mass) {
  const int64_t total = n_items + n_items * group_count;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total;
       x += (int64_t)gridDim.x * blockDim.x) {
    if (x < n_items) {
      double m = 0;
      for (int s = 0; s < slices; s++) m += part_mass[x * slices + s];
      mass[x] = m;
      continue;
    }
    const int64_t y = x - n_items, item = y / group_count;
    const PauliGroup &g = groups[group_begin + (int32_t)(y % group_count)];
    for (int k = 0; k < g.count; k++) {
      const int64_t idx = items[item].numer_off + g.first + k;
      double v = 0;
      for (int s = 0; s < slices; s++) v += part[idx * slices + s];
      numer[idx] = v;
    }
  }
}

// |0...0> in each of the consecutive states of len (a power of two) amplitudes.
template <typename T>
__global__ void k_batch_basis(T *__restrict__ s, uint64_t total, uint64_t len) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    T a;
    a.re = (i & (len - 1)) == 0 ? 1 : 0;
    a.im = 0;
    s[i] = a;
  }
}

__global__ void k_zero_i32(int32_t *__restrict__ v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0;
}

void launch_zero_i32(Ctx *c, int32_t *v, int64_t n) {
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024));
  k_zero_i32<<<blocks, 256, 0, c->stream>>>(v, n);
  QM_HIP(hipGetLastError());
}

void launch_batch_basis(Ctx *c, int precision, void *states, int64_t items, uint64_t len) {
  const uint64_t total = (uint64_t)items * len;
  const unsigned blocks = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((total + 255) / 256, (uint64_t)c->compute_units * 8));
  if (precision == QM_C128)
    k_batch_basis<C128><<<blocks, 256, 0, c->stream>>>(static_cast<C128 *>(states), total, len);
  else
    k_batch_basis<C64><<<blocks, 256, 0, c->stream>>>(static_cast<C64 *>(states), total, len);
  QM_HIP(hipGetLastError());
}

namespace {

template <typename T, int L, int R>
void launch_batch_t(Ctx *c, const BatchArgs &args, int64_t *grid_out) {
  constexpr int K = L + R;
  const size_t lds = sizeof(T) << K;
  // Workgroups per CU for this instance, set up once per process (thread-safe
  // initialisation: several lanes may launch concurrently).
  static const int resident = [&] {
    QM_HIP(hipFuncSetAttribute((const void *)k_batch<T, L, R>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    QM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<T, L, R>, 1 << L, lds));
    return std::max(per_cu, 1);
  }();
  int64_t blocks = std::min<int64_t>(args.n_items, (int64_t)c->compute_units * resident);
  if (!args.keep && *grid_out > 0) blocks = std::min<int64_t>(blocks, *grid_out);
  blocks = std::max<int64_t>(blocks, 1);
  *grid_out = blocks;
  k_batch<T, L, R><<<(int)blocks, 1 << L, lds, c->stream>>>(args);
  QM_HIP(hipGetLastError());
}

}  // namespace

// Compiled batch shapes (lane bits L, register bits R), largest tile first.
#define QM_BATCH_SHAPES(X) \
  X(C128, 8, 4) X(C128, 6, 4) X(C128, 6, 2) X(C64, 8, 5) X(C64, 6, 4) X(C64, 6, 2)

FusedShape batch_shape(int precision, int qubits) {
  const int contiguous = precision == QM_C128 ? 3 : 4;
  FusedShape best{0, 0, 0, 0};
#define QM_PICK(T, l, r)                                                               \
  if ((precision == QM_C128) == std::is_same<T, C128>::value && l + r <= qubits &&   \
      l + r > best.K)                                                                  \
    best = FusedShape{l + r, l, r, std::min(contiguous, l + r)};
  QM_BATCH_SHAPES(QM_PICK)
#undef QM_PICK
  return best;
}

void launch_batch(Ctx *c, int precision, const FusedShape &shape, const BatchArgs &args,
                  int64_t *grid) {
#define QM_LAUNCH(T, l, r)                                                           \
  if ((precision == QM_C128) == std::is_same<T, C128>::value && shape.L == l && \
      shape.R == r) {                                                                \
    launch_batch_t<T, l, r>(c, args, grid);                                          \
    return;                                                                          \
  }
  QM_BATCH_SHAPES(QM_LAUNCH)
#undef QM_LAUNCH
  fail(QM_ERR_UNSUPPORTED, "batch tile shape is not compiled");
}

void launch_pauli_slices(Ctx *c, int precision, const void *states, const BatchItem *items, int64_t n_items,
                         bool tile_pairs, const PauliGroup *groups, int32_t group_begin, int32_t group_count,
                         int32_t n_terms, int32_t slices, double *part, double *part_mass, double *numer,
                         double *mass) {
  // Two workgroups per CU: the grid is on about 2 x CUs / slices states at a
  // time (8 states of 2^20 amplitudes, 128 MiB in complex128, at 64 slices).
  const int64_t units = n_items * slices;
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>(units, (int64_t)c->compute_units * 2));
  if (tile_pairs) {
    // Tile pairs (k_pauli_tiles): slices = tiles of 2^13 amplitudes.
    const size_t lds = (size_t)n_terms * 4 * sizeof(double);
    if (precision == QM_C128)
      k_pauli_tiles<C128><<<blocks, 256, lds, c->stream>>>(static_cast<const C128 *>(states), items, n_items, groups,
                                                           group_begin, group_count, n_terms, part, part_mass);
    else
      k_pauli_tiles<C64><<<blocks, 256, lds, c->stream>>>(static_cast<const C64 *>(states), items, n_items, groups,
                                                          group_begin, group_count, n_terms, part, part_mass);
  } else if (precision == QM_C128)
    k_pauli_slices<C128><<<blocks, 256, 0, c->stream>>>(static_cast<const C128 *>(states), items, n_items, groups,
                                                         group_begin, group_count, slices, part, part_mass);
  else
    k_pauli_slices<C64><<<blocks, 256, 0, c->stream>>>(static_cast<const C64 *>(states), items, n_items, groups,
                                                        group_begin, group_count, slices, part, part_mass);
  QM_HIP(hipGetLastError());
  const int64_t total = n_items * (1 + (int64_t)group_count);
  const unsigned fold_blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 1024));
  k_pauli_fold<<<fold_blocks, 256, 0, c->stream>>>(items, n_items, groups, group_begin, group_count, slices, part,
                                                    part_mass, numer, mass);
  QM_HIP(hipGetLastError());
}

}  // namespace qm
