// Gate preparation and fused-pass planning (host only).
#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <exception>
#include <stdexcept>
#include <thread>

#include "gate_math.hpp"

namespace qm {

struct Status;
[[noreturn]] void fail(int code, const std::string &msg);

static inline int popc(uint64_t x) { return __builtin_popcountll(x); }

uint32_t lds_swizzle_mode(uint32_t j, int mode) {
  switch (mode) {
    case 0:
      return j;
    case 1: {
      uint32_t h = j >> 3;
      h ^= h >> 3;
      h ^= h >> 6;
      h ^= h >> 9;
      return j ^ (h & 7);
    }
    case 2:
      return j ^ ((j >> 3) & 7);
    default:
      return j ^ (((j >> 3) ^ (j >> 6) ^ (j >> 9) ^ (j >> 12)) & 7);
  }
}

int lds_swizzle_default() {
  static const int mode = [] {
    const char *e = std::getenv("QM_LDS_SWIZZLE");  // tuning knob
    return e ? std::atoi(e) : 3;
  }();
  return mode;
}

void validate_op(const qm_op &op, int n) {
  // QuantumCircuit.Add (QuantumCircuit.cs:97-126): range, distinctness, finite angle.
  if (op.kind < 0 || op.kind > 16) fail(QM_ERR_UNSUPPORTED, "Unsupported gate kind.");
  if (op.target < 0 || op.target >= n) fail(QM_ERR_RANGE, "qubit: target out of range.");
  if (op.control != -1) {
    if (op.control < 0 || op.control >= n) fail(QM_ERR_RANGE, "qubit: control out of range.");
    if (op.control == op.target) fail(QM_ERR_ARG, "The two qubits must differ.");
  }
  if (op.second_control != -1) {
    if (op.second_control < 0 || op.second_control >= n) fail(QM_ERR_RANGE, "qubit: second control out of range.");
    if (op.second_control == op.target || op.second_control == op.control)
      fail(QM_ERR_ARG, "All three qubits must differ.");
  }
  if (!std::isfinite(op.angle)) fail(QM_ERR_RANGE, "angle must be finite.");
  if (op.kind == 10 && op.control < 0) fail(QM_ERR_ARG, "Swap needs two qubits.");
  if ((op.kind == 9 || op.kind == 16) && op.control < 0) fail(QM_ERR_ARG, "Controlled X needs a control.");
  if (op.kind == 16 && op.second_control < 0) fail(QM_ERR_ARG, "Toffoli needs two controls.");
  if (op.kind == 15)
    for (double v : op.unitary)
      if (!std::isfinite(v)) fail(QM_ERR_RANGE, "Matrix entries must be finite.");
}

bool prepare_gate(const qm_op &op, DevGate *g) {
  std::memset(g, 0, sizeof(*g));
  uint64_t mask = op.control >= 0 ? (uint64_t)1 << op.control : 0;
  if (op.second_control >= 0) mask |= (uint64_t)1 << op.second_control;
  g->target = op.target;
  g->ctrl_mask = mask;
  g->swap_other = -1;
  auto diag1 = [&](double re, double im) {
    g->kind = G_DIAG1;
    g->c[0] = 1, g->c[1] = 0, g->c[6] = re, g->c[7] = im;
  };
  switch (op.kind) {
    case 11:  // I
      return false;
    case 1:   // X
    case 9:   // CNot
    case 16:  // Toffoli
      g->kind = G_X;
      return true;
    case 10:  // Swap(a,b): control = a, target = b; prepared without a control mask
      g->kind = G_SWAP;
      g->ctrl_mask = 0;
      g->swap_other = op.control;
      return true;
    case 0:
      g->kind = G_H;
      g->c[0] = std::sqrt(0.5);
      return true;
    case 2:
      g->kind = G_Y;
      return true;
    case 3:
      diag1(-1, 0);
      return true;
    case 4:
      diag1(0, 1);
      return true;
    case 12:
      diag1(-0.0, -1);
      return true;
    case 5:
      diag1(std::cos(M_PI / 4), std::sin(M_PI / 4));
      return true;
    case 13:
      diag1(std::cos(-M_PI / 4), std::sin(-M_PI / 4));
      return true;
    case 14:
      diag1(std::cos(op.angle), std::sin(op.angle));
      return true;
    case 6:
      g->kind = G_RX;
      g->c[0] = std::cos(op.angle / 2);
      g->c[1] = std::sin(op.angle / 2);
      return true;
    case 7:
      g->kind = G_RY;
      g->c[0] = std::cos(op.angle / 2);
      g->c[1] = std::sin(op.angle / 2);
      return true;
    case 8:
      g->kind = G_DIAG;
      g->c[0] = std::cos(-op.angle / 2), g->c[1] = std::sin(-op.angle / 2);
      g->c[6] = std::cos(op.angle / 2), g->c[7] = std::sin(op.angle / 2);
      return true;
    case 15:
      g->kind = G_MAT;
      for (int i = 0; i < 8; i++) g->c[i] = op.unitary[i];
      return true;
    default:
      fail(QM_ERR_UNSUPPORTED, "Unsupported gate kind.");
  }
}

double touched_fraction(const DevGate &g) {
  double f = 1.0 / (double)((uint64_t)1 << popc(g.ctrl_mask));
  if (g.kind == G_DIAG1 || g.kind == G_SWAP) f *= 0.5;
  return f;
}

// ---------------------------------------------------------------------------
// Planner.
// ---------------------------------------------------------------------------
namespace {

struct PGate {
  DevGate g;
  bool pair;        // needs its target as a register bit
  uint64_t touch;   // target and control qubits
};

std::vector<PGate> expand(const std::vector<DevGate> &in) {
  std::vector<PGate> out;
  out.reserve(in.size() * 2);
  for (const DevGate &g : in) {
    if (g.kind == G_SWAP) {
      // SWAP(a,b) = CX(a->b) CX(b->a) CX(a->b): three exact exchanges.
      int a = g.swap_other, b = g.target;
      DevGate x{};
      x.kind = G_X;
      x.swap_other = -1;
      x.target = b, x.ctrl_mask = (uint64_t)1 << a;
      out.push_back({x, true, ((uint64_t)1 << a) | ((uint64_t)1 << b)});
      x.target = a, x.ctrl_mask = (uint64_t)1 << b;
      out.push_back({x, true, ((uint64_t)1 << a) | ((uint64_t)1 << b)});
      x.target = b, x.ctrl_mask = (uint64_t)1 << a;
      out.push_back({x, true, ((uint64_t)1 << a) | ((uint64_t)1 << b)});
      continue;
    }
    bool pair = !(g.kind == G_DIAG || g.kind == G_DIAG1 || g.kind == G_PROTZ);
    out.push_back({g, pair, ((uint64_t)1 << g.target) | g.ctrl_mask});
  }
  return out;
}

// The reference's general coefficients (A, B, C, D) of a pair kind
// (PreparedStateVectorGate.cs:34-65).
void to_matrix(const DevGate &g, double *m) {
  for (int i = 0; i < 8; i++) m[i] = 0;
  switch (g.kind) {
    case G_RX:  // (c, -i s; -i s, c)
      m[0] = g.c[0], m[3] = -g.c[1], m[5] = -g.c[1], m[6] = g.c[0];
      break;
    case G_RY:  // (c, -s; s, c)
      m[0] = g.c[0], m[2] = -g.c[1], m[4] = g.c[1], m[6] = g.c[0];
      break;
    case G_H:
      m[0] = m[2] = m[4] = g.c[0], m[6] = -g.c[0];
      break;
    case G_Y:  // (0, -i; i, 0)
      m[3] = -1, m[5] = 1;
      break;
    default:
      for (int i = 0; i < 8; i++) m[i] = g.c[i];
      break;
  }
}

// Moves the diagonal ops of one phase whose target and controls are lane bits
// or outside the tile to the end of the phase, behind a G_PHASEPROD header, when
// no later pair op of the phase targets any qubit they touch (a diagonal gate
// commutes with every gate that does not change its qubits).  The kernel then
// forms their per-lane product once instead of touching every register per op.
void merge_phase_diagonals(std::vector<TileOp> &ops, const std::vector<uint64_t> &touch,
                           const std::vector<uint64_t> &pair_target, int begin,
                           std::vector<TileOp> &tail, std::vector<int> &headers) {
  const int n = (int)ops.size() - begin;
  if (n < 2) return;
  uint64_t later_targets = 0;
  std::vector<char> move(n, 0);
  for (int i = n - 1; i >= 0; i--) {
    const TileOp &t = ops[begin + i];
    const bool diag = (t.kind == G_DIAG || t.kind == G_DIAG1) && t.rbit < 0;
    if (diag && t.dreg == 0 && t.reg_cmask == 0 && !(touch[i] & later_targets)) move[i] = 1;
    later_targets |= pair_target[i];
  }
  int count = 0;
  for (int i = 0; i < n; i++) count += move[i];
  if (count < 2) return;
  std::vector<TileOp> kept;
  headers.push_back((int)tail.size());  // this phase's first merged op in the tail
  headers.push_back(count);
  for (int i = 0; i < n; i++) {
    if (move[i])
      tail.push_back(ops[begin + i]);
    else
      kept.push_back(ops[begin + i]);
  }
  ops.resize(begin);
  ops.insert(ops.end(), kept.begin(), kept.end());
}

// Rank of bit p among the set bits of m (number of set bits below p).
inline int rank_in(uint32_t m, int p) { return popc(m & ((1u << p) - 1)); }

// Register set rmask, thread bits in ascending local position after the
// `nfirst` positions `first` (which must not be register positions).
void fill_mapping(DevPass &d, int slot, uint32_t rmask, int K, int R, const int *first = nullptr,
                  int nfirst = 0) {
  int k = 0;
  uint32_t placed = 0;
  for (int i = 0; i < nfirst; i++) {
    d.lane_swz[slot][k] = lds_swizzle(1u << first[i]);
    d.lanepos[slot][k++] = (uint8_t)first[i];
    placed |= 1u << first[i];
  }
  for (int p = 0; p < K; p++) {
    if (!(rmask & (1u << p)) && !(placed & (1u << p))) {
      d.lane_swz[slot][k] = lds_swizzle(1u << p);
      d.lanepos[slot][k++] = (uint8_t)p;
    }
  }
  for (int reg = 0; reg < (1 << R); reg++) {
    uint32_t j = 0;
    int b = 0;
    for (int p = 0; p < K; p++)
      if (rmask & (1u << p)) {
        if ((reg >> b) & 1) j |= 1u << p;
        b++;
      }
    d.regloc[slot][reg] = (uint16_t)j;
    d.regswz[slot][reg] = (uint16_t)lds_swizzle(j);
  }
}

}  // namespace

std::vector<TileOp> FusedPlan::flat_ops() {
  std::vector<TileOp> all;
  for (FusedPass &p : passes) {
    p.dev.op_offset = (int32_t)all.size();
    all.insert(all.end(), p.ops.begin(), p.ops.end());
  }
  return all;
}

FusedPlan plan_fused(const std::vector<DevGate> &gates_in, int n, const FusedShape &sh, const PlanTrace *replay,
                     PlanTrace *record) {
  FusedPlan plan;
  plan.shape = sh;
  plan.n = n;
  plan.source_gates = (int64_t)gates_in.size();
  if (n < sh.K) fail(QM_ERR_ARG, "fused passes need at least K qubits");
  const int K = sh.K, L = sh.L, R = sh.R, c = sh.contiguous;
  const bool remap = sh.remap != 0 && c > 0 && 2 * c <= K;
  const uint64_t all = n >= 64 ? ~0ull : (((uint64_t)1 << n) - 1);
  const uint32_t low_local = (1u << c) - 1;
  // Assignment of qubits to index bits: logi[x] is the qubit at index bit x.
  // Invariant: for p < c, logi[p] == p, or logi[p] = h >= c and logi[h] = p
  // (a transposition); every other bit holds its own qubit.
  int logi[64];
  for (int x = 0; x < 64; x++) logi[x] = x;
  // Lane layout (FusedShape::lane_layout): the index bit each bit of the
  // assignment above is stored at (identity below `contiguous`), for the
  // state as the current pass loads it.
  int mload[64];
  for (int x = 0; x < 64; x++) mload[x] = x;
  auto displaced = [&] {
    for (int p = 0; p < c; p++)
      if (logi[p] != p) return true;
    return false;
  };

  std::vector<PGate> G = expand(gates_in);
  std::vector<int> pending(G.size());
  for (size_t i = 0; i < G.size(); i++) pending[i] = (int)i;
  // Perturbation stream of a restarted tile search (xorshift64 from the seed).
  uint64_t rng = sh.search_seed * 0x9E3779B97F4A7C15ull + 1;
  if (rng == 0) rng = 1;
  auto perturb = [&rng] {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };

  // A last pass with no gates restores the natural order if the final gate
  // pass could not.
  while (!pending.empty() || displaced()) {
    // 1. Tile qubits: greedy scan with dependency blocking; spare tile slots
    //    go to the targets of the deferred gates, then the scan is repeated.
    //    The qubits at the low index bits are always in the tile.
    uint64_t Q = 0;
    for (int p = 0; p < c; p++) Q |= (uint64_t)1 << logi[p];
    std::vector<int> included, deferred;
    const size_t pass_index = plan.passes.size();
    const bool replaying = replay && pass_index < replay->tiles.size() && pass_index < replay->rsets.size();
    auto scan_tile = [&](uint64_t &q, bool grow) {
      uint64_t blocked = 0;
      included.clear();
      deferred.clear();
      for (size_t k = 0; k < pending.size(); k++) {
        int idx = pending[k];
        const PGate &x = G[idx];
        if (blocked == all) {
          deferred.insert(deferred.end(), pending.begin() + k, pending.end());
          break;
        }
        if (x.touch & blocked) {
          deferred.push_back(idx);
          blocked |= x.touch;
          continue;
        }
        uint64_t tb = (uint64_t)1 << x.g.target;
        if (x.pair && !(q & tb)) {
          if (grow && popc(q) < K) {
            q |= tb;
          } else {
            deferred.push_back(idx);
            blocked |= x.touch;
            continue;
          }
        }
        included.push_back(idx);
      }
    };
    // Spare slots: qubits displaced from the low bits first (so their
    // transpositions can be undone), then the lowest qubits.
    auto fill = [&](uint64_t &q) {
      for (int p = 0; p < c && popc(q) < K; p++)
        if (logi[p] != p) q |= (uint64_t)1 << p;
      for (int x = 0; popc(q) < K && x < n; x++) q |= (uint64_t)1 << x;
    };
    const uint64_t forced = Q;
    if (replaying) {
      Q = replay->tiles[pass_index];
      scan_tile(Q, false);
    } else {
      scan_tile(Q, true);
      if (popc(Q) < K && !deferred.empty()) {
        for (int idx : deferred) {
          if (popc(Q) >= K) break;
          if (G[idx].pair) Q |= (uint64_t)1 << G[idx].g.target;
        }
        fill(Q);
        scan_tile(Q, false);
      }
      fill(Q);
      if (sh.search) {
        // Tile search: grow the tile one qubit at a time, each time by the qubit
        // whose addition lets the most pending gates run in this pass (ties: the
        // qubit a pending gate needs first); keep it if it beats the in-order
        // choice above.
        auto runnable = [&](uint64_t q) {
          uint64_t blocked = 0;
          int count = 0;
          for (size_t k = 0; k < pending.size() && blocked != all; k++) {
            const PGate &x = G[pending[k]];
            if ((x.touch & blocked) || (x.pair && !(q & ((uint64_t)1 << x.g.target)))) {
              blocked |= x.touch;
              continue;
            }
            count++;
          }
          return count;
        };
        int first_need[64];
        for (int x = 0; x < 64; x++) first_need[x] = INT32_MAX;
        for (size_t k = 0; k < pending.size(); k++) {
          const PGate &x = G[pending[k]];
          if (x.pair && first_need[x.g.target] == INT32_MAX) first_need[x.g.target] = (int)k;
        }
        uint64_t S = forced;
        while (popc(S) < K) {
          int best = -1, best_count = -1;
          for (int x = 0; x < n; x++) {
            if (S & ((uint64_t)1 << x)) continue;
            int cnt = runnable(S | ((uint64_t)1 << x));
            if (sh.search_seed) cnt += (int)(perturb() % (uint64_t)(kSearchNoise + 1));
            if (cnt > best_count || (cnt == best_count && first_need[x] < first_need[best])) {
              best = x;
              best_count = cnt;
            }
          }
          S |= (uint64_t)1 << best;
        }
        if (runnable(S) > (int)included.size()) {
          Q = S;
          scan_tile(Q, false);
        }
      }
      if (remap && deferred.empty() && displaced()) {
        // Last pass: the qubits displaced from the low index bits must be in
        // the tile for the store to restore the natural order (else a pass
        // without gates follows just for that).  Tile qubits that no pair gate
        // of the pass targets give them their slots; every gate still runs.
        uint64_t need = forced, want = 0;
        for (int idx : included)
          if (G[idx].pair) need |= (uint64_t)1 << G[idx].g.target;
        for (int p = 0; p < c; p++)
          if (logi[p] != p) want |= (uint64_t)1 << p;
        uint64_t missing = want & ~Q, spare = Q & ~need & ~want;
        while (missing && spare) {
          Q &= ~(spare & (~spare + 1));
          Q |= missing & (~missing + 1);
          spare &= spare - 1;
          missing &= missing - 1;
        }
        scan_tile(Q, false);
      }
    }
    if (record) {
      record->tiles.push_back(Q);
      record->rsets.emplace_back();
    }

    int phys[64];
    for (int x = 0; x < 64; x++) phys[logi[x]] = x;
    uint64_t P = 0;  // index bits of the tile
    for (int q = 0; q < n; q++)
      if (Q & ((uint64_t)1 << q)) P |= (uint64_t)1 << phys[q];

    FusedPass pass;
    DevPass &d = pass.dev;
    std::memset(&d, 0, sizeof(d));
    d.qmask = P;
    d.n = n;
    // Local positions follow the index bits in ascending order; local_of is
    // indexed by qubit.
    int qpos[64], local_of[64];
    for (int q = 0; q < 64; q++) local_of[q] = -1;
    {
      int k = 0;
      for (int x = 0; x < n; x++)
        if (P & ((uint64_t)1 << x)) {
          qpos[k] = x;
          local_of[logi[x]] = k;
          d.qpos[k] = (uint8_t)x;
          k++;
        }
    }
    // Runs of non-tile bits, ascending (index bits under the load's lane
    // layout); tile index bit k goes to the k-th lowest non-tile index bit.
    int nt_virt[64], n_nt = 0;  // non-tile bits (before the lane layout) in tile index bit order
    {
      uint64_t Pl = 0;
      for (int x = 0; x < n; x++)
        if (P & ((uint64_t)1 << x)) Pl |= (uint64_t)1 << mload[x];
      int virt_at[64];
      for (int x = 0; x < 64; x++) virt_at[mload[x]] = x;
      int src = 0, nr = 0;
      for (int q = 0; q < n;) {
        if (Pl & ((uint64_t)1 << q)) {
          q++;
          continue;
        }
        int start = q;
        while (q < n && !(Pl & ((uint64_t)1 << q))) nt_virt[n_nt++] = virt_at[q++];
        d.run_src[nr] = (uint8_t)src;
        d.run_dst[nr] = (uint8_t)start;
        d.run_len[nr] = (uint8_t)(q - start);
        src += q - start;
        nr++;
      }
      d.nruns = nr;
    }

    // 2. Phases: register sets inside the tile, same blocking rule.
    std::vector<int> remaining = included, overflow;
    std::vector<uint32_t> rsets;
    std::vector<std::vector<int>> phase_gates;
    while (!remaining.empty()) {
      if ((int)rsets.size() == (sh.max_phases > 0 ? std::min(sh.max_phases, kMaxPhases) : kMaxPhases)) {
        overflow = remaining;
        break;
      }
      uint32_t rset = 0;
      std::vector<int> here, next;
      const size_t phase_index = rsets.size();
      auto scan_phase = [&](uint32_t &rs, bool grow) {
        uint64_t pblocked = 0;
        here.clear();
        next.clear();
        for (int idx : remaining) {
          const PGate &x = G[idx];
          if (x.touch & pblocked) {
            next.push_back(idx);
            pblocked |= x.touch;
            continue;
          }
          if (x.pair) {
            uint32_t lb = 1u << local_of[x.g.target];
            if (!(rs & lb)) {
              if (grow && popc(rs) < R) {
                rs |= lb;
              } else {
                next.push_back(idx);
                pblocked |= x.touch;
                continue;
              }
            }
          }
          here.push_back(idx);
        }
      };
      if (replaying && phase_index < replay->rsets[pass_index].size()) {
        rset = replay->rsets[pass_index][phase_index];
        scan_phase(rset, false);
      } else {
        scan_phase(rset, true);
        if (popc(rset) < R && !next.empty()) {
          // Spare register bits go to the targets of the deferred gates (in
          // order), and the phase is rescanned: fewer phases, fewer transposes.
          for (int idx : next) {
            if (popc(rset) >= R) break;
            if (G[idx].pair) rset |= 1u << local_of[G[idx].g.target];
          }
          scan_phase(rset, false);
        }
        if (sh.phase_search && !next.empty()) {
          // Phase search: grow a register set position by position for the most
          // gates runnable in this phase (same blocking rule; ties: the position
          // a pending gate needs first), optionally improve it by single
          // exchanges, and keep it if it runs more gates than the in-order set.
          const int penalty = rsets.empty() ? sh.low_penalty : 0;
          auto score = [&](uint32_t rs) {
            uint64_t pb = 0;
            int cnt = (rs & low_local) ? -penalty : 0;
            for (int idx : remaining) {
              const PGate &x = G[idx];
              if ((x.touch & pb) || (x.pair && !(rs & (1u << local_of[x.g.target])))) {
                pb |= x.touch;
                continue;
              }
              cnt++;
            }
            return cnt;
          };
          int first_need[32];
          for (int i = 0; i < 32; i++) first_need[i] = INT32_MAX;
          for (size_t k = 0; k < remaining.size(); k++) {
            const PGate &x = G[remaining[k]];
            const int lp = local_of[x.g.target];
            if (x.pair && first_need[lp] == INT32_MAX) first_need[lp] = (int)k;
          }
          uint32_t S = 0;
          while (popc(S) < R) {
            int best = -1, best_score = INT32_MIN;
            for (int q = 0; q < K; q++) {
              if (S & (1u << q)) continue;
              const int v = score(S | (1u << q));
              if (v > best_score || (v == best_score && first_need[q] < first_need[best])) {
                best = q;
                best_score = v;
              }
            }
            S |= 1u << best;
          }
          int cur_score = score(S);
          for (bool improved = sh.phase_search >= 2; improved;) {
            improved = false;
            for (int a = 0; a < K && !improved; a++) {
              if (!(S & (1u << a))) continue;
              for (int b = 0; b < K && !improved; b++) {
                if (S & (1u << b)) continue;
                const uint32_t T = (S & ~(1u << a)) | (1u << b);
                const int v = score(T);
                if (v > cur_score) {
                  S = T;
                  cur_score = v;
                  improved = true;
                }
              }
            }
          }
          if (cur_score > (int)here.size() - ((rset & low_local) ? penalty : 0)) {
            rset = S;
            scan_phase(rset, false);
          }
        }
      }
      if (record) record->rsets[pass_index].push_back(rset);
      // Fill the register set from the top local positions (keeps the
      // contiguous low bits on lanes, so loads and stores stay coalesced).
      for (int p = K - 1; popc(rset) < R && p >= 0; p--)
        if (!(rset & (1u << p)) && !(low_local & (1u << p))) rset |= 1u << p;
      for (int p = K - 1; popc(rset) < R && p >= 0; p--) rset |= 1u << p;
      rsets.push_back(rset);
      phase_gates.push_back(here);
      remaining.swap(next);
    }
    // Gates not placed go back to pending, in original order.
    if (!overflow.empty()) {
      deferred.insert(deferred.end(), overflow.begin(), overflow.end());
      std::sort(deferred.begin(), deferred.end());
    }
    uint32_t io_rmask = 0;
    for (int p = K - 1; popc(io_rmask) < R; p--) io_rmask |= 1u << p;
    if (rsets.empty()) {  // a pass without gates (restores the natural order)
      rsets.push_back(io_rmask);
      phase_gates.emplace_back();
    }

    // 3. Remapping: the tile's qubits whose next use as a pair target comes
    //    first among the remaining gates are stored at the low index bits.
    //    They are taken among the thread bits of the last phase, whose thread
    //    order is then chosen so the store needs no transpose: thread bit i
    //    holds the qubit stored to index bit i.  When the last phase is also
    //    the load phase its thread order is fixed and nothing moves.
    const int last = (int)rsets.size() - 1;
    const bool fixed_lanes = last == 0 && (rsets[0] & low_local) == 0;
    auto on_lane = [&](int q) { return (rsets[last] & (1u << local_of[q])) == 0; };
    int newlogi[64];
    for (int x = 0; x < 64; x++) newlogi[x] = logi[x];
    bool lane_order = false;  // last phase reordered for a direct store
    auto undoable = [&](int p) { return newlogi[p] == p || (P & ((uint64_t)1 << newlogi[p])); };
    if (remap && deferred.empty()) {
      // No gates left: restore the natural order as far as the tile allows.
      lane_order = !fixed_lanes;
      for (int p = 0; p < c; p++) {
        if (newlogi[p] != p && undoable(p)) {
          const int h = newlogi[p];
          newlogi[p] = p;
          newlogi[h] = h;
        }
        lane_order = lane_order && on_lane(newlogi[p]);
      }
    } else if (remap && !fixed_lanes) {
      int first_use[64];
      for (int q = 0; q < 64; q++) first_use[q] = INT32_MAX;
      for (size_t k = 0; k < deferred.size(); k++) {
        const PGate &x = G[deferred[k]];
        if (x.pair && first_use[x.g.target] == INT32_MAX) first_use[x.g.target] = (int)k;
      }
      std::vector<std::pair<int, int>> cand, spare;
      for (int q = 0; q < n; q++)
        if ((Q & ((uint64_t)1 << q)) && on_lane(q))
          (first_use[q] != INT32_MAX ? cand : spare).push_back({first_use[q], q});
      std::sort(cand.begin(), cand.end());
      if ((int)cand.size() > c) {
        spare.insert(spare.begin(), cand.begin() + c, cand.end());
        cand.resize((size_t)c);
      }
      bool reserved[64] = {};
      auto undo = [&](int p) {
        const int h = newlogi[p];
        newlogi[p] = p;
        newlogi[h] = h;
      };
      // Low qubits wanted low go home (their partner bit is in the tile).
      for (const auto &e : cand)
        if (e.second < c) {
          if (newlogi[e.second] != e.second) undo(e.second);
          reserved[e.second] = true;
        }
      // Wanted qubits already at a low bit stay there.
      std::vector<int> want;
      for (const auto &e : cand) {
        const int b = e.second;
        if (b < c) continue;
        int at = -1;
        for (int p = 0; p < c; p++)
          if (newlogi[p] == b) at = p;
        if (at >= 0)
          reserved[at] = true;
        else
          want.push_back(b);
      }
      // The others (at their own bit b) take a free low bit p, preferring one
      // whose qubit is a register of the last phase: qubit b to p, qubit p to
      // bit b (p's transposition undone first when its partner is in the tile).
      auto place = [&](int b, int p) {
        if (newlogi[p] != p) undo(p);
        newlogi[p] = b;
        newlogi[b] = p;
        reserved[p] = true;
      };
      for (int b : want) {
        int p = -1;
        for (int x = 0; x < c && p < 0; x++)
          if (!reserved[x] && !on_lane(newlogi[x]) && undoable(x)) p = x;
        for (int x = 0; x < c && p < 0; x++)
          if (!reserved[x] && newlogi[x] == x) p = x;
        for (int x = 0; x < c && p < 0; x++)
          if (!reserved[x] && undoable(x)) p = x;
        if (p >= 0) place(b, p);
      }
      // Other transpositions are undone where the tile holds both bits and the
      // returning qubit is a thread bit of the last phase.
      for (int p = 0; p < c; p++)
        if (!reserved[p] && newlogi[p] != p && undoable(p) && on_lane(p)) undo(p);
      // Low bits still holding a register qubit of the last phase take a spare
      // thread-bit qubit (earliest next use first).
      lane_order = true;
      size_t next_spare = 0;
      for (int p = 0; p < c && lane_order; p++) {
        if (on_lane(newlogi[p])) continue;
        int b = -1;
        while (next_spare < spare.size() && b < 0) {
          const int q = spare[next_spare++].second;
          if (q >= c && newlogi[q] == q) b = q;
        }
        if (b < 0 || !undoable(p)) {
          lane_order = false;
          break;
        }
        place(b, p);
      }
    }
    if (sh.remap == 2 && !lane_order && !deferred.empty()) {
      // Moves that would need a transpose before the store are dropped.
      for (int x = 0; x < 64; x++) newlogi[x] = logi[x];
    }
    int newphys[64];
    for (int x = 0; x < 64; x++) newphys[newlogi[x]] = x;
    int store_pos[16];  // index bit each local position is stored to
    bool moved = false;
    for (int k = 0; k < K; k++) {
      store_pos[k] = newphys[logi[qpos[k]]];
      moved |= store_pos[k] != qpos[k];
    }

    // 4. Mappings and ops.
    d.nphases = (int32_t)rsets.size();
    for (int ph = 0; ph < d.nphases; ph++) fill_mapping(d, ph, rsets[ph], K, R);
    fill_mapping(d, kIo, io_rmask, K, R);
    d.load_direct = (rsets.front() & low_local) == 0;
    fill_mapping(d, kIoStore, io_rmask, K, R);
    if (!moved) {
      d.store_direct = (rsets.back() & low_local) == 0;
    } else {
      // Thread bits 0..c-1 must hold the local positions stored to index bits
      // 0..c-1, in that order, for coalesced stores.
      int low_src[16];
      uint32_t low_set = 0;
      for (int k = 0; k < K; k++)
        if (store_pos[k] < c) {
          low_src[store_pos[k]] = k;
          low_set |= 1u << k;
        }
      if (lane_order) {
        fill_mapping(d, last, rsets[last], K, R, low_src, c);
        d.store_direct = 1;
      } else {
        uint32_t st_rmask = 0;
        for (int p = K - 1; popc(st_rmask) < R; p--)
          if (!(low_set & (1u << p))) st_rmask |= 1u << p;
        fill_mapping(d, kIoStore, st_rmask, K, R, low_src, c);
        d.store_direct = 0;
      }
    }
    // Thread bit of each local position per phase mapping.
    int lane_of[kSlots][16];
    for (int ph = 0; ph < d.nphases; ph++)
      for (int k = 0; k < L; k++) lane_of[ph][d.lanepos[ph][k]] = k;
    pass.transposes = (d.nphases - 1) + (d.load_direct ? 0 : 1) + (d.store_direct ? 0 : 1);
    auto set_io = [&](int slot, const int *pos, uint64_t *reg_off, uint8_t *lane_gbit,
                      uint64_t *lane_off) {
      for (int reg = 0; reg < (1 << R); reg++) {
        uint32_t j = d.regloc[slot][reg];
        uint64_t g = 0;
        for (int p = 0; p < K; p++)
          if (j & (1u << p)) g |= (uint64_t)1 << pos[p];
        reg_off[reg] = g;
      }
      for (int k = 0; k < L; k++) {
        lane_gbit[k] = (uint8_t)pos[d.lanepos[slot][k]];
        lane_off[k] = (uint64_t)1 << lane_gbit[k];
      }
    };
    // Lane layout of the store: the next pass's load lanes contiguous..+2
    // (recorded in the replayed trace) to index bits contiguous..+2, this
    // pass's store lanes contiguous..+2 to the next three free bits, every
    // other bit keeping its order; the natural order when no next pass is
    // known (the last pass, or planning without a trace).
    const int st_slot = d.store_direct ? d.nphases - 1 : kIoStore;
    int mstore[64];
    for (int x = 0; x < 64; x++) mstore[x] = x;
    d.out_place = sh.lane_layout && replaying && pass_index < replay->lanes.size() ? 1 : 0;
    if (d.out_place && pass_index + 1 < replay->lanes.size()) {
      bool assigned[64] = {}, taken[64] = {};
      for (int x = 0; x < c; x++) assigned[x] = taken[x] = true;
      // Bits placed in a given order on the lowest free index bits from `from`.
      auto by_load_order = [&](int *v, int nv) {
        std::sort(v, v + nv, [&](int a, int b) { return mload[phys[newlogi[a]]] < mload[phys[newlogi[b]]]; });
      };
      auto place = [&](const int *v, int nv, int from, int limit) {
        for (int i = 0; i < nv; i++) {
          if (assigned[v[i]]) continue;
          while (from < limit && taken[from]) from++;
          if (from >= limit) return;
          mstore[v[i]] = from;
          assigned[v[i]] = taken[from] = true;
        }
      };
      // 1. The next pass's load lanes on bits c, c+1, c+2.
      const uint32_t nl = replay->lanes[pass_index + 1];
      for (int k = 0; k < 3; k++) {
        const int q = (int)((nl >> (8 * k)) & 0xffu);
        if (q >= n || c + k >= n) continue;
        const int v = newphys[q];
        if (v < c || assigned[v] || taken[c + k]) continue;
        mstore[v] = c + k;
        assigned[v] = taken[c + k] = true;
      }
      // 2. This pass's store lanes c..c+2 on the next free bits (below).
      int own[3], n_own = 0;
      for (int k = c; k < c + 3 && k < L; k++) own[n_own++] = store_pos[d.lanepos[st_slot][k]];
      // 3. The next pass's other tile qubits, in their load order, after them:
      //    the next pass's tile is then within a few low index bits' reach.
      int nxt[64], n_nxt = 0;
      if (pass_index + 1 < replay->tiles.size()) {
        const uint64_t nt = replay->tiles[pass_index + 1];
        for (int v = c; v < n; v++)
          if (!assigned[v] && (nt & ((uint64_t)1 << newlogi[v]))) nxt[n_nxt++] = v;
        by_load_order(nxt, n_nxt);
      }
      // Measured on the 30-qubit layer circuit (gate-layers/s): this order
      // 266-268; the next tile before this pass's store lanes 264-266; this
      // pass's whole tile first 264-267; without step 3, 257.
      place(own, n_own, c, n);
      place(nxt, n_nxt, c, n);
      // 4. The other bits fill the free index bits in the order they are
      // loaded from, so the tile index bits that vary fastest across the grid
      // stay on the lowest non-tile index bits for the store too.
      int rest[64], n_rest = 0;
      for (int v = c; v < n; v++)
        if (!assigned[v]) rest[n_rest++] = v;
      by_load_order(rest, n_rest);
      place(rest, n_rest, c, n);
    }
    int load_phys[16], store_phys[16];
    for (int k = 0; k < K; k++) {
      load_phys[k] = mload[qpos[k]];
      store_phys[k] = mstore[store_pos[k]];
    }
    set_io(d.load_direct ? 0 : kIo, load_phys, d.load_reg_off, d.load_lane_gbit, d.load_lane_off);
    set_io(st_slot, store_phys, d.store_reg_off, d.store_lane_gbit, d.store_lane_off);
    // Store runs: tile index bit k (the non-tile bit nt_virt[k]) to its index
    // bit under the store's lane layout (non-tile bits keep their place in
    // the assignment, only the lane layout moves them).
    {
      bool same = true;
      for (int k = 0; k < n_nt; k++) same = same && mstore[nt_virt[k]] == mload[nt_virt[k]];
      int nr = 0;
      for (int k = 0; k < n_nt && !same;) {
        const int start = k;
        while (k + 1 < n_nt && mstore[nt_virt[k + 1]] == mstore[nt_virt[k]] + 1) k++;
        k++;
        d.store_run_src[nr] = (uint8_t)start;
        d.store_run_dst[nr] = (uint8_t)mstore[nt_virt[start]];
        d.store_run_len[nr] = (uint8_t)(k - start);
        nr++;
      }
      d.store_nruns = nr;
    }

    std::vector<uint64_t> op_touch, op_pair_target;
    std::vector<TileOp> tail;  // merged diagonal ops, kept after every phase's ops
    std::vector<int> headers;
    for (int ph = 0; ph < d.nphases; ph++) {
      uint32_t rset = rsets[ph];
      d.phase_op_begin[ph] = (int32_t)pass.ops.size();
      for (int idx : phase_gates[ph]) {
        const PGate &x = G[idx];
        TileOp t;
        std::memset(&t, 0, sizeof(t));
        t.kind = x.g.kind;
        t.aux0 = x.g.aux0;
        t.aux1 = x.g.aux1;
        t.rbit = -1;
        for (int i = 0; i < 8; i++) t.c[i] = x.g.c[i];
        for (int q = 0; q < n; q++) {
          if (!(x.g.ctrl_mask & ((uint64_t)1 << q))) continue;
          int lp = local_of[q];
          if (lp < 0)
            t.out_cmask |= (uint64_t)1 << mload[phys[q]];
          else if (rset & (1u << lp))
            t.reg_cmask |= 1u << rank_in(rset, lp);
          else
            t.lane_cmask |= 1u << lane_of[ph][lp];
        }
        int lp = local_of[x.g.target];
        if (x.pair) {
          t.rbit = rank_in(rset, lp);
        } else if (lp < 0) {
          t.dout = (uint64_t)1 << mload[phys[x.g.target]];
        } else if (rset & (1u << lp)) {
          t.dreg = 1u << rank_in(rset, lp);
        } else {
          t.dlane = 1u << lane_of[ph][lp];
        }
        if (t.rbit >= 0 && t.reg_cmask != 0 && t.kind != G_X && t.kind < G_PHASEPROD) {
          // Register-controlled: the kernel runs the general 2x2 form, selecting the
          // identity for pairs whose control bits are clear (exact either way).
          to_matrix(x.g, t.c);
          t.kind = G_MAT;
        }
        pass.ops.push_back(t);
        op_touch.push_back(x.touch);
        op_pair_target.push_back(x.pair ? ((uint64_t)1 << x.g.target) : 0);
        pass.gates++;
      }
      headers.clear();
      merge_phase_diagonals(pass.ops, op_touch, op_pair_target, d.phase_op_begin[ph], tail,
                            headers);
      if (!headers.empty()) {
        d.phase_diag_begin[ph] = headers[0];  // rebased below
        d.phase_diag_count[ph] = headers[1];
      }
      op_touch.clear();
      op_pair_target.clear();
      d.phase_op_count[ph] = (int32_t)pass.ops.size() - d.phase_op_begin[ph];
    }
    const int32_t tail_base = (int32_t)pass.ops.size();
    for (int ph = 0; ph < d.nphases; ph++)
      if (d.phase_diag_count[ph]) d.phase_diag_begin[ph] += tail_base;
    pass.ops.insert(pass.ops.end(), tail.begin(), tail.end());
    pass.rsets = rsets;
    if (record) {
      // Load thread bits contiguous..contiguous+2 of this pass, as qubits.
      const int ld_slot = d.load_direct ? 0 : kIo;
      uint32_t lanes = 0xffffffu;
      for (int k = 0; k < 3; k++)
        if (c + k < L) {
          const uint32_t q = (uint32_t)logi[qpos[d.lanepos[ld_slot][c + k]]];
          lanes = (lanes & ~(0xffu << (8 * k))) | (q << (8 * k));
        }
      record->lanes.push_back(lanes);
    }
    plan.passes.push_back(std::move(pass));
    pending.swap(deferred);
    for (int x = 0; x < 64; x++) logi[x] = newlogi[x];
    for (int x = 0; x < 64; x++) mload[x] = mstore[x];
  }
  return plan;
}

FusedPlan plan_fused_searched(const std::vector<DevGate> &gates, int n, const std::vector<FusedShape> &bases,
                              const PlanSearch &opt, double (*cost)(const FusedPlan &), FusedShape *chosen,
                              PlanTrace *trace) {
  if (bases.empty()) fail(QM_ERR_ARG, "no tile shape to plan with");
  struct Try {
    FusedShape shape;
    PlanTrace prefix;
    bool use_prefix = false;
    FusedPlan plan;
    PlanTrace rec;
    double c = 0;
    std::exception_ptr err;
  };
  int threads = opt.threads > 0 ? opt.threads : (int)std::thread::hardware_concurrency();
  threads = std::max(1, std::min(threads, 16));
  // Every try is planned independently (plan_fused keeps no shared state), so
  // the result does not depend on which thread ran which try.
  auto run = [&](std::vector<Try> &tries) {
    std::atomic<size_t> next{0};
    auto work = [&] {
      for (size_t i; (i = next.fetch_add(1)) < tries.size();) {
        Try &t = tries[i];
        try {
          t.plan = plan_fused(gates, n, t.shape, t.use_prefix ? &t.prefix : nullptr, &t.rec);
          t.c = cost(t.plan);
        } catch (...) {
          t.err = std::current_exception();
        }
      }
    };
    const int helpers = std::min<int>(threads, (int)tries.size()) - 1;
    std::vector<std::thread> pool;
    for (int i = 0; i < helpers; i++) pool.emplace_back(work);
    work();
    for (std::thread &th : pool) th.join();
  };
  Try best;
  bool have = false;
  // Lowest cost first, then fewer passes; ties keep the earlier try.
  auto better = [&](const Try &t) {
    if (t.err) return false;
    if (!have) return true;
    if (t.c != best.c) return t.c < best.c;
    return t.plan.passes.size() < best.plan.passes.size();
  };
  uint64_t mix = opt.seed;
  auto next_seed = [&mix] {
    uint64_t z = (mix += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z ? z : 1;
  };
  // (1) the base shapes, deterministic; (2) perturbed restarts over them.
  std::vector<Try> tries(bases.size() + (size_t)std::max(0, opt.restarts));
  for (size_t i = 0; i < tries.size(); i++) {
    tries[i].shape = bases[i % bases.size()];
    tries[i].shape.search_seed = i < bases.size() ? 0 : next_seed();
  }
  run(tries);
  for (size_t i = 0; i < bases.size(); i++)
    if (tries[i].err) std::rethrow_exception(tries[i].err);  // a base shape must plan
  for (Try &t : tries)
    if (better(t)) best = std::move(t), have = true;
  // (3) iterated greedy moves from the best plan.
  for (int round = 0; round < opt.rounds && best.rec.tiles.size() > 1; round++) {
    std::vector<Try> moves((size_t)std::max(0, opt.moves));
    for (Try &m : moves) {
      const size_t k = (size_t)(next_seed() % best.rec.tiles.size());
      m.shape = best.shape;
      m.shape.search_seed = next_seed();
      m.prefix.tiles.assign(best.rec.tiles.begin(), best.rec.tiles.begin() + (long)k);
      m.prefix.rsets.assign(best.rec.rsets.begin(), best.rec.rsets.begin() + (long)k);
      m.use_prefix = true;
    }
    run(moves);
    for (Try &m : moves)
      if (better(m)) best = std::move(m);
  }
  if (chosen) *chosen = best.shape;
  if (trace) *trace = best.rec;
  return std::move(best.plan);
}

namespace {

// Marks pass p to accumulate the masses of `qubit`'s zero and one halves after
// its last phase, in that phase's register/lane/outside coordinates.
void set_mass_epilogue(FusedPass &p, int qubit) {
  DevPass &d = p.dev;
  const int K = popc(d.qmask);
  int local = -1;
  for (int k = 0; k < K; k++)
    if (d.qpos[k] == qubit) local = k;
  d.mass_on = 1;
  d.mass_dreg = d.mass_dlane = 0;
  d.mass_dout = 0;
  if (local < 0) {
    d.mass_dout = (uint64_t)1 << qubit;
    return;
  }
  const uint32_t rset = p.rsets.back();
  if (rset & (1u << local))
    d.mass_dreg = 1u << rank_in(rset, local);
  else
    d.mass_dlane = 1u << rank_in(~rset & ((1u << K) - 1), local);
}

}  // namespace

BatchProgram plan_batch(const std::vector<BatchStep> &steps, int n, const FusedShape &shape) {
  BatchProgram prog;
  prog.shape = shape;
  prog.n = n;
  std::vector<DevGate> seg;
  int32_t draws = 0;
  auto close = [&](int32_t damp_target, double gamma) {
    if (seg.empty()) {
      // A unit phase keeps the pass (the state is still read, reset and written).
      DevGate g{};
      g.kind = G_DIAG1;
      g.swap_other = -1;
      g.c[0] = 1, g.c[6] = 1;
      seg.push_back(g);
    }
    FusedPlan plan = plan_fused(seg, n, shape);
    if (damp_target >= 0) set_mass_epilogue(plan.passes.back(), damp_target);
    BatchSeg bs{};
    bs.pass_begin = (int32_t)prog.passes.size();
    bs.pass_count = (int32_t)plan.passes.size();
    bs.draws = draws;
    bs.damp_target = damp_target;
    bs.gamma = gamma;
    for (FusedPass &p : plan.passes) {
      p.dev.op_offset = (int32_t)prog.ops.size();
      prog.passes.push_back(p.dev);
      prog.ops.insert(prog.ops.end(), p.ops.begin(), p.ops.end());
    }
    prog.source_gates += plan.source_gates;
    prog.segs.push_back(bs);
    prog.draws_static += draws;
    seg.clear();
    draws = 0;
  };
  for (const BatchStep &st : steps) {
    if (st.damp_target >= 0) {
      close(st.damp_target, st.gamma);
      prog.damps++;
      DevGate g{};
      g.kind = G_DAMP;
      g.target = st.damp_target;
      g.swap_other = -1;
      seg.push_back(g);
      continue;
    }
    DevGate g = st.gate;
    if (g.kind == G_NOISE && g.aux0 >= 0) g.aux0 = draws++;
    seg.push_back(g);
  }
  close(-1, 0);
  return prog;
}

// ---------------------------------------------------------------------------
// Dense blocks.
// ---------------------------------------------------------------------------
namespace {

uint64_t gate_qubits(const DevGate &g) {
  uint64_t q = ((uint64_t)1 << g.target) | g.ctrl_mask;
  if (g.kind == G_SWAP) q |= (uint64_t)1 << g.swap_other;
  return q;
}

// The gate with its qubits renamed to positions within the block.
DevGate localise(const DevGate &g, uint64_t qmask) {
  auto pos = [&](int q) { return popc(qmask & (((uint64_t)1 << q) - 1)); };
  DevGate l = g;
  l.target = pos(g.target);
  if (g.kind == G_SWAP) l.swap_other = pos(g.swap_other);
  l.ctrl_mask = 0;
  for (int q = 0; q < 64; q++)
    if (g.ctrl_mask & ((uint64_t)1 << q)) l.ctrl_mask |= (uint64_t)1 << pos(q);
  return l;
}

}  // namespace

std::vector<DenseBlock> plan_dense(const std::vector<DevGate> &gates, int n, int kmax) {
  if (kmax < 1 || kmax > 5) throw std::invalid_argument("dense block width must be 1..5");
  std::vector<DenseBlock> out;
  uint64_t open = 0;
  std::vector<DevGate> run;
  auto flush = [&] {
    if (run.empty()) return;
    DenseBlock b;
    b.qmask = open;
    b.k = popc(open);
    b.gates = (int)run.size();
    const int D = 1 << b.k, M = 2 * D;
    std::vector<double> col(2 * D);
    b.a.assign((size_t)M * M, 0.0);
    for (int c = 0; c < D; c++) {
      std::fill(col.begin(), col.end(), 0.0);
      col[2 * c] = 1.0;
      for (const DevGate &g : run) apply_gate_host(localise(g, open), col.data(), b.k);
      for (int r = 0; r < D; r++) {
        const double re = col[2 * r], im = col[2 * r + 1];
        b.a[(size_t)r * M + c] = re;
        b.a[(size_t)r * M + D + c] = -im;
        b.a[(size_t)(D + r) * M + c] = im;
        b.a[(size_t)(D + r) * M + D + c] = re;
      }
    }
    out.push_back(std::move(b));
    run.clear();
    open = 0;
  };
  for (const DevGate &g : gates) {
    const uint64_t q = gate_qubits(g);
    if (popc(open | q) > kmax) flush();
    open |= q;
    run.push_back(g);
    if (popc(open) > kmax) flush();  // a gate wider than kmax is a block of its own
  }
  flush();
  (void)n;
  return out;
}

void apply_dense_host(const DenseBlock &b, double *s, int n) {
  const int D = 1 << b.k, M = 2 * D;
  std::vector<uint64_t> loc(D);
  std::vector<int> qpos;
  for (int q = 0; q < n; q++)
    if (b.qmask & ((uint64_t)1 << q)) qpos.push_back(q);
  for (int l = 0; l < D; l++) {
    uint64_t off = 0;
    for (int i = 0; i < b.k; i++)
      if (l & (1 << i)) off |= (uint64_t)1 << qpos[i];
    loc[l] = off;
  }
  std::vector<double> x(M), y(M);
  const uint64_t len = (uint64_t)1 << n;
  for (uint64_t base = 0; base < len; base++) {
    if (base & b.qmask) continue;
    for (int l = 0; l < D; l++) {
      x[l] = s[2 * (base | loc[l])];
      x[D + l] = s[2 * (base | loc[l]) + 1];
    }
    for (int r = 0; r < M; r++) {
      double acc = 0;
      for (int c = 0; c < M; c++) acc += b.a[(size_t)r * M + c] * x[c];
      y[r] = acc;
    }
    for (int l = 0; l < D; l++) {
      s[2 * (base | loc[l])] = y[l];
      s[2 * (base | loc[l]) + 1] = y[D + l];
    }
  }
}

// ---------------------------------------------------------------------------
// Host references.
// ---------------------------------------------------------------------------
namespace {

inline void apply_pair_host(int kind, const double *c, double &xr, double &xi, double &yr,
                            double &yi) {
  switch (kind) {
    case G_MAT: pair_mat(c, xr, xi, yr, yi); break;
    case G_RX: pair_rx(c[0], c[1], xr, xi, yr, yi); break;
    case G_RY: pair_ry(c[0], c[1], xr, xi, yr, yi); break;
    case G_H: pair_h(c[0], xr, xi, yr, yi); break;
    case G_Y: pair_y(xr, xi, yr, yi); break;
    case G_X: std::swap(xr, yr), std::swap(xi, yi); break;
    default: break;
  }
}

// The fused kernel's arithmetic (fused multiply-add forms).
inline void apply_pair_fused(int kind, const double *c, double &xr, double &xi, double &yr,
                             double &yi) {
  switch (kind) {
    case G_MAT: pair_mat_fma(c, xr, xi, yr, yi); break;
    case G_RX: pair_rx_fma(c[0], c[1], xr, xi, yr, yi); break;
    case G_RY: pair_ry_fma(c[0], c[1], xr, xi, yr, yi); break;
    case G_H: pair_h_fma(c[0], xr, xi, yr, yi); break;
    case G_Y: pair_y(xr, xi, yr, yi); break;
    case G_X: std::swap(xr, yr), std::swap(xi, yi); break;
    default: break;
  }
}

}  // namespace

void apply_gate_host(const DevGate &g, double *s, int n) {
  const uint64_t len = (uint64_t)1 << n;
  const uint64_t bit = (uint64_t)1 << g.target;
  if (g.kind == G_DIAG || g.kind == G_DIAG1) {
    for (uint64_t i = 0; i < len; i++) {
      if ((i & g.ctrl_mask) != g.ctrl_mask) continue;
      if (i & bit)
        mul_phase(g.c[6], g.c[7], s[2 * i], s[2 * i + 1]);
      else if (g.kind == G_DIAG)
        mul_phase(g.c[0], g.c[1], s[2 * i], s[2 * i + 1]);
    }
    return;
  }
  for (uint64_t i = 0; i < len; i++) {
    if (i & bit) continue;
    if ((i & g.ctrl_mask) != g.ctrl_mask) continue;
    uint64_t j = i | bit;
    if (g.kind == G_SWAP) {
      uint64_t ob = (uint64_t)1 << g.swap_other;
      if (!(i & ob)) continue;
      j ^= ob;
      std::swap(s[2 * i], s[2 * j]);
      std::swap(s[2 * i + 1], s[2 * j + 1]);
      continue;
    }
    apply_pair_host(g.kind, g.c, s[2 * i], s[2 * i + 1], s[2 * j], s[2 * j + 1]);
  }
}

void apply_plan_host(const FusedPlan &plan, double *state, int n) {
  const int K = plan.shape.K, L = plan.shape.L, R = plan.shape.R;
  const int regs = 1 << R, threads = 1 << L;
  std::vector<double> lds(2 * ((size_t)1 << K));
  std::vector<double> v(2 * (size_t)regs * threads);
  // Out-of-place passes (lane layout) alternate between the state and a
  // second buffer, as on the device.
  std::vector<double> other;
  double *s = state;
  for (const FusedPass &p : plan.passes) {
    const DevPass &d = p.dev;
    double *out = s;
    if (d.out_place) {
      if (other.empty()) other.resize(2 * ((size_t)1 << n));
      out = s == state ? other.data() : state;
    }
    const uint64_t tiles = (uint64_t)1 << (n - K);
    for (uint64_t tau = 0; tau < tiles; tau++) {
      uint64_t base = 0;
      for (int r = 0; r < d.nruns; r++)
        base |= ((tau >> d.run_src[r]) & (((uint64_t)1 << d.run_len[r]) - 1)) << d.run_dst[r];
      uint64_t sbase = base;
      if (d.store_nruns) {
        sbase = 0;
        for (int r = 0; r < d.store_nruns; r++)
          sbase |= ((tau >> d.store_run_src[r]) & (((uint64_t)1 << d.store_run_len[r]) - 1))
                   << d.store_run_dst[r];
      }
      auto lane_local = [&](int slot, int t) {
        uint32_t j = 0;
        for (int k = 0; k < L; k++)
          if ((t >> k) & 1) j |= 1u << d.lanepos[slot][k];
        return j;
      };
      auto transpose = [&](int from, int to) {
        for (int t = 0; t < threads; t++)
          for (int r = 0; r < regs; r++) {
            uint32_t j = lane_local(from, t) | d.regloc[from][r];
            lds[2 * j] = v[2 * (t * regs + r)];
            lds[2 * j + 1] = v[2 * (t * regs + r) + 1];
          }
        for (int t = 0; t < threads; t++)
          for (int r = 0; r < regs; r++) {
            uint32_t j = lane_local(to, t) | d.regloc[to][r];
            v[2 * (t * regs + r)] = lds[2 * j];
            v[2 * (t * regs + r) + 1] = lds[2 * j + 1];
          }
      };
      for (int t = 0; t < threads; t++) {
        uint64_t lo = 0;
        for (int k = 0; k < L; k++)
          if ((t >> k) & 1) lo |= (uint64_t)1 << d.load_lane_gbit[k];
        for (int r = 0; r < regs; r++) {
          uint64_t g = base | lo | d.load_reg_off[r];
          v[2 * (t * regs + r)] = s[2 * g];
          v[2 * (t * regs + r) + 1] = s[2 * g + 1];
        }
      }
      int cur = d.load_direct ? 0 : kIo;
      for (int ph = 0; ph < d.nphases; ph++) {
        if (cur != ph) transpose(cur, ph);
        cur = ph;
        for (int o = 0; o < d.phase_op_count[ph]; o++) {
          const TileOp &op = p.ops[d.phase_op_begin[ph] + o];
          bool out_ok = (base & op.out_cmask) == op.out_cmask;
          if (!out_ok) continue;
          for (int t = 0; t < threads; t++) {
            if ((t & op.lane_cmask) != op.lane_cmask) continue;
            double *vt = &v[2 * (size_t)t * regs];
            if (op.rbit >= 0) {
              for (int r = 0; r < regs; r++) {
                if (r & (1 << op.rbit)) continue;
                if ((r & op.reg_cmask) != op.reg_cmask) continue;
                int r1 = r | (1 << op.rbit);
                apply_pair_fused(op.kind, op.c, vt[2 * r], vt[2 * r + 1], vt[2 * r1], vt[2 * r1 + 1]);
              }
            } else {
              bool hi_fixed = ((t & op.dlane) != 0) || ((base & op.dout) != 0);
              for (int r = 0; r < regs; r++) {
                if ((r & op.reg_cmask) != op.reg_cmask) continue;
                bool hi = hi_fixed || ((r & op.dreg) != 0);
                if (hi)
                  mul_phase_fma(op.c[6], op.c[7], vt[2 * r], vt[2 * r + 1]);
                else if (op.kind == G_DIAG)
                  mul_phase_fma(op.c[0], op.c[1], vt[2 * r], vt[2 * r + 1]);
              }
            }
          }
        }
        if (d.phase_diag_count[ph]) {
          const TileOp *q = &p.ops[d.phase_diag_begin[ph]];
          for (int t = 0; t < threads; t++) {
            double pr, pi;
            phase_product(q, d.phase_diag_count[ph], (uint32_t)t, base, pr, pi);
            double *vt = &v[2 * (size_t)t * regs];
            for (int r = 0; r < regs; r++) mul_phase_fma(pr, pi, vt[2 * r], vt[2 * r + 1]);
          }
        }
      }
      if (!d.store_direct) transpose(cur, kIoStore);
      for (int t = 0; t < threads; t++) {
        uint64_t lo = 0;
        for (int k = 0; k < L; k++)
          if ((t >> k) & 1) lo |= (uint64_t)1 << d.store_lane_gbit[k];
        for (int r = 0; r < regs; r++) {
          uint64_t g = sbase | lo | d.store_reg_off[r];
          out[2 * g] = v[2 * (t * regs + r)];
          out[2 * g + 1] = v[2 * (t * regs + r) + 1];
        }
      }
    }
    s = out;
  }
  if (s != state) std::memcpy(state, s, sizeof(double) * 2 * ((size_t)1 << n));
}

}  // namespace qm
