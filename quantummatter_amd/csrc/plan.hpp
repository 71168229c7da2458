// Host-side gate preparation and fused-pass planning for libqmhip.so.
//
// Pure C++ (no HIP types) so the planner is unit-tested on the CPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "qm_hip.h"

namespace qm {

// Device gate kinds.  Each is an arithmetic specialisation of the reference's
// 2x2 update (CpuStateVectorBackend.cs:93-177) that rounds exactly as the
// general complex form does for that coefficient pattern (no FMA anywhere).
enum GKind : int32_t {
  G_MAT = 0,    // general 2x2 complex matrix (A,B,C,D)
  G_RX = 1,     // (c, -i s; -i s, c): c[0]=c, c[1]=s
  G_RY = 2,     // (c, -s; s, c):      c[0]=c, c[1]=s
  G_H = 3,      // (h, h; h, -h):      c[0]=h
  G_Y = 4,      // (0, -i; i, 0): exact permutation with sign
  G_X = 5,      // exchange (X, CNOT, Toffoli)
  G_DIAG = 6,   // diag(A, D) on both halves: A = c[0..1], D = c[6..7]
  G_DIAG1 = 7,  // diag(1, D): A is exactly 1, only the one-half changes
  G_SWAP = 8,   // exchange of (a=1,t=0) with (a=0,t=1); swap_other = a
  // Fused passes only: the next `rbit` ops are diagonal gates whose targets
  // and controls are lane bits or outside the tile; their per-lane product is
  // formed once and applied to every register in one multiplication.
  G_PHASEPROD = 9,
  // Batched passes only: ops resolved per state at run time.
  G_NOISE = 10,  // Pauli channel branch drawn from the state's uniform pool (aux0 = draw offset
                 // within the segment or -1 for no draw, aux1 = noise kind, c[0] = probability)
  G_DAMP = 11,   // amplitude-damping Kraus branch decided on the device before the pass
  G_PROT = 12,   // rotation whose (c, s) come from the state's angle table (aux0 = table slot,
                 // aux1 = G_RX or G_RY)
  G_PROTZ = 13,  // Rz whose D = (cos, sin)(theta/2) comes from the angle table (A = conj(D))
};

struct DevGate {
  int32_t kind;
  int32_t target;
  int32_t swap_other;  // G_SWAP: the other qubit; else -1
  int32_t pad;
  uint64_t ctrl_mask;
  double c[8];  // coefficients, layout per kind
  int32_t aux0;  // dynamic kinds: see GKind
  int32_t aux1;
};

// Validates one GateOperation as QuantumCircuit.Add does (QuantumCircuit.cs:97-126);
// throws qm::Status on failure.
void validate_op(const qm_op &op, int qubits);

// PreparedStateVectorGate.Prepare (PreparedStateVectorGate.cs:13-70), refined
// into a device gate kind.  Returns false for identity (skipped, as every
// reference backend skips StateVectorKernelKind.Identity).
bool prepare_gate(const qm_op &op, DevGate *out);

// Fraction of the state a gate reads and writes.
double touched_fraction(const DevGate &g);

// ---------------------------------------------------------------------------
// Fused tile passes.
//
// A pass reads every amplitude once, applies a run of gates in registers, and
// writes it back.  The state is cut into tiles of 2^K amplitudes spanned by K
// "tile qubits" Q, which always include the lowest `contiguous` qubits so
// every global access is a >=128-byte contiguous run.  Each of the 2^L threads
// of a tile holds 2^R amplitudes (K = L + R); which R tile bits are register
// bits changes between phases by an LDS transpose.  A pair gate's target must
// be a register bit of its phase; a diagonal gate's target and every control
// may be a register bit, a lane bit or a qubit outside the tile.
//
// Remapping (FusedShape::remap): a pass may store its tile under a different
// assignment of qubits to its index bits than it loaded it with, so the low
// `contiguous` index bits (the contiguous runs) hold qubits the next pass
// needs instead of always qubits 0..contiguous-1.  The assignment is kept as
// disjoint transpositions of a low index bit p with a higher one h (qubit h at
// index bit p, qubit p at index bit h), and every plan ends with the identity
// assignment, so states outside a plan are always in natural order.
// ---------------------------------------------------------------------------
constexpr int kMaxK = 14;
constexpr int kMaxPhases = 12;
constexpr int kMaxRegs = 32;
constexpr int kIo = kMaxPhases;             // mapping slot of the load layout
constexpr int kIoStore = kMaxPhases + 1;     // mapping slot of the store layout
constexpr int kSlots = kMaxPhases + 2;

struct TileOp {
  int32_t kind;         // GKind (swaps are planned as three exchanges)
  int32_t rbit;         // register bit of the target for pair kinds, else -1
  uint32_t reg_cmask;   // control bits that are register bits (mask on register index)
  uint32_t lane_cmask;  // control bits that are lane bits (mask on thread index)
  uint64_t out_cmask;   // control bits outside the tile (global mask)
  uint32_t dreg;        // diagonal target as a register-index mask (0 if not a register bit)
  uint32_t dlane;       // diagonal target as a thread-index mask (0 if not a lane bit)
  uint64_t dout;        // diagonal target as a global mask outside the tile (0 if inside)
  double c[8];
  int32_t aux0;         // dynamic kinds: see GKind
  int32_t aux1;
};
static_assert(sizeof(TileOp) == 112, "TileOp layout");

struct DevPass {
  uint64_t qmask;                     // global mask of the tile qubits
  uint64_t load_reg_off[kMaxRegs];    // global offset of register r under the load mapping
  uint64_t store_reg_off[kMaxRegs];   // ... under the store mapping
  int32_t n;
  int32_t nphases;
  int32_t load_direct;   // 1: load straight into phase 0's mapping
  int32_t store_direct;  // 1: store straight from the last phase's mapping
  int32_t nruns;         // runs of non-tile bits (tile index -> global base)
  int32_t op_offset;     // first TileOp of this pass in the program
  int32_t phase_op_begin[kMaxPhases];  // relative to op_offset
  int32_t phase_op_count[kMaxPhases];
  // Merged lane/outside diagonal ops applied at the end of each phase as one
  // per-lane phase product (indices relative to op_offset; count 0 = none).
  int32_t phase_diag_begin[kMaxPhases];
  int32_t phase_diag_count[kMaxPhases];
  // Batched passes: masses of one qubit's zero and one halves accumulated after
  // the last phase (mass_on), its position relative to the last phase's mapping.
  int32_t mass_on;
  uint32_t mass_dreg;
  uint32_t mass_dlane;
  uint32_t mass_pad;
  uint64_t mass_dout;
  uint16_t regloc[kSlots][kMaxRegs];  // local index of register r per mapping
  uint16_t regswz[kSlots][kMaxRegs];  // lds_swizzle(regloc), see lds_swizzle()
  uint32_t lane_swz[kSlots][16];      // lds_swizzle(1 << lanepos) of thread bit k
  uint64_t load_lane_off[16];                 // global offset of thread bit k (load mapping)
  uint64_t store_lane_off[16];                // ... (store mapping)
  uint8_t lanepos[kSlots][16];        // local position of thread bit k per mapping
  uint8_t load_lane_gbit[16];                 // global bit of thread bit k under the load mapping
  uint8_t store_lane_gbit[16];
  uint8_t run_src[16];
  uint8_t run_dst[16];
  uint8_t run_len[16];
  uint8_t qpos[16];  // tile index bits (before the lane layout), ascending
  // Runs of the store's tile index -> global base when the pass stores under
  // another lane layout than it loaded with (FusedShape::lane_layout);
  // store_nruns == 0: the load's runs.
  int32_t store_nruns;
  int32_t out_place;  // 1: the pass reads one state buffer and writes another (lane layout)
  uint8_t store_run_src[32];
  uint8_t store_run_dst[32];
  uint8_t store_run_len[32];
};

// LDS address swizzle: XOR of the low 3 bits of a 16-byte-slot index with a
// fold of its higher bits.  Linear over GF(2), so the swizzle of a disjoint OR
// is the XOR of the swizzles (lane part and register part are swizzled apart).
// The planner precomputes swizzled addresses into the pass tables, so the
// swizzle can be chosen per plan without touching the kernel: any map that is
// linear over GF(2) works.  Mode (default 3): 0 none; 1 cascaded fold;
// 2 XOR with bits 3..5; 3 XOR with the 3-bit groups 3..5, 6..8, 9..11, 12..14.
uint32_t lds_swizzle_mode(uint32_t j, int mode);
int lds_swizzle_default();
inline uint32_t lds_swizzle(uint32_t j) { return lds_swizzle_mode(j, lds_swizzle_default()); }

// Per-lane product of the diagonal ops q[0..count) that follow a G_PHASEPROD
// header (lane/outside targets and controls only), shared by the kernel and
// the host reference.
template <typename F>
#if defined(__HIPCC__)
__host__ __device__
#endif
inline void phase_product(const TileOp *q, int count, uint32_t t, uint64_t base, F &pr, F &pi) {
  pr = 1;
  pi = 0;
  for (int k = 0; k < count; k++) {
    const bool on = ((base & q[k].out_cmask) == q[k].out_cmask) &&
                    ((t & q[k].lane_cmask) == q[k].lane_cmask);
    const bool hi = ((base & q[k].dout) != 0) || ((t & q[k].dlane) != 0);
    const bool full = q[k].kind == 6;  // G_DIAG
    const F fr = on ? (F)(hi ? q[k].c[6] : (full ? q[k].c[0] : 1.0)) : (F)1;
    const F fi = on ? (F)(hi ? q[k].c[7] : (full ? q[k].c[1] : 0.0)) : (F)0;
    const F nr = pr * fr - pi * fi;
    const F ni = pr * fi + pi * fr;
    pr = nr;
    pi = ni;
  }
}

struct FusedShape {
  int K;           // tile bits
  int L;           // lane bits (threads = 2^L)
  int R;           // register bits
  int contiguous;  // lowest index bits always in the tile
  int variant = -1;  // code variant: interpreter mode or generator variant (-1: default)
  // 0: the lowest `contiguous` qubits stay at the low index bits; 1: passes
  // move qubits into and out of them (a transpose before the store where the
  // last phase cannot store directly); 2: as 1, but only moves that keep the
  // store direct (and the final restore).
  int remap = 0;
  // 1: each pass's tile is also grown qubit by qubit for the most runnable
  // gates, and the better of the two choices is taken (slower planning).
  int search = 0;
  // Phases per pass at most (0: kMaxPhases).  Bounds the work of one pass, and
  // with it the registers its generated kernel needs.
  int max_phases = 0;
  // Compiled schedule only (0: off): the circuit is also planned with this
  // many low bits kept in the tiles, each with at most 9 phases per pass as
  // well, and the plan with the lowest modelled time runs (see run_gates).
  // The greedy tile search finds fewer passes with one or the other value
  // circuit by circuit.
  int alt_contiguous = 0;
  // Register sets of the phases: 0 in gate order (the targets of the first
  // pending gates); 1 also grown one register position at a time by the
  // position that lets the most pending gates of the pass run in the phase,
  // the better of the two taken; 2 as 1, then improved by exchanging one
  // position for another while that lets more gates run.
  int phase_search = 2;
  // With phase_search: gates a first phase's register set is charged for
  // holding a contiguous low bit (the load then needs a transpose).
  int low_penalty = 0;
  // Tile search only (0: deterministic): seed of a perturbation of the tile
  // growth, which adds a pseudo-random 0..kSearchNoise to each candidate
  // qubit's count of runnable gates, so a restart explores other tiles.
  uint64_t search_seed = 0;
  // Lane layout (1: on; takes effect when a plan is replayed from a trace
  // that holds each pass's load lanes).  Every pass stores the state under a
  // permutation of the index bits above the `contiguous` low bits chosen so
  // that the next pass's thread bits contiguous..contiguous+2 (the lane bits
  // of a wave above one contiguous run) load from index bits
  // contiguous..contiguous+2, this pass's own store lanes go to the next
  // three bits and the next pass's other tile qubits right above: each load
  // and store instruction of a wave then covers 2^(contiguous+3) consecutive
  // amplitudes or a few nearby runs instead of eight runs far apart, and the
  // next pass's tiles are compact.  The last pass stores in natural order.
  int lane_layout = 0;
  // Batched compiled passes (> 0): the pass kernels run over consecutive
  // states of 2^state_bits amplitudes (tile index = state << (state_bits - K)
  // | tile), each state with its own block of the pass's coefficients.
  int state_bits = 0;
};
constexpr int kSearchNoise = 2;

struct FusedPass {
  DevPass dev;
  std::vector<TileOp> ops;  // phases contiguous
  int gates = 0;            // source gates applied (after swap expansion)
  int transposes = 0;       // LDS round trips
  std::vector<uint32_t> rsets;  // register set (tile-local mask) of each phase
};

struct FusedPlan {
  FusedShape shape{};
  int n = 0;
  std::vector<FusedPass> passes;
  int64_t source_gates = 0;
  // Concatenated ops of all passes; fills DevPass::op_offset.
  std::vector<TileOp> flat_ops();
};

// The searched decisions of a plan: each pass's tile (qubit mask) and each
// phase's register set (tile-local mask, before unused slots are filled).
// The planner's decisions depend on the gates' kinds, targets and controls
// only, never on their coefficients, so a trace recorded for one circuit
// replays exactly for every circuit of the same structure (new angles).
struct PlanTrace {
  std::vector<uint64_t> tiles;
  std::vector<std::vector<uint32_t>> rsets;
  // Qubits of each pass's load thread bits contiguous..contiguous+2, 8 bits
  // each (0xff: none), for the lane layout of a replay.
  std::vector<uint32_t> lanes;
};

// Plans prepared gates (identity removed) for an n-qubit state, n >= shape.K.
// With `replay`, the tile and phase searches are skipped and the recorded
// decisions are taken; with `record`, the decisions are written there.
FusedPlan plan_fused(const std::vector<DevGate> &gates, int n, const FusedShape &shape,
                     const PlanTrace *replay = nullptr, PlanTrace *record = nullptr);

// Searched planning (compiled schedule).  The greedy tile search fills each
// pass with the most runnable gates, which is myopic: on brick circuits the
// tiles of one pass decide how much the next passes can run.  The search
// (1) plans every base shape deterministically, (2) restarts the planner with
// perturbed tile growth (FusedShape::search_seed) over the bases, and (3)
// improves the best plan by iterated greedy moves: keep the best plan's first
// k tiles (k drawn per move), replan the rest with a perturbed search, and
// keep the result if it costs less.  Restarts and each round of moves run on
// `threads` host threads (0: hardware threads, at most 16); every plan is
// recorded, so the chosen plan replays exactly from its trace.  Returns the
// plan of least `cost`; its shape and trace go to *chosen and *trace.
struct PlanSearch {
  int restarts = 48;   // perturbed restarts over the base shapes
  int rounds = 6;      // rounds of iterated greedy moves
  int moves = 16;      // moves per round (run in parallel)
  int threads = 0;
  uint64_t seed = 0x5EED;
};
FusedPlan plan_fused_searched(const std::vector<DevGate> &gates, int n, const std::vector<FusedShape> &bases,
                              const PlanSearch &opt, double (*cost)(const FusedPlan &), FusedShape *chosen,
                              PlanTrace *trace);

// ---------------------------------------------------------------------------
// Batched programs: one program run on many same-width states, each state by
// one workgroup.  A program is a list of segments; each segment is a fused
// plan whose passes may hold ops resolved per state (noise branches drawn from
// the state's own uniform pool, rotations from its angle table).  A segment
// that ends in amplitude damping accumulates the masses of the damping target
// in its last pass; the branch is decided on the device and applied by a
// G_DAMP op at the start of the next segment.
// ---------------------------------------------------------------------------
struct BatchSeg {
  int32_t pass_begin;
  int32_t pass_count;
  int32_t draws;        // static draws of the segment's noise ops
  int32_t damp_target;  // -1: no damping after the segment
  double gamma;
};

struct BatchStep {
  // One gate (static or dynamic), or a damping channel when damp_target >= 0.
  DevGate gate;
  int32_t damp_target = -1;
  double gamma = 0;
};

struct BatchProgram {
  FusedShape shape{};
  int n = 0;
  std::vector<DevPass> passes;
  std::vector<TileOp> ops;
  std::vector<BatchSeg> segs;
  int64_t draws_static = 0;   // draws of all noise ops (damping draws are dynamic)
  int64_t damps = 0;
  int64_t source_gates = 0;
};

BatchProgram plan_batch(const std::vector<BatchStep> &steps, int n, const FusedShape &shape);

// ---------------------------------------------------------------------------
// Dense blocks: consecutive gates whose qubits fit in k are multiplied into one
// 2^k x 2^k complex matrix, applied to the state in one pass as a real
// (2^(k+1) x 2^(k+1)) by (2^(k+1) x columns) product on the matrix cores.
// ---------------------------------------------------------------------------
struct DenseBlock {
  uint64_t qmask = 0;     // block qubits (local bit b = b-th lowest qubit of the mask)
  int k = 0;              // popcount(qmask), 1..5
  int gates = 0;          // gates multiplied into the block
  std::vector<double> a;  // real form [[Re U, -Im U], [Im U, Re U]], row-major, 2^(k+1) square
};

// Greedy blocking in gate order: a gate joins the open block while the union of
// qubits stays within kmax (kmax 1..5); gates acting on more qubits than kmax
// form a block of their own.
std::vector<DenseBlock> plan_dense(const std::vector<DevGate> &gates, int n, int kmax);

// Host reference of the dense pass (the kernel's index mapping, FP64 sums).
void apply_dense_host(const DenseBlock &b, double *state, int n);

// Host reference of the kernels, used by the planner's CPU tests: applies a
// plan / one gate to an interleaved complex128 state with the kernels'
// arithmetic and indexing.
void apply_plan_host(const FusedPlan &plan, double *state, int n);
void apply_gate_host(const DevGate &g, double *state, int n);

}  // namespace qm
