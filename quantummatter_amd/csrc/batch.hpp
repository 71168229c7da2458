// Batched program execution (batch.hip): kernel arguments and launcher.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "dev_reduce.hpp"
#include "plan.hpp"

namespace qm {

struct Ctx;

// One state of a batch: where it lives, which segments of the program it runs
// and what is measured on it afterwards.
struct BatchItem {
  int64_t state_off;    // amplitude offset of the state (kept states only)
  uint64_t len;         // amplitudes, 2^n with n >= K
  uint64_t tiles;       // 2^(n - K)
  int32_t seg_begin;    // program segments run by this item
  int32_t seg_count;
  int32_t group_begin;  // Pauli groups evaluated after the program
  int32_t group_count;
  int64_t numer_off;    // first numerator slot of the item
  int64_t shots;        // terminal shots
  int64_t shot_off;     // first uniform / sample slot of the item
  int32_t measure;      // 1: a MeasureAll draw from the pool after the shots
  int32_t want_mass;    // 1: write the total mass
};

struct BatchArgs {
  void *states;          // state buffers: per item (keep) or per workgroup slot
  int64_t slot_len;      // amplitudes per workgroup slot (keep == 0)
  int32_t keep;          // 1: item states live at states + item.state_off
  int32_t slots;         // angle-table entries per item
  const BatchItem *items;
  int64_t n_items;
  const DevPass *passes;
  const TileOp *ops;
  const BatchSeg *segs;
  const double *pools;   // [items][stride] uniform pools, or null
  int64_t stride;
  const double *angles;  // [items][slots][2] rotation coefficients, or null
  const PauliGroup *groups;
  double *numer;         // Pauli numerators at item.numer_off + group.first + k
  double *mass;          // [items]
  const double *shot_u;  // uniforms at item.shot_off
  int64_t *samples;      // samples at item.shot_off
  int64_t *outcomes;     // [items]
  int64_t *consumed;     // [items] draws consumed from the pool
  int32_t *err;          // [items] status per item
  unsigned long long *fallbacks;  // margin fallbacks (sequential re-decisions)
};

// Pauli terms grouped by flip mask (stable in term order, at most
// kTermsPerGroup per group); order[k] = original index of sorted term k.
struct TermGroups {
  std::vector<PauliGroup> groups;
  std::vector<int64_t> order;
};
TermGroups group_terms(const qm_pauli *terms, int64_t n_terms);

// Host description of one batched launch.  Inputs are set by the caller;
// run_batch_launch uploads them in one arena, launches once, and fills the
// outputs (per item status, numerators in group order, masses, samples,
// outcomes, draw counts).  With `keep`, item states live at
// items[i].state_off in one buffer of keep_len amplitudes and `on_states` is
// called with it and the device array of item masses (on the context's
// stream, after the batch kernel) before the arena is released.
struct BatchLaunch {
  int precision = QM_C128;
  FusedShape shape{};
  BatchProgram prog;
  std::vector<PauliGroup> groups;
  std::vector<BatchItem> items;
  int64_t slot_len = 0;
  bool keep = false;
  int64_t keep_len = 0;
  int64_t n_numer = 0;
  int64_t n_shots = 0;
  const double *pools = nullptr;  // [items][stride]
  int64_t stride = 0;
  std::vector<double> angles;     // [items][slots][2]
  int32_t slots = 0;
  const double *shot_u = nullptr;  // [n_shots]
  bool timed = false;
  std::function<void(const void *states, const double *mass, hipStream_t stream)> on_states;
  // Replaces the batch kernel when the states are kept for the reduction
  // after it (batched compiled passes): evolves every item's state at
  // states + item.state_off from |0...0>.
  std::function<void(void *states, hipStream_t stream)> evolve;
  bool evolved = false;  // set by run_batch_launch: evolve ran instead of the batch kernel
  std::vector<int32_t> err;
  std::vector<double> numer, mass;
  std::vector<int64_t> samples, outcomes, consumed;
  int64_t fallbacks = 0;
  double device_ms = 0;
};

void run_batch_launch(Ctx *c, BatchLaunch &L);

// Zeroes n 32-bit values (a kernel, so it can be captured into a graph).
void launch_zero_i32(Ctx *c, int32_t *v, int64_t n);
// Sets each of `items` consecutive states of `len` amplitudes to |0...0>.
void launch_batch_basis(Ctx *c, int precision, void *states, int64_t items, uint64_t len);

// Largest compiled batch tile shape with K <= qubits (K = 0: none).
FusedShape batch_shape(int precision, int qubits);

// Launches the batch kernel; *grid in: cap on workgroups for slot buffers
// (0: none), out: workgroups launched.
void launch_batch(Ctx *c, int precision, const FusedShape &shape, const BatchArgs &args,
                  int64_t *grid);

// Mass and Pauli numerators of kept states of one width (items' state_off),
// many workgroups per state, then the partial sums added in slice order
// (k_pauli_fold) into mass[item] and numer[numer_off + first + k].  With
// tile_pairs, slices = len / 2^kPauliTileBitsHost tile units (k_pauli_tiles,
// n_terms * 32 bytes of LDS), else len / slices-amplitude slices
// (k_pauli_slices).  part holds n_numer x slices doubles, part_mass
// n_items x slices.
constexpr int kPauliTileBitsHost = 12;
void launch_pauli_slices(Ctx *c, int precision, const void *states, const BatchItem *items, int64_t n_items,
                         bool tile_pairs, const PauliGroup *groups, int32_t group_begin, int32_t group_count,
                         int32_t n_terms, int32_t slices, double *part, double *part_mass, double *numer,
                         double *mass);

}  // namespace qm
