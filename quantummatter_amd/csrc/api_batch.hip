// C ABI of the batched entry points of libqmhip.so: noisy trajectories and
// parameter-shift / multi-parameter energy evaluation, each one kernel launch
// over many states (batch.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "jit.hpp"
#include "qm_internal.hpp"

namespace qm {

TermGroups group_terms(const qm_pauli *terms, int64_t n_terms) {
  TermGroups g;
  g.order.resize((size_t)n_terms);
  std::iota(g.order.begin(), g.order.end(), 0);
  std::stable_sort(g.order.begin(), g.order.end(), [&](int64_t a, int64_t b) {
    return terms[a].flip_mask < terms[b].flip_mask;
  });
  for (int64_t k = 0; k < n_terms;) {
    PauliGroup pg{};
    pg.flip = terms[g.order[k]].flip_mask;
    pg.first = (int32_t)k;
    while (k < n_terms && pg.count < kTermsPerGroup && terms[g.order[k]].flip_mask == pg.flip) {
      pg.sign[pg.count] = terms[g.order[k]].sign_mask;
      pg.y[pg.count] = terms[g.order[k]].y_count;
      pg.count++;
      k++;
    }
    g.groups.push_back(pg);
  }
  return g;
}

namespace {

// One device allocation for everything a batched call uploads and reads back.
struct Arena {
  size_t size = 0;
  void *base = nullptr;
  size_t add(size_t bytes) {
    const size_t off = (size + 255) & ~(size_t)255;
    size = off + bytes;
    return off;
  }
  template <typename P>
  P *at(size_t off) const {
    return reinterpret_cast<P *>(static_cast<char *>(base) + off);
  }
  ~Arena() {
    if (base) (void)hipFree(base);
  }
};

// Welford update of NoisyStateVectorSimulator.EstimateProbabilities
// (NoisyStateVectorSimulator.cs:72-90) over `count` consecutive trajectories
// whose states sit at states + t * len, trajectory `first` being the first of
// them.  Thread j owns basis index j and walks the trajectories in order, so
// every index sees the reference's sequence of updates; the probability is
// re*re + im*im without contraction (StateVector.cs:56-60).
template <typename T>
__global__ __launch_bounds__(256) void k_welford(const T *__restrict__ states, int64_t len,
                                                 int64_t count, const double *__restrict__ mass,
                                                 int64_t first, double *__restrict__ mean,
                                                 double *__restrict__ m2) {
#pragma clang fp contract(off)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += stride) {
    double mu = mean[j], v = m2[j];
    for (int64_t t = 0; t < count; t++) {
      const T a = states[t * len + j];
      const double re = a.x, im = a.y;
      const double p = (re * re + im * im) / mass[t];
      const double delta = p - mu;
      mu += delta / (double)(first + t + 1);
      v += delta * (p - mu);
    }
    mean[j] = mu;
    m2[j] = v;
  }
}

// Standard errors in place of M2: sqrt(max(0, M2) / (M (M - 1)))
// (NoisyStateVectorSimulator.cs:92-101); a NaN M2 stays NaN, as Math.Max keeps it.
__global__ __launch_bounds__(256) void k_std_error(double *__restrict__ m2, int64_t len,
                                                   double divisor) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += stride) {
    const double v = m2[j];
    m2[j] = sqrt((v < 0 ? 0.0 : v) / divisor);
  }
}

// Clears the fallback counter inside a captured graph (a kernel node rather
// than a memset node).
__global__ void k_zero_counter(unsigned long long *p) {
  if (threadIdx.x == 0) *p = 0;
}

}  // namespace

// Device arena of one batched launch: every input the kernel reads and every
// output it writes, at fixed offsets, in one allocation.
struct BatchArena {
  Arena ar;
  int64_t slots = 0;
  size_t o_states = 0, o_items = 0, o_passes = 0, o_ops = 0, o_segs = 0, o_pools = 0, o_angles = 0,
         o_groups = 0, o_numer = 0, o_mass = 0, o_shotu = 0, o_samples = 0, o_outcomes = 0,
         o_consumed = 0, o_err = 0, o_fb = 0, o_part = 0, o_part_mass = 0;
  int32_t slices = 0;  // > 0: partial sums of launch_pauli_slices (set before setup)

  // Sizes the arena for L (workgroup slots within the budget unless L.keep),
  // allocates it and uploads the inputs on the context's stream.
  void setup(Ctx *c, const BatchLaunch &L) {
    const int64_t n_items = (int64_t)L.items.size();
    const int64_t ab = amp_bytes(L.precision);
    if (!L.keep) {
      // Workgroup slots: as many states as can be resident at once, within the budget.
      slots = std::min<int64_t>(n_items, (int64_t)c->compute_units * 2);
      const int64_t limit = c->device_budget ? c->device_budget : c->device_memory;
      const int64_t room = limit - c->resident_bytes - (64ll << 20);
      slots = std::min<int64_t>(slots, room / std::max<int64_t>(L.slot_len * ab, 1));
      if (slots < 1) fail(QM_ERR_BUDGET, "HIP batch exceeds the device memory budget.");
    }
    const int64_t state_amps = L.keep ? L.keep_len : slots * L.slot_len;
    o_states = ar.add((size_t)(state_amps * ab));
    o_items = ar.add(L.items.size() * sizeof(BatchItem));
    o_passes = ar.add(L.prog.passes.size() * sizeof(DevPass));
    o_ops = ar.add(std::max<size_t>(L.prog.ops.size(), 1) * sizeof(TileOp));
    o_segs = ar.add(L.prog.segs.size() * sizeof(BatchSeg));
    o_pools = ar.add(L.pools ? (size_t)(n_items * L.stride) * 8 : 0);
    o_angles = ar.add(L.angles.size() * 8);
    o_groups = ar.add(std::max<size_t>(L.groups.size(), 1) * sizeof(PauliGroup));
    o_numer = ar.add((size_t)std::max<int64_t>(L.n_numer, 1) * 8);
    o_mass = ar.add((size_t)n_items * 8);
    o_shotu = ar.add((size_t)std::max<int64_t>(L.n_shots, 1) * 8);
    o_samples = ar.add((size_t)std::max<int64_t>(L.n_shots, 1) * 8);
    o_outcomes = ar.add((size_t)n_items * 8);
    o_consumed = ar.add((size_t)n_items * 8);
    o_err = ar.add((size_t)n_items * 4);
    o_fb = ar.add(8);
    o_part = ar.add(slices ? (size_t)(std::max<int64_t>(L.n_numer, 1) * slices) * 8 : 0);
    o_part_mass = ar.add(slices ? (size_t)(n_items * slices) * 8 : 0);
    if ((c->device_budget && c->resident_bytes + (int64_t)ar.size > c->device_budget) ||
        c->resident_bytes + (int64_t)ar.size > c->device_memory)
      fail(QM_ERR_BUDGET, "HIP batch exceeds the device memory budget.");
    hipError_t e = hipMalloc(&ar.base, ar.size);
    if (e != hipSuccess) {
      ar.base = nullptr;
      (void)hipGetLastError();
      fail(QM_ERR_BUDGET, "HIP batch exceeds the free device memory.");
    }
    hipStream_t st = c->stream;
    auto up = [&](size_t off, const void *src, size_t bytes) {
      if (bytes) QM_HIP(hipMemcpyAsync(ar.at<char>(off), src, bytes, hipMemcpyHostToDevice, st));
    };
    up(o_items, L.items.data(), L.items.size() * sizeof(BatchItem));
    up(o_passes, L.prog.passes.data(), L.prog.passes.size() * sizeof(DevPass));
    up(o_ops, L.prog.ops.data(), L.prog.ops.size() * sizeof(TileOp));
    up(o_segs, L.prog.segs.data(), L.prog.segs.size() * sizeof(BatchSeg));
    if (L.pools) up(o_pools, L.pools, (size_t)(n_items * L.stride) * 8);
    up(o_angles, L.angles.data(), L.angles.size() * 8);
    up(o_groups, L.groups.data(), L.groups.size() * sizeof(PauliGroup));
    if (L.n_shots) up(o_shotu, L.shot_u, (size_t)L.n_shots * 8);
  }

  BatchArgs args(const BatchLaunch &L) const {
    BatchArgs a{};
    a.states = ar.at<void>(o_states);
    a.slot_len = L.slot_len;
    a.keep = L.keep ? 1 : 0;
    a.slots = L.slots;
    a.items = ar.at<BatchItem>(o_items);
    a.n_items = (int64_t)L.items.size();
    a.passes = ar.at<DevPass>(o_passes);
    a.ops = ar.at<TileOp>(o_ops);
    a.segs = ar.at<BatchSeg>(o_segs);
    a.pools = L.pools ? ar.at<double>(o_pools) : nullptr;
    a.stride = L.stride;
    a.angles = L.angles.empty() ? nullptr : ar.at<double>(o_angles);
    a.groups = ar.at<PauliGroup>(o_groups);
    a.numer = ar.at<double>(o_numer);
    a.mass = ar.at<double>(o_mass);
    a.shot_u = ar.at<double>(o_shotu);
    a.samples = ar.at<int64_t>(o_samples);
    a.outcomes = ar.at<int64_t>(o_outcomes);
    a.consumed = ar.at<int64_t>(o_consumed);
    a.err = ar.at<int32_t>(o_err);
    a.fallbacks = ar.at<unsigned long long>(o_fb);
    return a;
  }
};

// Slices per state when a launch's mass and Pauli numerators are better
// reduced after the batch kernel by launch_pauli_slices, else 0.  Taken when
// every item has the same width, of at least QM_PAULI_SLICE_QUBITS (tuning
// knob, default 15; 0 turns it off) qubits, and the same Pauli groups, and
// when all states can stay resident for the reduction.  One workgroup per
// state would read its state from HBM once per flip group with every state of
// the launch in flight; slicing each state over many workgroups keeps the few
// states being reduced in the last-level cache.  With QM_PAULI_TILES (tuning
// knob, default 1) and the terms' LDS slots within 32 KiB, a slice is a tile
// of 2^kPauliTileBitsHost amplitudes and the tile-pair kernel shares each read
// among the groups of one partner tile (*tile_pairs set); else a slice is at
// least 2048 pairs (8 per thread), at most 64 slices per state.
static int pauli_slices(const Ctx *c, const BatchLaunch &L, bool *tile_pairs, int32_t *n_terms) {
  static const int min_qubits = [] {
    const char *e = std::getenv("QM_PAULI_SLICE_QUBITS");
    return e ? std::atoi(e) : 15;
  }();
  static const bool tiles_on = [] {
    const char *e = std::getenv("QM_PAULI_TILES");
    return e ? std::atoi(e) != 0 : true;
  }();
  if (min_qubits <= 0 || L.keep || L.items.empty() || L.groups.empty()) return 0;
  const BatchItem &f = L.items.front();
  if (f.group_count <= 0 || f.len < (1ull << min_qubits)) return 0;
  for (const BatchItem &it : L.items)
    if (it.len != f.len || it.group_begin != f.group_begin || it.group_count != f.group_count) return 0;
  const int64_t n_items = (int64_t)L.items.size();
  int32_t terms = 0;
  for (int32_t k = f.group_begin; k < f.group_begin + f.group_count; k++)
    terms = std::max(terms, L.groups[(size_t)k].first + L.groups[(size_t)k].count);
  *n_terms = terms;
  *tile_pairs = tiles_on && f.len >= (1ull << kPauliTileBitsHost) && (int64_t)terms * 4 * 8 <= (32 << 10);
  const int slices = *tile_pairs ? (int)(f.len >> kPauliTileBitsHost)
                                 : (int)std::min<uint64_t>(64, std::max<uint64_t>(1, f.len / 2 / 2048));
  // Resident states plus partials within the budget (the slot arena is the fallback).
  const int64_t limit = c->device_budget ? c->device_budget : c->device_memory;
  const int64_t need = n_items * (int64_t)f.len * amp_bytes(L.precision) +
                       (std::max<int64_t>(L.n_numer, 1) + n_items) * slices * 8;
  if (need > limit - c->resident_bytes - (256ll << 20)) return 0;
  return slices;
}

// Turns L into the sliced form when pauli_slices takes it: states kept, the
// batch kernel's own mass and numerators off (each item's group count and
// mass flag zeroed; the groups they had go to *group_begin, *group_count).
static int make_sliced(const Ctx *c, BatchLaunch &L, int32_t *group_begin, int32_t *group_count,
                       bool *tile_pairs, int32_t *n_terms) {
  const int slices = pauli_slices(c, L, tile_pairs, n_terms);
  if (!slices) return 0;
  *group_begin = L.items.front().group_begin;
  *group_count = L.items.front().group_count;
  L.keep = true;
  L.keep_len = 0;
  L.on_states = nullptr;
  for (BatchItem &it : L.items) {
    it.state_off = L.keep_len;
    L.keep_len += (int64_t)it.len;
    it.group_count = 0;
    it.want_mass = 0;
  }
  return slices;
}

void run_batch_launch(Ctx *c, BatchLaunch &L) {
  const int64_t n_items = (int64_t)L.items.size();
  int32_t gb = 0, gc = 0, n_terms = 0;
  bool tile_pairs = false;
  const int slices = make_sliced(c, L, &gb, &gc, &tile_pairs, &n_terms);
  BatchArena A;
  A.slices = slices;
  A.setup(c, L);
  hipStream_t st = c->stream;
  QM_HIP(hipMemsetAsync(A.ar.at<char>(A.o_fb), 0, 8, st));
  const BatchArgs a = A.args(L);
  int64_t grid = L.keep ? 0 : A.slots;
  if (L.timed) QM_HIP(hipEventRecord(c->ev_start, st));
  L.evolved = L.evolve && slices;
  if (L.evolved) {
    QM_HIP(hipMemsetAsync(a.err, 0, (size_t)n_items * sizeof(int32_t), st));
    L.evolve(a.states, st);
  } else {
    launch_batch(c, L.precision, L.shape, a, &grid);
  }
  if (slices)
    launch_pauli_slices(c, L.precision, a.states, a.items, n_items, tile_pairs, a.groups, gb, gc, n_terms, slices,
                        A.ar.at<double>(A.o_part), A.ar.at<double>(A.o_part_mass), a.numer, a.mass);
  if (L.timed) QM_HIP(hipEventRecord(c->ev_stop, st));

  L.err.assign((size_t)n_items, 0);
  L.numer.assign((size_t)L.n_numer, 0.0);
  L.mass.assign((size_t)n_items, 0.0);
  L.samples.assign((size_t)L.n_shots, 0);
  L.outcomes.assign((size_t)n_items, 0);
  L.consumed.assign((size_t)n_items, 0);
  unsigned long long fb = 0;
  auto down = [&](void *dst, size_t off, size_t bytes) {
    if (bytes) QM_HIP(hipMemcpyAsync(dst, A.ar.at<char>(off), bytes, hipMemcpyDeviceToHost, st));
  };
  down(L.err.data(), A.o_err, L.err.size() * 4);
  down(L.numer.data(), A.o_numer, L.numer.size() * 8);
  down(L.mass.data(), A.o_mass, L.mass.size() * 8);
  down(L.samples.data(), A.o_samples, L.samples.size() * 8);
  down(L.outcomes.data(), A.o_outcomes, L.outcomes.size() * 8);
  down(L.consumed.data(), A.o_consumed, L.consumed.size() * 8);
  down(&fb, A.o_fb, 8);
  if (L.keep && L.on_states) L.on_states(A.ar.at<void>(A.o_states), A.ar.at<double>(A.o_mass), st);
  QM_HIP(hipStreamSynchronize(st));
  if (L.timed) {
    float ms = 0;
    QM_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    L.device_ms = ms;
  }
  L.fallbacks = (int64_t)fb;
}

namespace {

void validate_terms(const qm_pauli *terms, int64_t n_terms, int n) {
  if (n_terms < 0 || (n_terms > 0 && !terms)) fail(QM_ERR_ARG, "terms are null");
  const uint64_t full = n >= 64 ? ~0ull : ((1ull << n) - 1);
  for (int64_t i = 0; i < n_terms; i++)
    if ((terms[i].flip_mask | terms[i].sign_mask) & ~full)
      fail(QM_ERR_ARG, "Pauli string width must match the state.");
}

// A batch over `items` same-width states running one program.
struct BatchJob {
  int qubits = 0;
  int precision = QM_C128;
  std::vector<BatchStep> steps;
  int64_t items = 0;
  const double *pools = nullptr;
  int64_t stride = 0;
  std::vector<double> angles;  // [items][slots][2]
  int32_t slots = 0;
  const qm_pauli *terms = nullptr;
  int64_t n_terms = 0;
  bool want_mass = false;
  const double *shot_u = nullptr;
  int64_t shots = 0;
  bool measure = false;
};

struct BatchOut {
  std::vector<double> mass;   // [items]
  std::vector<double> numer;  // [items][n_terms], original term order
  int64_t fallbacks = 0;
  double device_ms = 0;
  int passes = 0;
  double algorithmic_bytes = 0;
  int64_t state_bytes = 0;
};

// The launch description of a job: plan, items, groups and inputs.
TermGroups build_launch(const BatchJob &job, BatchLaunch &L) {
  const int n = job.qubits;
  L.precision = job.precision;
  L.shape = batch_shape(job.precision, n);
  if (L.shape.K == 0) fail(QM_ERR_UNSUPPORTED, "Batched execution needs at least 8 qubits.");
  L.prog = plan_batch(job.steps, n, L.shape);
  TermGroups g = group_terms(job.terms, job.n_terms);
  L.groups = g.groups;
  L.slot_len = (int64_t)1 << n;
  L.items.resize((size_t)job.items);
  for (int64_t i = 0; i < job.items; i++) {
    BatchItem &it = L.items[(size_t)i];
    it = BatchItem{};
    it.len = 1ull << n;
    it.tiles = 1ull << (n - L.shape.K);
    it.seg_begin = 0;
    it.seg_count = (int32_t)L.prog.segs.size();
    it.group_begin = 0;
    it.group_count = (int32_t)g.groups.size();
    it.numer_off = i * job.n_terms;
    it.shots = job.shots;
    it.shot_off = i * job.shots;
    it.measure = job.measure ? 1 : 0;
    it.want_mass = job.want_mass ? 1 : 0;
  }
  L.n_numer = job.items * job.n_terms;
  L.n_shots = job.items * job.shots;
  L.pools = job.pools;
  L.stride = job.stride;
  L.angles = job.angles;
  L.slots = job.slots;
  L.shot_u = job.shot_u;
  return g;
}

// Per-item statuses to errors; masses and numerators (original term order) to `out`.
void collect_out(const BatchJob &job, const TermGroups &g, const int32_t *err, const double *numer,
                 const double *mass, int passes, BatchOut &out) {
  for (int64_t i = 0; i < job.items; i++) {
    const int code = err[i];
    if (code == QM_ERR_POOL) fail(code, "The uniform pool is exhausted.");
    if (code == QM_ERR_MASS) fail(code, "The state has no finite positive probability mass.");
    if (code) fail(code, "batched state failed");
  }
  out.mass.assign(mass, mass + job.items);
  out.numer.resize((size_t)(job.items * job.n_terms));
  for (int64_t i = 0; i < job.items; i++)
    for (int64_t k = 0; k < job.n_terms; k++)
      out.numer[(size_t)(i * job.n_terms + g.order[(size_t)k])] = numer[i * job.n_terms + k];
  const int64_t state_bytes = amp_bytes(job.precision) << job.qubits;
  out.passes = passes;
  out.state_bytes = state_bytes;
  // Each pass reads and writes the state; the final stage reads it once per
  // mass, per Pauli group and for the CDF.
  const double finals = (job.want_mass || job.n_terms ? 1.0 : 0.0) + (double)g.groups.size() +
                        (job.shots || job.measure ? 1.0 : 0.0);
  out.algorithmic_bytes = (double)job.items * (double)state_bytes * (2.0 * (double)passes + finals);
}

// Batched compiled passes: the energy and gradient jobs of states of at least
// QM_BATCH_COMPILED_QUBITS qubits (tuning knob, default 15, where the states
// are kept for the reduction after the passes; 0 turns it off)
// run the compiled schedule's pass kernels over all their states at once
// instead of the interpreting batch kernel.  Every item runs the same gates
// but its own rotation angles: the plan is searched once for the structure,
// each item's plan is replayed from its trace with its angles, and the pass
// kernels (generated without the angle-dependent factored forms, so one
// source serves every item) read a coefficient block per item, filled from
// the op coefficients each slot copies.  The states are kept for the
// reduction after the passes.
static int compiled_batch_qubits() {
  static const int q = [] {
    const char *e = std::getenv("QM_BATCH_COMPILED_QUBITS");
    return e ? std::atoi(e) : 15;
  }();
  return q;
}

static bool compiled_batch_fits(const BatchJob &job) {
  const int q = compiled_batch_qubits();
  if (q <= 0 || job.qubits < q || job.shots || job.measure || job.pools || job.items < 1) return false;
  for (const BatchStep &st : job.steps)
    if (st.damp_target >= 0 || st.gate.kind == G_NOISE || st.gate.kind == G_DAMP) return false;
  return true;
}

// Item i's gates: the dynamic rotations resolved from its angle table
// (angles: [items][slots][2]).
static std::vector<DevGate> item_gates(const BatchJob &job, const double *angles, int64_t i) {
  std::vector<DevGate> g;
  g.reserve(job.steps.size());
  for (const BatchStep &st : job.steps) {
    DevGate d = st.gate;
    if (d.kind == G_PROT || d.kind == G_PROTZ) {
      const double *a = &angles[(size_t)((i * job.slots + d.aux0) * 2)];
      for (double &v : d.c) v = 0;
      if (d.kind == G_PROT) {
        d.kind = d.aux1;  // G_RX or G_RY: c[0] = cos, c[1] = sin of half the angle
        d.c[0] = a[0];
        d.c[1] = a[1];
      } else {
        d.kind = G_DIAG;  // Rz = diag(conj(D), D), D = (cos, sin) of half the angle
        d.c[0] = a[0];
        d.c[1] = -a[1];
        d.c[6] = a[0];
        d.c[7] = a[1];
      }
      d.aux0 = d.aux1 = 0;
    }
    g.push_back(d);
  }
  return g;
}

struct CompiledBatch {
  FusedShape shape;            // the kernels' shape (state_bits set)
  FusedShape chosen;           // the planning shape, replayed per item with `trace`
  PlanTrace trace;
  JitProgram jp;
  std::vector<size_t> op_count;  // ops of each pass (every item's plan has the same)
  std::vector<double> coef;    // per pass: [items][slots of the pass]
  std::vector<int64_t> off;    // first coefficient of each pass's blocks
  int passes = 0;
};

// Every item's coefficient blocks for the angles at `angles`: each item's plan
// replayed with its angles (the same passes and ops), its coefficients copied
// into its blocks; items split over host threads.
static void fill_compiled_coef(const BatchJob &job, const double *angles, const CompiledBatch &r, double *dst) {
  const int n = job.qubits;
  const int threads = (int)std::max<int64_t>(1, std::min<int64_t>(
                          std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())), job.items));
  std::vector<std::string> errs((size_t)threads);
  auto work = [&](int w) {
    try {
      for (int64_t i = w; i < job.items; i += threads) {
        const FusedPlan pi = plan_fused(item_gates(job, angles, i), n, r.chosen, &r.trace);
        if (pi.passes.size() != r.op_count.size()) fail(QM_ERR_DEVICE, "a batch item's plan differs");
        for (size_t p = 0; p < pi.passes.size(); p++) {
          const std::vector<int32_t> &pv = r.jp.prov[p];
          const size_t ns = pv.size() / 2;
          const std::vector<TileOp> &ops = pi.passes[p].ops;
          if (ops.size() != r.op_count[p]) fail(QM_ERR_DEVICE, "a batch item's plan differs");
          double *d = dst + r.off[p] + i * (int64_t)ns;
          for (size_t k = 0; k < ns; k++) d[k] = ops[(size_t)pv[2 * k]].c[pv[2 * k + 1]];
        }
      }
    } catch (const Status &e) {
      errs[(size_t)w] = e.msg.empty() ? "batch item planning failed" : e.msg;
    } catch (...) {
      errs[(size_t)w] = "batch item planning failed";
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < threads; w++) pool.emplace_back(work, w);
  work(0);
  for (std::thread &t : pool) t.join();
  for (const std::string &e : errs)
    if (!e.empty()) fail(QM_ERR_DEVICE, e);
}

static bool prepare_compiled_batch(Ctx *c, const BatchJob &job, CompiledBatch &r) {
  const int n = job.qubits;
  FusedShape base = c->shape[job.precision];
  apply_tuning_knobs(base);
  base.lane_layout = 0;
  base.variant = jit_default_variant(job.precision) & ~(16 | 8192);  // no factored or merged diagonal forms
  if (n < base.K) return false;
  const std::vector<DevGate> g0 = item_gates(job, job.angles.data(), 0);
  FusedPlan plan = plan_compiled_trace(g0, n, job.precision, base, &r.chosen, &r.trace);
  r.chosen.variant = base.variant;  // a cached plan of the structure may come from another variant
  r.chosen.lane_layout = 0;
  if (plan.passes.empty() || !jit_supported(plan)) return false;
  plan = plan_fused(g0, n, r.chosen, &r.trace);
  plan.shape.state_bits = n;
  r.shape = plan.shape;
  r.jp = jit_prepare(plan, job.precision, c->device);
  r.passes = (int)plan.passes.size();
  r.off.assign(plan.passes.size(), 0);
  r.op_count.clear();
  int64_t total = 0;
  for (size_t p = 0; p < plan.passes.size(); p++) {
    r.off[p] = total;
    r.op_count.push_back(plan.passes[p].ops.size());
    total += job.items * (int64_t)(r.jp.prov[p].size() / 2);
  }
  r.coef.assign((size_t)total, 0.0);
  fill_compiled_coef(job, job.angles.data(), r, r.coef.data());
  return true;
}

// The compiled batch's stream work on the kept states: coefficients from
// `h_coef` (pinned when under graph capture), statuses cleared, |0...0>,
// every pass.
static void issue_compiled_batch(Ctx *c, const BatchJob &job, const CompiledBatch &cb, const double *h_coef,
                                 double *d_coef, int32_t *err, void *states, hipStream_t st) {
  if (!cb.coef.empty())
    QM_HIP(hipMemcpyAsync(d_coef, h_coef, cb.coef.size() * sizeof(double), hipMemcpyHostToDevice, st));
  launch_zero_i32(c, err, job.items);
  launch_batch_basis(c, job.precision, states, job.items, 1ull << job.qubits);
  for (int p = 0; p < cb.passes; p++)
    jit_launch_batch(c->compute_units, job.precision, cb.shape, cb.jp, (size_t)p, d_coef + cb.off[(size_t)p],
                     job.items, states, st);
}

void run_batch(Ctx *c, const BatchJob &job, int64_t *samples, int64_t *outcomes, int64_t *consumed,
               bool timed, BatchOut &out) {
  BatchLaunch L;
  const TermGroups g = build_launch(job, L);
  L.timed = timed;
  CompiledBatch cb;
  double *d_coef = nullptr;
  if (compiled_batch_fits(job) && prepare_compiled_batch(c, job, cb)) {
    const int64_t bytes = (int64_t)(cb.coef.size() * sizeof(double));
    d_coef = (double *)ctx_program(c, bytes + 256);
    if (bytes)
      QM_HIP(hipMemcpyAsync(d_coef, cb.coef.data(), (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    L.evolve = [&](void *states, hipStream_t st) {
      launch_batch_basis(c, job.precision, states, job.items, 1ull << job.qubits);
      for (int p = 0; p < cb.passes; p++)
        jit_launch_batch(c->compute_units, job.precision, cb.shape, cb.jp, (size_t)p, d_coef + cb.off[(size_t)p],
                         job.items, states, st);
    };
  }
  run_batch_launch(c, L);
  collect_out(job, g, L.err.data(), L.numer.data(), L.mass.data(),
              L.evolved ? cb.passes : (int)L.prog.passes.size(), out);
  if (samples && job.shots) std::memcpy(samples, L.samples.data(), L.samples.size() * 8);
  if (outcomes && job.measure) std::memcpy(outcomes, L.outcomes.data(), L.outcomes.size() * 8);
  if (consumed) std::memcpy(consumed, L.consumed.data(), L.consumed.size() * 8);
  out.fallbacks = L.fallbacks;
  out.device_ms = L.device_ms;
}

void fill_profile(qm_profile *p, const BatchOut &o) {
  if (!p) return;
  p->device_ms = o.device_ms;
  p->state_bytes = o.state_bytes;
  p->kernel_launches = 1;
  p->passes = o.passes;
  p->algorithmic_bytes = o.algorithmic_bytes;
}

bool valid_mass(double m) { return std::isfinite(m) && m > 0; }

// ExpectationValue of one item's term (StateVector.Observations.cs:107-141).
double expectation(const BatchOut &o, int64_t item, int64_t n_terms, int64_t t) {
  const double m = o.mass[(size_t)item];
  if (!valid_mass(m)) fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
  const double e = o.numer[(size_t)(item * n_terms + t)];
  if (!std::isfinite(e)) fail(QM_ERR_NONFINITE, "The expectation value is nonfinite.");
  return e / m;
}

// ExpectationEnergy (StateVector.Energy.cs:10-46): Neumaier sum in term order.
double energy_of(const BatchOut &o, int64_t item, const double *coeffs, int64_t n_terms) {
  if (!valid_mass(o.mass[(size_t)item]))
    fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
  double sum = 0, corr = 0;
  for (int64_t t = 0; t < n_terms; t++) {
    if (coeffs[t] == 0) continue;
    const double value = coeffs[t] * expectation(o, item, n_terms, t);
    const double next = sum + value;
    corr += std::fabs(sum) >= std::fabs(value) ? (sum - next) + value : (value - next) + sum;
    sum = next;
    if (!std::isfinite(sum) || !std::isfinite(corr))
      fail(QM_ERR_NONFINITE, "The energy accumulation is nonfinite.");
  }
  const double e = sum + corr;
  if (!std::isfinite(e)) fail(QM_ERR_NONFINITE, "The energy is nonfinite.");
  return e;
}

// NoisyCircuit.Execute (NoisyCircuit.cs:48-65) as batch steps: segments to
// gates, channels with p != 0 to branches (CpuNoiseChannelBackend.cs:10-83).
std::vector<BatchStep> noisy_steps(int qubits, const qm_step *steps, int64_t n_steps,
                                   const qm_op *ops, int64_t n_ops) {
  if (n_steps < 0 || (n_steps > 0 && !steps)) fail(QM_ERR_ARG, "steps are null");
  std::vector<BatchStep> out;
  for (int64_t k = 0; k < n_steps; k++) {
    const qm_step &s = steps[k];
    if (!s.is_channel) {
      if (s.op_begin < 0 || s.op_count < 0 || s.op_begin + s.op_count > n_ops)
        fail(QM_ERR_ARG, "bad segment");
      for (const DevGate &g : prepare_all(ops + s.op_begin, s.op_count, qubits)) {
        BatchStep bs;
        bs.gate = g;
        out.push_back(bs);
      }
      continue;
    }
    const double p = s.probability;
    if (p == 0) continue;
    if (s.target < 0 || s.target >= qubits) fail(QM_ERR_RANGE, "target");
    if (!std::isfinite(p) || p < 0 || p > 1)
      fail(QM_ERR_RANGE, "Channel probability must be finite and in [0,1].");
    if (s.noise_kind < 0 || s.noise_kind > 3) fail(QM_ERR_ARG, "unknown noise kind");
    BatchStep bs;
    if (s.noise_kind == 3) {
      bs.damp_target = s.target;
      bs.gamma = p;
    } else {
      DevGate &g = bs.gate;
      std::memset(&g, 0, sizeof(g));
      g.kind = G_NOISE;
      g.target = s.target;
      g.swap_other = -1;
      g.aux0 = (s.noise_kind != 2 && p == 1) ? -1 : 0;  // flips at p == 1 draw nothing
      g.aux1 = s.noise_kind;
      g.c[0] = p;
    }
    out.push_back(bs);
  }
  return out;
}

// EstimateProbabilities over batched trajectories: chunks of trajectories
// are evolved in one launch each (states kept), then one Welford pass per
// chunk folds their normalised probabilities into the running mean and M2 in
// trajectory order.
struct EstimateOut {
  int64_t fallbacks = 0;
  int passes = 0;
  int64_t chunks = 0;
  double device_ms = 0;
};

void run_estimate(Ctx *c, int n, int precision, const std::vector<BatchStep> &steps,
                  int64_t trajectories, const double *pools, int64_t stride, double *means,
                  double *std_errors, int64_t *consumed, bool timed, EstimateOut &out) {
  const FusedShape shape = batch_shape(precision, n);
  if (shape.K == 0) fail(QM_ERR_UNSUPPORTED, "Batched execution needs at least 8 qubits.");
  const BatchProgram prog = plan_batch(steps, n, shape);
  const int64_t len = (int64_t)1 << n;
  const int64_t ab = amp_bytes(precision);
  const int64_t limit = c->device_budget ? c->device_budget : c->device_memory;
  const int64_t room = limit - c->resident_bytes - 16 * len - (256ll << 20);
  const int64_t chunk = std::min<int64_t>({trajectories, 2048, room / (len * ab)});
  if (chunk < 1) fail(QM_ERR_BUDGET, "HIP batch exceeds the device memory budget.");
  double *acc = nullptr;  // [mean | m2], 2 * len doubles
  if (hipMalloc(&acc, (size_t)(16 * len)) != hipSuccess) {
    (void)hipGetLastError();
    fail(QM_ERR_BUDGET, "HIP batch exceeds the free device memory.");
  }
  struct Free {
    double *p;
    ~Free() { (void)hipFree(p); }
  } guard{acc};
  hipStream_t st = c->stream;
  QM_HIP(hipMemsetAsync(acc, 0, (size_t)(16 * len), st));
  const unsigned grid = (unsigned)std::min<int64_t>((len + 255) / 256, (int64_t)c->compute_units * 8);
  if (timed) QM_HIP(hipEventRecord(c->ev_start, st));
  for (int64_t first = 0; first < trajectories; first += chunk) {
    const int64_t count = std::min(chunk, trajectories - first);
    BatchLaunch L;
    L.precision = precision;
    L.shape = shape;
    L.prog = prog;
    L.keep = true;
    L.keep_len = count * len;
    L.items.resize((size_t)count);
    for (int64_t i = 0; i < count; i++) {
      BatchItem &it = L.items[(size_t)i];
      it = BatchItem{};
      it.state_off = i * len;
      it.len = (uint64_t)len;
      it.tiles = (uint64_t)len >> shape.K;
      it.seg_count = (int32_t)prog.segs.size();
      it.want_mass = 1;
    }
    L.pools = pools ? pools + first * stride : nullptr;
    L.stride = stride;
    L.on_states = [&](const void *states, const double *mass, hipStream_t s) {
      if (precision == QM_C128)
        k_welford<double2><<<grid, 256, 0, s>>>(static_cast<const double2 *>(states), len, count,
                                                mass, first, acc, acc + len);
      else
        k_welford<float2><<<grid, 256, 0, s>>>(static_cast<const float2 *>(states), len, count,
                                               mass, first, acc, acc + len);
      QM_HIP(hipGetLastError());
    };
    run_batch_launch(c, L);
    for (int64_t i = 0; i < count; i++) {
      const int code = L.err[(size_t)i];
      if (code == QM_ERR_POOL) fail(code, "The uniform pool is exhausted.");
      if (code == QM_ERR_MASS || (!code && !valid_mass(L.mass[(size_t)i])))
        fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
      if (code) fail(code, "batched state failed");
    }
    if (consumed) std::memcpy(consumed + first, L.consumed.data(), (size_t)count * 8);
    out.fallbacks += L.fallbacks;
    out.chunks++;
  }
  const double divisor = (double)trajectories * (double)(trajectories - 1);
  k_std_error<<<grid, 256, 0, st>>>(acc + len, len, divisor);
  QM_HIP(hipGetLastError());
  if (timed) QM_HIP(hipEventRecord(c->ev_stop, st));
  QM_HIP(hipMemcpyAsync(means, acc, (size_t)(8 * len), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(std_errors, acc + len, (size_t)(8 * len), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  if (timed) {
    float ms = 0;
    QM_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    out.device_ms = ms;
  }
  out.passes = (int)prog.passes.size();
}

// Template ops of a ParameterizedCircuit (ParameterizedCircuit.cs:68-96) as
// batch steps: parameter rotations read their coefficients from the angle table.
struct Template {
  std::vector<BatchStep> steps;
  std::vector<int32_t> rot_kind;   // per occurrence: 6 Rx, 7 Ry, 8 Rz
  std::vector<int32_t> rot_param;  // per occurrence: parameter index
};

Template make_template(int n, const qm_op *ops, const int32_t *param_index, int64_t n_ops,
                       int64_t n_params) {
  if (n_ops < 0 || (n_ops > 0 && (!ops || !param_index))) fail(QM_ERR_ARG, "operations are null");
  Template tp;
  for (int64_t i = 0; i < n_ops; i++) {
    const qm_op &op = ops[i];
    const int32_t p = param_index[i];
    BatchStep bs;
    if (p < 0) {
      validate_op(op, n);
      if (!prepare_gate(op, &bs.gate)) continue;
      tp.steps.push_back(bs);
      continue;
    }
    if (p >= n_params) fail(QM_ERR_RANGE, "parameterIndex");
    if (op.kind != 6 && op.kind != 7 && op.kind != 8) fail(QM_ERR_ARG, "Parameters bind Rx, Ry or Rz only.");
    if (op.target < 0 || op.target >= n) fail(QM_ERR_RANGE, "target");
    DevGate &g = bs.gate;
    std::memset(&g, 0, sizeof(g));
    g.target = op.target;
    g.swap_other = -1;
    g.kind = op.kind == 8 ? G_PROTZ : G_PROT;
    g.aux0 = (int32_t)tp.rot_kind.size();
    g.aux1 = op.kind == 6 ? G_RX : G_RY;
    tp.rot_kind.push_back(op.kind);
    tp.rot_param.push_back(p);
    tp.steps.push_back(bs);
  }
  return tp;
}

// Coefficients of one rotation at angle a, as PreparedStateVectorGate.cs:26-45
// forms them on the host: (cos, sin)(a/2).
void put_angle(double *dst, double a) {
  dst[0] = std::cos(a / 2);
  dst[1] = std::sin(a / 2);
}

void check_parameters(const double *parameters, int64_t n_params) {
  if (n_params < 0 || (n_params > 0 && !parameters)) fail(QM_ERR_ARG, "parameters are null");
  for (int64_t i = 0; i < n_params; i++)
    if (!std::isfinite(parameters[i])) fail(QM_ERR_RANGE, "Parameters must be finite.");
}

// The 1 + 2R items of a parameter-shift gradient: item 0 the bound circuit,
// items 1 + 2r / 2 + 2r rotation occurrence r shifted by +pi/2 / -pi/2
// (VariationalEnergyEvaluator.cs:50-80).  The reference appends a second
// same-axis rotation of angle +-pi/2 (ParameterizedCircuit.cs:49-51,85-89) so
// that the shift survives a huge angle; here the two prepared rotations are
// multiplied on the host into one (see shifted_angle), which keeps the shift
// for every finite angle.
BatchJob gradient_job(int qubits, int precision, const Template &tp, const qm_pauli *terms,
                      int64_t n_terms) {
  const int64_t R = (int64_t)tp.rot_kind.size();
  BatchJob job;
  job.qubits = qubits;
  job.precision = precision;
  job.steps = tp.steps;
  job.items = 1 + 2 * R;
  job.slots = (int32_t)std::max<int64_t>(R, 1);
  job.angles.assign((size_t)(job.items * job.slots * 2), 0.0);
  job.terms = terms;
  job.n_terms = n_terms;
  job.want_mass = true;
  return job;
}

// Coefficients of R(shift) R(a) for one axis, shift = +-pi/2.  Rx, Ry and Rz
// prepared at angle a are all determined by (c, s) = (cos, sin)(a/2)
// (PreparedStateVectorGate.cs:26-45), and the product of two of them is the
// rotation with (c cb - s sb, s cb + c sb), (cb, sb) = (cos, sin)(shift/2):
// the shifted rotation is formed from the prepared coefficients of both
// gates, never from a + shift, which rounds to a once |a| exceeds ~2^53 pi/2
// (a = 1e100 + pi/2 == 1e100) and loses ~|a| 2^-53 of the shift before that.
void shifted_angle(double *dst, double a, double shift) {
  const double c = std::cos(a / 2), s = std::sin(a / 2);
  const double cb = std::cos(shift / 2), sb = std::sin(shift / 2);
  dst[0] = c * cb - s * sb;
  dst[1] = s * cb + c * sb;
}

void gradient_angles(const Template &tp, const double *parameters, int32_t slots, double *angles) {
  const int64_t R = (int64_t)tp.rot_kind.size();
  for (int64_t i = 0; i < 1 + 2 * R; i++)
    for (int64_t r = 0; r < R; r++) {
      const double a = parameters[tp.rot_param[(size_t)r]];
      double *dst = &angles[(size_t)((i * slots + r) * 2)];
      if (i > 0 && (i - 1) / 2 == r)
        shifted_angle(dst, a, (i % 2 == 1) ? M_PI / 2 : -M_PI / 2);
      else
        put_angle(dst, a);
    }
}

// n_sets parameter vectors, one item each.
BatchJob energy_job(int qubits, int precision, const Template &tp, const qm_pauli *terms,
                    int64_t n_terms, int64_t n_sets) {
  BatchJob job;
  job.qubits = qubits;
  job.precision = precision;
  job.steps = tp.steps;
  job.items = n_sets;
  job.slots = (int32_t)std::max<int64_t>((int64_t)tp.rot_kind.size(), 1);
  job.angles.assign((size_t)(job.items * job.slots * 2), 0.0);
  job.terms = terms;
  job.n_terms = n_terms;
  job.want_mass = true;
  return job;
}

void energy_angles(const Template &tp, const double *sets, int64_t n_params, int64_t n_sets,
                   int32_t slots, double *angles) {
  const int64_t R = (int64_t)tp.rot_kind.size();
  for (int64_t i = 0; i < n_sets; i++)
    for (int64_t r = 0; r < R; r++)
      put_angle(&angles[(size_t)((i * slots + r) * 2)], sets[i * n_params + tp.rot_param[(size_t)r]]);
}

// Host assembly in the reference's order (VariationalEnergyEvaluator.cs:50-86):
// per rotation occurrence and term, c * (0.5 plus - 0.5 minus) Neumaier-summed
// into its parameter, then the corrections added.
void assemble_gradient(const BatchOut &out, const Template &tp, const double *coeffs, int64_t n_terms,
                       int64_t n_params, double *energy, double *gradient) {
  const int64_t R = (int64_t)tp.rot_kind.size();
  *energy = energy_of(out, 0, coeffs, n_terms);
  std::vector<double> grad((size_t)n_params, 0.0), corr((size_t)n_params, 0.0);
  std::vector<double> plus((size_t)n_terms);
  for (int64_t r = 0; r < R; r++) {
    const int64_t param = tp.rot_param[(size_t)r];
    const int64_t ip = 1 + 2 * r, im = 2 + 2 * r;
    if (n_terms == 0 && !valid_mass(out.mass[(size_t)ip]))
      fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
    for (int64_t t = 0; t < n_terms; t++) plus[(size_t)t] = expectation(out, ip, n_terms, t);
    if (n_terms == 0 && !valid_mass(out.mass[(size_t)im]))
      fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
    for (int64_t t = 0; t < n_terms; t++) {
      const double minus = expectation(out, im, n_terms, t);
      const double contribution = coeffs[t] * (0.5 * plus[(size_t)t] - 0.5 * minus);
      const double prev = grad[(size_t)param];
      const double next = prev + contribution;
      corr[(size_t)param] += std::fabs(prev) >= std::fabs(contribution) ? (prev - next) + contribution
                                                                          : (contribution - next) + prev;
      grad[(size_t)param] = next;
      if (!std::isfinite(next) || !std::isfinite(corr[(size_t)param]))
        fail(QM_ERR_NONFINITE, "The gradient accumulation is nonfinite.");
    }
  }
  for (int64_t p = 0; p < n_params; p++) {
    grad[(size_t)p] += corr[(size_t)p];
    if (!std::isfinite(grad[(size_t)p])) fail(QM_ERR_NONFINITE, "The gradient is nonfinite.");
    gradient[p] = grad[(size_t)p];
  }
}

// One batched job captured as a hipGraph and replayed: the arena stays
// allocated, the angle table is uploaded from pinned memory by the graph's
// first node and the per-item outputs land in pinned memory.  Only the angles
// change between replays.
struct GraphedBatch {
  Ctx *c = nullptr;
  BatchJob job;
  BatchLaunch L;
  TermGroups g;
  BatchArena A;
  double *h_angles = nullptr;
  char *h_out = nullptr;  // err | numer | mass | fallbacks
  size_t off_numer = 0, off_mass = 0, off_fb = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int slices = 0;          // sliced mass and numerators (make_sliced), else 0
  int32_t group_begin = 0, group_count = 0, n_terms = 0;
  bool tile_pairs = false;  // tile-pair reduction (pauli_slices)
  int64_t held = 0;        // arena bytes counted in the context's resident bytes
  uint64_t last_use = 0;   // program tick of the last evaluation (cache eviction)
  // Batched compiled passes (compiled_batch_fits): coefficient blocks formed
  // on the host from each evaluation's angles, copied from pinned memory.
  bool compiled = false;
  CompiledBatch cb;
  double *h_coef = nullptr, *d_coef = nullptr;

  ~GraphedBatch() {
    if (c && held) c->resident_bytes -= held;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_angles) (void)hipHostFree(h_angles);
    if (h_out) (void)hipHostFree(h_out);
    if (h_coef) (void)hipHostFree(h_coef);
    if (d_coef) (void)hipFree(d_coef);
  }

  // The stream work of one evaluation, issued eagerly or under capture.
  void issue() {
    hipStream_t st = c->stream;
    const BatchArgs a = A.args(L);
    QM_HIP(hipMemcpyAsync(A.ar.at<char>(A.o_angles), h_angles, L.angles.size() * 8,
                          hipMemcpyHostToDevice, st));
    k_zero_counter<<<1, 64, 0, st>>>(A.ar.at<unsigned long long>(A.o_fb));
    QM_HIP(hipGetLastError());
    int64_t grid = L.keep ? 0 : A.slots;
    if (compiled)
      issue_compiled_batch(c, job, cb, h_coef, d_coef, a.err, a.states, st);
    else
      launch_batch(c, L.precision, L.shape, a, &grid);
    const int64_t n = (int64_t)L.items.size();
    if (slices)
      launch_pauli_slices(c, L.precision, a.states, a.items, n, tile_pairs, a.groups, group_begin, group_count,
                          n_terms, slices,
                          A.ar.at<double>(A.o_part), A.ar.at<double>(A.o_part_mass), a.numer, a.mass);
    QM_HIP(hipMemcpyAsync(h_out, A.ar.at<char>(A.o_err), (size_t)n * 4, hipMemcpyDeviceToHost, st));
    QM_HIP(hipMemcpyAsync(h_out + off_numer, A.ar.at<char>(A.o_numer), (size_t)L.n_numer * 8,
                          hipMemcpyDeviceToHost, st));
    QM_HIP(hipMemcpyAsync(h_out + off_mass, A.ar.at<char>(A.o_mass), (size_t)n * 8, hipMemcpyDeviceToHost, st));
    QM_HIP(hipMemcpyAsync(h_out + off_fb, A.ar.at<char>(A.o_fb), 8, hipMemcpyDeviceToHost, st));
  }

  void build(Ctx *ctx, BatchJob &&j) {
    c = ctx;
    job = std::move(j);
    g = build_launch(job, L);
    slices = make_sliced(c, L, &group_begin, &group_count, &tile_pairs, &n_terms);
    A.slices = slices;
    A.setup(c, L);
    // The arena stays allocated while the program lives: later state and batch
    // budget checks see it.
    held = (int64_t)A.ar.size;
    c->resident_bytes += held;
    if (slices && compiled_batch_fits(job) && prepare_compiled_batch(c, job, cb)) {
      const size_t bytes = std::max<size_t>(cb.coef.size(), 1) * sizeof(double);
      if (hipHostMalloc((void **)&h_coef, bytes) != hipSuccess || hipMalloc((void **)&d_coef, bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(QM_ERR_BUDGET, "memory for a graphed batch's coefficients is not available");
      }
      std::memcpy(h_coef, cb.coef.data(), cb.coef.size() * sizeof(double));
      held += (int64_t)bytes;
      c->resident_bytes += (int64_t)bytes;
      compiled = true;
    }
    const int64_t n = (int64_t)L.items.size();
    if (hipHostMalloc((void **)&h_angles, std::max<size_t>(L.angles.size(), 1) * 8) != hipSuccess ||
        (off_numer = ((size_t)n * 4 + 7) & ~(size_t)7, off_mass = off_numer + (size_t)L.n_numer * 8,
         off_fb = off_mass + (size_t)n * 8,
         hipHostMalloc((void **)&h_out, off_fb + 8) != hipSuccess)) {
      (void)hipGetLastError();
      fail(QM_ERR_BUDGET, "pinned host memory for a graphed batch is not available");
    }
    std::memcpy(h_angles, L.angles.data(), L.angles.size() * 8);
    // One eager evaluation first: it loads the kernel and sets its launch
    // attributes outside the capture.
    issue();
    QM_HIP(hipStreamSynchronize(c->stream));
    QM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    issue();
    QM_HIP(hipStreamEndCapture(c->stream, &graph));
    QM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }

  // Replays with the angles already in h_angles; fills `out` like run_batch.
  void replay(bool timed, BatchOut &out) {
    hipStream_t st = c->stream;
    if (compiled) fill_compiled_coef(job, h_angles, cb, h_coef);
    if (timed) QM_HIP(hipEventRecord(c->ev_start, st));
    QM_HIP(hipGraphLaunch(exec, st));
    if (timed) QM_HIP(hipEventRecord(c->ev_stop, st));
    QM_HIP(hipStreamSynchronize(st));
    collect_out(job, g, reinterpret_cast<const int32_t *>(h_out),
                reinterpret_cast<const double *>(h_out + off_numer),
                reinterpret_cast<const double *>(h_out + off_mass),
                compiled ? cb.passes : (int)L.prog.passes.size(), out);
    out.fallbacks = (int64_t) * reinterpret_cast<const unsigned long long *>(h_out + off_fb);
    if (timed) {
      float ms = 0;
      QM_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
      out.device_ms = ms;
    }
  }
};

// A variational program: one template and Hamiltonian, evaluated many times
// (a VQE loop) through graphed batches.
struct VqeProgram {
  Ctx *c = nullptr;
  int qubits = 0, precision = QM_C128;
  int64_t n_params = 0;
  Template tp;
  std::vector<qm_pauli> terms;
  std::vector<double> coeffs;
  std::unique_ptr<GraphedBatch> grad;
  std::map<int64_t, std::unique_ptr<GraphedBatch>> energy;  // by number of parameter sets
  uint64_t tick = 0;
};

// Graphed energy batches kept per program (distinct n_sets values); the least
// recently used one is freed to make room for another.
constexpr size_t kMaxEnergyGraphs = 4;

Ctx *batch_ctx(qm_ctx *h, qm_profile *profile) {
  Ctx *c = reinterpret_cast<Ctx *>(h);
  if (!c) fail(QM_ERR_ARG, "context is null");
  use_device(c);
  if (profile && !c->profiling)
    fail(QM_ERR_STATE, "EnableProfiling must be selected when creating the HIP backend.");
  return c;
}

void check_width(int32_t qubits, int32_t precision) {
  if (qubits < 1 || qubits > 34) fail(QM_ERR_RANGE, "qubits must be between 1 and 34.");
  if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

int qm_batch_min_qubits(int32_t precision, int32_t *qubits) {
  return guarded([&] {
    if (!qubits) fail(QM_ERR_ARG, "qubits is null");
    if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    int q = 1;
    while (q <= 34 && batch_shape(precision, q).K == 0) q++;
    *qubits = q;
  });
}

int qm_trajectories_run(qm_ctx *h, int32_t qubits, int32_t precision, const qm_step *steps,
                        int64_t n_steps, const qm_op *ops, int64_t n_ops, int64_t trajectories,
                        const double *pools, int64_t pool_stride, int64_t shots,
                        const double *shot_uniforms, int32_t measure_all, int64_t *samples,
                        int64_t *outcomes, int64_t *consumed, int64_t *fallbacks,
                        qm_profile *profile) {
  return guarded([&] {
    Ctx *c = batch_ctx(h, profile);
    check_width(qubits, precision);
    if (trajectories < 1) fail(QM_ERR_RANGE, "trajectories must be positive.");
    if (pool_stride < 0 || (pool_stride > 0 && !pools)) fail(QM_ERR_ARG, "pools are null");
    if (shots < 0 || (shots > 0 && (!shot_uniforms || !samples))) fail(QM_ERR_ARG, "shot buffers are null");
    if (measure_all && !outcomes) fail(QM_ERR_ARG, "outcomes are null");
    std::lock_guard<std::recursive_mutex> lk(c->lock);
    BatchJob job;
    job.qubits = qubits;
    job.precision = precision;
    job.items = trajectories;
    job.pools = pool_stride > 0 ? pools : nullptr;
    job.stride = pool_stride;
    job.shot_u = shot_uniforms;
    job.shots = shots;
    job.measure = measure_all != 0;
    job.steps = noisy_steps(qubits, steps, n_steps, ops, n_ops);
    BatchOut out;
    run_batch(c, job, samples, outcomes, consumed, profile != nullptr, out);
    if (fallbacks) *fallbacks = out.fallbacks;
    fill_profile(profile, out);
  });
}

int qm_trajectories_estimate(qm_ctx *h, int32_t qubits, int32_t precision, const qm_step *steps,
                             int64_t n_steps, const qm_op *ops, int64_t n_ops,
                             int64_t trajectories, const double *pools, int64_t pool_stride,
                             double *means, double *std_errors, int64_t *consumed,
                             int64_t *fallbacks, qm_profile *profile) {
  return guarded([&] {
    Ctx *c = batch_ctx(h, profile);
    check_width(qubits, precision);
    if (trajectories < 2) fail(QM_ERR_RANGE, "trajectories must be at least 2.");
    if (pool_stride < 0 || (pool_stride > 0 && !pools)) fail(QM_ERR_ARG, "pools are null");
    if (!means || !std_errors) fail(QM_ERR_ARG, "outputs are null");
    std::lock_guard<std::recursive_mutex> lk(c->lock);
    const std::vector<BatchStep> bsteps = noisy_steps(qubits, steps, n_steps, ops, n_ops);
    EstimateOut out;
    run_estimate(c, qubits, precision, bsteps, trajectories, pool_stride > 0 ? pools : nullptr,
                 pool_stride, means, std_errors, consumed, profile != nullptr, out);
    if (fallbacks) *fallbacks = out.fallbacks;
    if (profile) {
      const int64_t state_bytes = amp_bytes(precision) << qubits;
      profile->device_ms = out.device_ms;
      profile->state_bytes = state_bytes;
      profile->kernel_launches = (int32_t)(2 * out.chunks + 1);
      profile->passes = out.passes;
      // Per trajectory: every pass reads and writes the state, the mass and the
      // Welford pass read it; per chunk the running mean and M2 are read and
      // written; the standard-error pass reads and writes M2.
      const double len = std::ldexp(1.0, qubits);
      profile->algorithmic_bytes = (double)trajectories * (double)state_bytes * (2.0 * out.passes + 2.0) +
                                   (double)out.chunks * 32.0 * len + 16.0 * len;
    }
  });
}

int qm_energy_gradient_batch(qm_ctx *h, int32_t qubits, int32_t precision, const qm_op *ops,
                             const int32_t *param_index, int64_t n_ops, const double *parameters,
                             int64_t n_params, const qm_pauli *terms, const double *coeffs,
                             int64_t n_terms, double *energy, double *gradient,
                             qm_profile *profile) {
  return guarded([&] {
    Ctx *c = batch_ctx(h, profile);
    check_width(qubits, precision);
    check_parameters(parameters, n_params);
    validate_terms(terms, n_terms, qubits);
    if (n_terms > 0 && !coeffs) fail(QM_ERR_ARG, "coefficients are null");
    if (!energy || (n_params > 0 && !gradient)) fail(QM_ERR_ARG, "outputs are null");
    std::lock_guard<std::recursive_mutex> lk(c->lock);
    Template tp = make_template(qubits, ops, param_index, n_ops, n_params);
    BatchJob job = gradient_job(qubits, precision, tp, terms, n_terms);
    gradient_angles(tp, parameters, job.slots, job.angles.data());
    BatchOut out;
    run_batch(c, job, nullptr, nullptr, nullptr, profile != nullptr, out);
    assemble_gradient(out, tp, coeffs, n_terms, n_params, energy, gradient);
    fill_profile(profile, out);
  });
}

int qm_energy_batch(qm_ctx *h, int32_t qubits, int32_t precision, const qm_op *ops,
                    const int32_t *param_index, int64_t n_ops, const double *parameter_sets,
                    int64_t n_params, int64_t n_sets, const qm_pauli *terms, const double *coeffs,
                    int64_t n_terms, double *energies, qm_profile *profile) {
  return guarded([&] {
    Ctx *c = batch_ctx(h, profile);
    check_width(qubits, precision);
    if (n_sets < 1) fail(QM_ERR_RANGE, "n_sets must be positive.");
    check_parameters(parameter_sets, n_params * n_sets);
    validate_terms(terms, n_terms, qubits);
    if (n_terms > 0 && !coeffs) fail(QM_ERR_ARG, "coefficients are null");
    if (!energies) fail(QM_ERR_ARG, "energies is null");
    std::lock_guard<std::recursive_mutex> lk(c->lock);
    Template tp = make_template(qubits, ops, param_index, n_ops, n_params);
    BatchJob job = energy_job(qubits, precision, tp, terms, n_terms, n_sets);
    energy_angles(tp, parameter_sets, n_params, n_sets, job.slots, job.angles.data());
    BatchOut out;
    run_batch(c, job, nullptr, nullptr, nullptr, profile != nullptr, out);
    for (int64_t i = 0; i < n_sets; i++) energies[i] = energy_of(out, i, coeffs, n_terms);
    fill_profile(profile, out);
  });
}

int qm_vqe_program_create(qm_ctx *h, int32_t qubits, int32_t precision, const qm_op *ops,
                          const int32_t *param_index, int64_t n_ops, int64_t n_params,
                          const qm_pauli *terms, const double *coeffs, int64_t n_terms,
                          qm_vqe_program **out) {
  return guarded([&] {
    Ctx *c = batch_ctx(h, nullptr);
    if (!out) fail(QM_ERR_ARG, "out is null");
    check_width(qubits, precision);
    if (batch_shape(precision, qubits).K == 0)
      fail(QM_ERR_UNSUPPORTED, "Batched execution needs at least 8 qubits.");
    if (n_params < 0) fail(QM_ERR_RANGE, "parameterCount");
    validate_terms(terms, n_terms, qubits);
    if (n_terms > 0 && !coeffs) fail(QM_ERR_ARG, "coefficients are null");
    auto p = std::make_unique<VqeProgram>();
    p->c = c;
    p->qubits = qubits;
    p->precision = precision;
    p->n_params = n_params;
    p->tp = make_template(qubits, ops, param_index, n_ops, n_params);
    p->terms.assign(terms, terms + n_terms);
    p->coeffs.assign(coeffs, coeffs + n_terms);
    *out = reinterpret_cast<qm_vqe_program *>(p.release());
  });
}

int qm_vqe_program_gradient(qm_vqe_program *h, const double *parameters, double *energy,
                            double *gradient, qm_profile *profile) {
  return guarded([&] {
    VqeProgram *p = reinterpret_cast<VqeProgram *>(h);
    if (!p) fail(QM_ERR_ARG, "program is null");
    batch_ctx(reinterpret_cast<qm_ctx *>(p->c), profile);
    check_parameters(parameters, p->n_params);
    if (!energy || (p->n_params > 0 && !gradient)) fail(QM_ERR_ARG, "outputs are null");
    std::lock_guard<std::recursive_mutex> lk(p->c->lock);
    const int64_t n_terms = (int64_t)p->terms.size();
    if (!p->grad) {
      auto gb = std::make_unique<GraphedBatch>();
      BatchJob job = gradient_job(p->qubits, p->precision, p->tp, p->terms.data(), n_terms);
      gradient_angles(p->tp, parameters, job.slots, job.angles.data());
      gb->build(p->c, std::move(job));
      p->grad = std::move(gb);
    }
    GraphedBatch &gb = *p->grad;
    gradient_angles(p->tp, parameters, gb.job.slots, gb.h_angles);
    BatchOut out;
    gb.replay(profile != nullptr, out);
    assemble_gradient(out, p->tp, p->coeffs.data(), n_terms, p->n_params, energy, gradient);
    fill_profile(profile, out);
  });
}

int qm_vqe_program_energies(qm_vqe_program *h, const double *parameter_sets, int64_t n_sets,
                            double *energies, qm_profile *profile) {
  return guarded([&] {
    VqeProgram *p = reinterpret_cast<VqeProgram *>(h);
    if (!p) fail(QM_ERR_ARG, "program is null");
    batch_ctx(reinterpret_cast<qm_ctx *>(p->c), profile);
    if (n_sets < 1) fail(QM_ERR_RANGE, "n_sets must be positive.");
    check_parameters(parameter_sets, p->n_params * n_sets);
    if (!energies) fail(QM_ERR_ARG, "energies is null");
    std::lock_guard<std::recursive_mutex> lk(p->c->lock);
    const int64_t n_terms = (int64_t)p->terms.size();
    if (!p->energy.count(n_sets) && p->energy.size() >= kMaxEnergyGraphs) {
      auto lru = p->energy.begin();
      for (auto it = p->energy.begin(); it != p->energy.end(); ++it)
        if (it->second->last_use < lru->second->last_use) lru = it;
      QM_HIP(hipStreamSynchronize(p->c->stream));
      p->energy.erase(lru);
    }
    std::unique_ptr<GraphedBatch> &slot = p->energy[n_sets];
    if (!slot) {
      auto gb = std::make_unique<GraphedBatch>();
      BatchJob job = energy_job(p->qubits, p->precision, p->tp, p->terms.data(), n_terms, n_sets);
      energy_angles(p->tp, parameter_sets, p->n_params, n_sets, job.slots, job.angles.data());
      gb->build(p->c, std::move(job));
      slot = std::move(gb);
    }
    slot->last_use = ++p->tick;
    energy_angles(p->tp, parameter_sets, p->n_params, n_sets, slot->job.slots, slot->h_angles);
    BatchOut out;
    slot->replay(profile != nullptr, out);
    for (int64_t i = 0; i < n_sets; i++) energies[i] = energy_of(out, i, p->coeffs.data(), n_terms);
    fill_profile(profile, out);
  });
}

int qm_vqe_program_destroy(qm_vqe_program *h) {
  return guarded([&] {
    VqeProgram *p = reinterpret_cast<VqeProgram *>(h);
    if (!p) return;
    use_device(p->c);
    {
      std::lock_guard<std::recursive_mutex> lk(p->c->lock);
      QM_HIP(hipStreamSynchronize(p->c->stream));
      p->grad.reset();
      p->energy.clear();
    }
    delete p;
  });
}

}  // extern "C"
