// C ABI of libqmhip.so (include/qm_hip.h): contexts, resident states and the
// host runtime around the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <exception>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "dense.hpp"
#include "jit.hpp"
#include "qm_internal.hpp"

namespace qm {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
void clear_error() { g_error.clear(); }

void fail(int code, const std::string &msg) { throw Status{code, msg}; }

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
      fail(QM_ERR_BUDGET, std::string("device allocation failed: ") + what);
    fail(QM_ERR_DEVICE, std::string(hipGetErrorString(e)) + " in " + what);
  }
}

double *ctx_partials(Ctx *c, int64_t n) {
  if (n > c->partial_cap) {
    if (c->d_partial) {
      QM_HIP(hipStreamSynchronize(c->stream));
      QM_HIP(hipFree(c->d_partial));
      c->d_partial = nullptr;
    }
    int64_t cap = std::max<int64_t>(n, 1 << 16);
    QM_HIP(hipMalloc(&c->d_partial, (size_t)cap * sizeof(double)));
    c->partial_cap = cap;
  }
  return c->d_partial;
}

int64_t *ctx_ipartials(Ctx *c, int64_t n) {
  if (n > c->ipartial_cap) {
    if (c->d_ipartial) {
      QM_HIP(hipStreamSynchronize(c->stream));
      QM_HIP(hipFree(c->d_ipartial));
      c->d_ipartial = nullptr;
    }
    int64_t cap = std::max<int64_t>(n, 1 << 14);
    QM_HIP(hipMalloc(&c->d_ipartial, (size_t)cap * sizeof(int64_t)));
    c->ipartial_cap = cap;
  }
  return c->d_ipartial;
}

double *ctx_pinned(Ctx *c, int64_t n) {
  if (n > c->pinned_cap) {
    if (c->h_pinned) {
      QM_HIP(hipStreamSynchronize(c->stream));
      QM_HIP(hipHostFree(c->h_pinned));
      c->h_pinned = nullptr;
    }
    int64_t cap = std::max<int64_t>(n, 4096);
    QM_HIP(hipHostMalloc((void **)&c->h_pinned, (size_t)cap * sizeof(double), hipHostMallocDefault));
    c->pinned_cap = cap;
  }
  return c->h_pinned;
}

void *ctx_program(Ctx *c, int64_t bytes) {
  if (bytes > c->prog_cap) {
    if (c->d_prog) {
      QM_HIP(hipStreamSynchronize(c->stream));
      QM_HIP(hipFree(c->d_prog));
      c->d_prog = nullptr;
    }
    int64_t cap = std::max<int64_t>(bytes, 1 << 20);
    QM_HIP(hipMalloc(&c->d_prog, (size_t)cap));
    c->prog_cap = cap;
  }
  return c->d_prog;
}

// Host <-> device copies of large buffers through two pinned staging chunks of
// the context: the DMA of one chunk overlaps the host-side copy of the other,
// which runs the PCIe link faster than copies from pageable caller memory.
// Synchronous on the context's stream; small buffers are copied directly.
namespace {
constexpr size_t kStageChunk = (size_t)8 << 20;   // two of them pinned per context
constexpr size_t kDirectCopy = (size_t)1 << 20;   // smaller buffers are copied directly
}

void copy_to_device(Ctx *c, void *dst, const void *src, size_t bytes) {
  if (bytes <= kDirectCopy) {
    QM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    QM_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  double *pin = ctx_pinned(c, (int64_t)(2 * kStageChunk / sizeof(double)));
  char *stage[2] = {reinterpret_cast<char *>(pin), reinterpret_cast<char *>(pin) + kStageChunk};
  hipEvent_t done[2];
  QM_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
  QM_HIP(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
  const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
  try {
    for (size_t k = 0; k < chunks; k++) {
      const size_t off = k * kStageChunk, n = std::min(kStageChunk, bytes - off);
      if (k >= 2) QM_HIP(hipEventSynchronize(done[k & 1]));  // the chunk's buffer is free again
      std::memcpy(stage[k & 1], static_cast<const char *>(src) + off, n);
      QM_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, stage[k & 1], n, hipMemcpyHostToDevice, c->stream));
      QM_HIP(hipEventRecord(done[k & 1], c->stream));
    }
    QM_HIP(hipStreamSynchronize(c->stream));
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipEventDestroy(done[0]);
    (void)hipEventDestroy(done[1]);
    throw;
  }
  QM_HIP(hipEventDestroy(done[0]));
  QM_HIP(hipEventDestroy(done[1]));
}

void copy_to_host(Ctx *c, void *dst, const void *src, size_t bytes) {
  if (bytes <= kDirectCopy) {
    QM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    QM_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  double *pin = ctx_pinned(c, (int64_t)(2 * kStageChunk / sizeof(double)));
  char *stage[2] = {reinterpret_cast<char *>(pin), reinterpret_cast<char *>(pin) + kStageChunk};
  hipEvent_t done[2];
  QM_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
  QM_HIP(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
  const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t k) {
    const size_t off = k * kStageChunk, n = std::min(kStageChunk, bytes - off);
    QM_HIP(hipMemcpyAsync(stage[k & 1], static_cast<const char *>(src) + off, n, hipMemcpyDeviceToHost,
                          c->stream));
    QM_HIP(hipEventRecord(done[k & 1], c->stream));
  };
  try {
    issue(0);
    for (size_t k = 0; k < chunks; k++) {
      QM_HIP(hipEventSynchronize(done[k & 1]));
      if (k + 1 < chunks) issue(k + 1);
      const size_t off = k * kStageChunk, n = std::min(kStageChunk, bytes - off);
      std::memcpy(static_cast<char *>(dst) + off, stage[k & 1], n);
    }
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipEventDestroy(done[0]);
    (void)hipEventDestroy(done[1]);
    throw;
  }
  QM_HIP(hipEventDestroy(done[0]));
  QM_HIP(hipEventDestroy(done[1]));
}

void use_device(Ctx *c) {
  if (c == nullptr) fail(QM_ERR_ARG, "context is null");
  if (c->disposed) fail(QM_ERR_DISPOSED, "The HIP backend has been disposed.");
  QM_HIP(hipSetDevice(c->device));
}

void check_cancel(const volatile int32_t *cancel) {
  if (cancel && *cancel) fail(QM_ERR_CANCELLED, "The operation was canceled.");
}

std::vector<DevGate> prepare_all(const qm_op *ops, int64_t n_ops, int n) {
  if (n_ops < 0 || (n_ops > 0 && ops == nullptr)) fail(QM_ERR_ARG, "operations are null");
  std::vector<DevGate> gates;
  gates.reserve((size_t)n_ops);
  for (int64_t i = 0; i < n_ops; i++) {
    validate_op(ops[i], n);
    DevGate g;
    if (prepare_gate(ops[i], &g)) gates.push_back(g);
  }
  return gates;
}


// Tuning knobs read from the environment (unset: no change): QM_CONTIGUOUS
// (lowest index bits kept in every tile), QM_REMAP (remapping mode 0..2),
// QM_SEARCH (tile search 0/1), QM_MAX_PHASES (phases per pass at most),
// QM_PHASE_SEARCH (register sets of the phases 0..2) and QM_PHASE_LOW_PENALTY
// (charge, in gates, of a contiguous low bit in a pass's first register set)
// and QM_LANE_LAYOUT (0 turns the compiled schedule's lane layout off).
void apply_tuning_knobs(FusedShape &shape) {
  auto knob = [](const char *name) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : -1;
  };
  static const int contiguous = knob("QM_CONTIGUOUS"), remap = knob("QM_REMAP"), search = knob("QM_SEARCH"),
                   max_phases = knob("QM_MAX_PHASES"), phase_search = knob("QM_PHASE_SEARCH"),
                   low_penalty = knob("QM_PHASE_LOW_PENALTY"), lane_layout = knob("QM_LANE_LAYOUT");
  if (lane_layout == 0 || lane_layout == 1) shape.lane_layout = shape.lane_layout && lane_layout;
  if (phase_search >= 0 && phase_search <= 2) shape.phase_search = phase_search;
  if (low_penalty >= 0) shape.low_penalty = low_penalty;
  if (contiguous >= 0 && contiguous < shape.K) shape.contiguous = contiguous, shape.alt_contiguous = 0;
  if (remap >= 0 && remap <= 2) shape.remap = remap;
  if (search == 0 || search == 1) shape.search = search;
  if (max_phases >= 1) shape.max_phases = max_phases, shape.alt_contiguous = 0;
}

// Plans by gate structure.  The planner's decisions (each pass's tile, each
// phase's register set, and for the compiled schedule the chosen candidate)
// depend on the gates' kinds, targets and controls only, so they are kept per
// structure, width, precision and shape, and a circuit of a known structure
// (new angles: a parameter sweep, a VQE loop, a repeated served circuit) is
// planned by replaying them instead of searching again.  At most kPlanCacheMax
// structures are kept, the oldest dropped first.
namespace {
constexpr size_t kPlanCacheMax = 256;
struct CachedPlan {
  FusedShape shape;
  PlanTrace trace;
};
struct PlanCache {
  std::mutex mu;
  std::unordered_map<std::string, CachedPlan> map;
  std::deque<std::string> order;
};
PlanCache &plan_cache() {
  static PlanCache *c = new PlanCache;  // outlives every context
  return *c;
}
std::string plan_key(const std::vector<DevGate> &gates, int n, int precision, const FusedShape &sh,
                     int schedule) {
  std::string k;
  auto put = [&](const void *p, size_t bytes) { k.append((const char *)p, bytes); };
  const int32_t head[] = {n,           precision,      schedule,         sh.K,         sh.L,
                          sh.R,        sh.contiguous,  sh.remap,         sh.search,    sh.max_phases,
                          sh.alt_contiguous, sh.phase_search, sh.low_penalty, sh.lane_layout};
  put(head, sizeof head);
  k.reserve(k.size() + gates.size() * 28);
  for (const DevGate &g : gates) {
    const int32_t v[] = {g.kind, g.target, g.swap_other, g.aux0, g.aux1};
    put(v, sizeof v);
    put(&g.ctrl_mask, sizeof g.ctrl_mask);
  }
  return k;
}
bool plan_cache_get(const std::string &key, CachedPlan *out) {
  PlanCache &c = plan_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.map.find(key);
  if (it == c.map.end()) return false;
  *out = it->second;
  return true;
}
void plan_cache_put(const std::string &key, CachedPlan v) {
  PlanCache &c = plan_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.map.count(key)) return;
  while (c.order.size() >= kPlanCacheMax) {
    c.map.erase(c.order.front());
    c.order.pop_front();
  }
  c.order.push_back(key);
  c.map.emplace(key, std::move(v));
}
// plan_fused through the cache: replays the structure's decisions when known.
FusedPlan plan_cached(const std::vector<DevGate> &gates, int n, int precision, const FusedShape &shape,
                      int schedule) {
  const std::string key = plan_key(gates, n, precision, shape, schedule);
  CachedPlan hit;
  if (plan_cache_get(key, &hit)) return plan_fused(gates, n, hit.shape, &hit.trace);
  CachedPlan rec{shape, {}};
  FusedPlan plan = plan_fused(gates, n, shape, nullptr, &rec.trace);
  plan_cache_put(key, std::move(rec));
  return plan;
}
// Searched planning of the compiled schedule (plan_fused_searched), cached by
// structure.  Base shapes: the shape's low-bit count and its alternative, each
// with the full phase budget and with at most 9 phases per pass.  Cost: the
// pass-time model fitted on nine plans of the 30-qubit c128 layer circuit at
// the default 8192-amplitude tiles with two workgroups per CU (0.077 ms per
// gate, the same in every plan, + 1.8 ms per pass + 0.26 ms per LDS exchange
// over a 16 GiB state; within 2.5 %).  Search budget: 96 perturbed restarts
// and 8 rounds of 16 iterated greedy moves below 28 qubits; from 28 qubits,
// where one pass of a state costs ~2.5 ms or more and the plan is cached per
// structure, 1000 restarts and 60 rounds (30-qubit 40-layer layer circuit:
// 17 -> 16 passes, 226 -> 237 gate-layers/s, ~1 s of planning on 16 host
// threads).  Knobs: QM_PLAN_RESTARTS (perturbed restarts; 0 plans the base
// shapes only) and QM_PLAN_ROUNDS (rounds of moves) override both budgets.
double compiled_plan_cost(const FusedPlan &p) {
  double t = 0;
  for (const FusedPass &x : p.passes) t += 1.8 + 0.26 * x.transposes;
  return t;
}
FusedPlan plan_compiled(const std::vector<DevGate> &gates, int n, const FusedShape &shape, const std::string &key,
                        CachedPlan *out = nullptr) {
  static const int env_restarts = [] {
    const char *e = std::getenv("QM_PLAN_RESTARTS");
    return e ? std::max(0, std::atoi(e)) : -1;
  }();
  static const int env_rounds = [] {
    const char *e = std::getenv("QM_PLAN_ROUNDS");
    return e ? std::max(0, std::atoi(e)) : -1;
  }();
  PlanSearch opt;
  opt.restarts = n >= 28 ? 1000 : 96;
  opt.rounds = n >= 28 ? 60 : 8;
  if (env_restarts >= 0) opt.restarts = env_restarts;
  if (env_rounds >= 0) opt.rounds = env_rounds;
  if (opt.restarts == 0) opt.rounds = 0;
  std::vector<int> lows{shape.contiguous}, caps{shape.max_phases};
  if (shape.alt_contiguous > 0 && shape.alt_contiguous < shape.K && shape.alt_contiguous != shape.contiguous)
    lows.push_back(shape.alt_contiguous);
  if (shape.max_phases != 9) caps.push_back(9);
  std::vector<FusedShape> bases;
  for (int low : lows)
    for (int cap : caps) {
      FusedShape b = shape;
      b.contiguous = low;
      b.max_phases = cap;
      bases.push_back(b);
    }
  if (bases.empty()) bases.push_back(shape);
  CachedPlan rec{shape, {}};
  FusedPlan plan = plan_fused_searched(gates, n, bases, opt, compiled_plan_cost, &rec.shape, &rec.trace);
  // The lane layout needs each next pass's load lanes: the chosen plan is
  // replayed from its trace, which holds them (same tiles, phases and gates).
  if (rec.shape.lane_layout) plan = plan_fused(gates, n, rec.shape, &rec.trace);
  if (out) *out = rec;
  plan_cache_put(key, std::move(rec));
  return plan;
}
}  // namespace

FusedPlan plan_compiled_trace(const std::vector<DevGate> &gates, int n, int precision, const FusedShape &shape,
                              FusedShape *chosen, PlanTrace *trace) {
  const std::string key = plan_key(gates, n, precision, shape, QM_SCHED_COMPILED);
  CachedPlan hit;
  if (!plan_cache_get(key, &hit)) (void)plan_compiled(gates, n, shape, key, &hit);
  *chosen = hit.shape;
  *trace = hit.trace;
  return plan_fused(gates, n, hit.shape, &hit.trace);
}

// Runs prepared gates on a resident state with the chosen schedule.
// Whether QM_SCHED_AUTO compiles a structure's kernels in the background on
// its first (interpreted) execution; tuning knob QM_AUTO_COMPILE=0 turns it off.
static bool auto_compile() {
  static const bool on = [] {
    const char *e = std::getenv("QM_AUTO_COMPILE");
    return !e || std::atoi(e) != 0;
  }();
  return on;
}

ExecStats run_gates(State *s, const std::vector<DevGate> &gates, int schedule_in,
                    const volatile int32_t *cancel, const FusedShape *force) {
  Ctx *c = s->ctx;
  ExecStats st;
  const double state_bytes = (double)s->bytes;
  // Tile shape of a schedule: forced, tuned for this width, or the context's
  // default adjusted for the schedule.
  auto shape_for = [&](int sched) {
    FusedShape shape = c->shape[s->precision];
    const int tuned_slot = sched == QM_SCHED_FUSED ? 0 : sched == QM_SCHED_COMPILED ? 1 : -1;
    if (force) {
      shape = *force;
    } else if (tuned_slot >= 0 && c->tuned[tuned_slot][s->precision][s->qubits].K) {
      shape = c->tuned[tuned_slot][s->precision][s->qubits];
    } else if (sched == QM_SCHED_COMPILED && s->qubits < shape.K && s->qubits >= 12) {
      const FusedShape wide = shape;
      // Narrower states: the widest tile of 256 threads that fits (8192 or 4096
      // amplitudes).
      const int k = std::min(s->qubits, 13);
      shape = FusedShape{k, 8, k - 8, wide.contiguous};
      shape.remap = wide.remap;
      shape.search = wide.search;
      shape.max_phases = wide.max_phases;
      shape.lane_layout = wide.lane_layout;
    }
    if (!force && sched == QM_SCHED_FUSED && !(tuned_slot >= 0 && c->tuned[0][s->precision][s->qubits].K)) {
      // The interpreting kernel keeps the lowest qubits in place (its passes are
      // bound by the kernel's own work, not by the access pattern); its tiles are
      // searched like the compiled schedule's.
      shape.contiguous = s->precision == QM_C128 ? 3 : 4;
      shape.remap = 0;
      shape.search = 1;
    }
    // The interpreting kernel runs every pass in place.
    if (sched == QM_SCHED_FUSED) shape.lane_layout = 0;
    if (!force) apply_tuning_knobs(shape);
    return shape;
  };
  // QM_SCHED_AUTO: the compiled kernels when this structure's are all in the
  // cache, else the interpreting kernel now and the compilation in the
  // background, so a structure run again (a parameter sweep, a VQE loop, a
  // served circuit family) runs compiled from then on.
  const bool automatic = schedule_in == QM_SCHED_AUTO;
  int schedule = automatic ? QM_SCHED_COMPILED : schedule_in;
  FusedShape shape = shape_for(schedule);
  if (schedule == QM_SCHED_COMPILED && s->qubits >= shape.K && gates.size() > 1) {
    const std::string key = plan_key(gates, s->qubits, s->precision, shape, QM_SCHED_COMPILED);
    CachedPlan hit;
    const bool known = plan_cache_get(key, &hit);
    FusedPlan plan;
    bool have_plan = false;
    if (known) {
      plan = plan_fused(gates, s->qubits, hit.shape, &hit.trace);
      have_plan = true;
    } else if (!automatic) {
      plan = plan_compiled(gates, s->qubits, shape, key);
      have_plan = true;
    } else if (auto_compile()) {
      // The search and the compilation run in the background; this call runs
      // interpreted, and a later call of the structure replays the searched plan.
      const int n = s->qubits;
      const FusedShape base = shape;
      jit_plan_async(key, [gates, n, base, key] { return plan_compiled(gates, n, base, key); }, s->precision,
                     c->device);
    }
    JitProgram jp;
    bool ready = false;
    if (have_plan && jit_supported(plan)) {
      if (!automatic) {
        jp = jit_prepare(plan, s->precision, c->device);
        ready = true;
      } else {
        ready = jit_lookup(plan, s->precision, c->device, &jp);
        if (!ready && auto_compile()) jit_prepare_async(plan, s->precision, c->device);
      }
    }
    bool out_place = false;
    for (const FusedPass &p : plan.passes) out_place = out_place || p.dev.out_place;
    // Passes that store under another lane layout than they load read one
    // buffer and write the other; without room for the second buffer this
    // call runs interpreted.
    if (ready && out_place && !state_scratch(s)) ready = false;
    if (ready) {
      const int64_t bytes = (int64_t)(jp.coef.size() * sizeof(double));
      double *d_coef = (double *)ctx_program(c, bytes + 256);
      if (bytes)
        QM_HIP(hipMemcpyAsync(d_coef, jp.coef.data(), (size_t)bytes, hipMemcpyHostToDevice, c->stream));
      // The buffer holding the newest state becomes the state's own, also
      // when a cancellation ends the run between passes.
      struct Current {
        State *s;
        void *cur;
        ~Current() {
          if (cur != s->data) std::swap(s->data, s->scratch);
        }
      } cur{s, s->data};
      for (size_t i = 0; i < plan.passes.size(); i++) {
        check_cancel(cancel);
        if (plan.passes[i].dev.out_place) {
          void *other = cur.cur == s->data ? s->scratch : s->data;
          jit_launch(s, shape, jp, i, d_coef, c->stream, cur.cur, other);
          cur.cur = other;
        } else {
          jit_launch(s, shape, jp, i, d_coef, c->stream, cur.cur, cur.cur);
        }
        st.launches++;
        st.passes++;
        st.bytes += 2 * state_bytes;
        if (cancel && (st.launches & 31) == 0) QM_HIP(hipStreamSynchronize(c->stream));
      }
      QM_HIP(hipStreamSynchronize(c->stream));
      return st;
    }
    schedule = QM_SCHED_FUSED;
    shape = shape_for(schedule);
  } else if (automatic) {
    schedule = QM_SCHED_FUSED;
    shape = shape_for(schedule);
  }
  if (schedule == QM_SCHED_FUSED && !fused_shape_supported(s->precision, shape.L, shape.R)) {
    const int search = shape.search;
    shape = default_fused_shape(s->precision);
    shape.search = search;
  }
  if (schedule == QM_SCHED_FUSED && s->qubits >= shape.K && gates.size() > 1) {
    FusedPlan plan = plan_cached(gates, s->qubits, s->precision, shape, QM_SCHED_FUSED);
    std::vector<TileOp> flat = plan.flat_ops();
    const int64_t pass_bytes = (int64_t)(plan.passes.size() * sizeof(DevPass));
    const int64_t op_bytes = (int64_t)(flat.size() * sizeof(TileOp));
    char *prog = (char *)ctx_program(c, pass_bytes + op_bytes + 256);
    std::vector<char> host((size_t)(pass_bytes + op_bytes));
    for (size_t i = 0; i < plan.passes.size(); i++)
      std::memcpy(host.data() + i * sizeof(DevPass), &plan.passes[i].dev, sizeof(DevPass));
    if (op_bytes) std::memcpy(host.data() + pass_bytes, flat.data(), (size_t)op_bytes);
    QM_HIP(hipMemcpyAsync(prog, host.data(), host.size(), hipMemcpyHostToDevice, c->stream));
    for (size_t i = 0; i < plan.passes.size(); i++) {
      check_cancel(cancel);
      launch_fused_pass(s, shape, prog + i * sizeof(DevPass), prog + pass_bytes, c->stream);
      st.launches++;
      st.passes++;
      st.bytes += 2 * state_bytes;
      if (cancel && (st.launches & 31) == 0) QM_HIP(hipStreamSynchronize(c->stream));
    }
    // The program buffer must outlive the launches: wait before returning.
    QM_HIP(hipStreamSynchronize(c->stream));
    return st;
  }
  for (const DevGate &g : gates) {
    check_cancel(cancel);
    launch_gate(s, g, c->stream);
    st.launches++;
    st.passes++;
    st.bytes += 2 * state_bytes * touched_fraction(g);
    // Bound queued work and cancellation latency (CudaStateVectorBackend.cs:155-159).
    if (cancel && (st.launches & 31) == 0) QM_HIP(hipStreamSynchronize(c->stream));
  }
  return st;
}

static State *as_state(qm_state *h) {
  if (h == nullptr) fail(QM_ERR_ARG, "state is null");
  State *s = reinterpret_cast<State *>(h);
  use_device(s->ctx);
  return s;
}

static double take(const double *u, int64_t n, int64_t *cursor);

// The caller's uniform draws (the reference's System.Random): a pool read at a
// cursor, or a callback asked for one value each time the reference calls
// Random.NextDouble(), so a caller's own generator stays in step draw for draw.
struct DrawSource {
  const double *u = nullptr;
  int64_t n = 0;
  int64_t *cursor = nullptr;
  qm_draw_fn fn = nullptr;
  void *user = nullptr;
  double next() {
    if (!fn) return take(u, n, cursor);
    double v = 0;
    if (fn(user, &v) != 0) fail(QM_ERR_POOL, "The draw source failed.");
    return v;
  }
};
static DrawSource pool_draws(const double *u, int64_t n, int64_t *cursor) {
  DrawSource d;
  d.u = u;
  d.n = n;
  d.cursor = cursor;
  return d;
}
static DrawSource callback_draws(qm_draw_fn fn, void *user) {
  if (!fn) fail(QM_ERR_ARG, "draw callback is null");
  DrawSource d;
  d.fn = fn;
  d.user = user;
  return d;
}

static void noise_apply(State *s, int kind, double p, int target, DrawSource &draws);
static void noisy_execute(State *s, const qm_step *steps, int64_t n_steps, const qm_op *ops,
                          DrawSource &draws);
static void dynamic_execute(State *s, int32_t classical_bits, const qm_dyn_step *steps, int64_t n_steps,
                            const qm_op *ops, int64_t n_ops, const int32_t *sel_qubits,
                            const int32_t *sel_cbits, int64_t n_sel, int32_t schedule, DrawSource &draws,
                            int32_t *classical, const int32_t *cancel);

}  // namespace qm

using namespace qm;

extern "C" {

const char *qm_last_error(void) { return g_error.c_str(); }
int32_t qm_abi_version(void) { return QM_ABI_VERSION; }

int qm_device_count(int32_t *count) {
  return guarded([&] {
    if (!count) fail(QM_ERR_ARG, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
  });
}

int qm_device_get_info(int32_t index, qm_device_info *out) {
  return guarded([&] {
    if (!out) fail(QM_ERR_ARG, "info is null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      fail(QM_ERR_NODEVICE, "No compatible AMD HIP device/driver was found.");
    if (index < 0 || index >= n) fail(QM_ERR_RANGE, "HIP device index is not available.");
    hipDeviceProp_t p;
    QM_HIP(hipGetDeviceProperties(&p, index));
    std::memset(out, 0, sizeof(*out));
    out->index = index;
    out->compute_units = p.multiProcessorCount;
    out->memory_bytes = (int64_t)p.totalGlobalMem;
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    int drv = 0;
    (void)hipDriverGetVersion(&drv);
    std::snprintf(out->driver, sizeof(out->driver), "%d", drv);
  });
}

int qm_ctx_create(const qm_ctx_options *opt, qm_ctx **out) {
  return guarded([&] {
    if (!out) fail(QM_ERR_ARG, "out is null");
    *out = nullptr;
    qm_ctx_options o{};
    if (opt) o = *opt;
    // CudaBackendOptions.Validate (CudaBackendOptions.cs:14-19); 0 = no budget.
    if (o.device_index < 0) fail(QM_ERR_RANGE, "DeviceIndex must be nonnegative.");
    if (o.device_budget_bytes != 0 && o.device_budget_bytes < 32)
      fail(QM_ERR_RANGE, "DeviceMemoryBudgetBytes must be at least 32.");
    if (o.staging_budget_bytes != 0 && o.staging_budget_bytes < 32)
      fail(QM_ERR_RANGE, "StagingMemoryBudgetBytes must be at least 32.");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      fail(QM_ERR_NODEVICE,
           "No compatible AMD HIP device/driver was found. Select a CPU backend explicitly.");
    if (o.device_index >= n) fail(QM_ERR_RANGE, "HIP device index is not available.");
    Ctx *c = new Ctx();
    try {
      c->device = o.device_index;
      c->profiling = o.enable_profiling != 0;
      c->device_budget = o.device_budget_bytes;
      c->staging_budget = o.staging_budget_bytes;
      QM_HIP(hipSetDevice(c->device));
      hipDeviceProp_t p;
      QM_HIP(hipGetDeviceProperties(&p, c->device));
      c->device_memory = (int64_t)p.totalGlobalMem;
      c->compute_units = p.multiProcessorCount;
      QM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      QM_HIP(hipEventCreate(&c->ev_start));
      QM_HIP(hipEventCreate(&c->ev_stop));
      noalloc_setup(c->device);  // the scratch array's address on this device (once per device)
    } catch (...) {
      delete c;
      throw;
    }
    *out = reinterpret_cast<qm_ctx *>(c);
  });
}

int qm_ctx_destroy(qm_ctx *h) {
  return guarded([&] {
    if (!h) return;
    Ctx *c = reinterpret_cast<Ctx *>(h);
    std::lock_guard<std::recursive_mutex> g(c->lock);
    if (c->disposed) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->d_partial) (void)hipFree(c->d_partial);
    if (c->d_ipartial) (void)hipFree(c->d_ipartial);
    if (c->d_prog) (void)hipFree(c->d_prog);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->disposed = true;
    // The Ctx object itself stays allocated so late calls observe
    // QM_ERR_DISPOSED (CudaStateVectorBackend.cs:212-214).
  });
}

int qm_ctx_resident_bytes(qm_ctx *h, int64_t *bytes) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    use_device(c);
    if (!bytes) fail(QM_ERR_ARG, "bytes is null");
    std::lock_guard<std::recursive_mutex> g(c->lock);
    *bytes = c->resident_bytes;
  });
}

}  // extern "C"

namespace qm {

State *state_new(Ctx *c, int qubits, int precision) {
  if (qubits < 1 || qubits > 34) fail(QM_ERR_RANGE, "qubits must be between 1 and 34.");
  if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
  const int64_t bytes = amp_bytes(precision) << qubits;
  if ((c->device_budget && c->resident_bytes + bytes > c->device_budget) ||
      c->resident_bytes + bytes > c->device_memory)
    fail(QM_ERR_BUDGET, "HIP state exceeds the device memory budget.");
  State *s = new State();
  s->ctx = c;
  s->qubits = qubits;
  s->precision = precision;
  s->length = (int64_t)1 << qubits;
  s->bytes = bytes;
  hipError_t e = hipMalloc(&s->data, (size_t)bytes);
  if (e != hipSuccess) {
    delete s;
    (void)hipGetLastError();
    fail(QM_ERR_BUDGET, "HIP state exceeds the free device memory.");
  }
  c->resident_bytes += bytes;
  return s;
}

void state_free(State *s) {
  Ctx *c = s->ctx;
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(s->data);
  c->resident_bytes -= s->bytes;
  if (s->scratch) {
    (void)hipFree(s->scratch);
    c->resident_bytes -= s->bytes;
  }
  delete s;
}

bool state_scratch(State *s) {
  if (s->scratch) return true;
  Ctx *c = s->ctx;
  if ((c->device_budget && c->resident_bytes + s->bytes > c->device_budget) ||
      c->resident_bytes + s->bytes > c->device_memory)
    return false;
  if (hipMalloc(&s->scratch, (size_t)s->bytes) != hipSuccess) {
    (void)hipGetLastError();
    s->scratch = nullptr;
    return false;
  }
  c->resident_bytes += s->bytes;
  return true;
}

}  // namespace qm

extern "C" {

int qm_execute(qm_ctx *h, int32_t qubits, const qm_op *ops, int64_t n_ops, double *amps,
               int64_t n_amps, const volatile int32_t *cancel, qm_profile *profile) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    if (!c) fail(QM_ERR_ARG, "context is null");
    if (profile && !c->profiling)
      fail(QM_ERR_STATE, "EnableProfiling must be selected when creating the HIP backend.");
    if (qubits < 1 || qubits > 34) fail(QM_ERR_RANGE, "qubits must be between 1 and 34.");
    if (!amps || n_amps != ((int64_t)1 << qubits))
      fail(QM_ERR_ARG, "Buffer length must equal 2^qubitCount.");
    check_cancel(cancel);
    std::lock_guard<std::recursive_mutex> g(c->lock);
    use_device(c);
    check_cancel(cancel);
    const int64_t bytes = 16 * n_amps;
    if ((c->device_budget && bytes > c->device_budget) || bytes > c->device_memory ||
        (c->staging_budget && bytes > c->staging_budget))
      fail(QM_ERR_BUDGET, "HIP state exceeds the device or host staging memory budget.");
    std::vector<DevGate> gates = prepare_all(ops, n_ops, qubits);
    State *s = state_new(c, qubits, QM_C128);
    try {
      copy_to_device(c, s->data, amps, (size_t)bytes);
      if (profile) QM_HIP(hipEventRecord(c->ev_start, c->stream));
      ExecStats es = run_gates(s, gates, QM_SCHED_AUTO, cancel);
      if (profile) QM_HIP(hipEventRecord(c->ev_stop, c->stream));
      QM_HIP(hipStreamSynchronize(c->stream));
      check_cancel(cancel);
      // Commit only after the device work and the cancellation check (:166-176).
      copy_to_host(c, amps, s->data, (size_t)bytes);
      if (profile) {
        float ms = 0;
        QM_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
        profile->device_ms = ms;
        profile->state_bytes = bytes;
        profile->kernel_launches = es.launches;
        profile->passes = es.passes;
        profile->algorithmic_bytes = es.bytes;
      }
    } catch (...) {
      state_free(s);
      throw;
    }
    state_free(s);
  });
}

int qm_state_create(qm_ctx *h, int32_t qubits, int32_t precision, qm_state **out) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    use_device(c);
    if (!out) fail(QM_ERR_ARG, "out is null");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> g(c->lock);
    State *s = state_new(c, qubits, precision);
    try {
      launch_set_basis(s, 0, c->stream);
      QM_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      state_free(s);
      throw;
    }
    *out = reinterpret_cast<qm_state *>(s);
  });
}

int qm_state_destroy(qm_state *h) {
  return guarded([&] {
    if (!h) return;
    State *s = reinterpret_cast<State *>(h);
    Ctx *c = s->ctx;
    std::lock_guard<std::recursive_mutex> g(c->lock);
    if (!c->disposed) (void)hipSetDevice(c->device);
    state_free(s);
  });
}

int qm_state_info(qm_state *h, int32_t *qubits, int32_t *precision) {
  return guarded([&] {
    State *s = as_state(h);
    if (qubits) *qubits = s->qubits;
    if (precision) *precision = s->precision;
  });
}

int qm_state_reset(qm_state *h) {
  return guarded([&] {
    State *s = as_state(h);
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    launch_set_basis(s, 0, s->ctx->stream);
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
  });
}

int qm_state_upload(qm_state *h, const double *amps, int64_t n) {
  return guarded([&] {
    State *s = as_state(h);
    if (!amps || n != s->length) fail(QM_ERR_ARG, "Buffer length must equal 2^qubitCount.");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    Ctx *c = s->ctx;
    if (s->precision == QM_C128) {
      copy_to_device(c, s->data, amps, (size_t)s->bytes);
    } else {
      // Round through a temporary complex128 state on the device.
      State *tmp = state_new(c, s->qubits, QM_C128);
      try {
        copy_to_device(c, tmp->data, amps, (size_t)tmp->bytes);
        launch_convert(tmp, s, c->stream);
        QM_HIP(hipStreamSynchronize(c->stream));
      } catch (...) {
        state_free(tmp);
        throw;
      }
      state_free(tmp);
    }
    QM_HIP(hipStreamSynchronize(c->stream));
  });
}

int qm_state_download(qm_state *h, double *amps, int64_t n) {
  return guarded([&] {
    State *s = as_state(h);
    if (!amps || n != s->length) fail(QM_ERR_ARG, "Buffer length must equal 2^qubitCount.");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    Ctx *c = s->ctx;
    if (s->precision == QM_C128) {
      copy_to_host(c, amps, s->data, (size_t)s->bytes);
    } else {
      State *tmp = state_new(c, s->qubits, QM_C128);
      try {
        launch_convert(s, tmp, c->stream);
        copy_to_host(c, amps, tmp->data, (size_t)tmp->bytes);
        QM_HIP(hipStreamSynchronize(c->stream));
      } catch (...) {
        state_free(tmp);
        throw;
      }
      state_free(tmp);
    }
    QM_HIP(hipStreamSynchronize(c->stream));
  });
}

int qm_state_copy(qm_state *dst, qm_state *src) {
  return guarded([&] {
    State *d = as_state(dst);
    State *s = as_state(src);
    if (d->ctx != s->ctx) fail(QM_ERR_ARG, "states belong to different contexts");
    if (d->qubits != s->qubits) fail(QM_ERR_ARG, "state widths differ");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    launch_convert(s, d, s->ctx->stream);
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
  });
}

int qm_state_execute(qm_state *h, const qm_op *ops, int64_t n_ops, int32_t schedule,
                     const volatile int32_t *cancel, qm_profile *profile) {
  return guarded([&] {
    State *s = as_state(h);
    Ctx *c = s->ctx;
    if (profile && !c->profiling)
      fail(QM_ERR_STATE, "EnableProfiling must be selected when creating the HIP backend.");
    if (schedule != QM_SCHED_PER_GATE && schedule != QM_SCHED_FUSED && schedule != QM_SCHED_COMPILED &&
        schedule != QM_SCHED_AUTO)
      fail(QM_ERR_ARG, "unknown schedule");
    check_cancel(cancel);
    std::lock_guard<std::recursive_mutex> g(c->lock);
    std::vector<DevGate> gates = prepare_all(ops, n_ops, s->qubits);
    if (profile) QM_HIP(hipEventRecord(c->ev_start, c->stream));
    ExecStats es = run_gates(s, gates, schedule, cancel);
    if (profile) QM_HIP(hipEventRecord(c->ev_stop, c->stream));
    QM_HIP(hipStreamSynchronize(c->stream));
    if (profile) {
      float ms = 0;
      QM_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
      profile->device_ms = ms;
      profile->state_bytes = s->bytes;
      profile->kernel_launches = es.launches;
      profile->passes = es.passes;
      profile->algorithmic_bytes = es.bytes;
    }
    check_cancel(cancel);
  });
}

int qm_state_mass(qm_state *h, double *mass) {
  return guarded([&] {
    State *s = as_state(h);
    if (!mass) fail(QM_ERR_ARG, "mass is null");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    *mass = reduce_mass(s, s->ctx->stream).total;
  });
}

static void validate_mass(double m) {
  if (!std::isfinite(m) || !(m > 0))
    fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
}

int qm_state_measure_all(qm_state *h, double u, int64_t *outcome) {
  return guarded([&] {
    State *s = as_state(h);
    if (!outcome) fail(QM_ERR_ARG, "outcome is null");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    int64_t out = -1;
    s->boundary_fallbacks += cdf_search(s, &u, 1, &out, s->ctx->stream);
    // Collapse to |outcome> with amplitude +1 (StateVector.cs:150-151).
    launch_set_basis(s, out, s->ctx->stream);
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
    *outcome = out;
  });
}

static uint64_t selection_mask(const State *s, const int32_t *sel, int k) {
  // ValidateSelection (StateVector.Observations.cs:143-168).
  if (!sel || k < 1 || k > s->qubits) fail(QM_ERR_ARG, "Select between one and QubitCount distinct qubits.");
  uint64_t m = 0;
  for (int i = 0; i < k; i++) {
    if (sel[i] < 0 || sel[i] >= s->qubits) fail(QM_ERR_RANGE, "Selected qubits must be in range.");
    uint64_t b = 1ull << sel[i];
    if (m & b) fail(QM_ERR_ARG, "Selected qubits must be distinct.");
    m |= b;
  }
  return m;
}

// StateVector.Measure (StateVector.Observations.cs:45-102): one full-basis
// draw with the sequential-CDF rule, the matching branch kept and divided by
// the square root of its mass.  Returns the packed outcome (bit b = sel[b]).
static int64_t measure_subset(State *s, const int32_t *sel, int k, double u) {
  const uint64_t mask = selection_mask(s, sel, k);
  int64_t basis = -1;
  s->boundary_fallbacks += cdf_search(s, &u, 1, &basis, s->ctx->stream);
  const uint64_t pattern = (uint64_t)basis & mask;
  const double branch = reduce_masked_mass(s, mask, pattern, s->ctx->stream);
  validate_mass(branch);
  launch_collapse_subset(s, mask, pattern, std::sqrt(branch), s->ctx->stream);
  int64_t out = 0;
  for (int b = 0; b < k; b++) out |= (((uint64_t)basis >> sel[b]) & 1) << b;
  return out;
}

int qm_state_measure(qm_state *h, const int32_t *sel, int32_t k, double u, int64_t *outcome) {
  return guarded([&] {
    State *s = as_state(h);
    if (!outcome) fail(QM_ERR_ARG, "outcome is null");
    selection_mask(s, sel, k);
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    const int64_t out = measure_subset(s, sel, k, u);
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
    *outcome = out;
  });
}

int qm_state_marginal(qm_state *h, const int32_t *sel, int32_t k, int64_t outcome, double *p) {
  return guarded([&] {
    State *s = as_state(h);
    if (!p) fail(QM_ERR_ARG, "probability is null");
    const uint64_t mask = selection_mask(s, sel, k);
    if (outcome < 0 || outcome >= ((int64_t)1 << k)) fail(QM_ERR_RANGE, "outcome");
    uint64_t pattern = 0;
    for (int b = 0; b < k; b++) pattern |= (((uint64_t)outcome >> b) & 1) << sel[b];
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    const double total = reduce_mass(s, s->ctx->stream).total;
    const double selected = reduce_masked_mass(s, mask, pattern, s->ctx->stream);
    validate_mass(total);
    *p = selected / total;
  });
}

int qm_state_sample(qm_state *h, const double *u, int64_t shots, int64_t *outcomes) {
  return guarded([&] {
    State *s = as_state(h);
    if (shots < 1) fail(QM_ERR_RANGE, "shots must be positive");
    if (!u || !outcomes) fail(QM_ERR_ARG, "buffers are null");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    s->boundary_fallbacks += cdf_search(s, u, shots, outcomes, s->ctx->stream);
  });
}

int qm_state_expectations(qm_state *h, const qm_pauli *terms, int64_t n, double *values) {
  return guarded([&] {
    State *s = as_state(h);
    if (n < 0 || (n > 0 && (!terms || !values))) fail(QM_ERR_ARG, "buffers are null");
    const uint64_t all = s->qubits >= 64 ? ~0ull : ((1ull << s->qubits) - 1);
    for (int64_t i = 0; i < n; i++)
      if ((terms[i].flip_mask | terms[i].sign_mask) & ~all)
        fail(QM_ERR_ARG, "Observable and state sizes must match.");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    const double total = reduce_mass(s, s->ctx->stream).total;
    validate_mass(total);
    std::vector<double> num((size_t)n);
    reduce_pauli(s, terms, n, num.data(), s->ctx->stream);
    for (int64_t i = 0; i < n; i++) {
      if (!std::isfinite(num[i])) fail(QM_ERR_NONFINITE, "The observable accumulation is nonfinite.");
      values[i] = num[i] / total;
    }
  });
}

// Validation of qm_state_energy's arguments (StateVector.Energy.cs:10-25).
static void check_energy_args(const State *s, const qm_pauli *terms, const double *coeffs, int64_t n,
                              const double *energy) {
  if (!energy || n < 0 || (n > 0 && (!terms || !coeffs))) fail(QM_ERR_ARG, "buffers are null");
  const uint64_t all = s->qubits >= 64 ? ~0ull : ((1ull << s->qubits) - 1);
  for (int64_t i = 0; i < n; i++) {
    if ((terms[i].flip_mask | terms[i].sign_mask) & ~all)
      fail(QM_ERR_ARG, "Hamiltonian and state sizes must match.");
    if (!std::isfinite(coeffs[i])) fail(QM_ERR_RANGE, "coefficient");
  }
}

// Neumaier summation in term order (StateVector.Energy.cs:26-45) of
// coeffs[idx[k]] * num[k] / total over the terms with nonzero coefficients.
static double neumaier_energy(const double *num, const int64_t *idx, int64_t n_used, const double *coeffs,
                              double total) {
  double sum = 0, corr = 0;
  for (int64_t k = 0; k < n_used; k++) {
    if (!std::isfinite(num[k])) fail(QM_ERR_NONFINITE, "The observable accumulation is nonfinite.");
    const double value = coeffs[idx[k]] * (num[k] / total);
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

int qm_state_energy(qm_state *h, const qm_pauli *terms, const double *coeffs, int64_t n,
                    double *energy) {
  return guarded([&] {
    State *s = as_state(h);
    check_energy_args(s, terms, coeffs, n, energy);
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    const double total = reduce_mass(s, s->ctx->stream).total;
    validate_mass(total);
    std::vector<qm_pauli> used;
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < n; i++)
      if (coeffs[i] != 0) {
        used.push_back(terms[i]);
        idx.push_back(i);
      }
    std::vector<double> num(used.size());
    reduce_pauli(s, used.data(), (int64_t)used.size(), num.data(), s->ctx->stream);
    *energy = neumaier_energy(num.data(), idx.data(), (int64_t)used.size(), coeffs, total);
  });
}

// The same energy with every device buffer in the library's static scratch
// array and every host table on the stack: no allocation on any path that
// succeeds (the terms with nonzero coefficients are limited to
// QM_NOALLOC_MAX_TERMS).
int qm_state_energy_noalloc(qm_state *h, const qm_pauli *terms, const double *coeffs, int64_t n,
                            double *energy) {
  return guarded([&] {
    State *s = as_state(h);
    check_energy_args(s, terms, coeffs, n, energy);
    qm_pauli used[QM_NOALLOC_MAX_TERMS];
    int64_t idx[QM_NOALLOC_MAX_TERMS];
    int64_t n_used = 0;
    for (int64_t i = 0; i < n; i++)
      if (coeffs[i] != 0) {
        if (n_used == QM_NOALLOC_MAX_TERMS)
          fail(QM_ERR_RANGE, "more Hamiltonian terms than the scratch buffer holds (QM_NOALLOC_MAX_TERMS)");
        used[n_used] = terms[i];
        idx[n_used++] = i;
      }
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    double total = 0, num[QM_NOALLOC_MAX_TERMS];
    reduce_energy_noalloc(s, used, n_used, &total, num, s->ctx->stream);
    validate_mass(total);
    *energy = neumaier_energy(num, idx, n_used, coeffs, total);
  });
}

int qm_state_inner(qm_state *a, qm_state *b, double *re, double *im) {
  return guarded([&] {
    State *x = as_state(a);
    State *y = as_state(b);
    if (x->ctx != y->ctx || x->qubits != y->qubits) fail(QM_ERR_ARG, "states differ in width or context");
    if (!re || !im) fail(QM_ERR_ARG, "outputs are null");
    std::lock_guard<std::recursive_mutex> g(x->ctx->lock);
    reduce_inner(x, y, re, im, x->ctx->stream);
  });
}

int qm_state_execute_dense(qm_state *h, const qm_op *ops, int64_t n_ops, int32_t max_block_qubits,
                           qm_profile *profile) {
  return guarded([&] {
    State *s = as_state(h);
    Ctx *c = s->ctx;
    if (profile && !c->profiling)
      fail(QM_ERR_STATE, "EnableProfiling must be selected when creating the HIP backend.");
    if (max_block_qubits < 1 || max_block_qubits > 5) fail(QM_ERR_RANGE, "block width must be 1..5");
    std::lock_guard<std::recursive_mutex> g(c->lock);
    std::vector<DevGate> gates = prepare_all(ops, n_ops, s->qubits);
    std::vector<DenseBlock> blocks = plan_dense(gates, s->qubits, std::min<int>(max_block_qubits, s->qubits));
    // Block matrices in the state's precision: FP64 products for complex128,
    // FP32 (the block multiplied in FP64 on the host, then rounded) for complex64.
    const bool c64 = s->precision == QM_C64;
    const size_t elem = c64 ? sizeof(float) : sizeof(double);
    std::vector<char> host;
    std::vector<size_t> at;
    for (const DenseBlock &b : blocks) {
      at.push_back(host.size());
      const size_t off = host.size();
      host.resize(off + b.a.size() * elem);
      if (c64) {
        float *f = reinterpret_cast<float *>(host.data() + off);
        for (size_t i = 0; i < b.a.size(); i++) f[i] = (float)b.a[i];
      } else {
        std::memcpy(host.data() + off, b.a.data(), b.a.size() * sizeof(double));
      }
    }
    const int64_t bytes = (int64_t)host.size();
    char *d = (char *)ctx_program(c, bytes + 256);
    if (bytes) QM_HIP(hipMemcpyAsync(d, host.data(), (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    if (profile) QM_HIP(hipEventRecord(c->ev_start, c->stream));
    double flops = 0;
    for (size_t i = 0; i < blocks.size(); i++) {
      launch_dense(s, blocks[i], d + at[i], c->stream);
      const double m = 2.0 * (double)(1 << blocks[i].k);
      flops += 2.0 * m * m * (double)(s->length >> blocks[i].k);
    }
    if (profile) QM_HIP(hipEventRecord(c->ev_stop, c->stream));
    QM_HIP(hipStreamSynchronize(c->stream));
    if (profile) {
      float ms = 0;
      QM_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
      profile->device_ms = ms;
      profile->state_bytes = s->bytes;
      profile->kernel_launches = (int32_t)blocks.size();
      profile->passes = (int32_t)blocks.size();
      // Algorithmic bytes of the executed schedule; the matrix-core work is
      // reported by qm_dense_flops for the same call.
      profile->algorithmic_bytes = 2.0 * (double)s->bytes * (double)blocks.size();
    }
    c->last_dense_flops = flops;
  });
}

int qm_dense_flops(qm_ctx *h, double *flops) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    if (!c || !flops) fail(QM_ERR_ARG, "null argument");
    *flops = c->last_dense_flops;
  });
}

int qm_plan_dense_apply_host(int32_t qubits, const qm_op *ops, int64_t n_ops, int32_t max_block_qubits,
                             double *amplitudes, int32_t *blocks_out) {
  return guarded([&] {
    if (qubits < 1 || qubits > 26) fail(QM_ERR_RANGE, "qubits must be between 1 and 26 for the host reference");
    if (!amplitudes) fail(QM_ERR_ARG, "amplitudes is null");
    if (max_block_qubits < 1 || max_block_qubits > 5) fail(QM_ERR_RANGE, "block width must be 1..5");
    std::vector<DevGate> gates = prepare_all(ops, n_ops, qubits);
    std::vector<DenseBlock> blocks = plan_dense(gates, qubits, std::min<int>(max_block_qubits, qubits));
    for (const DenseBlock &b : blocks) apply_dense_host(b, amplitudes, qubits);
    if (blocks_out) *blocks_out = (int32_t)blocks.size();
  });
}

int qm_jit_compile_check(int32_t qubits, int32_t precision, const qm_op *ops, int64_t n_ops,
                         int32_t tile_bits, int32_t lane_bits, int32_t *passes) {
  return guarded([&] {
    if (!passes) fail(QM_ERR_ARG, "passes is null");
    if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    const int R = tile_bits - lane_bits;
    const int max_k = 14;
    if (!(lane_bits >= 6 && lane_bits <= 10 && R >= 2 && R <= 5 && tile_bits <= max_k))
      fail(QM_ERR_UNSUPPORTED, "the compiled schedule does not generate this tile shape");
    if (qubits < tile_bits || qubits > 34) fail(QM_ERR_RANGE, "qubits");
    std::vector<DevGate> gates = prepare_all(ops, n_ops, qubits);
    FusedShape shape = compiled_shape(tile_bits, lane_bits);
    // Debugging aid QM_JIT_CHECK_SEARCHED=1: the plan the compiled schedule
    // runs (the context default shape when tile_bits and lane_bits are the
    // default's, searched planning), instead of the greedy plan of the shape.
    static const bool searched = [] {
      const char *e = std::getenv("QM_JIT_CHECK_SEARCHED");
      return e && std::atoi(e) != 0;
    }();
    if (searched) {
      const Ctx defaults;
      const FusedShape &d = defaults.shape[precision];
      if (d.K == tile_bits && d.L == lane_bits) shape = d;
    }
    apply_tuning_knobs(shape);
    FusedPlan plan = searched ? plan_compiled(gates, qubits, shape,
                                              plan_key(gates, qubits, precision, shape, QM_SCHED_COMPILED))
                              : plan_fused(gates, qubits, shape);
    *passes = jit_compile_check(plan, precision);
  });
}

int qm_jit_stats(int64_t *compiled, int64_t *cache_hits, double *compile_seconds) {
  return guarded([&] { jit_stats(compiled, cache_hits, compile_seconds); });
}

int qm_jit_drain(void) {
  return guarded([&] { jit_drain(); });
}

int qm_state_boundary_fallbacks(qm_state *h, int64_t *count) {
  return guarded([&] {
    State *s = as_state(h);
    if (!count) fail(QM_ERR_ARG, "count is null");
    *count = s->boundary_fallbacks;
  });
}

int qm_ctx_set_fused_shape(qm_ctx *h, int32_t precision, int32_t tile_bits, int32_t lane_bits) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    use_device(c);
    if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    const int R = tile_bits - lane_bits;
    // Any shape the compiled schedule can generate (its LDS exchange buffer holds
    // 2^K reals: at most 128 KiB of the CU's 160 KiB); the interpreting kernel
    // uses this shape when it is one of its instantiations and its default shape
    // otherwise.
    const int max_k = 14;
    if (!fused_shape_supported(precision, lane_bits, R) &&
        !(lane_bits >= 6 && lane_bits <= 10 && R >= 2 && R <= 5 && tile_bits <= max_k))
      fail(QM_ERR_UNSUPPORTED, "fused tile shape is not supported");
    std::lock_guard<std::recursive_mutex> g(c->lock);
    c->shape[precision] = compiled_shape(tile_bits, lane_bits);
    for (auto &per_schedule : c->tuned)  // an explicit shape replaces tuned ones
      for (FusedShape &t : per_schedule[precision]) t = FusedShape{};
  });
}

int qm_ctx_get_fused_shape(qm_ctx *h, int32_t precision, int32_t *tile_bits, int32_t *lane_bits) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    use_device(c);
    if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    if (tile_bits) *tile_bits = c->shape[precision].K;
    if (lane_bits) *lane_bits = c->shape[precision].L;
  });
}

int qm_autotune(qm_ctx *h, int32_t qubits, int32_t precision, int32_t schedule, const qm_op *ops,
                int64_t n_ops, qm_tune_entry *entries, int32_t capacity, int32_t *count,
                int32_t *best) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    use_device(c);
    if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    if (schedule != QM_SCHED_FUSED && schedule != QM_SCHED_COMPILED)
      fail(QM_ERR_ARG, "only the fused and compiled schedules have tile shapes");
    if (qubits < 1 || qubits > 34) fail(QM_ERR_RANGE, "qubits must be between 1 and 34.");
    if (capacity < 0 || (capacity > 0 && !entries)) fail(QM_ERR_ARG, "entries are null");
    std::lock_guard<std::recursive_mutex> g(c->lock);
    std::vector<DevGate> gates = prepare_all(ops, n_ops, qubits);
    // A prefix of a few layers is representative and keeps tuning short.
    if (gates.size() > (size_t)(32 * qubits)) gates.resize((size_t)(32 * qubits));
    if (gates.size() < 2) fail(QM_ERR_ARG, "the circuit has fewer than two gates to tune on");
    const int contiguous = precision == QM_C128 ? 3 : 4;
    std::vector<FusedShape> cands;
    if (schedule == QM_SCHED_FUSED) {
      for (const FusedShape &f : fused_shapes(precision))
        if (f.K <= qubits) cands.push_back(f);
    } else {
      // Generated kernels: tiles of 2^11..2^14 (c128) or 2^12..2^14 (c64)
      // amplitudes with 128..1024 threads, in the default code variant.
      const int kl[][2] = {{11, 7}, {12, 7}, {12, 8}, {12, 9}, {13, 8}, {13, 9}, {14, 9}, {14, 10}};
      for (const auto &p : kl) {
        const int K = p[0], L = p[1], R = K - L;
        if (K > qubits || R < 2 || R > 5 || K < 12 - (precision == QM_C128))
          continue;
        cands.push_back(compiled_shape(K, L));
        if ((K == 14 && L == 9) || (K == 13 && L == 8))  // the defaults with 3 and 5 low bits kept
          for (int low : {3, 5}) cands.push_back(compiled_shape(K, L, low));
        if (K == 13 && L == 8) cands.push_back(FusedShape{K, L, R, contiguous});  // without remapping
      }
    }
    if (cands.empty()) fail(QM_ERR_UNSUPPORTED, "no tile shape fits this width");
    State *s = state_new(c, qubits, precision);
    struct Free {
      State *s;
      ~Free() { state_free(s); }
    } guard{s};
    std::vector<qm_tune_entry> out;
    int32_t winner = 0;
    for (size_t i = 0; i < cands.size(); i++) {
      const FusedShape &f = cands[i];
      launch_set_basis(s, 0, c->stream);
      ExecStats st = run_gates(s, gates, schedule, nullptr, &f);  // compiles, warms caches
      double ms = 0;
      for (int rep = 0; rep < 2; rep++) {
        launch_set_basis(s, 0, c->stream);
        QM_HIP(hipStreamSynchronize(c->stream));
        const auto t0 = std::chrono::steady_clock::now();
        run_gates(s, gates, schedule, nullptr, &f);
        const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ms = rep == 0 ? t : std::min(ms, t);
      }
      qm_tune_entry e{};
      e.tile_bits = f.K;
      e.lane_bits = f.L;
      e.variant = f.variant;
      e.contiguous = f.contiguous;
      e.remap = f.remap;
      e.passes = st.passes;
      e.ms = ms;
      e.ms_per_gate = ms / (double)gates.size();
      out.push_back(e);
      if (ms < out[(size_t)winner].ms) winner = (int32_t)i;
    }
    c->tuned[schedule == QM_SCHED_FUSED ? 0 : 1][precision][qubits] = cands[(size_t)winner];
    for (int32_t i = 0; i < std::min<int32_t>(capacity, (int32_t)out.size()); i++) entries[i] = out[(size_t)i];
    if (count) *count = (int32_t)out.size();
    if (best) *best = winner;
  });
}

int qm_plan_fused(int32_t qubits, int32_t precision, const qm_op *ops, int64_t n_ops,
                  int32_t *passes, int32_t *transposes, int64_t *gates) {
  return guarded([&] {
    if (qubits < 1 || qubits > 34) fail(QM_ERR_RANGE, "qubits must be between 1 and 34.");
    if (precision != QM_C128 && precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    std::vector<DevGate> g = prepare_all(ops, n_ops, qubits);
    const FusedShape shape = default_fused_shape(precision);
    int32_t np = 0, nt = 0;
    if (qubits >= shape.K && g.size() > 1) {
      FusedPlan plan = plan_fused(g, qubits, shape);
      np = (int32_t)plan.passes.size();
      for (const FusedPass &p : plan.passes) nt += p.transposes;
    } else {
      np = (int32_t)g.size();
    }
    if (passes) *passes = np;
    if (transposes) *transposes = nt;
    if (gates) *gates = (int64_t)g.size();
  });
}

int qm_plan_apply_host(int32_t qubits, const qm_op *ops, int64_t n_ops, int32_t tile_bits,
                       int32_t lane_bits, int32_t contiguous, int32_t flags, double *amps) {
  return guarded([&] {
    if (qubits < 1 || qubits > 26) fail(QM_ERR_RANGE, "host reference is limited to 26 qubits");
    if (!amps) fail(QM_ERR_ARG, "amplitudes are null");
    std::vector<DevGate> g = prepare_all(ops, n_ops, qubits);
    FusedShape shape{tile_bits, lane_bits, tile_bits - lane_bits, contiguous};
    shape.remap = flags & 3;
    shape.search = (flags >> 2) & 1;
    shape.lane_layout = (flags >> 5) & 1;
    if (flags < 0 || flags > 63 || shape.remap == 3) fail(QM_ERR_RANGE, "unsupported planning flags");
    if (shape.R < 1 || shape.R > 5 || shape.L < 1 || shape.L > 10 || shape.K > kMaxK ||
        contiguous < 0 || contiguous >= shape.K)
      fail(QM_ERR_RANGE, "unsupported tile shape");
    if (qubits < shape.K) {
      for (const DevGate &x : g) apply_gate_host(x, amps, qubits);
      return;
    }
    if (flags & 16) {
      // Searched plan (plan_fused_searched, a small budget): the plan replayed
      // from its trace must equal the searched plan, and is the one applied.
      PlanSearch opt;
      opt.restarts = 12;
      opt.rounds = 2;
      opt.moves = 6;
      FusedShape capped = shape;
      capped.max_phases = 9;
      FusedShape chosen;
      PlanTrace trace;
      FusedPlan searched = plan_fused_searched(g, qubits, {shape, capped}, opt, compiled_plan_cost, &chosen, &trace);
      FusedPlan replayed = plan_fused(g, qubits, chosen, &trace);
      bool same = replayed.passes.size() == searched.passes.size();
      for (size_t i = 0; same && i < searched.passes.size(); i++) {
        const FusedPass &a = replayed.passes[i], &b = searched.passes[i];
        if (chosen.lane_layout) {
          // The replay adds the lane layout: the same tiles, phases and gates.
          same = a.rsets == b.rsets && a.gates == b.gates && a.ops.size() == b.ops.size() &&
                 a.dev.qmask == b.dev.qmask && a.dev.out_place == 1;
        } else {
          same = std::memcmp(&a.dev, &b.dev, sizeof(DevPass)) == 0 && a.ops.size() == b.ops.size() &&
                 (a.ops.empty() || std::memcmp(a.ops.data(), b.ops.data(), a.ops.size() * sizeof(TileOp)) == 0);
        }
      }
      if (!same) fail(QM_ERR_DEVICE, "a replayed searched plan differs from the searched plan");
      apply_plan_host(replayed, amps, qubits);
      return;
    }
    if (flags & 8) {
      // Replay: decisions recorded on the same structure with other
      // coefficients, then replayed on these gates; the replayed plan must
      // equal a fresh plan of these gates in every table and op.
      std::vector<DevGate> other = g;
      for (DevGate &x : other)
        for (double &v : x.c) v = 0.5 * v + 0.25;
      PlanTrace trace, fresh_trace;
      (void)plan_fused(other, qubits, shape, nullptr, &trace);
      FusedPlan replayed = plan_fused(g, qubits, shape, &trace);
      FusedPlan fresh = plan_fused(g, qubits, shape, nullptr, &fresh_trace);
      if (shape.lane_layout) fresh = plan_fused(g, qubits, shape, &fresh_trace);
      bool same = replayed.passes.size() == fresh.passes.size();
      for (size_t i = 0; same && i < fresh.passes.size(); i++) {
        const FusedPass &a = replayed.passes[i], &b = fresh.passes[i];
        same = std::memcmp(&a.dev, &b.dev, sizeof(DevPass)) == 0 && a.ops.size() == b.ops.size() &&
               (a.ops.empty() || std::memcmp(a.ops.data(), b.ops.data(), a.ops.size() * sizeof(TileOp)) == 0);
      }
      if (!same) fail(QM_ERR_DEVICE, "a replayed plan differs from the fresh plan");
      apply_plan_host(replayed, amps, qubits);
      return;
    }
    PlanTrace trace;
    FusedPlan plan = plan_fused(g, qubits, shape, nullptr, &trace);
    if (shape.lane_layout) plan = plan_fused(g, qubits, shape, &trace);
    apply_plan_host(plan, amps, qubits);
  });
}

int qm_state_noise(qm_state *h, int32_t kind, double p, int32_t target, const double *u,
                   int64_t n_u, int64_t *cursor) {
  return guarded([&] {
    State *s = as_state(h);
    if (!cursor) fail(QM_ERR_ARG, "cursor is null");
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    DrawSource d = pool_draws(u, n_u, cursor);
    noise_apply(s, kind, p, target, d);
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
  });
}

int qm_state_noise_draw(qm_state *h, int32_t kind, double p, int32_t target, qm_draw_fn draw,
                        void *user) {
  return guarded([&] {
    State *s = as_state(h);
    DrawSource d = callback_draws(draw, user);
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    noise_apply(s, kind, p, target, d);
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
  });
}

int qm_noise_apply(qm_ctx *h, int32_t qubits, int32_t kind, double p, int32_t target, double *amps,
                   int64_t n_amps, qm_draw_fn draw, void *user, const volatile int32_t *cancel) {
  return guarded([&] {
    Ctx *c = reinterpret_cast<Ctx *>(h);
    if (!c) fail(QM_ERR_ARG, "context is null");
    // CpuNoiseChannelBackend.Apply's checks (CpuNoiseChannelBackend.cs:13-22).
    if (!amps || qubits < 1 || qubits > 34 || n_amps != ((int64_t)1 << qubits))
      fail(QM_ERR_ARG, "State length must be a power of two, at least two.");
    if (target < 0 || target >= qubits) fail(QM_ERR_RANGE, "target");
    if (!std::isfinite(p) || p < 0 || p > 1) fail(QM_ERR_RANGE, "Channel probability must be finite and in [0,1].");
    if (kind < 0 || kind > 3) fail(QM_ERR_ARG, "unknown noise kind");
    DrawSource d = callback_draws(draw, user);
    check_cancel(cancel);
    if (p == 0) return;  // no draw and no change (:24-27)
    std::lock_guard<std::recursive_mutex> g(c->lock);
    use_device(c);
    const int64_t bytes = 16 * n_amps;
    if ((c->device_budget && bytes > c->device_budget) || bytes > c->device_memory ||
        (c->staging_budget && bytes > c->staging_budget))
      fail(QM_ERR_BUDGET, "HIP state exceeds the device or host staging memory budget.");
    State *s = state_new(c, qubits, QM_C128);
    try {
      copy_to_device(c, s->data, amps, (size_t)bytes);
      noise_apply(s, kind, p, target, d);
      QM_HIP(hipStreamSynchronize(c->stream));
      check_cancel(cancel);
      // The caller's buffer is written only after the branch is applied.
      copy_to_host(c, amps, s->data, (size_t)bytes);
    } catch (...) {
      state_free(s);
      throw;
    }
    state_free(s);
  });
}

int qm_state_noisy_execute(qm_state *h, const qm_step *steps, int64_t n_steps, const qm_op *ops,
                           const double *u, int64_t n_u, int64_t *cursor) {
  return guarded([&] {
    State *s = as_state(h);
    if (!cursor) fail(QM_ERR_ARG, "buffers are null");
    DrawSource d = pool_draws(u, n_u, cursor);
    noisy_execute(s, steps, n_steps, ops, d);
  });
}

int qm_state_noisy_execute_draw(qm_state *h, const qm_step *steps, int64_t n_steps, const qm_op *ops,
                                qm_draw_fn draw, void *user) {
  return guarded([&] {
    State *s = as_state(h);
    DrawSource d = callback_draws(draw, user);
    noisy_execute(s, steps, n_steps, ops, d);
  });
}

int qm_state_dynamic_execute(qm_state *h, int32_t classical_bits, const qm_dyn_step *steps,
                             int64_t n_steps, const qm_op *ops, int64_t n_ops,
                             const int32_t *sel_qubits, const int32_t *sel_cbits, int64_t n_sel,
                             int32_t schedule, const double *u, int64_t n_u, int64_t *cursor,
                             int32_t *classical, const int32_t *cancel) {
  return guarded([&] {
    State *s = as_state(h);
    if (!cursor) fail(QM_ERR_ARG, "buffers are null");
    DrawSource d = pool_draws(u, n_u, cursor);
    dynamic_execute(s, classical_bits, steps, n_steps, ops, n_ops, sel_qubits, sel_cbits, n_sel, schedule, d,
                    classical, cancel);
  });
}

int qm_state_dynamic_execute_draw(qm_state *h, int32_t classical_bits, const qm_dyn_step *steps,
                                  int64_t n_steps, const qm_op *ops, int64_t n_ops,
                                  const int32_t *sel_qubits, const int32_t *sel_cbits, int64_t n_sel,
                                  int32_t schedule, qm_draw_fn draw, void *user, int32_t *classical,
                                  const int32_t *cancel) {
  return guarded([&] {
    State *s = as_state(h);
    DrawSource d = callback_draws(draw, user);
    dynamic_execute(s, classical_bits, steps, n_steps, ops, n_ops, sel_qubits, sel_cbits, n_sel, schedule, d,
                    classical, cancel);
  });
}

}  // extern "C"

namespace qm {

// NoisyCircuit.Execute (NoisyCircuit.cs:48-65) on a resident state.
static void noisy_execute(State *s, const qm_step *steps, int64_t n_steps, const qm_op *ops,
                          DrawSource &draws) {
  if (n_steps < 0 || (n_steps > 0 && !steps)) fail(QM_ERR_ARG, "buffers are null");
  std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
  // Consecutive unitary segments run as one fused sequence.
  std::vector<DevGate> run;
  auto flush = [&] {
    if (!run.empty()) run_gates(s, run, QM_SCHED_FUSED, nullptr);
    run.clear();
  };
  for (int64_t k = 0; k < n_steps; k++) {
    const qm_step &st = steps[k];
    if (!st.is_channel) {
      if (st.op_begin < 0 || st.op_count < 0) fail(QM_ERR_ARG, "bad segment");
      std::vector<DevGate> g2 = prepare_all(ops + st.op_begin, st.op_count, s->qubits);
      run.insert(run.end(), g2.begin(), g2.end());
    } else if (st.probability != 0) {
      flush();
      noise_apply(s, st.noise_kind, st.probability, st.target, draws);
    } else if (!std::isfinite(st.probability) || st.probability < 0 || st.probability > 1) {
      fail(QM_ERR_RANGE, "Channel probability must be finite and in [0,1].");
    }
  }
  flush();
  QM_HIP(hipStreamSynchronize(s->ctx->stream));
}

// DynamicCircuit.Execute / NoisyDynamicCircuit.Execute on a resident state.
static void dynamic_execute(State *s, int32_t classical_bits, const qm_dyn_step *steps, int64_t n_steps,
                            const qm_op *ops, int64_t n_ops, const int32_t *sel_qubits,
                            const int32_t *sel_cbits, int64_t n_sel, int32_t schedule, DrawSource &draws,
                            int32_t *classical, const int32_t *cancel) {
    if (!classical || n_steps < 0 || (n_steps > 0 && !steps)) fail(QM_ERR_ARG, "buffers are null");
    if (classical_bits < 1 || classical_bits > 30) fail(QM_ERR_RANGE, "classicalBits");
    if (schedule < QM_SCHED_PER_GATE || schedule > QM_SCHED_AUTO) fail(QM_ERR_ARG, "unknown schedule");
    const int n = s->qubits;
    const int32_t reg_mask = (int32_t)((1u << classical_bits) - 1);
    // Every instruction is validated before any device work, as the builders
    // (DynamicCircuit.cs:27-75, NoisyDynamicCircuit.cs:64-77) validate on append.
    std::vector<std::vector<DevGate>> seg((size_t)n_steps);
    for (int64_t k = 0; k < n_steps; k++) {
      const qm_dyn_step &st = steps[k];
      switch (st.kind) {
        case QM_DYN_SEGMENT:
          if (st.begin < 0 || st.op_count < 0 || st.begin + st.op_count > n_ops || (st.op_count > 0 && !ops))
            fail(QM_ERR_ARG, "bad segment");
          if (st.cond_mask < 0 || (st.cond_mask & ~reg_mask)) fail(QM_ERR_RANGE, "classicalMask");
          if (st.cond_value < 0 || (st.cond_value & ~st.cond_mask)) fail(QM_ERR_RANGE, "expectedValue");
          seg[(size_t)k] = prepare_all(ops + st.begin, st.op_count, n);
          break;
        case QM_DYN_MEASURE: {
          if (st.count < 1 || st.count > n || st.begin < 0 || st.begin + st.count > n_sel || !sel_qubits ||
              !sel_cbits)
            fail(QM_ERR_ARG, "Measurement requires equally sized nonempty selections.");
          selection_mask(s, sel_qubits + st.begin, st.count);
          int32_t seen = 0;
          for (int b = 0; b < st.count; b++) {
            const int32_t c = sel_cbits[st.begin + b];
            if (c < 0 || c >= classical_bits) fail(QM_ERR_RANGE, "classicalBits");
            if (seen & (1 << c)) fail(QM_ERR_ARG, "Indices must be distinct.");
            seen |= 1 << c;
          }
          break;
        }
        case QM_DYN_RESET:
          if (st.target < 0 || st.target >= n) fail(QM_ERR_RANGE, "qubit");
          break;
        case QM_DYN_NOISE:
          if (st.target < 0 || st.target >= n) fail(QM_ERR_RANGE, "target");
          if (!std::isfinite(st.probability) || st.probability < 0 || st.probability > 1)
            fail(QM_ERR_RANGE, "Channel probability must be finite and in [0,1].");
          if (st.noise_kind < 0 || st.noise_kind > 3) fail(QM_ERR_ARG, "unknown noise kind");
          break;
        default:
          fail(QM_ERR_ARG, "unknown dynamic instruction");
      }
    }
    check_cancel(cancel);
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    // DynamicCircuit.Execute (DynamicCircuit.cs:77-112) with NoisyDynamicCircuit's
    // channel placements inline (NoisyDynamicCircuit.cs:79-97).  The state stays
    // on the device throughout; the register lives on the host, where each
    // measurement's outcome (one word) arrives.  Segments that run between two
    // measurements, resets or channels are executed as one gate sequence.
    int32_t reg = 0;
    std::vector<DevGate> run;
    auto flush = [&] {
      if (!run.empty()) run_gates(s, run, schedule, cancel);
      run.clear();
    };
    for (int64_t k = 0; k < n_steps; k++) {
      check_cancel(cancel);
      const qm_dyn_step &st = steps[k];
      if (st.kind == QM_DYN_SEGMENT) {
        if ((reg & st.cond_mask) == st.cond_value)
          run.insert(run.end(), seg[(size_t)k].begin(), seg[(size_t)k].end());
      } else if (st.kind == QM_DYN_MEASURE) {
        flush();
        const double draw = draws.next();
        const int64_t out = measure_subset(s, sel_qubits + st.begin, st.count, draw);
        int32_t written = 0, write_mask = 0;
        for (int b = 0; b < st.count; b++) {
          written |= (int32_t)((out >> b) & 1) << sel_cbits[st.begin + b];
          write_mask |= 1 << sel_cbits[st.begin + b];
        }
        reg = (reg & ~write_mask) | written;
      } else if (st.kind == QM_DYN_RESET) {
        flush();
        const double draw = draws.next();
        const int32_t q = st.target;
        if (measure_subset(s, &q, 1, draw) != 0) {
          qm_op x{};
          x.kind = 1;  // GateKind.X
          x.target = q;
          x.control = -1;
          x.second_control = -1;
          std::vector<DevGate> g2 = prepare_all(&x, 1, n);
          run.insert(run.end(), g2.begin(), g2.end());
        }
      } else if (st.probability != 0) {
        flush();
        noise_apply(s, st.noise_kind, st.probability, st.target, draws);
      }
    }
    flush();
    QM_HIP(hipStreamSynchronize(s->ctx->stream));
    check_cancel(cancel);
    *classical = reg;
}

static double take(const double *u, int64_t n, int64_t *cursor) {
  if (!u || *cursor >= n) fail(QM_ERR_POOL, "The uniform pool is exhausted.");
  return u[(*cursor)++];
}

// CpuNoiseChannelBackend.Apply (CpuNoiseChannelBackend.cs:10-145) on a resident state.
static void noise_apply(State *s, int kind, double p, int target, DrawSource &draws) {
  if (target < 0 || target >= s->qubits) fail(QM_ERR_RANGE, "target");
  if (!std::isfinite(p) || p < 0 || p > 1)
    fail(QM_ERR_RANGE, "Channel probability must be finite and in [0,1].");
  if (kind < 0 || kind > 3) fail(QM_ERR_ARG, "unknown noise kind");
  if (p == 0) return;
  hipStream_t st = s->ctx->stream;
  if (kind == 3) {
    double m0, m1;
    reduce_bit_masses(s, target, &m0, &m1, st);
    double jump = p * m1, stay = m0 + (1 - p) * m1, total = jump + stay;
    if (!std::isfinite(total) || total <= 0)
      fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
    bool do_jump;
    if (!(jump > 0)) {
      do_jump = false;
    } else if (stay == 0) {
      do_jump = true;
    } else {
      const double r = draws.next();
      const double ratio = jump / total;
      if (std::fabs(r - ratio) <= 1e-12 * ratio) {
        // Within the summation-order margin: redo the masses in the
        // reference's Kahan order before deciding.
        serial_bit_masses(s, target, &m0, &m1, st);
        jump = p * m1, stay = m0 + (1 - p) * m1, total = jump + stay;
        s->boundary_fallbacks++;
      }
      do_jump = r < jump / total;
    }
    const double branch = do_jump ? jump : stay;
    const double scale = 1 / std::sqrt(branch);
    launch_scale_pairs(s, target, do_jump, scale, std::sqrt(p) * scale, std::sqrt(1 - p) * scale, st);
    return;
  }
  int pauli;
  if (kind == 2) {
    const double r = draws.next();
    const double quarter = p / 4;
    pauli = r < quarter ? 1 : r < 2 * quarter ? 2 : r < 3 * quarter ? 3 : 0;
  } else {
    const bool apply = p == 1 || draws.next() < p;
    pauli = !apply ? 0 : kind == 0 ? 1 : 3;
  }
  if (pauli == 0) return;
  DevGate g{};
  g.target = target;
  g.swap_other = -1;
  if (pauli == 1) {
    g.kind = G_X;
  } else if (pauli == 2) {
    g.kind = G_Y;  // (one.im, -one.re), (-zero.im, zero.re): CpuNoiseChannelBackend.cs:77-78
  } else {
    g.kind = G_DIAG1;  // second = -one
    g.c[0] = 1, g.c[6] = -1, g.c[7] = 0;
  }
  launch_gate(s, g, st);
}

}  // namespace qm
