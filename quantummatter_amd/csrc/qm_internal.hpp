// Internal types shared by the host runtime and the kernels of libqmhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <mutex>
#include <string>
#include <vector>

#include "qm_hip.h"
#include "plan.hpp"

namespace qm {

// Thread-local error text behind qm_last_error().
void set_error(const std::string &msg);
void clear_error();

// A status with its message, carried through the runtime by exception and
// turned into a return code at the C boundary.
struct Status {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg);
void hip_check(hipError_t e, const char *what);

#define QM_HIP(call) ::qm::hip_check((call), #call)

// Runs fn, turning a thrown Status (or any exception) into a status code and
// the thread's last-error text: nothing is thrown across the C boundary.
template <typename Fn>
int guarded(Fn &&fn) {
  try {
    clear_error();
    fn();
    return QM_OK;
  } catch (const Status &s) {
    set_error(s.msg);
    return s.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return QM_ERR_BUDGET;
  } catch (const std::exception &e) {
    set_error(e.what());
    return QM_ERR_DEVICE;
  } catch (...) {
    set_error("unknown failure");
    return QM_ERR_DEVICE;
  }
}

// Default form of a compiled-schedule shape: 2^K-amplitude tiles over 2^L
// threads, the lowest `contiguous` index bits (default 4) in every tile,
// remapped between passes, tiles chosen by the tile search, lane layout on
// (searched plans run out of place between the state and a scratch buffer).
inline FusedShape compiled_shape(int K, int L, int contiguous = 4) {
  FusedShape s{K, L, K - L, contiguous};
  s.remap = 1;
  s.search = 1;
  s.lane_layout = 1;
  return s;
}

inline FusedShape alt_low_bits(FusedShape s, int alt) {
  s.alt_contiguous = alt;
  return s;
}

struct Ctx {
  int device = 0;
  bool profiling = false;
  int64_t device_budget = 0;   // 0: device capacity
  int64_t staging_budget = 0;  // 0: unlimited
  int64_t device_memory = 0;
  int compute_units = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  std::recursive_mutex lock;  // serialises calls on one context
  int64_t resident_bytes = 0;
  // Reduction scratch (grown on demand, never shrunk while the context lives).
  double *d_partial = nullptr;
  int64_t partial_cap = 0;  // doubles
  int64_t *d_ipartial = nullptr;
  int64_t ipartial_cap = 0;
  double *h_pinned = nullptr;  // small pinned readback area
  int64_t pinned_cap = 0;
  // Device copy of fused-pass programs.
  void *d_prog = nullptr;
  int64_t prog_cap = 0;
  bool disposed = false;
  double last_dense_flops = 0;  // matrix-core flops of the last dense execution
  // Tile shape per precision (qm_ctx_set_fused_shape).  The compiled schedule
  // uses it as given; the interpreting kernel uses it when it is one of its
  // instantiations and its own default otherwise.  Defaults: complex128
  // 8192-amplitude tiles, 256 threads x 32 registers (two workgroups per CU,
  // each with a 64 KiB LDS exchange buffer, so one tile's loads and stores
  // overlap the other's arithmetic), the low 3 index bits (or 4 when that plan
  // models faster) remapped between passes; complex64 16384-amplitude tiles,
  // 512 threads x 32 registers (two workgroups per CU), the low 4 bits (128-byte
  // runs; or 5 when that plan models faster: on the 30-qubit layer circuit 15
  // passes against 17 with 5, 444-447 gate-layers/s against 436).  Over
  // six layer circuits of 26-30 qubits the complex128 shape runs 1-23 %
  // faster than 16384-amplitude tiles with one workgroup per CU (18 passes of
  // the 30-qubit 40-layer circuit against 14, each shorter); in complex64 the
  // 16384-amplitude shape is the faster (19 passes).
  FusedShape shape[2] = {alt_low_bits(compiled_shape(13, 8, 3), 4), alt_low_bits(compiled_shape(14, 9, 4), 5)};
  // Shapes chosen by qm_autotune, [fused 0 / compiled 1][precision][qubits];
  // K == 0: not tuned.  They take precedence over `shape` for their width.
  FusedShape tuned[2][2][35] = {};
};

struct State {
  Ctx *ctx = nullptr;
  int qubits = 0;
  int precision = QM_C128;
  int64_t length = 0;  // amplitudes
  void *data = nullptr;
  int64_t bytes = 0;
  int64_t boundary_fallbacks = 0;
  // Second buffer of `bytes` for passes that read one buffer and write the
  // other (lane-layout plans); allocated on first use, counted in the
  // context's resident bytes, freed with the state.
  void *scratch = nullptr;
};

inline int64_t amp_bytes(int precision) { return precision == QM_C128 ? 16 : 8; }

// Shared host helpers (api.hip).
void use_device(Ctx *c);  // checks the context and makes its device current
void check_cancel(const volatile int32_t *cancel);
std::vector<DevGate> prepare_all(const qm_op *ops, int64_t n_ops, int n);  // validated, identity dropped

// Resident states and gate execution (api.hip).
State *state_new(Ctx *c, int qubits, int precision);  // checks the context's budget
void state_free(State *s);
// Allocates the state's scratch buffer if it has none; false when it does not fit
// the budget or the device.
bool state_scratch(State *s);
struct ExecStats {
  int launches = 0;
  int passes = 0;
  double bytes = 0;
};
// Runs prepared gates on a state with the chosen schedule (synchronous).
// `force` (optional) replaces the context's tile shape for this call.
ExecStats run_gates(State *s, const std::vector<DevGate> &gates, int schedule,
                    const volatile int32_t *cancel, const FusedShape *force = nullptr);

// Scratch management on the context's stream.
double *ctx_partials(Ctx *c, int64_t n);
int64_t *ctx_ipartials(Ctx *c, int64_t n);
double *ctx_pinned(Ctx *c, int64_t n);
void *ctx_program(Ctx *c, int64_t bytes);
// The compiled schedule's searched plan of a structure, through the plan cache
// (api.hip); *chosen and *trace replay it for other coefficients.
FusedPlan plan_compiled_trace(const std::vector<DevGate> &gates, int n, int precision, const FusedShape &shape,
                              FusedShape *chosen, PlanTrace *trace);
void apply_tuning_knobs(FusedShape &shape);
// Synchronous host <-> device copies staged through two pinned chunks of the
// context (the DMA of one chunk overlaps the host copy of the other).
void copy_to_device(Ctx *c, void *dst, const void *src, size_t bytes);
void copy_to_host(Ctx *c, void *dst, const void *src, size_t bytes);

// ---- kernel launchers (kernels.hip) ---------------------------------------
void launch_gate(State *s, const DevGate &g, hipStream_t st);
void launch_fused_pass(State *s, const FusedShape &shape, const void *d_pass, const void *d_ops,
                       hipStream_t st);
FusedShape default_fused_shape(int precision);
std::vector<FusedShape> fused_shapes(int precision);  // every interpreter instantiation
bool fused_shape_supported(int precision, int L, int R);
void launch_set_basis(State *s, int64_t index, hipStream_t st);  // |index>
void launch_scale_pairs(State *s, int target, bool jump, double scale, double jump_scale,
                        double stay_scale, hipStream_t st);
void launch_collapse_subset(State *s, uint64_t mask, uint64_t pattern, double norm, hipStream_t st);
void launch_convert(const State *src, State *dst, hipStream_t st);

// Reductions.  Results land in host memory (synchronous on the stream).
struct MassResult {
  double total;
  int64_t last_supported;
};
MassResult reduce_mass(State *s, hipStream_t st);
// Masked sum of probabilities: sum over i with (i & mask) == pattern.
double reduce_masked_mass(State *s, uint64_t mask, uint64_t pattern, hipStream_t st);
// Probability masses of target bit 0 and 1.
void reduce_bit_masses(State *s, int target, double *m0, double *m1, hipStream_t st);
// Kahan-compensated sequential masses in the reference's order (one lane).
void serial_bit_masses(State *s, int target, double *m0, double *m1, hipStream_t st);
// Un-normalised expectation numerators for a set of Pauli terms.
void reduce_pauli(State *s, const qm_pauli *terms, int64_t n_terms, double *numerators,
                  hipStream_t st);
void reduce_inner(State *a, State *b, double *re, double *im, hipStream_t st);

// Energy reductions without allocation: the mass and the un-normalised
// numerators of up to QM_NOALLOC_MAX_TERMS terms, bit-identical to
// reduce_mass / reduce_pauli, in the device's static scratch array.
// noalloc_setup (device current) finds that array's device address once per
// device; calls on one device are serialised by a per-device lock.
void noalloc_setup(int device);
size_t noalloc_scratch_bytes();
void reduce_energy_noalloc(State *s, const qm_pauli *terms, int64_t n_terms, double *total,
                           double *numerators, hipStream_t st);

// CDF search for a batch of draws with the strict-upper-bound rule; resolves
// draws within the rounding margin of a boundary with the sequential order.
// Returns the number of margin fallbacks.
int64_t cdf_search(State *s, const double *uniforms, int64_t n, int64_t *outcomes, hipStream_t st,
                   double *total_out = nullptr);

}  // namespace qm
