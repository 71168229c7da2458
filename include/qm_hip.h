/*
 * qm_hip.h -- C ABI of libqmhip.so, the MI355X (gfx950) state-vector backend
 * for QuantumMatter.
 *
 * Plain C types only: pointers, sizes, POD records.  No C++ exceptions cross
 * this boundary.  Every entry point returns a QM_* status; qm_last_error()
 * returns a thread-local message for the last failing call on that thread.
 * The caller owns every host buffer; the library keeps no host pointer after
 * a call returns.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the QuantumMatter repository).  The C# P/Invoke binding that a
 * QuantumMatter.Infrastructure maintainer adds is given in INTEGRATION.md.
 */
#ifndef QM_HIP_H
#define QM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QM_ABI_VERSION 1

/* Status codes.  The shim maps them to the .NET exceptions the reference
 * raises for the same condition. */
#define QM_OK 0
#define QM_ERR_ARG 1         /* ArgumentException / ArgumentNullException */
#define QM_ERR_RANGE 2       /* ArgumentOutOfRangeException */
#define QM_ERR_MASS 3        /* InvalidOperationException: no finite positive mass */
#define QM_ERR_NONFINITE 4   /* InvalidOperationException: nonfinite accumulation */
#define QM_ERR_POOL 5        /* uniform pool exhausted (the System.Random stand-in) */
#define QM_ERR_UNSUPPORTED 6 /* NotSupportedException */
#define QM_ERR_BUDGET 7      /* InvalidOperationException: device/staging budget */
#define QM_ERR_CANCELLED 8   /* OperationCanceledException */
#define QM_ERR_NODEVICE 9    /* PlatformNotSupportedException */
#define QM_ERR_DISPOSED 10   /* ObjectDisposedException */
#define QM_ERR_DEVICE 11     /* HIP runtime failure */
#define QM_ERR_STATE 12      /* InvalidOperationException (API misuse, e.g. profiling off) */

/* Amplitude precision of a device state. */
#define QM_C128 0 /* complex128: interleaved (re,im) doubles, 16 B per amplitude */
#define QM_C64 1  /* complex64: interleaved (re,im) floats, 8 B; FP64 reductions */

/* Execution schedules for qm_state_execute. */
#define QM_SCHED_PER_GATE 0 /* one launch per non-identity gate (the reference's CUDA shape) */
#define QM_SCHED_FUSED 1    /* register/LDS tile passes: many gates per HBM pass (default) */
#define QM_SCHED_COMPILED 2 /* the same passes compiled to straight-line kernels (hipRTC), cached
                               by circuit structure: the first run of a structure compiles */
#define QM_SCHED_AUTO 3     /* compiled kernels when the structure's are all cached; otherwise the
                               interpreting kernel now, and the structure's kernels compiled in the
                               background for its next run (QM_AUTO_COMPILE=0: no background
                               compilation).  qm_execute uses this schedule. */

/* GateOperation (src/QuantumMatter.Core/QuantumCircuit.cs:49-61).  kind uses
 * the GateKind numbering (:4-40); control/second_control are -1 when absent;
 * Swap(a,b) stores control=a, target=b; unitary is the row-major 2x2 matrix
 * m00,m01,m10,m11 as (re,im) pairs, read only for kind 15 (Unitary). */
typedef struct qm_op {
  int32_t kind;
  int32_t target;
  int32_t control;
  int32_t second_control;
  double angle;
  double unitary[8];
} qm_op;

/* PauliString masks (src/QuantumMatter.Core/PauliString.cs:15-26). */
typedef struct qm_pauli {
  uint64_t flip_mask;
  uint64_t sign_mask;
  int32_t y_count;
  int32_t pad;
} qm_pauli;

/* One NoisyCircuit step (src/QuantumMatter.Core/Noise/NoisyCircuit.cs:48-65):
 * a unitary segment ops[op_begin, op_begin+op_count) or, when is_channel, a
 * NoiseKind channel (Noise/NoiseChannel.cs:4-13) of parameter `probability`. */
typedef struct qm_step {
  int32_t is_channel;
  int32_t noise_kind;
  int32_t target;
  int32_t pad;
  double probability;
  int64_t op_begin;
  int64_t op_count;
} qm_step;

/* CudaDeviceInfo (src/QuantumMatter.Infrastructure/CudaBackendOptions.cs:29). */
typedef struct qm_device_info {
  int32_t index;
  int32_t compute_units;
  int64_t memory_bytes;
  char name[128];
  char arch[64];
  char driver[64];
} qm_device_info;

/* CudaBackendOptions (CudaBackendOptions.cs:4-20).  Budgets >= 32 bytes. */
typedef struct qm_ctx_options {
  int32_t device_index;
  int32_t enable_profiling;
  int64_t device_budget_bytes;
  int64_t staging_budget_bytes;
} qm_ctx_options;

/* CudaExecutionProfile (CudaBackendOptions.cs:35) plus the schedule's shape. */
typedef struct qm_profile {
  double device_ms;          /* hipEvent interval around the gate sequence */
  int64_t state_bytes;       /* bytes of the device state */
  int32_t kernel_launches;   /* launches issued for the sequence */
  int32_t passes;            /* full-state HBM passes issued */
  double algorithmic_bytes;  /* bytes the executed schedule must move */
} qm_profile;

typedef struct qm_ctx qm_ctx;
typedef struct qm_state qm_state;

/* Thread-local message of the last failure on this thread ("" if none). */
const char *qm_last_error(void);
int32_t qm_abi_version(void);

/* ---- devices and contexts ---------------------------------------------- */
/* CudaStateVectorBackend.GetDevices (CudaStateVectorBackend.cs:70-85). */
int qm_device_count(int32_t *count);
int qm_device_get_info(int32_t index, qm_device_info *out);

/* CudaStateVectorBackend ctor (CudaStateVectorBackend.cs:21-64): validates
 * options, QM_ERR_NODEVICE without a device, QM_ERR_RANGE for a bad index. */
int qm_ctx_create(const qm_ctx_options *options, qm_ctx **out);
/* CudaStateVectorBackend.Dispose (:191-216): waits for work; idempotent per
 * handle value is the caller's concern (the shim keeps a disposed flag). */
int qm_ctx_destroy(qm_ctx *ctx);
/* Device memory currently held by this context's states (bytes). */
int qm_ctx_resident_bytes(qm_ctx *ctx, int64_t *bytes);
/* No reference counterpart (the reference has no fusion): the fused schedule's
 * tile shape per precision -- 2^tile_bits amplitudes per tile, 2^lane_bits
 * threads per workgroup.  Only compiled shapes are accepted (QM_ERR_UNSUPPORTED
 * otherwise): 6 <= lane_bits <= 10, 2 <= tile_bits - lane_bits <= 5,
 * tile_bits <= 14.  Default 14/9 for both precisions.  Set by the caller's
 * tuning or by qm_autotune. */
int qm_ctx_set_fused_shape(qm_ctx *ctx, int32_t precision, int32_t tile_bits, int32_t lane_bits);
int qm_ctx_get_fused_shape(qm_ctx *ctx, int32_t precision, int32_t *tile_bits, int32_t *lane_bits);

/* No reference counterpart: the autotuner.  Times every tile shape the chosen
 * schedule (QM_SCHED_FUSED: each interpreter instantiation and its code mode;
 * QM_SCHED_COMPILED: generated kernels with 2^11..2^14-amplitude tiles) on a
 * scratch state of `qubits` qubits running a prefix of the given circuit (at
 * most 32 x qubits gates), and keeps the fastest for that schedule, precision
 * and width in the context: later executions of that width use it, until
 * qm_ctx_set_fused_shape replaces it.  entries[0..min(capacity, *count)) get
 * one row per candidate (wall milliseconds of the prefix, best of two runs),
 * *best the winning row.  Results do not depend on the shape beyond rounding
 * (fused and compiled schedules stay within 1e-12 of the per-gate one). */
typedef struct qm_tune_entry {
  int32_t tile_bits;
  int32_t lane_bits;
  int32_t variant;  /* interpreter mode / code-generation variant, -1 = default */
  int32_t passes;
  double ms;
  double ms_per_gate;
  int32_t contiguous;  /* lowest index bits kept in every tile */
  int32_t remap;       /* 1: passes reassign qubits to those bits (compiled schedule) */
} qm_tune_entry;
int qm_autotune(qm_ctx *ctx, int32_t qubits, int32_t precision, int32_t schedule, const qm_op *ops,
                int64_t n_ops, qm_tune_entry *entries, int32_t capacity, int32_t *count,
                int32_t *best);

/* ---- drop-in whole-circuit execution ------------------------------------ */
/* IStateVectorBackend.Execute (src/QuantumMatter.Core/IStateVectorBackend.cs:13)
 * as implemented by CudaStateVectorBackend.ExecuteCore (:103-188): validates
 * n_amplitudes == 2^qubits, checks budgets before allocation, uploads,
 * evolves in insertion order, downloads and commits into `amplitudes` only
 * after success.  On any failure or cancellation the input is unchanged.
 * cancel (nullable) is polled between launches.  profile (nullable) requires
 * enable_profiling (else QM_ERR_STATE, like ExecuteProfiled :92-101). */
int qm_execute(qm_ctx *ctx, int32_t qubits, const qm_op *ops, int64_t n_ops, double *amplitudes,
               int64_t n_amplitudes, const volatile int32_t *cancel, qm_profile *profile);

/* ---- device-resident states ---------------------------------------------- */
/* StateVector ctor (src/QuantumMatter.Core/StateVector.cs:14-25) on the
 * device: |0...0>, 1 <= qubits <= 34, 64-bit indexing.  Checked against the
 * context's device budget before allocation. */
int qm_state_create(qm_ctx *ctx, int32_t qubits, int32_t precision, qm_state **out);
int qm_state_destroy(qm_state *state);
int qm_state_info(qm_state *state, int32_t *qubits, int32_t *precision);
/* StateVector.Reset (StateVector.cs:66-71). */
int qm_state_reset(qm_state *state);
/* Host <-> device copies of interleaved (re,im) doubles; a C64 state rounds on
 * upload and widens on download. */
int qm_state_upload(qm_state *state, const double *amplitudes, int64_t n_amplitudes);
int qm_state_download(qm_state *state, double *amplitudes, int64_t n_amplitudes);
/* Copy (and convert precision) between states of equal width. */
int qm_state_copy(qm_state *dst, qm_state *src);

/* StateVector.Evolve (StateVector.cs:91-102) on the resident state. */
int qm_state_execute(qm_state *state, const qm_op *ops, int64_t n_ops, int32_t schedule,
                     const volatile int32_t *cancel, qm_profile *profile);

/* INoiseChannelBackend.Apply (src/QuantumMatter.Core/Noise/INoiseChannelBackend.cs:13)
 * as CpuNoiseChannelBackend.Apply (src/QuantumMatter.Infrastructure/CpuNoiseChannelBackend.cs:10-145).
 * Draws are taken from uniforms[*cursor ...] under the reference's rules. */
int qm_state_noise(qm_state *state, int32_t kind, double probability, int32_t target,
                   const double *uniforms, int64_t n_uniforms, int64_t *cursor);
/* StateVector.EvolveTrajectory -> NoisyCircuit.Execute (NoisyCircuit.cs:48-65). */
int qm_state_noisy_execute(qm_state *state, const qm_step *steps, int64_t n_steps,
                           const qm_op *ops, const double *uniforms, int64_t n_uniforms,
                           int64_t *cursor);

/* The caller's random source, for the channel and measurement entry points
 * that take the reference's System.Random (INoiseChannelBackend.cs:13-14):
 * called once, on the calling thread, each time the reference calls
 * Random.NextDouble() -- for amplitude damping only when the jump mass is
 * positive and the stay mass nonzero (CpuNoiseChannelBackend.cs:117), never at
 * p == 0 or for a flip at p == 1 -- so the caller's generator advances exactly
 * as it would under the reference.  Writes one uniform in [0,1) to *value and
 * returns 0; a nonzero return aborts the call with QM_ERR_POOL. */
typedef int (*qm_draw_fn)(void *user, double *value);

/* INoiseChannelBackend.Apply (Noise/INoiseChannelBackend.cs:13-14) as
 * CpuNoiseChannelBackend.Apply (CpuNoiseChannelBackend.cs:10-145) on a resident
 * state, drawing from the caller's source. */
int qm_state_noise_draw(qm_state *state, int32_t kind, double probability, int32_t target,
                        qm_draw_fn draw, void *user);
/* The drop-in INoiseChannelBackend.Apply on host amplitudes: validates the
 * length (2^qubits, at least two) and target as CpuNoiseChannelBackend.cs:13-22
 * does, uploads, applies the channel with the caller's draws, and writes the
 * conditional state back into `amplitudes` only after success (on failure or
 * cancellation the buffer is unchanged).  p == 0 returns at once, no draw. */
int qm_noise_apply(qm_ctx *ctx, int32_t qubits, int32_t kind, double probability, int32_t target,
                   double *amplitudes, int64_t n_amplitudes, qm_draw_fn draw, void *user,
                   const volatile int32_t *cancel);
/* StateVector.EvolveTrajectory (StateVector.cs:74-88) with the caller's draws. */
int qm_state_noisy_execute_draw(qm_state *state, const qm_step *steps, int64_t n_steps,
                                const qm_op *ops, qm_draw_fn draw, void *user);

/* Born mass sum |a|^2 (StateVector.cs:109-124), FP64 on device. */
int qm_state_mass(qm_state *state, double *mass);
/* StateVector.MeasureAll (StateVector.cs:106-153) with draw u. */
int qm_state_measure_all(qm_state *state, double u, int64_t *outcome);
/* StateVector.Measure (StateVector.Observations.cs:45-102). */
int qm_state_measure(qm_state *state, const int32_t *qubits, int32_t k, double u,
                     int64_t *outcome);
/* StateVector.MarginalProbability (StateVector.Observations.cs:10-36). */
int qm_state_marginal(qm_state *state, const int32_t *qubits, int32_t k, int64_t outcome,
                      double *probability);
/* StateVectorSimulator.Run terminal sampling (src/QuantumMatter.Application/StateVectorSimulator.cs:44-90):
 * one draw per shot, strict-upper-bound CDF search, no collapse. */
int qm_state_sample(qm_state *state, const double *uniforms, int64_t shots, int64_t *outcomes);
/* StateVector.ExpectationValue (StateVector.Observations.cs:107-141) for T
 * strings at once; terms sharing a flip mask share one state read. */
int qm_state_expectations(qm_state *state, const qm_pauli *terms, int64_t n_terms,
                          double *values);
/* StateVector.ExpectationEnergy (src/QuantumMatter.Core/StateVector.Energy.cs:10-46). */
int qm_state_energy(qm_state *state, const qm_pauli *terms, const double *coefficients,
                    int64_t n_terms, double *energy);
/* No counterpart in the reference: qm_state_energy for callers that cannot
 * allocate (code between hipStreamBeginCapture/EndCapture elsewhere on the
 * device, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and result bit for bit, but it allocates nothing: its
 * partial sums, term tables and results live in a static device array of
 * libqmhip.so (about 17 MB per device, held from the first context on that
 * device until the process ends), its host tables on the stack, and its
 * copies use the caller's memory.  At most QM_NOALLOC_MAX_TERMS terms with
 * nonzero coefficients (QM_ERR_RANGE beyond).  All calls of a process on one
 * device share that device's array and would race on it if they overlapped:
 * the library serialises them with a per-device lock held until the call's
 * last copy completes. */
#define QM_NOALLOC_MAX_TERMS 512
int qm_state_energy_noalloc(qm_state *state, const qm_pauli *terms, const double *coefficients,
                            int64_t n_terms, double *energy);
/* <a|b> over two states of equal width, FP64 accumulation (re, im). */
int qm_state_inner(qm_state *a, qm_state *b, double *re, double *im);

/* ---- schedule inspection (host only, no device needed) -------------------- */
/* No reference counterpart (the reference has no fusion, docs/gpu-execution.md:108):
 * reports the fused schedule QM_SCHED_FUSED would run for these operations. */
int qm_plan_fused(int32_t qubits, int32_t precision, const qm_op *ops, int64_t n_ops,
                  int32_t *passes, int32_t *transposes, int64_t *gates);
/* Runs the fused schedule for an arbitrary tile shape on the host with the
 * kernels' exact arithmetic and indexing (complex128, qubits <= 26), so the
 * planner is verified without a device.  flags: bits 0-1 the remapping mode
 * (1 or 2: passes move qubits into and out of the `contiguous` lowest index
 * bits; the state is back in natural order at the end), bit 2 the tile search
 * (each pass's tile grown for the most runnable gates), bit 3 the replay check
 * (the plan is built from decisions recorded on the same gates with other
 * coefficients, must equal a fresh plan, and is the one applied), bit 4 the
 * compiled schedule's searched planning (perturbed restarts and iterated
 * greedy moves over the tile choices, a small budget; the plan replayed from
 * its recorded trace must equal the searched plan, and is the one applied),
 * bit 5 the lane layout (the plan is replayed from its trace with every pass
 * storing under the index-bit permutation that puts the next pass's load
 * lanes above the contiguous run; passes then alternate between the state and
 * a second buffer, as the compiled schedule runs them). */
int qm_plan_apply_host(int32_t qubits, const qm_op *ops, int64_t n_ops, int32_t tile_bits,
                       int32_t lane_bits, int32_t contiguous, int32_t flags, double *amplitudes);

/* Decisions on CDF and branch boundaries: draws within the rigorous rounding
 * margin of a boundary were re-decided with the reference's sequential
 * summation order.  Returns the count accumulated on this state. */
int qm_state_boundary_fallbacks(qm_state *state, int64_t *count);

/* Gates fused into dense blocks of at most max_block_qubits (1..5) qubits,
 * each block one pass of a 2^(k+1)-square real matrix product on the matrix
 * cores: FP64 MFMA on a complex128 state, FP32 MFMA on a complex64 state (the
 * block's matrix formed in FP64 on the host and rounded to FP32).  The
 * reference has no fused form; this
 * is the dense-block variant of IStateVectorBackend.Execute
 * (src/QuantumMatter.Core/IStateVectorBackend.cs:13) on a resident state. */
int qm_state_execute_dense(qm_state *state, const qm_op *ops, int64_t n_ops, int32_t max_block_qubits,
                           qm_profile *profile);
/* Matrix-core flops of the context's last dense execution. */
int qm_dense_flops(qm_ctx *ctx, double *flops);
/* Host reference of the dense-block schedule (qubits <= 26); *blocks = passes. */
int qm_plan_dense_apply_host(int32_t qubits, const qm_op *ops, int64_t n_ops, int32_t max_block_qubits,
                             double *amplitudes, int32_t *blocks);

/* Pass kernels compiled for QM_SCHED_COMPILED so far in this process, cache
 * hits, and the wall time spent compiling (no counterpart in the reference). */
int qm_jit_stats(int64_t *compiled, int64_t *cache_hits, double *compile_seconds);
/* Blocks until the background compilations that QM_SCHED_AUTO started have
 * finished (no counterpart in the reference: a caller that wants the next run
 * of a structure to be compiled, e.g. before timing it). */
int qm_jit_drain(void);
/* Generates and compiles (without loading) the pass kernels of a plan for the
 * given tile shape; *passes = kernels compiled.  Needs no device. */
int qm_jit_compile_check(int32_t qubits, int32_t precision, const qm_op *ops, int64_t n_ops,
                         int32_t tile_bits, int32_t lane_bits, int32_t *passes);

/* One instruction of a dynamic program (src/QuantumMatter.Core/Dynamic/
 * DynamicCircuit.cs:4-137, Dynamic/NoisyDynamicCircuit.cs:7-99). */
#define QM_DYN_SEGMENT 0 /* ops[begin, begin+op_count), run when (classical & cond_mask) == cond_value */
#define QM_DYN_MEASURE 1 /* measure sel_qubits[begin..begin+count) into sel_cbits[...] (one draw) */
#define QM_DYN_RESET 2   /* measure `target` (one draw) and apply X on outcome 1; no classical write */
#define QM_DYN_NOISE 3   /* channel noise_kind of `probability` on `target` (p == 0: nothing) */
typedef struct qm_dyn_step {
  int32_t kind;
  int32_t cond_mask;  /* 0 = unconditional */
  int32_t cond_value;
  int32_t target;
  int32_t noise_kind;
  int32_t count;
  int64_t begin;
  int64_t op_count;
  double probability;
} qm_dyn_step;

/* StateVector.EvolveDynamic / EvolveNoisyDynamic (src/QuantumMatter.Core/
 * StateVector.Dynamic.cs:11-25, StateVector.NoisyDynamic.cs:11-27) via
 * DynamicCircuit.Execute (Dynamic/DynamicCircuit.cs:77-112) and
 * NoisyDynamicCircuit.Execute (Dynamic/NoisyDynamicCircuit.cs:79-97) on a
 * resident state: the classical register (classical_bits wide) starts at 0;
 * draws come from u[*cursor...] in the reference's order (one per Measure,
 * one per Reset, the channel rules of qm_state_noise).  Every instruction is
 * validated before the state is touched.  The state never leaves the device;
 * between measurements, segments run as one gate sequence under `schedule`.
 * On failure or cancellation the state may be partly evolved and *classical
 * is not written. */
int qm_state_dynamic_execute(qm_state *state, int32_t classical_bits, const qm_dyn_step *steps,
                             int64_t n_steps, const qm_op *ops, int64_t n_ops,
                             const int32_t *sel_qubits, const int32_t *sel_cbits, int64_t n_sel,
                             int32_t schedule, const double *uniforms, int64_t n_uniforms,
                             int64_t *cursor, int32_t *classical, const int32_t *cancel);
/* The same program drawing from the caller's source (one call per Measure and
 * Reset, the channel rules of qm_state_noise_draw), for a caller whose
 * System.Random must stay in step with the reference's. */
int qm_state_dynamic_execute_draw(qm_state *state, int32_t classical_bits, const qm_dyn_step *steps,
                                  int64_t n_steps, const qm_op *ops, int64_t n_ops,
                                  const int32_t *sel_qubits, const int32_t *sel_cbits, int64_t n_sel,
                                  int32_t schedule, qm_draw_fn draw, void *user, int32_t *classical,
                                  const int32_t *cancel);

/* ---------------------------------------------------------------------------
 * Batched execution: one program over many same-width states, one kernel
 * launch, no host round trip between the first gate and the last result.
 * Each state is owned by one workgroup, so per-state decisions (noise
 * branches, damping masses, CDF searches) stay on the device.  Widths below
 * qm_batch_min_qubits() are not batched (QM_ERR_UNSUPPORTED).
 * ------------------------------------------------------------------------- */
int qm_batch_min_qubits(int32_t precision, int32_t *qubits);

/* NoisyStateVectorSimulator.Run (src/QuantumMatter.Application/Noise/
 * NoisyStateVectorSimulator.cs:36-62) over `trajectories` independent
 * trajectories of NoisyCircuit.Execute (src/QuantumMatter.Core/Noise/
 * NoisyCircuit.cs:48-65) with CpuNoiseChannelBackend.Apply's branch rules
 * (src/QuantumMatter.Infrastructure/CpuNoiseChannelBackend.cs:10-153).
 * Trajectory t starts from |0...0> and draws from its own pool
 * pools[t * pool_stride ...] in step order; after the circuit it takes `shots`
 * terminal samples with StateVectorSimulator.Run's CDF rule
 * (src/QuantumMatter.Application/StateVectorSimulator.cs:44-90) from
 * shot_uniforms[t * shots ...] into samples[t * shots ...], then, when
 * measure_all, one MeasureAll draw from its pool (StateVector.cs:106-153) into
 * outcomes[t].  consumed[t] = draws taken from its pool.  Damping and CDF
 * decisions within the rounding margin are re-decided in the reference's
 * sequential order; *fallbacks counts them.  Errors: QM_ERR_POOL when a pool
 * runs out, QM_ERR_MASS for a state without finite positive mass. */
int qm_trajectories_run(qm_ctx *ctx, int32_t qubits, int32_t precision, const qm_step *steps,
                        int64_t n_steps, const qm_op *ops, int64_t n_ops, int64_t trajectories,
                        const double *pools, int64_t pool_stride, int64_t shots,
                        const double *shot_uniforms, int32_t measure_all, int64_t *samples,
                        int64_t *outcomes, int64_t *consumed, int64_t *fallbacks,
                        qm_profile *profile);

/* NoisyStateVectorSimulator.EstimateProbabilities (src/QuantumMatter.Application/
 * Noise/NoisyStateVectorSimulator.cs:64-106) over `trajectories` (>= 2)
 * trajectories, trajectory t drawing from pools[t * pool_stride ...] as in
 * qm_trajectories_run (no terminal measurement).  means[j] is the Welford mean
 * of the normalised probability of basis index j over the trajectories in
 * order, std_errors[j] = sqrt(max(0, M2_j) / (M (M - 1))); both are 2^qubits
 * doubles.  consumed[t] (optional) = draws taken from pool t.  Errors:
 * QM_ERR_RANGE for fewer than 2 trajectories, QM_ERR_POOL, QM_ERR_MASS. */
int qm_trajectories_estimate(qm_ctx *ctx, int32_t qubits, int32_t precision, const qm_step *steps,
                             int64_t n_steps, const qm_op *ops, int64_t n_ops,
                             int64_t trajectories, const double *pools, int64_t pool_stride,
                             double *means, double *std_errors, int64_t *consumed,
                             int64_t *fallbacks, qm_profile *profile);

/* VariationalEnergyEvaluator.EvaluateWithGradient (src/QuantumMatter.Application/
 * Variational/VariationalEnergyEvaluator.cs:34-86): the energy of the bound
 * circuit and the parameter-shift gradient, all 1 + 2R circuits (R = rotation
 * occurrences, src/QuantumMatter.Core/Variational/ParameterizedCircuit.cs:68-96) in one launch.  ops/param_index
 * are the template (param_index[i] >= 0 for a parameter rotation, which must
 * be Rx, Ry or Rz).  Expectations use FP64 reductions for both precisions. */
int qm_energy_gradient_batch(qm_ctx *ctx, int32_t qubits, int32_t precision, const qm_op *ops,
                             const int32_t *param_index, int64_t n_ops, const double *parameters,
                             int64_t n_params, const qm_pauli *terms, const double *coeffs,
                             int64_t n_terms, double *energy, double *gradient,
                             qm_profile *profile);

/* VariationalEnergyEvaluator.Evaluate (VariationalEnergyEvaluator.cs:23-32) for
 * n_sets parameter vectors (parameter_sets[s * n_params ...]) in one launch;
 * energies[s] as StateVector.ExpectationEnergy (StateVector.Energy.cs:10-46). */
int qm_energy_batch(qm_ctx *ctx, int32_t qubits, int32_t precision, const qm_op *ops,
                    const int32_t *param_index, int64_t n_ops, const double *parameter_sets,
                    int64_t n_params, int64_t n_sets, const qm_pauli *terms, const double *coeffs,
                    int64_t n_terms, double *energies, qm_profile *profile);

/* A variational program for repeated evaluation (a VQE loop): the template
 * and Hamiltonian of qm_energy_gradient_batch / qm_energy_batch, planned once.
 * Its first gradient call (and the first energy call for each n_sets) builds
 * a batch whose device arena stays allocated and captures it as a hipGraph
 * (angle upload from pinned memory, the batch kernel, result downloads to
 * pinned memory); later calls write the new angles and replay the graph.
 * Results equal those of the one-shot entry points for the same parameters.
 * The program holds its arenas until qm_vqe_program_destroy; calls on one
 * program are serialised by its context.  Terms and coefficients are copied. */
typedef struct qm_vqe_program qm_vqe_program;
int qm_vqe_program_create(qm_ctx *ctx, int32_t qubits, int32_t precision, const qm_op *ops,
                          const int32_t *param_index, int64_t n_ops, int64_t n_params,
                          const qm_pauli *terms, const double *coeffs, int64_t n_terms,
                          qm_vqe_program **out);
/* VariationalEnergyEvaluator.EvaluateWithGradient (VariationalEnergyEvaluator.cs:34-86). */
int qm_vqe_program_gradient(qm_vqe_program *program, const double *parameters, double *energy,
                            double *gradient, qm_profile *profile);
/* VariationalEnergyEvaluator.Evaluate (VariationalEnergyEvaluator.cs:23-32) for n_sets vectors. */
int qm_vqe_program_energies(qm_vqe_program *program, const double *parameter_sets, int64_t n_sets,
                            double *energies, qm_profile *profile);
int qm_vqe_program_destroy(qm_vqe_program *program);

/* ---------------------------------------------------------------------------
 * Request server (no counterpart in the reference, which runs one circuit per
 * call): requests of 1..34 qubits are admitted in arrival order when their
 * state fits the HBM budget beside everything running; widths up to
 * batch_max_qubits are batched (one workgroup per state, one launch per tile
 * shape, continuously re-formed from whatever waits), wider ones run on
 * concurrent full-chip lanes.  Each lane has its own stream, so work overlaps.
 * Results use the same kernels and rules as qm_state_sample /
 * qm_state_expectations / a downloaded resident state.
 * ------------------------------------------------------------------------- */
#define QM_REQ_AMPLITUDES 0  /* out: 2^qubits amplitudes in the request's precision */
#define QM_REQ_SHOTS 1       /* out: int64_t[shots], StateVectorSimulator.Run's CDF rule */
#define QM_REQ_EXPECTATION 2 /* out: double[n_terms], StateVector.ExpectationValue per term */

typedef struct qm_server_options {
  int32_t device_index;
  int32_t full_lanes;          /* concurrent full-chip lanes (0: 3) */
  int32_t batch_max_qubits;    /* widths batched (0: 14) */
  int32_t max_batch;           /* states per batched launch (0: 1024) */
  int64_t device_budget_bytes; /* admission budget (0: device memory less 4 GiB) */
  int32_t schedule;            /* full-chip lanes: 0 or QM_SCHED_FUSED, or QM_SCHED_AUTO for
                                  streams that repeat circuit structures */
  int32_t pad;
} qm_server_options;

typedef struct qm_request {
  int32_t qubits;
  int32_t precision;
  int32_t kind;
  int32_t pad;
  const qm_op *ops;
  int64_t n_ops;
  int64_t shots;
  const double *uniforms; /* [shots], copied at submit */
  const qm_pauli *terms;  /* [n_terms], copied at submit */
  int64_t n_terms;
  void *out; /* written before the request is reported done; valid until then */
} qm_request;

typedef struct qm_request_status {
  int64_t id;
  int32_t status;
  int32_t lane;      /* 0: batch lane, 1..: full-chip lanes */
  double submit_ms;  /* server clock */
  double start_ms;
  double done_ms;
  int64_t fallbacks; /* margin re-decisions of the launch that served it */
} qm_request_status;

typedef struct qm_server_stats {
  int64_t submitted, completed, failed, pending;
  int64_t reserved_bytes, peak_reserved_bytes, budget_bytes;
  int64_t batches, batched_requests;
  double p50_ms, p99_ms; /* latency, done - submit, over completed requests */
  double span_ms;        /* first submit to last completion */
} qm_server_stats;

typedef struct qm_server qm_server;

int qm_server_create(const qm_server_options *options, qm_server **server);
/* Validates and copies the request (ops prepared now); *id identifies it. */
int qm_server_submit(qm_server *server, const qm_request *request, int64_t *id);
/* Blocks until the request is done; returns its status (message in qm_last_error). */
int qm_server_wait(qm_server *server, int64_t id, qm_request_status *status);
int qm_server_drain(qm_server *server);
int qm_server_get_stats(qm_server *server, qm_server_stats *stats);
int qm_server_reset_stats(qm_server *server);
/* Waits for every submitted request, stops the lanes and frees the contexts. */
int qm_server_destroy(qm_server *server);

#ifdef __cplusplus
}
#endif
#endif
