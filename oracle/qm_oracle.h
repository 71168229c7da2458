/*
 * qm_oracle.h -- CPU restatement of QuantumMatter's double-precision CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * backend: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (libqmhip.so) never links or calls it.
 *
 * Every function restates one reference routine line by line (file:line given
 * at each declaration, paths relative to the QuantumMatter repository).  It is
 * built with -ffp-contract=off so that every product and sum rounds exactly as
 * the .NET runtime rounds System.Numerics.Complex arithmetic (no FMA
 * contraction), which makes results bit-comparable with the reference.
 *
 * System.Random is replaced by a caller-supplied pool of uniforms in [0,1)
 * consumed through a cursor; each API consumes draws in exactly the order and
 * under exactly the conditions of the reference (see each function).
 */
#ifndef QM_ORACLE_H
#define QM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, shared numbering with include/qm_hip.h. */
#define QMO_OK 0
#define QMO_ERR_ARG 1       /* ArgumentException */
#define QMO_ERR_RANGE 2     /* ArgumentOutOfRangeException */
#define QMO_ERR_MASS 3      /* InvalidOperationException: no finite positive mass */
#define QMO_ERR_NONFINITE 4 /* InvalidOperationException: nonfinite accumulation */
#define QMO_ERR_POOL 5      /* uniform pool exhausted */
#define QMO_ERR_UNSUPPORTED 6

/* GateKind numbering of src/QuantumMatter.Core/QuantumCircuit.cs:4-40. */
enum {
  QMO_H = 0, QMO_X = 1, QMO_Y = 2, QMO_Z = 3, QMO_S = 4, QMO_T = 5,
  QMO_RX = 6, QMO_RY = 7, QMO_RZ = 8, QMO_CNOT = 9, QMO_SWAP = 10, QMO_I = 11,
  QMO_SDG = 12, QMO_TDG = 13, QMO_PHASE = 14, QMO_UNITARY = 15, QMO_TOFFOLI = 16
};

/* NoiseKind numbering of src/QuantumMatter.Core/Noise/NoiseChannel.cs:4-13. */
enum { QMO_BITFLIP = 0, QMO_PHASEFLIP = 1, QMO_DEPOLARIZING = 2, QMO_DAMPING = 3 };

/* GateOperation (QuantumCircuit.cs:49-61).  Swap(a,b) stores control=a and
 * target=b.  unitary[] is the row-major 2x2 matrix m00,m01,m10,m11 as
 * interleaved (re,im) pairs, used only when kind == UNITARY.
 * Layout is identical to qm_op in include/qm_hip.h. */
typedef struct {
  int32_t kind;
  int32_t target;
  int32_t control;
  int32_t second_control;
  double angle;
  double unitary[8];
} qmo_op;

/* PreparedStateVectorGate (PreparedStateVectorGate.cs:6-11). */
enum { QMO_K_IDENTITY = 0, QMO_K_MATRIX = 1, QMO_K_HADAMARD = 2, QMO_K_EXCHANGE = 3,
       QMO_K_DIAGONAL = 4, QMO_K_SWAP = 5 };
typedef struct {
  int32_t kernel;
  int32_t target;
  int64_t control_mask;
  int32_t swap_control;
  int32_t pad;
  double coeff[8]; /* A, B, C, D as (re, im) */
} qmo_prepared;

/* Pauli string masks (PauliString.cs:15-26). */
typedef struct {
  uint64_t flip_mask;
  uint64_t sign_mask;
  int32_t y_count;
  int32_t pad;
} qmo_pauli;

/* One step of a NoisyCircuit (Noise/NoisyCircuit.cs:48-65): either a unitary
 * segment ops[op_begin, op_begin+op_count) or a channel on target. */
typedef struct {
  int32_t is_channel;
  int32_t noise_kind;
  int32_t target;
  int32_t pad;
  double probability;
  int64_t op_begin;
  int64_t op_count;
} qmo_step;

/* PreparedStateVectorGate.Prepare (PreparedStateVectorGate.cs:13-70). */
int qmo_prepare(const qmo_op *op, qmo_prepared *out);

/* CpuStateVectorBackend.Execute (CpuStateVectorBackend.cs:10-91) on an
 * interleaved (re,im) buffer of 2^qubits amplitudes. */
int qmo_execute(int qubits, const qmo_op *ops, int64_t n_ops, double *amps);

/* CpuNoiseChannelBackend.Apply (CpuNoiseChannelBackend.cs:10-83), ApplyDamping
 * (:85-145).  Draws come from uniforms[*cursor...]. */
int qmo_noise_apply(int kind, double probability, int target, int qubits, double *amps,
                    const double *uniforms, int64_t n_uniforms, int64_t *cursor);

/* NoisyCircuit.Execute (Noise/NoisyCircuit.cs:48-65). */
int qmo_noisy_execute(int qubits, const qmo_step *steps, int64_t n_steps, const qmo_op *ops,
                      double *amps, const double *uniforms, int64_t n_uniforms, int64_t *cursor);

/* StateVector.MeasureAll (StateVector.cs:106-153). */
int qmo_measure_all(int qubits, double *amps, double u, int64_t *outcome);

/* StateVector.Measure (StateVector.Observations.cs:45-102). */
int qmo_measure(int qubits, double *amps, const int32_t *sel, int k, double u, int64_t *outcome);

/* StateVector.MarginalProbability (StateVector.Observations.cs:10-36). */
int qmo_marginal(int qubits, const double *amps, const int32_t *sel, int k, int64_t outcome,
                 double *probability);

/* StateVector.ExpectationValue (StateVector.Observations.cs:107-141). */
int qmo_expectation(int qubits, const double *amps, const qmo_pauli *p, double *value);

/* StateVector.ExpectationEnergy (StateVector.Energy.cs:10-46). */
int qmo_energy(int qubits, const double *amps, const qmo_pauli *terms, const double *coeffs,
               int64_t n_terms, double *energy);

/* StateVectorSimulator.Run terminal sampling (StateVectorSimulator.cs:44-90):
 * sequential CDF, one draw per shot, strict upper bound search. */
int qmo_sample(int qubits, const double *amps, const double *uniforms, int64_t shots,
               int64_t *outcomes);

/* Mass of the state as a sequential left-to-right sum (StateVector.cs:109-124). */
double qmo_mass(int qubits, const double *amps);

/* NoisyStateVectorSimulator.Run per trajectory (NoisyStateVectorSimulator.cs:44-50):
 * Reset, noisy evolution, MeasureAll -- with one uniform pool per trajectory of
 * length pool_stride.  Optionally also `shots` terminal samples of the
 * pre-measurement trajectory state (Run's CDF rule) from shot_uniforms.
 * outcome[t] and samples[t*shots + s] are written.  Used for batched parity. */
int qmo_trajectories(int qubits, const qmo_step *steps, int64_t n_steps, const qmo_op *ops,
                     int64_t trajectories, const double *pools, int64_t pool_stride,
                     int64_t *consumed, int64_t *outcomes, int64_t shots,
                     const double *shot_uniforms, int64_t *samples);

/* NoisyStateVectorSimulator.EstimateProbabilities (:64-106) with one pool per
 * trajectory.  means and stderr are 2^qubits doubles each. */
int qmo_estimate_probabilities(int qubits, const qmo_step *steps, int64_t n_steps,
                               const qmo_op *ops, int64_t trajectories, const double *pools,
                               int64_t pool_stride, double *means, double *stderr_out);

/* ParameterizedCircuit.BindCore (Variational/ParameterizedCircuit.cs:68-96):
 * template ops with param_index[i] >= 0 take angle parameters[param_index[i]];
 * when shifted_rotation >= 0 a same-axis rotation of angle `shift` is appended
 * after that rotation occurrence.  Writes up to n_ops+1 ops, returns count in
 * *n_out. */
int qmo_bind(const qmo_op *ops, const int32_t *param_index, int64_t n_ops,
             const double *parameters, int64_t n_params, int64_t shifted_rotation, double shift,
             qmo_op *out, int64_t *n_out);

/* VariationalEnergyEvaluator.EvaluateWithGradient (VariationalEnergyEvaluator.cs:34-86). */
int qmo_energy_gradient(int qubits, const qmo_op *ops, const int32_t *param_index, int64_t n_ops,
                        const double *parameters, int64_t n_params, const qmo_pauli *terms,
                        const double *coeffs, int64_t n_terms, double *energy, double *gradient);

/* VariationalEnergyEvaluator.Evaluate (:23-30). */
int qmo_energy_bound(int qubits, const qmo_op *ops, const int32_t *param_index, int64_t n_ops,
                     const double *parameters, int64_t n_params, const qmo_pauli *terms,
                     const double *coeffs, int64_t n_terms, double *energy);

/* One instruction of a dynamic program (Dynamic/DynamicCircuit.cs:4-137,
 * Dynamic/NoisyDynamicCircuit.cs:7-99).  Layout is identical to qm_dyn_step
 * in include/qm_hip.h. */
enum { QMO_DYN_SEGMENT = 0, QMO_DYN_MEASURE = 1, QMO_DYN_RESET = 2, QMO_DYN_NOISE = 3 };
typedef struct {
  int32_t kind;
  int32_t cond_mask;   /* segment: runs when (classical & cond_mask) == cond_value */
  int32_t cond_value;
  int32_t target;      /* reset qubit / noise target */
  int32_t noise_kind;
  int32_t count;       /* measure: selection size */
  int64_t begin;       /* segment: first op; measure: first selection entry */
  int64_t op_count;    /* segment: op count */
  double probability;  /* noise */
} qmo_dyn_step;

/* DynamicCircuit.Execute (Dynamic/DynamicCircuit.cs:77-112) with the channel
 * placements of NoisyDynamicCircuit.Execute (Dynamic/NoisyDynamicCircuit.cs:
 * 79-97) inline: the classical register starts at 0; a segment runs when its
 * condition holds; Measure draws once (StateVector.Observations.cs:45-102)
 * and overwrites its destination bits; Reset measures its qubit (one draw)
 * and applies X on outcome 1; a channel applies CpuNoiseChannelBackend's
 * rules (p == 0: nothing).  Draws come from uniforms[*cursor...]. */
int qmo_dynamic_execute(int qubits, const qmo_dyn_step *steps, int64_t n_steps, const qmo_op *ops,
                        const int32_t *sel_qubits, const int32_t *sel_cbits, double *amps,
                        const double *uniforms, int64_t n_uniforms, int64_t *cursor,
                        int32_t *classical);

#ifdef __cplusplus
}
#endif
#endif
