/*
 * qm_oracle.c -- scalar CPU restatement of QuantumMatter's reference path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  Never linked by the product.
 * Build with -ffp-contract=off: System.Numerics.Complex arithmetic in .NET is
 * evaluated without fused multiply-add, and so is every expression here.
 *
 * Citations are to the QuantumMatter repository (src/...).
 */
#include "qm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double re, im;
} cx;

/* System.Numerics.Complex operator* (complex x complex). */
static inline cx cmul(cx a, cx b) {
  cx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.im * b.re + a.re * b.im;
  return r;
}
static inline cx cadd(cx a, cx b) {
  cx r = {a.re + b.re, a.im + b.im};
  return r;
}
/* double * Complex and Complex * double scale both components. */
static inline cx cscale(cx a, double s) {
  cx r = {a.re * s, a.im * s};
  return r;
}
static inline cx mk(double re, double im) {
  cx r = {re, im};
  return r;
}
static inline double prob(const cx *s, int64_t i) { return s[i].re * s[i].re + s[i].im * s[i].im; }

static int is_finite(double x) { return isfinite(x) != 0; }

static int draw(const double *u, int64_t n, int64_t *cursor, double *out) {
  if (*cursor >= n) return QMO_ERR_POOL;
  *out = u[(*cursor)++];
  return QMO_OK;
}

/* ---------------------------------------------------------------------------
 * Gate kernels: CpuStateVectorBackend.ApplyMatrix/ApplyX/ApplyDiagonal/ApplySwap
 * (src/QuantumMatter.Infrastructure/CpuStateVectorBackend.cs:93-177).
 * Indices are 64-bit so the oracle also covers registers above 2^31 amplitudes.
 * ------------------------------------------------------------------------- */
static void apply_matrix(cx *s, int64_t len, int target, cx a, cx b, cx c, cx d, int64_t mask) {
  int64_t bit = (int64_t)1 << target;
  for (int64_t pair = 0; pair < len / 2; pair++) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low;
    if ((first & mask) != mask) continue;
    int64_t second = first | bit;
    cx x = s[first], y = s[second];
    s[first] = cadd(cmul(a, x), cmul(b, y));
    s[second] = cadd(cmul(c, x), cmul(d, y));
  }
}

static void apply_x(cx *s, int64_t len, int target, int64_t mask) {
  int64_t bit = (int64_t)1 << target;
  for (int64_t pair = 0; pair < len / 2; pair++) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low;
    if ((first & mask) != mask) continue;
    int64_t second = first | bit;
    cx t = s[first];
    s[first] = s[second];
    s[second] = t;
  }
}

static void apply_diagonal(cx *s, int64_t len, int target, cx zero, cx one, int64_t mask) {
  int64_t bit = (int64_t)1 << target;
  for (int64_t i = 0; i < len; i++) {
    if ((i & mask) == mask) s[i] = cmul(s[i], (i & bit) == 0 ? zero : one);
  }
}

static void apply_swap(cx *s, int64_t len, int control, int target) {
  int64_t cb = (int64_t)1 << control, tb = (int64_t)1 << target;
  for (int64_t i = 0; i < len; i++) {
    if ((i & cb) != 0 && (i & tb) == 0) {
      int64_t o = i ^ cb ^ tb;
      cx t = s[i];
      s[i] = s[o];
      s[o] = t;
    }
  }
}

static int validate_op(const qmo_op *op, int qubits) {
  if (op->kind < 0 || op->kind > QMO_TOFFOLI) return QMO_ERR_UNSUPPORTED;
  if (op->target < 0 || op->target >= qubits) return QMO_ERR_RANGE;
  if (op->control >= qubits || op->second_control >= qubits) return QMO_ERR_RANGE;
  return QMO_OK;
}

/* CpuStateVectorBackend.Execute (CpuStateVectorBackend.cs:10-91). */
int qmo_execute(int qubits, const qmo_op *ops, int64_t n_ops, double *amps) {
  if (qubits < 1 || qubits > 40 || (n_ops > 0 && ops == NULL) || amps == NULL) return QMO_ERR_ARG;
  cx *s = (cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  for (int64_t k = 0; k < n_ops; k++) {
    const qmo_op *op = &ops[k];
    int rc = validate_op(op, qubits);
    if (rc) return rc;
    int64_t mask = op->control >= 0 ? (int64_t)1 << op->control : 0;
    if (op->second_control >= 0) mask |= (int64_t)1 << op->second_control;
    switch (op->kind) {
      case QMO_I:
        break;
      case QMO_CNOT:
      case QMO_TOFFOLI:
      case QMO_X:
        apply_x(s, len, op->target, mask);
        break;
      case QMO_SWAP:
        apply_swap(s, len, op->control, op->target);
        break;
      case QMO_H: {
        double h = sqrt(0.5);
        apply_matrix(s, len, op->target, mk(h, 0), mk(h, 0), mk(h, 0), mk(-h, 0), mask);
        break;
      }
      case QMO_Y: /* (0, -i; i, 0); -Complex.ImaginaryOne is (-0, -1). */
        apply_matrix(s, len, op->target, mk(0, 0), mk(-0.0, -1), mk(0, 1), mk(0, 0), mask);
        break;
      case QMO_Z:
        apply_diagonal(s, len, op->target, mk(1, 0), mk(-1, 0), mask);
        break;
      case QMO_S:
        apply_diagonal(s, len, op->target, mk(1, 0), mk(0, 1), mask);
        break;
      case QMO_SDG:
        apply_diagonal(s, len, op->target, mk(1, 0), mk(-0.0, -1), mask);
        break;
      case QMO_T:
      case QMO_TDG:
      case QMO_PHASE: {
        double phase = op->kind == QMO_T ? M_PI / 4 : op->kind == QMO_TDG ? -M_PI / 4 : op->angle;
        /* Complex.FromPolarCoordinates(1, phase) = (1*cos, 1*sin). */
        apply_diagonal(s, len, op->target, mk(1, 0), mk(1.0 * cos(phase), 1.0 * sin(phase)), mask);
        break;
      }
      case QMO_RX:
      case QMO_RY: {
        double cosine = cos(op->angle / 2);
        double sine = sin(op->angle / 2);
        /* Rx: upper = -ImaginaryOne * sine = (-0*sine, -1*sine); Ry: upper = -sine. */
        cx upper = op->kind == QMO_RX ? mk(-0.0 * sine, -1.0 * sine) : mk(-sine, 0);
        cx lower = op->kind == QMO_RX ? upper : mk(sine, 0);
        apply_matrix(s, len, op->target, mk(cosine, 0), upper, lower, mk(cosine, 0), mask);
        break;
      }
      case QMO_RZ:
        apply_diagonal(s, len, op->target, mk(cos(-op->angle / 2), sin(-op->angle / 2)),
                       mk(cos(op->angle / 2), sin(op->angle / 2)), mask);
        break;
      case QMO_UNITARY: {
        const double *u = op->unitary;
        apply_matrix(s, len, op->target, mk(u[0], u[1]), mk(u[2], u[3]), mk(u[4], u[5]),
                     mk(u[6], u[7]), mask);
        break;
      }
      default:
        return QMO_ERR_UNSUPPORTED;
    }
  }
  return QMO_OK;
}

/* PreparedStateVectorGate.Prepare (PreparedStateVectorGate.cs:13-70). */
int qmo_prepare(const qmo_op *op, qmo_prepared *g) {
  memset(g, 0, sizeof(*g));
  int64_t mask = op->control >= 0 ? (int64_t)1 << op->control : 0;
  if (op->second_control >= 0) mask |= (int64_t)1 << op->second_control;
  g->target = op->target;
  g->control_mask = mask;
  g->swap_control = -1;
  cx a = mk(0, 0), b = mk(0, 0), c = mk(0, 0), d = mk(0, 0);
  int kernel = QMO_K_MATRIX;
  switch (op->kind) {
    case QMO_I:
      kernel = QMO_K_IDENTITY;
      break;
    case QMO_X:
    case QMO_CNOT:
    case QMO_TOFFOLI:
      kernel = QMO_K_EXCHANGE;
      break;
    case QMO_SWAP:
      kernel = QMO_K_SWAP;
      g->control_mask = 0;
      g->swap_control = op->control;
      break;
    case QMO_H: {
      double h = sqrt(0.5);
      kernel = QMO_K_HADAMARD;
      a = mk(h, 0), b = mk(h, 0), c = mk(h, 0), d = mk(-h, 0);
      break;
    }
    case QMO_Y:
      a = mk(0, 0), b = mk(-0.0, -1), c = mk(0, 1), d = mk(0, 0);
      break;
    case QMO_Z:
      kernel = QMO_K_DIAGONAL, a = mk(1, 0), d = mk(-1, 0);
      break;
    case QMO_S:
      kernel = QMO_K_DIAGONAL, a = mk(1, 0), d = mk(0, 1);
      break;
    case QMO_SDG:
      kernel = QMO_K_DIAGONAL, a = mk(1, 0), d = mk(-0.0, -1);
      break;
    case QMO_T:
      kernel = QMO_K_DIAGONAL, a = mk(1, 0), d = mk(cos(M_PI / 4), sin(M_PI / 4));
      break;
    case QMO_TDG:
      kernel = QMO_K_DIAGONAL, a = mk(1, 0), d = mk(cos(-M_PI / 4), sin(-M_PI / 4));
      break;
    case QMO_PHASE:
      kernel = QMO_K_DIAGONAL, a = mk(1, 0), d = mk(cos(op->angle), sin(op->angle));
      break;
    case QMO_RX: {
      double xc = cos(op->angle / 2);
      cx xs = mk(-0.0 * sin(op->angle / 2), -1.0 * sin(op->angle / 2));
      a = mk(xc, 0), b = xs, c = xs, d = mk(xc, 0);
      break;
    }
    case QMO_RY: {
      double yc = cos(op->angle / 2), ys = sin(op->angle / 2);
      a = mk(yc, 0), b = mk(-ys, 0), c = mk(ys, 0), d = mk(yc, 0);
      break;
    }
    case QMO_RZ:
      kernel = QMO_K_DIAGONAL;
      a = mk(cos(-op->angle / 2), sin(-op->angle / 2));
      d = mk(cos(op->angle / 2), sin(op->angle / 2));
      break;
    case QMO_UNITARY: {
      const double *u = op->unitary;
      a = mk(u[0], u[1]), b = mk(u[2], u[3]), c = mk(u[4], u[5]), d = mk(u[6], u[7]);
      break;
    }
    default:
      return QMO_ERR_UNSUPPORTED;
  }
  g->kernel = kernel;
  g->coeff[0] = a.re, g->coeff[1] = a.im, g->coeff[2] = b.re, g->coeff[3] = b.im;
  g->coeff[4] = c.re, g->coeff[5] = c.im, g->coeff[6] = d.re, g->coeff[7] = d.im;
  return QMO_OK;
}

/* ---------------------------------------------------------------------------
 * Noise: CpuNoiseChannelBackend (src/QuantumMatter.Infrastructure/CpuNoiseChannelBackend.cs).
 * ------------------------------------------------------------------------- */
/* AddMass (:147-153): Kahan compensated accumulation. */
static void add_mass(double value, double *sum, double *comp) {
  double corrected = value - *comp;
  double next = *sum + corrected;
  *comp = (next - *sum) - corrected;
  *sum = next;
}

/* ApplyDamping (:85-145). */
static int apply_damping(cx *s, int64_t len, int64_t bit, double gamma, const double *u,
                         int64_t n_u, int64_t *cursor) {
  double zero_mass = 0, one_mass = 0, zc = 0, oc = 0;
  for (int64_t i = 0; i < len; i++) {
    double m = prob(s, i);
    if ((i & bit) == 0)
      add_mass(m, &zero_mass, &zc);
    else
      add_mass(m, &one_mass, &oc);
  }
  double jump_mass = gamma * one_mass;
  double stay_mass = zero_mass + (1 - gamma) * one_mass;
  double total = jump_mass + stay_mass;
  if (!is_finite(total) || total <= 0) return QMO_ERR_MASS;
  int jump;
  if (!(jump_mass > 0)) {
    jump = 0;
  } else if (stay_mass == 0) {
    jump = 1;
  } else {
    double r;
    int rc = draw(u, n_u, cursor, &r);
    if (rc) return rc;
    jump = r < jump_mass / total;
  }
  double branch = jump ? jump_mass : stay_mass;
  double scale = 1 / sqrt(branch);
  double jump_scale = sqrt(gamma) * scale;
  double stay_scale = sqrt(1 - gamma) * scale;
  for (int64_t pair = 0; pair < len / 2; pair++) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low;
    int64_t second = first | bit;
    if (jump) {
      s[first] = cscale(s[second], jump_scale);
      s[second] = mk(0, 0);
    } else {
      s[first] = cscale(s[first], scale);
      s[second] = cscale(s[second], stay_scale);
    }
  }
  return QMO_OK;
}

/* Apply (:10-83). */
int qmo_noise_apply(int kind, double p, int target, int qubits, double *amps, const double *u,
                    int64_t n_u, int64_t *cursor) {
  if (qubits < 1 || amps == NULL || cursor == NULL) return QMO_ERR_ARG;
  if (target < 0 || target >= qubits) return QMO_ERR_RANGE;
  if (!is_finite(p) || p < 0 || p > 1) return QMO_ERR_RANGE;
  if (kind < 0 || kind > QMO_DAMPING) return QMO_ERR_ARG;
  if (p == 0) return QMO_OK;
  cx *s = (cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  int64_t bit = (int64_t)1 << target;
  if (kind == QMO_DAMPING) return apply_damping(s, len, bit, p, u, n_u, cursor);
  int pauli;
  if (kind == QMO_DEPOLARIZING) {
    double r;
    int rc = draw(u, n_u, cursor, &r);
    if (rc) return rc;
    double quarter = p / 4;
    pauli = r < quarter ? 1 : r < 2 * quarter ? 2 : r < 3 * quarter ? 3 : 0;
  } else {
    int apply;
    if (p == 1) {
      apply = 1;
    } else {
      double r;
      int rc = draw(u, n_u, cursor, &r);
      if (rc) return rc;
      apply = r < p;
    }
    pauli = !apply ? 0 : kind == QMO_BITFLIP ? 1 : 3;
  }
  if (pauli == 0) return QMO_OK;
  for (int64_t pair = 0; pair < len / 2; pair++) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low;
    int64_t second = first | bit;
    cx zero = s[first], one = s[second];
    if (pauli == 3) {
      s[second] = mk(-one.re, -one.im);
    } else if (pauli == 1) {
      s[first] = one;
      s[second] = zero;
    } else {
      s[first] = mk(one.im, -one.re);
      s[second] = mk(-zero.im, zero.re);
    }
  }
  return QMO_OK;
}

/* NoisyCircuit.Execute (Noise/NoisyCircuit.cs:48-65): p == 0 skips the channel. */
int qmo_noisy_execute(int qubits, const qmo_step *steps, int64_t n_steps, const qmo_op *ops,
                      double *amps, const double *u, int64_t n_u, int64_t *cursor) {
  for (int64_t k = 0; k < n_steps; k++) {
    const qmo_step *st = &steps[k];
    int rc;
    if (!st->is_channel) {
      rc = qmo_execute(qubits, ops + st->op_begin, st->op_count, amps);
    } else if (st->probability != 0) {
      rc = qmo_noise_apply(st->noise_kind, st->probability, st->target, qubits, amps, u, n_u,
                           cursor);
    } else {
      rc = QMO_OK;
    }
    if (rc) return rc;
  }
  return QMO_OK;
}

/* ---------------------------------------------------------------------------
 * Observations: StateVector.cs and StateVector.Observations.cs.
 * ------------------------------------------------------------------------- */
double qmo_mass(int qubits, const double *amps) {
  const cx *s = (const cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  double total = 0;
  for (int64_t i = 0; i < len; i++) total += prob(s, i);
  return total;
}

static int valid_mass(double m) { return is_finite(m) && m > 0; }

/* StateVector.MeasureAll (StateVector.cs:106-153). */
int qmo_measure_all(int qubits, double *amps, double u, int64_t *outcome) {
  cx *s = (cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  double total = 0;
  int64_t last = -1;
  for (int64_t i = 0; i < len; i++) {
    double p = prob(s, i);
    total += p;
    if (p > 0) last = i;
  }
  if (!valid_mass(total)) return QMO_ERR_MASS;
  double d = u * total;
  double cum = 0;
  int64_t out = last;
  for (int64_t i = 0; i < len; i++) {
    cum += prob(s, i);
    if (d < cum) {
      out = i;
      break;
    }
  }
  memset(s, 0, sizeof(cx) * (size_t)len);
  s[out] = mk(1, 0);
  *outcome = out;
  return QMO_OK;
}

/* ValidateSelection (StateVector.Observations.cs:143-168). */
static int selection_mask(int qubits, const int32_t *sel, int k, int64_t *mask) {
  if (k < 1 || k > qubits) return QMO_ERR_ARG;
  int64_t m = 0;
  for (int i = 0; i < k; i++) {
    if (sel[i] < 0 || sel[i] >= qubits) return QMO_ERR_RANGE;
    int64_t b = (int64_t)1 << sel[i];
    if (m & b) return QMO_ERR_ARG;
    m |= b;
  }
  *mask = m;
  return QMO_OK;
}

/* StateVector.Measure (StateVector.Observations.cs:45-102). */
int qmo_measure(int qubits, double *amps, const int32_t *sel, int k, double u, int64_t *outcome) {
  int64_t mask;
  int rc = selection_mask(qubits, sel, k, &mask);
  if (rc) return rc;
  cx *s = (cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  double total = 0;
  int64_t last = -1;
  for (int64_t i = 0; i < len; i++) {
    double p = prob(s, i);
    total += p;
    if (p > 0) last = i;
  }
  if (!valid_mass(total)) return QMO_ERR_MASS;
  double d = u * total;
  double cum = 0;
  int64_t basis = last;
  for (int64_t i = 0; i < len; i++) {
    cum += prob(s, i);
    if (d < cum) {
      basis = i;
      break;
    }
  }
  int64_t pattern = basis & mask;
  double branch = 0;
  for (int64_t i = 0; i < len; i++)
    if ((i & mask) == pattern) branch += prob(s, i);
  if (!valid_mass(branch)) return QMO_ERR_MASS;
  double norm = sqrt(branch);
  for (int64_t i = 0; i < len; i++) {
    cx v = s[i];
    s[i] = (i & mask) == pattern ? mk(v.re / norm, v.im / norm) : mk(0, 0);
  }
  int64_t out = 0;
  for (int b = 0; b < k; b++) out |= ((basis >> sel[b]) & 1) << b;
  *outcome = out;
  return QMO_OK;
}

/* StateVector.MarginalProbability (StateVector.Observations.cs:10-36). */
int qmo_marginal(int qubits, const double *amps, const int32_t *sel, int k, int64_t outcome,
                 double *probability) {
  int64_t mask;
  int rc = selection_mask(qubits, sel, k, &mask);
  if (rc) return rc;
  if (outcome < 0 || outcome >= ((int64_t)1 << k)) return QMO_ERR_RANGE;
  int64_t pattern = 0;
  for (int b = 0; b < k; b++) pattern |= ((outcome >> b) & 1) << sel[b];
  const cx *s = (const cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  double total = 0, selected = 0;
  for (int64_t i = 0; i < len; i++) {
    double p = prob(s, i);
    total += p;
    if ((i & mask) == pattern) selected += p;
  }
  if (!valid_mass(total)) return QMO_ERR_MASS;
  *probability = selected / total;
  return QMO_OK;
}

/* StateVector.ExpectationValue (StateVector.Observations.cs:107-141). */
int qmo_expectation(int qubits, const double *amps, const qmo_pauli *p, double *value) {
  const cx *s = (const cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  double total = 0, e = 0;
  for (int64_t i = 0; i < len; i++) {
    total += prob(s, i);
    cx partner = s[i ^ (int64_t)p->flip_mask];
    cx conj = mk(partner.re, -partner.im);
    cx overlap = cmul(conj, s[i]);
    double contribution;
    switch (p->y_count & 3) {
      case 0: contribution = overlap.re; break;
      case 1: contribution = -overlap.im; break;
      case 2: contribution = -overlap.re; break;
      default: contribution = overlap.im; break;
    }
    e += (__builtin_popcountll((uint64_t)i & p->sign_mask) & 1) == 0 ? contribution : -contribution;
  }
  if (!valid_mass(total)) return QMO_ERR_MASS;
  if (!is_finite(e)) return QMO_ERR_NONFINITE;
  *value = e / total;
  return QMO_OK;
}

/* StateVector.ExpectationEnergy (StateVector.Energy.cs:10-46): Neumaier sum in
 * term order, zero coefficients skipped. */
int qmo_energy(int qubits, const double *amps, const qmo_pauli *terms, const double *coeffs,
               int64_t n_terms, double *energy) {
  double mass = qmo_mass(qubits, amps);
  if (!valid_mass(mass)) return QMO_ERR_MASS;
  double sum = 0, corr = 0;
  for (int64_t t = 0; t < n_terms; t++) {
    if (coeffs[t] == 0) continue;
    double ev;
    int rc = qmo_expectation(qubits, amps, &terms[t], &ev);
    if (rc) return rc;
    double value = coeffs[t] * ev;
    double next = sum + value;
    corr += fabs(sum) >= fabs(value) ? (sum - next) + value : (value - next) + sum;
    sum = next;
    if (!is_finite(sum) || !is_finite(corr)) return QMO_ERR_NONFINITE;
  }
  double e = sum + corr;
  if (!is_finite(e)) return QMO_ERR_NONFINITE;
  *energy = e;
  return QMO_OK;
}

/* StateVectorSimulator.Run sampling (StateVectorSimulator.cs:44-90). */
int qmo_sample(int qubits, const double *amps, const double *u, int64_t shots, int64_t *outcomes) {
  const cx *s = (const cx *)amps;
  int64_t len = (int64_t)1 << qubits;
  double *cum = (double *)malloc(sizeof(double) * (size_t)len);
  if (!cum) return QMO_ERR_ARG;
  double total = 0;
  int64_t last = -1;
  for (int64_t i = 0; i < len; i++) {
    double p = prob(s, i);
    total += p;
    cum[i] = total;
    if (p > 0) last = i;
  }
  if (!valid_mass(total)) {
    free(cum);
    return QMO_ERR_MASS;
  }
  for (int64_t shot = 0; shot < shots; shot++) {
    double d = u[shot] * total;
    int64_t lo = 0, hi = len;
    while (lo < hi) {
      int64_t mid = lo + (hi - lo) / 2;
      if (cum[mid] <= d)
        lo = mid + 1;
      else
        hi = mid;
    }
    outcomes[shot] = lo < len ? lo : last;
  }
  free(cum);
  return QMO_OK;
}

/* ---------------------------------------------------------------------------
 * Trajectories: NoisyStateVectorSimulator (src/QuantumMatter.Application/Noise/).
 * ------------------------------------------------------------------------- */
static void reset_state(cx *s, int64_t len) {
  memset(s, 0, sizeof(cx) * (size_t)len);
  s[0] = mk(1, 0);
}

int qmo_trajectories(int qubits, const qmo_step *steps, int64_t n_steps, const qmo_op *ops,
                     int64_t trajectories, const double *pools, int64_t stride, int64_t *consumed,
                     int64_t *outcomes, int64_t shots, const double *shot_u, int64_t *samples) {
  int64_t len = (int64_t)1 << qubits;
  cx *s = (cx *)malloc(sizeof(cx) * (size_t)len);
  if (!s) return QMO_ERR_ARG;
  int rc = QMO_OK;
  for (int64_t t = 0; t < trajectories && rc == QMO_OK; t++) {
    const double *pool = pools + t * stride;
    int64_t cursor = 0;
    reset_state(s, len);
    rc = qmo_noisy_execute(qubits, steps, n_steps, ops, (double *)s, pool, stride, &cursor);
    if (rc) break;
    if (shots > 0 && samples)
      rc = qmo_sample(qubits, (const double *)s, shot_u + t * shots, shots, samples + t * shots);
    if (rc) break;
    if (outcomes) {
      double r;
      rc = draw(pool, stride, &cursor, &r);
      if (rc) break;
      rc = qmo_measure_all(qubits, (double *)s, r, &outcomes[t]);
    }
    if (consumed) consumed[t] = cursor;
  }
  free(s);
  return rc;
}

/* EstimateProbabilities (:64-106) and TotalMass (:125-148). */
int qmo_estimate_probabilities(int qubits, const qmo_step *steps, int64_t n_steps,
                               const qmo_op *ops, int64_t trajectories, const double *pools,
                               int64_t stride, double *means, double *var) {
  if (trajectories < 2) return QMO_ERR_RANGE;
  int64_t len = (int64_t)1 << qubits;
  cx *s = (cx *)malloc(sizeof(cx) * (size_t)len);
  if (!s) return QMO_ERR_ARG;
  memset(means, 0, sizeof(double) * (size_t)len);
  memset(var, 0, sizeof(double) * (size_t)len);
  int rc = QMO_OK;
  for (int64_t t = 0; t < trajectories; t++) {
    int64_t cursor = 0;
    reset_state(s, len);
    rc = qmo_noisy_execute(qubits, steps, n_steps, ops, (double *)s, pools + t * stride, stride,
                           &cursor);
    if (rc) break;
    double total = 0, comp = 0;
    for (int64_t i = 0; i < len; i++) {
      double corrected = prob(s, i) - comp;
      double next = total + corrected;
      comp = (next - total) - corrected;
      total = next;
    }
    if (!valid_mass(total)) {
      rc = QMO_ERR_MASS;
      break;
    }
    for (int64_t i = 0; i < len; i++) {
      double p = prob(s, i) / total;
      double delta = p - means[i];
      means[i] += delta / (double)(t + 1);
      var[i] += delta * (p - means[i]);
    }
  }
  if (rc == QMO_OK) {
    double divisor = (double)trajectories * (double)(trajectories - 1);
    for (int64_t i = 0; i < len; i++) var[i] = sqrt(fmax(0, var[i]) / divisor);
  }
  free(s);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Variational: ParameterizedCircuit.BindCore and VariationalEnergyEvaluator.
 * ------------------------------------------------------------------------- */
int qmo_bind(const qmo_op *ops, const int32_t *param_index, int64_t n_ops,
             const double *parameters, int64_t n_params, int64_t shifted, double shift,
             qmo_op *out, int64_t *n_out) {
  for (int64_t i = 0; i < n_params; i++)
    if (!is_finite(parameters[i])) return QMO_ERR_RANGE;
  int64_t w = 0, rotation = 0;
  for (int64_t i = 0; i < n_ops; i++) {
    qmo_op op = ops[i];
    if (param_index[i] >= 0) {
      if (param_index[i] >= n_params) return QMO_ERR_RANGE;
      op.angle = parameters[param_index[i]];
      out[w++] = op;
      if (rotation == shifted) {
        op.angle = shift;
        out[w++] = op;
      }
      rotation++;
    } else {
      out[w++] = op;
    }
  }
  *n_out = w;
  return QMO_OK;
}

static int prepare_bound(int qubits, const qmo_op *ops, int64_t n, cx *s, int64_t len) {
  reset_state(s, len);
  return qmo_execute(qubits, ops, n, (double *)s);
}

int qmo_energy_bound(int qubits, const qmo_op *ops, const int32_t *param_index, int64_t n_ops,
                     const double *parameters, int64_t n_params, const qmo_pauli *terms,
                     const double *coeffs, int64_t n_terms, double *energy) {
  int64_t len = (int64_t)1 << qubits;
  qmo_op *bound = (qmo_op *)malloc(sizeof(qmo_op) * (size_t)(n_ops + 1));
  cx *s = (cx *)malloc(sizeof(cx) * (size_t)len);
  int rc = (bound && s) ? QMO_OK : QMO_ERR_ARG;
  int64_t nb = 0;
  if (!rc) rc = qmo_bind(ops, param_index, n_ops, parameters, n_params, -1, 0, bound, &nb);
  if (!rc) rc = prepare_bound(qubits, bound, nb, s, len);
  if (!rc) rc = qmo_energy(qubits, (const double *)s, terms, coeffs, n_terms, energy);
  free(bound);
  free(s);
  return rc;
}

int qmo_energy_gradient(int qubits, const qmo_op *ops, const int32_t *param_index, int64_t n_ops,
                        const double *parameters, int64_t n_params, const qmo_pauli *terms,
                        const double *coeffs, int64_t n_terms, double *energy, double *gradient) {
  int64_t len = (int64_t)1 << qubits;
  qmo_op *bound = (qmo_op *)malloc(sizeof(qmo_op) * (size_t)(n_ops + 1));
  cx *s = (cx *)malloc(sizeof(cx) * (size_t)len);
  double *corr = (double *)calloc((size_t)(n_params > 0 ? n_params : 1), sizeof(double));
  double *plus = (double *)calloc((size_t)(n_terms > 0 ? n_terms : 1), sizeof(double));
  int rc = (bound && s && corr && plus) ? QMO_OK : QMO_ERR_ARG;
  int64_t nb = 0;
  for (int64_t i = 0; i < n_params; i++) gradient[i] = 0;
  if (!rc) rc = qmo_bind(ops, param_index, n_ops, parameters, n_params, -1, 0, bound, &nb);
  if (!rc) rc = prepare_bound(qubits, bound, nb, s, len);
  if (!rc) rc = qmo_energy(qubits, (const double *)s, terms, coeffs, n_terms, energy);
  int64_t rotations = 0;
  for (int64_t i = 0; i < n_ops; i++) rotations += param_index[i] >= 0;
  for (int64_t r = 0; r < rotations && !rc; r++) {
    /* The parameter of occurrence r (RotationParameterIndices[r]). */
    int64_t seen = 0, param = -1;
    for (int64_t i = 0; i < n_ops; i++) {
      if (param_index[i] < 0) continue;
      if (seen++ == r) {
        param = param_index[i];
        break;
      }
    }
    rc = qmo_bind(ops, param_index, n_ops, parameters, n_params, r, M_PI / 2, bound, &nb);
    if (!rc) rc = prepare_bound(qubits, bound, nb, s, len);
    if (!rc && n_terms == 0 && !valid_mass(qmo_mass(qubits, (const double *)s))) rc = QMO_ERR_MASS;
    for (int64_t t = 0; t < n_terms && !rc; t++)
      rc = qmo_expectation(qubits, (const double *)s, &terms[t], &plus[t]);
    if (!rc) rc = qmo_bind(ops, param_index, n_ops, parameters, n_params, r, -M_PI / 2, bound, &nb);
    if (!rc) rc = prepare_bound(qubits, bound, nb, s, len);
    if (!rc && n_terms == 0 && !valid_mass(qmo_mass(qubits, (const double *)s))) rc = QMO_ERR_MASS;
    for (int64_t t = 0; t < n_terms && !rc; t++) {
      double minus;
      rc = qmo_expectation(qubits, (const double *)s, &terms[t], &minus);
      if (rc) break;
      double contribution = coeffs[t] * (0.5 * plus[t] - 0.5 * minus);
      double prev = gradient[param];
      double next = prev + contribution;
      corr[param] += fabs(prev) >= fabs(contribution) ? (prev - next) + contribution
                                                      : (contribution - next) + prev;
      gradient[param] = next;
      if (!is_finite(next) || !is_finite(corr[param])) rc = QMO_ERR_NONFINITE;
    }
  }
  for (int64_t p = 0; p < n_params && !rc; p++) {
    gradient[p] += corr[p];
    if (!is_finite(gradient[p])) rc = QMO_ERR_NONFINITE;
  }
  free(bound);
  free(s);
  free(corr);
  free(plus);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Dynamic programs: Dynamic/DynamicCircuit.cs and Dynamic/NoisyDynamicCircuit.cs.
 * ------------------------------------------------------------------------- */
int qmo_dynamic_execute(int qubits, const qmo_dyn_step *steps, int64_t n_steps, const qmo_op *ops,
                        const int32_t *sel_q, const int32_t *sel_c, double *amps, const double *u,
                        int64_t n_u, int64_t *cursor, int32_t *classical) {
  int32_t reg = 0; /* DynamicCircuit.cs:81: each execution starts at zero */
  for (int64_t k = 0; k < n_steps; k++) {
    const qmo_dyn_step *st = &steps[k];
    int rc = QMO_OK;
    double draw_u;
    int64_t outcome;
    switch (st->kind) {
      case QMO_DYN_SEGMENT: /* :88-90 */
        if ((reg & st->cond_mask) == st->cond_value)
          rc = qmo_execute(qubits, ops + st->begin, st->op_count, amps);
        break;
      case QMO_DYN_MEASURE: { /* :91-100 */
        rc = draw(u, n_u, cursor, &draw_u);
        if (rc) break;
        rc = qmo_measure(qubits, amps, sel_q + st->begin, st->count, draw_u, &outcome);
        if (rc) break;
        int32_t written = 0, write_mask = 0;
        for (int b = 0; b < st->count; b++) {
          written |= (int32_t)((outcome >> b) & 1) << sel_c[st->begin + b];
          write_mask |= (int32_t)1 << sel_c[st->begin + b];
        }
        reg = (reg & ~write_mask) | written;
        break;
      }
      case QMO_DYN_RESET: { /* :101-107 */
        rc = draw(u, n_u, cursor, &draw_u);
        if (rc) break;
        const int32_t q = st->target;
        rc = qmo_measure(qubits, amps, &q, 1, draw_u, &outcome);
        if (rc || outcome == 0) break;
        qmo_op x;
        memset(&x, 0, sizeof x);
        x.kind = QMO_X;
        x.target = q;
        x.control = -1;
        x.second_control = -1;
        rc = qmo_execute(qubits, &x, 1, amps);
        break;
      }
      case QMO_DYN_NOISE: /* NoisyCircuit.cs:58: zero strength is skipped */
        if (st->probability != 0)
          rc = qmo_noise_apply(st->noise_kind, st->probability, st->target, qubits, amps, u, n_u,
                               cursor);
        break;
      default:
        rc = QMO_ERR_ARG;
    }
    if (rc) return rc;
  }
  *classical = reg;
  return QMO_OK;
}
