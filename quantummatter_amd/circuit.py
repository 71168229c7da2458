"""Circuit, observable, noise and template model of QuantumMatter, restated for the
HIP backend's host side.

Names, argument meaning, validation order and error kinds follow the reference
C# types so that code and tests written against them read the same:

* ``GateKind`` / ``GateOperation`` / ``QuantumCircuit``
  -- src/QuantumMatter.Core/QuantumCircuit.cs:4-219
* ``SingleQubitUnitary`` -- src/QuantumMatter.Core/SingleQubitUnitary.cs:6-80
* ``PauliString`` -- src/QuantumMatter.Core/PauliString.cs:4-27
* ``PauliTerm`` / ``PauliHamiltonian`` -- src/QuantumMatter.Core/Observables/*.cs
* ``NoiseKind`` / ``NoiseChannel`` / ``NoisyCircuit``
  -- src/QuantumMatter.Core/Noise/NoiseChannel.cs:4-45, NoisyCircuit.cs:8-66
* ``ParameterizedCircuit`` -- src/QuantumMatter.Core/Variational/ParameterizedCircuit.cs:8-99

Every type can flatten itself to the packed records of ``include/qm_hip.h``
(``qm_op``, ``qm_pauli``, ``qm_step``) that cross the C ABI.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, replace
from typing import Iterable, List, Optional, Sequence

import numpy as np

from .errors import ArgumentError, ArgumentOutOfRangeError

# Packed records; layouts match include/qm_hip.h exactly.
OP_DTYPE = np.dtype(
    [("kind", "<i4"), ("target", "<i4"), ("control", "<i4"), ("second_control", "<i4"),
     ("angle", "<f8"), ("unitary", "<f8", (8,))], align=True)
PAULI_DTYPE = np.dtype(
    [("flip_mask", "<u8"), ("sign_mask", "<u8"), ("y_count", "<i4"), ("pad", "<i4")], align=True)
STEP_DTYPE = np.dtype(
    [("is_channel", "<i4"), ("noise_kind", "<i4"), ("target", "<i4"), ("pad", "<i4"),
     ("probability", "<f8"), ("op_begin", "<i8"), ("op_count", "<i8")], align=True)
DYN_STEP_DTYPE = np.dtype(
    [("kind", "<i4"), ("cond_mask", "<i4"), ("cond_value", "<i4"), ("target", "<i4"),
     ("noise_kind", "<i4"), ("count", "<i4"), ("begin", "<i8"), ("op_count", "<i8"),
     ("probability", "<f8")], align=True)
assert OP_DTYPE.itemsize == 88 and PAULI_DTYPE.itemsize == 24 and STEP_DTYPE.itemsize == 40
assert DYN_STEP_DTYPE.itemsize == 48
DYN_SEGMENT, DYN_MEASURE, DYN_RESET, DYN_NOISE = 0, 1, 2, 3

#: Widest register the C ABI accepts (64-bit indexing; the reference stops at 30).
MAX_QUBITS = 34


class GateKind(enum.IntEnum):
    """Stable operation identifiers (QuantumCircuit.cs:4-40)."""
    H = 0
    X = 1
    Y = 2
    Z = 3
    S = 4
    T = 5
    Rx = 6
    Ry = 7
    Rz = 8
    CNot = 9
    Swap = 10
    I = 11
    Sdg = 12
    Tdg = 13
    Phase = 14
    Unitary = 15
    Toffoli = 16


_BUILTIN_1Q = {GateKind.H, GateKind.X, GateKind.Y, GateKind.Z, GateKind.S, GateKind.T,
               GateKind.Rx, GateKind.Ry, GateKind.Rz, GateKind.I, GateKind.Sdg, GateKind.Tdg,
               GateKind.Phase}


class SingleQubitUnitary:
    """Immutable row-major 2x2 unitary validated at construction (SingleQubitUnitary.cs:41-63)."""

    ValidationTolerance = 1e-12

    def __init__(self, m00: complex, m01: complex, m10: complex, m11: complex):
        for name, v in (("m00", m00), ("m01", m01), ("m10", m10), ("m11", m11)):
            v = complex(v)
            if not (math.isfinite(v.real) and math.isfinite(v.imag)):
                raise ArgumentOutOfRangeError(name, "Matrix entries must be finite.")
        m00, m01, m10, m11 = complex(m00), complex(m01), complex(m10), complex(m11)
        sq = lambda z: z.real * z.real + z.imag * z.imag  # noqa: E731
        if (abs(sq(m00) + sq(m10) - 1) > self.ValidationTolerance
                or abs(sq(m01) + sq(m11) - 1) > self.ValidationTolerance):
            raise ArgumentError("Matrix columns must have unit norm.")
        if abs(m00.conjugate() * m01 + m10.conjugate() * m11) > self.ValidationTolerance:
            raise ArgumentError("Matrix columns must be orthogonal.")
        self.M00, self.M01, self.M10, self.M11 = m00, m01, m10, m11

    @classmethod
    def Create(cls, m00, m01=None, m10=None, m11=None) -> "SingleQubitUnitary":
        if m01 is None:
            m = np.asarray(m00, dtype=complex)
            if m.shape != (2, 2):
                raise ArgumentError("A single-qubit matrix must have exactly two rows and two columns.")
            return cls(m[0, 0], m[0, 1], m[1, 0], m[1, 1])
        return cls(m00, m01, m10, m11)

    def Adjoint(self) -> "SingleQubitUnitary":
        return SingleQubitUnitary(self.M00.conjugate(), self.M10.conjugate(),
                                  self.M01.conjugate(), self.M11.conjugate())

    def packed(self) -> List[float]:
        out: List[float] = []
        for z in (self.M00, self.M01, self.M10, self.M11):
            out += [z.real, z.imag]
        return out


@dataclass(frozen=True)
class GateOperation:
    """Immutable operation record (QuantumCircuit.cs:49-61)."""
    Kind: GateKind
    Target: int
    Control: int = -1
    Angle: float = 0.0
    SecondControl: int = -1
    Unitary: Optional[SingleQubitUnitary] = None


class QuantumCircuit:
    """Mutable validated circuit; qubit 0 is the least significant index bit.

    Widths run from 1 to ``MAX_QUBITS`` (the reference caps at 30 because it
    indexes with 32-bit integers; the HIP backend indexes with 64 bits).
    """

    def __init__(self, qubits: int):
        if not isinstance(qubits, (int, np.integer)) or qubits < 1 or qubits > MAX_QUBITS:
            raise ArgumentOutOfRangeError("qubits", f"Number of qubits must be between 1 and {MAX_QUBITS}.")
        self.QubitCount = int(qubits)
        self._ops: List[GateOperation] = []
        self._packed = None  # (qm_op records, operation count they were formed at)

    @property
    def Operations(self) -> Sequence[GateOperation]:
        return tuple(self._ops)

    def Copy(self) -> "QuantumCircuit":
        c = QuantumCircuit(self.QubitCount)
        c._ops = list(self._ops)
        return c

    def _append_validated(self, op: GateOperation) -> None:
        self._ops.append(op)

    def _validate(self, q: int) -> None:
        if not isinstance(q, (int, np.integer)) or q < 0 or q >= self.QubitCount:
            raise ArgumentOutOfRangeError("qubit")

    def _add(self, kind, target, control=None, angle=0.0, second=None, unitary=None):
        # Validation order of QuantumCircuit.Add (QuantumCircuit.cs:97-126).
        self._validate(target)
        if control is not None:
            self._validate(control)
            if target == control:
                raise ArgumentError("The two qubits must differ.")
        if second is not None:
            self._validate(second)
            if second == target or second == control:
                raise ArgumentError("All three qubits must differ.")
        if not math.isfinite(angle):
            raise ArgumentOutOfRangeError("angle")
        self._ops.append(GateOperation(GateKind(kind), int(target),
                                       -1 if control is None else int(control), float(angle),
                                       -1 if second is None else int(second), unitary))
        return self

    def H(self, q): return self._add(GateKind.H, q)
    def X(self, q): return self._add(GateKind.X, q)
    def Y(self, q): return self._add(GateKind.Y, q)
    def Z(self, q): return self._add(GateKind.Z, q)
    def S(self, q): return self._add(GateKind.S, q)
    def T(self, q): return self._add(GateKind.T, q)
    def Rx(self, q, angle): return self._add(GateKind.Rx, q, angle=angle)
    def Ry(self, q, angle): return self._add(GateKind.Ry, q, angle=angle)
    def Rz(self, q, angle): return self._add(GateKind.Rz, q, angle=angle)
    def CNot(self, control, target): return self._add(GateKind.CNot, target, control)
    def Swap(self, a, b): return self._add(GateKind.Swap, b, a)
    def I(self, q): return self._add(GateKind.I, q)
    def Sdg(self, q): return self._add(GateKind.Sdg, q)
    def Tdg(self, q): return self._add(GateKind.Tdg, q)
    def Phase(self, q, angle): return self._add(GateKind.Phase, q, angle=angle)

    def Controlled(self, kind, control, target, angle=0.0):
        if GateKind(kind) not in _BUILTIN_1Q:
            raise ArgumentOutOfRangeError("kind", "A built-in single-qubit gate is required.")
        return self._add(kind, target, control, angle)

    def CH(self, c, t): return self.Controlled(GateKind.H, c, t)
    def CY(self, c, t): return self.Controlled(GateKind.Y, c, t)
    def CZ(self, c, t): return self.Controlled(GateKind.Z, c, t)
    def CS(self, c, t): return self.Controlled(GateKind.S, c, t)
    def CT(self, c, t): return self.Controlled(GateKind.T, c, t)
    def CSdg(self, c, t): return self.Controlled(GateKind.Sdg, c, t)
    def CTdg(self, c, t): return self.Controlled(GateKind.Tdg, c, t)
    def CRx(self, c, t, a): return self.Controlled(GateKind.Rx, c, t, a)
    def CRy(self, c, t, a): return self.Controlled(GateKind.Ry, c, t, a)
    def CRz(self, c, t, a): return self.Controlled(GateKind.Rz, c, t, a)
    def CPhase(self, c, t, a): return self.Controlled(GateKind.Phase, c, t, a)

    def U(self, target, unitary: SingleQubitUnitary):
        if unitary is None:
            raise ArgumentError("unitary is null.")
        return self._add(GateKind.Unitary, target, unitary=unitary)

    def ControlledU(self, control, target, unitary: SingleQubitUnitary):
        if unitary is None:
            raise ArgumentError("unitary is null.")
        return self._add(GateKind.Unitary, target, control, unitary=unitary)

    def Toffoli(self, first, second, target):
        return self._add(GateKind.Toffoli, target, first, second=second)

    def to_ops(self) -> np.ndarray:
        """The operations in qm_op layout.  Operations are only ever appended, so
        the records are formed once per operation count (callers must not
        modify the array)."""
        if self._packed is None or self._packed[1] != len(self._ops):
            self._packed = (ops_array(self._ops), len(self._ops))
        return self._packed[0]


_NO_UNITARY = (0.0,) * 8


def ops_array(ops: Iterable[GateOperation]) -> np.ndarray:
    """The operations in qm_op layout (one record per operation, built in one pass)."""
    return np.array([(int(op.Kind), op.Target, op.Control, op.SecondControl, op.Angle,
                      _NO_UNITARY if op.Unitary is None else tuple(op.Unitary.packed()))
                     for op in ops], dtype=OP_DTYPE)


class PauliString:
    """Tensor product of I/X/Y/Z with qubit 0 at the right (PauliString.cs:7-26)."""

    def __init__(self, operators: str):
        if operators is None:
            raise ArgumentError("operators is null.")
        if len(operators) < 1 or len(operators) > MAX_QUBITS:
            raise ArgumentOutOfRangeError("operators", "A Pauli string requires 1 through 30 labels.")
        flip = sign = y = 0
        for q in range(len(operators)):
            label = operators[len(operators) - 1 - q]
            if label not in "IXYZ":
                raise ArgumentError("Only uppercase I, X, Y and Z labels are supported.")
            if label in "XY":
                flip |= 1 << q
            if label in "ZY":
                sign |= 1 << q
            if label == "Y":
                y += 1
        self.Operators = operators
        self.FlipMask, self.SignMask, self.YCount = flip, sign, y

    @property
    def QubitCount(self) -> int:
        return len(self.Operators)


@dataclass(frozen=True)
class PauliTerm:
    """Finite real coefficient times a Pauli string (Observables/PauliTerm.cs:7-13)."""
    Coefficient: float
    Observable: PauliString

    def __post_init__(self):
        if self.Observable is None:
            raise ArgumentError("observable is null.")
        if not math.isfinite(self.Coefficient):
            raise ArgumentOutOfRangeError("coefficient")


class PauliHamiltonian:
    """Ordered real sum of register-sized Pauli terms (Observables/PauliHamiltonian.cs:8-26)."""

    def __init__(self, qubits: int, terms: Sequence[PauliTerm]):
        if qubits < 1 or qubits > MAX_QUBITS:
            raise ArgumentOutOfRangeError("qubits")
        if terms is None:
            raise ArgumentError("terms is null.")
        for t in terms:
            if t is None or t.Observable.QubitCount != qubits:
                raise ArgumentError("Every term must be nonnull and match the register size.")
        self.QubitCount = qubits
        self.Terms = tuple(terms)
        self._packed = None

    def packed(self):
        """(qm_pauli records, coefficients) of the terms; the terms are fixed, so
        the arrays are formed once (callers must not modify them)."""
        if self._packed is None:
            self._packed = (pauli_array([t.Observable for t in self.Terms]),
                            np.array([t.Coefficient for t in self.Terms], dtype=np.float64))
        return self._packed


def pauli_array(strings: Sequence[PauliString]) -> np.ndarray:
    return np.array([(p.FlipMask, p.SignMask, p.YCount, 0) for p in strings], dtype=PAULI_DTYPE)


class NoiseKind(enum.IntEnum):
    """Stable channel identifiers (Noise/NoiseChannel.cs:4-13)."""
    BitFlip = 0
    PhaseFlip = 1
    Depolarizing = 2
    AmplitudeDamping = 3


@dataclass(frozen=True)
class NoiseChannel:
    """Single-qubit channel with a parameter in [0,1] (Noise/NoiseChannel.cs:16-45)."""
    Kind: NoiseKind
    Probability: float

    def __post_init__(self):
        p = self.Probability
        if not math.isfinite(p) or p < 0 or p > 1:
            raise ArgumentOutOfRangeError("probability", "Channel probability must be finite and in [0,1].")

    @staticmethod
    def BitFlip(p): return NoiseChannel(NoiseKind.BitFlip, float(p))
    @staticmethod
    def PhaseFlip(p): return NoiseChannel(NoiseKind.PhaseFlip, float(p))
    @staticmethod
    def Depolarizing(p): return NoiseChannel(NoiseKind.Depolarizing, float(p))
    @staticmethod
    def AmplitudeDamping(g): return NoiseChannel(NoiseKind.AmplitudeDamping, float(g))


class NoisyCircuit:
    """Ordered unitary segments and channel placements (Noise/NoisyCircuit.cs:8-66)."""

    def __init__(self, qubits: int):
        if qubits < 1 or qubits > MAX_QUBITS:
            raise ArgumentOutOfRangeError("qubits")
        self.QubitCount = qubits
        self._steps: list = []

    @property
    def StepCount(self) -> int:
        return len(self._steps)

    def Append(self, circuit: QuantumCircuit) -> "NoisyCircuit":
        if circuit is None:
            raise ArgumentError("circuit is null.")
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit sizes must match.")
        self._steps.append(("segment", circuit.Operations))
        return self

    def ApplyNoise(self, channel: NoiseChannel, target: int) -> "NoisyCircuit":
        if channel is None:
            raise ArgumentError("channel is null.")
        if target < 0 or target >= self.QubitCount:
            raise ArgumentOutOfRangeError("target")
        self._steps.append(("channel", channel, int(target)))
        return self

    def packed(self):
        """Returns (steps, ops) arrays in qm_step / qm_op layout."""
        steps = np.zeros(len(self._steps), dtype=STEP_DTYPE)
        all_ops: List[GateOperation] = []
        for i, st in enumerate(self._steps):
            if st[0] == "segment":
                steps[i]["is_channel"] = 0
                steps[i]["op_begin"] = len(all_ops)
                steps[i]["op_count"] = len(st[1])
                all_ops.extend(st[1])
            else:
                steps[i]["is_channel"] = 1
                steps[i]["noise_kind"] = int(st[1].Kind)
                steps[i]["probability"] = st[1].Probability
                steps[i]["target"] = st[2]
        return steps, ops_array(all_ops)


class ParameterizedCircuit:
    """Template of constant gates and parameter-indexed Rx/Ry/Rz rotations
    (Variational/ParameterizedCircuit.cs:8-99)."""

    def __init__(self, qubits: int, parameter_count: int):
        if qubits < 1 or qubits > MAX_QUBITS:
            raise ArgumentOutOfRangeError("qubits")
        if parameter_count < 0:
            raise ArgumentOutOfRangeError("parameterCount")
        self.QubitCount = qubits
        self.ParameterCount = parameter_count
        self._ops: List[GateOperation] = []
        self._param_index: List[int] = []
        self._rotations: List[int] = []
        self._packed = None  # (ops, param_index, operation count they were formed at)

    @property
    def OperationCount(self) -> int:
        return len(self._ops)

    @property
    def RotationParameterIndices(self) -> Sequence[int]:
        return tuple(self._rotations)

    def Append(self, circuit: QuantumCircuit) -> "ParameterizedCircuit":
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit sizes must match.")
        for op in circuit.Operations:
            self._ops.append(op)
            self._param_index.append(-1)
        return self

    def _rotation(self, kind, target, p):
        if target < 0 or target >= self.QubitCount:
            raise ArgumentOutOfRangeError("target")
        if p < 0 or p >= self.ParameterCount:
            raise ArgumentOutOfRangeError("parameterIndex")
        self._ops.append(GateOperation(kind, int(target)))
        self._param_index.append(int(p))
        self._rotations.append(int(p))
        return self

    def Rx(self, target, p): return self._rotation(GateKind.Rx, target, p)
    def Ry(self, target, p): return self._rotation(GateKind.Ry, target, p)
    def Rz(self, target, p): return self._rotation(GateKind.Rz, target, p)

    def Bind(self, parameters: Sequence[float]) -> QuantumCircuit:
        return self._bind(parameters, -1, 0.0)

    def BindWithRotationShift(self, parameters, rotation_index: int, shift: float) -> QuantumCircuit:
        if rotation_index < 0 or rotation_index >= len(self._rotations):
            raise ArgumentOutOfRangeError("rotationIndex")
        if not math.isfinite(shift):
            raise ArgumentOutOfRangeError("shift")
        return self._bind(parameters, rotation_index, shift)

    def _bind(self, parameters, shifted, shift) -> QuantumCircuit:
        if len(parameters) != self.ParameterCount:
            raise ArgumentError("Binding size must match ParameterCount.")
        for v in parameters:
            if not math.isfinite(v):
                raise ArgumentOutOfRangeError("parameters", "Parameters must be finite.")
        out = QuantumCircuit(self.QubitCount)
        rotation = 0
        for op, p in zip(self._ops, self._param_index):
            if p >= 0:
                out._append_validated(replace(op, Angle=float(parameters[p])))
                if rotation == shifted:
                    out._append_validated(replace(op, Angle=float(shift)))
                rotation += 1
            else:
                out._append_validated(op)
        return out

    def packed(self):
        """Returns (template ops, parameter index per op) in qm_op / int32 layout,
        formed once per template state (callers must not modify them)."""
        if self._packed is None or self._packed[2] != len(self._ops):
            self._packed = (ops_array(self._ops), np.asarray(self._param_index, dtype=np.int32), len(self._ops))
        return self._packed[0], self._packed[1]


class DynamicCircuit:
    """Unitary snapshots, mid-circuit measurements, classical conditions and resets
    (Dynamic/DynamicCircuit.cs:4-137): builder validation as the reference's, each
    instruction copied when appended."""

    def __init__(self, qubits: int, classical_bits: int):
        if qubits < 1 or qubits > MAX_QUBITS:
            raise ArgumentOutOfRangeError("qubits")
        if classical_bits < 1 or classical_bits > 30:
            raise ArgumentOutOfRangeError("classicalBits")
        self.QubitCount = qubits
        self.ClassicalBitCount = classical_bits
        self._steps: list = []

    @property
    def StepCount(self) -> int:
        return len(self._steps)

    def _circuit(self, circuit: QuantumCircuit):
        if circuit is None:
            raise ArgumentError("circuit is null.")
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit sizes must match.")
        return tuple(circuit.Operations)

    def Append(self, circuit: QuantumCircuit) -> "DynamicCircuit":
        self._steps.append(("segment", self._circuit(circuit), 0, 0))
        return self

    def If(self, classical_mask: int, expected_value: int, circuit: QuantumCircuit) -> "DynamicCircuit":
        ops = self._circuit(circuit)
        register = (1 << self.ClassicalBitCount) - 1
        if classical_mask <= 0 or classical_mask & ~register:
            raise ArgumentOutOfRangeError("classicalMask")
        if expected_value < 0 or expected_value & ~classical_mask:
            raise ArgumentOutOfRangeError("expectedValue")
        self._steps.append(("segment", ops, int(classical_mask), int(expected_value)))
        return self

    @staticmethod
    def _indices(indices, count, name) -> None:
        seen = set()
        for i in indices:
            if i < 0 or i >= count:
                raise ArgumentOutOfRangeError(name)
            if i in seen:
                raise ArgumentError("Indices must be distinct.")
            seen.add(i)

    def Measure(self, qubits: Sequence[int], classical_bits: Sequence[int]) -> "DynamicCircuit":
        qubits, classical_bits = list(qubits), list(classical_bits)
        if not qubits or len(qubits) != len(classical_bits) or len(qubits) > self.QubitCount:
            raise ArgumentError("Measurement requires equally sized nonempty selections.")
        self._indices(qubits, self.QubitCount, "qubits")
        self._indices(classical_bits, self.ClassicalBitCount, "classicalBits")
        self._steps.append(("measure", tuple(qubits), tuple(classical_bits)))
        return self

    def Reset(self, qubit: int) -> "DynamicCircuit":
        if qubit < 0 or qubit >= self.QubitCount:
            raise ArgumentOutOfRangeError("qubit")
        self._steps.append(("reset", int(qubit)))
        return self

    def _noise(self, channel: "NoiseChannel", target: int) -> None:
        if channel is None:
            raise ArgumentError("channel is null.")
        if target < 0 or target >= self.QubitCount:
            raise ArgumentOutOfRangeError("target")
        self._steps.append(("noise", channel, int(target)))

    def packed(self):
        """(steps, ops, sel_qubits, sel_cbits) in the qm_dyn_step / qm_op layouts."""
        steps = np.zeros(len(self._steps), dtype=DYN_STEP_DTYPE)
        ops: List[GateOperation] = []
        sel_q: List[int] = []
        sel_c: List[int] = []
        for i, st in enumerate(self._steps):
            rec = steps[i]
            if st[0] == "segment":
                rec["kind"] = DYN_SEGMENT
                rec["begin"], rec["op_count"] = len(ops), len(st[1])
                rec["cond_mask"], rec["cond_value"] = st[2], st[3]
                ops.extend(st[1])
            elif st[0] == "measure":
                rec["kind"] = DYN_MEASURE
                rec["begin"], rec["count"] = len(sel_q), len(st[1])
                sel_q.extend(st[1])
                sel_c.extend(st[2])
            elif st[0] == "reset":
                rec["kind"] = DYN_RESET
                rec["target"] = st[1]
            else:
                rec["kind"] = DYN_NOISE
                rec["noise_kind"] = int(st[1].Kind)
                rec["probability"] = st[1].Probability
                rec["target"] = st[2]
        return (steps, ops_array(ops), np.array(sel_q, dtype=np.int32), np.array(sel_c, dtype=np.int32))


class NoisyDynamicCircuit(DynamicCircuit):
    """A dynamic program with explicit single-qubit channel placements between
    instructions (Dynamic/NoisyDynamicCircuit.cs:7-99).  A placement is
    unconditional and leaves classical memory alone; zero strength consumes no
    draw."""

    def ApplyNoise(self, channel: "NoiseChannel", target: int) -> "NoisyDynamicCircuit":
        self._noise(channel, target)
        return self
