"""Batched device entry points: noisy trajectories and variational energies.

``HipTrajectoryRunner`` runs many trajectories of one NoisyCircuit
(NoisyStateVectorSimulator.Run, src/QuantumMatter.Application/Noise/
NoisyStateVectorSimulator.cs:36-62) in one launch of ``qm_trajectories_run``;
``HipBatchedEnergyEvaluator`` evaluates VariationalEnergyEvaluator.Evaluate /
EvaluateWithGradient (src/QuantumMatter.Application/Variational/
VariationalEnergyEvaluator.cs:23-86) with all 1 + 2R circuits in one launch
of ``qm_energy_gradient_batch`` (or many parameter vectors with
``qm_energy_batch``); ``HipVqeProgram`` replays the same evaluations from
captured hipGraphs for a VQE loop.  Each state is owned by one workgroup; results match the
per-state session path (and the reference's CPU path) under the same
host-supplied uniforms.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import QmProfile, call, ptr
from .backend import C128, HipExecutionProfile, HipStateVectorBackend
from .circuit import NoisyCircuit, ParameterizedCircuit, PauliHamiltonian
from .errors import ArgumentError, ObjectDisposedError
from .variational import VariationalEnergyResult


def batch_min_qubits(precision: int = C128) -> int:
    """Smallest register width the batched kernels take (narrower states are
    not batched: use sessions)."""
    q = C.c_int32()
    call("qm_batch_min_qubits", precision, C.byref(q))
    return q.value


def _profile(prof: Optional[QmProfile]) -> Optional[HipExecutionProfile]:
    if prof is None:
        return None
    return HipExecutionProfile(prof.device_ms, prof.state_bytes, prof.kernel_launches, prof.passes,
                               prof.algorithmic_bytes)


@dataclass
class TrajectoryResult:
    samples: Optional[np.ndarray]    # [trajectories, shots] terminal samples
    outcomes: Optional[np.ndarray]   # [trajectories] MeasureAll outcomes
    consumed: np.ndarray             # [trajectories] draws taken from each pool
    boundary_fallbacks: int          # draws re-decided in the sequential order
    profile: Optional[HipExecutionProfile]


class HipTrajectoryRunner:
    """Noisy trajectories on the device, one workgroup per trajectory."""

    def __init__(self, backend: HipStateVectorBackend, precision: int = C128):
        if backend is None:
            raise ArgumentError("backend is null.")
        self._backend = backend
        self._precision = precision

    def Run(self, circuit: NoisyCircuit, pools: np.ndarray, shots: int = 0,
            shot_uniforms: Optional[np.ndarray] = None, measure_all: bool = True,
            profile: bool = False) -> TrajectoryResult:
        """Trajectory t starts from |0...0> and draws from pools[t] in step order;
        then `shots` samples from shot_uniforms[t] and one MeasureAll draw."""
        if circuit is None:
            raise ArgumentError("circuit is null.")
        pools = np.ascontiguousarray(pools, dtype=np.float64)
        if pools.ndim != 2:
            raise ArgumentError("pools must be [trajectories, stride].")
        T, stride = pools.shape
        steps, ops = circuit.packed()
        samples = np.zeros((T, shots), dtype=np.int64) if shots else None
        su = None
        if shots:
            su = np.ascontiguousarray(shot_uniforms, dtype=np.float64)
            if su.shape != (T, shots):
                raise ArgumentError("shot_uniforms must be [trajectories, shots].")
        outcomes = np.zeros(T, dtype=np.int64) if measure_all else None
        consumed = np.zeros(T, dtype=np.int64)
        fb = C.c_int64()
        prof = QmProfile() if profile else None
        call("qm_trajectories_run", self._backend._handle(), circuit.QubitCount, self._precision,
             ptr(steps), len(steps), ptr(ops), len(ops), T, ptr(pools), stride, shots, ptr(su),
             1 if measure_all else 0, ptr(samples), ptr(outcomes), ptr(consumed), C.byref(fb),
             None if prof is None else C.byref(prof))
        return TrajectoryResult(samples, outcomes, consumed, fb.value, _profile(prof))

    def EstimateProbabilities(self, circuit: NoisyCircuit, pools: np.ndarray,
                              profile: bool = False) -> ProbabilityEstimate:
        """EstimateProbabilities (NoisyStateVectorSimulator.cs:64-106): trajectory t
        draws from pools[t]; means and standard errors of the Born probabilities."""
        if circuit is None:
            raise ArgumentError("circuit is null.")
        pools = np.ascontiguousarray(pools, dtype=np.float64)
        if pools.ndim != 2:
            raise ArgumentError("pools must be [trajectories, stride].")
        T, stride = pools.shape
        steps, ops = circuit.packed()
        n = circuit.QubitCount
        means = np.zeros(1 << n, dtype=np.float64)
        se = np.zeros(1 << n, dtype=np.float64)
        consumed = np.zeros(T, dtype=np.int64)
        fb = C.c_int64()
        prof = QmProfile() if profile else None
        call("qm_trajectories_estimate", self._backend._handle(), n, self._precision,
             ptr(steps), len(steps), ptr(ops), len(ops), T, ptr(pools), stride, ptr(means), ptr(se),
             ptr(consumed), C.byref(fb), None if prof is None else C.byref(prof))
        return ProbabilityEstimate(n, T, means, se, consumed, fb.value, _profile(prof))


@dataclass
class ProbabilityEstimate:
    """NoisyProbabilityResult (Noise/NoisyProbabilityResult.cs) plus the draw accounting."""
    QubitCount: int
    Trajectories: int
    Means: np.ndarray            # [2^n] Welford means of the normalised probabilities
    StandardErrors: np.ndarray   # [2^n] sqrt(M2 / (M (M - 1)))
    consumed: np.ndarray         # [trajectories] draws taken from each pool
    boundary_fallbacks: int
    profile: Optional[HipExecutionProfile]


class HipBatchedEnergyEvaluator:
    """Variational energies and parameter-shift gradients, batched on the device."""

    def __init__(self, backend: HipStateVectorBackend, precision: int = C128):
        if backend is None:
            raise ArgumentError("backend is null.")
        self._backend = backend
        self._precision = precision
        self.last_profile: Optional[HipExecutionProfile] = None

    @staticmethod
    def _validate(circuit, hamiltonian):
        if circuit is None or hamiltonian is None:
            raise ArgumentError("circuit and hamiltonian are required.")
        if circuit.QubitCount != hamiltonian.QubitCount:
            raise ArgumentError("Circuit and Hamiltonian sizes must match.")

    def EvaluateWithGradient(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian,
                             parameters: Sequence[float], profile: bool = False) -> VariationalEnergyResult:
        self._validate(circuit, hamiltonian)
        ops, pidx = circuit.packed()
        paulis, coeffs = hamiltonian.packed()
        prm = np.ascontiguousarray(parameters, dtype=np.float64)
        if len(prm) != circuit.ParameterCount:
            raise ArgumentError("Binding size must match ParameterCount.")
        energy = C.c_double()
        grad = np.zeros(len(prm), dtype=np.float64)
        prof = QmProfile() if profile else None
        call("qm_energy_gradient_batch", self._backend._handle(), circuit.QubitCount, self._precision,
             ptr(ops), ptr(pidx), len(ops), ptr(prm), len(prm), ptr(paulis), ptr(coeffs), len(coeffs),
             C.byref(energy), ptr(grad), None if prof is None else C.byref(prof))
        self.last_profile = _profile(prof)
        return VariationalEnergyResult(energy.value, [float(g) for g in grad])

    def Evaluate(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian,
                 parameters: Sequence[float]) -> float:
        return float(self.EvaluateMany(circuit, hamiltonian, [parameters])[0])

    def EvaluateMany(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian,
                     parameter_sets, profile: bool = False) -> np.ndarray:
        """Energies of several parameter vectors (rows) in one launch."""
        self._validate(circuit, hamiltonian)
        ops, pidx = circuit.packed()
        paulis, coeffs = hamiltonian.packed()
        sets = np.ascontiguousarray(parameter_sets, dtype=np.float64)
        if sets.ndim != 2 or sets.shape[1] != circuit.ParameterCount:
            raise ArgumentError("Binding size must match ParameterCount.")
        out = np.zeros(sets.shape[0], dtype=np.float64)
        prof = QmProfile() if profile else None
        call("qm_energy_batch", self._backend._handle(), circuit.QubitCount, self._precision,
             ptr(ops), ptr(pidx), len(ops), ptr(sets), sets.shape[1], sets.shape[0], ptr(paulis),
             ptr(coeffs), len(coeffs), ptr(out), None if prof is None else C.byref(prof))
        self.last_profile = _profile(prof)
        return out


class HipVqeProgram:
    """One template and Hamiltonian evaluated many times (a VQE loop) through
    graphed batches (qm_vqe_program_*): the first gradient call, and the first
    energy call for each number of parameter sets, capture a hipGraph that
    later calls replay with new angles.  The evaluator methods take the
    circuit and Hamiltonian for interface parity and require the program's own."""

    def __init__(self, backend: HipStateVectorBackend, circuit: ParameterizedCircuit,
                 hamiltonian: PauliHamiltonian, precision: int = C128):
        if backend is None:
            raise ArgumentError("backend is null.")
        HipBatchedEnergyEvaluator._validate(circuit, hamiltonian)
        self._backend = backend
        self._circuit = circuit
        self._hamiltonian = hamiltonian
        self._n_params = circuit.ParameterCount
        ops, pidx = circuit.packed()
        paulis, coeffs = hamiltonian.packed()
        self._coeffs = coeffs
        h = C.c_void_p()
        call("qm_vqe_program_create", backend._handle(), circuit.QubitCount, precision, ptr(ops), ptr(pidx),
             len(ops), circuit.ParameterCount, ptr(paulis), ptr(coeffs), len(coeffs), C.byref(h))
        self._h = h
        self.last_profile: Optional[HipExecutionProfile] = None

    def _check(self, circuit, hamiltonian):
        if self._h is None:
            raise ObjectDisposedError("HipVqeProgram")
        if circuit is not self._circuit or hamiltonian is not self._hamiltonian:
            raise ArgumentError("The program was built for another circuit or Hamiltonian.")

    def EvaluateWithGradient(self, circuit, hamiltonian, parameters, profile: bool = False) -> VariationalEnergyResult:
        self._check(circuit, hamiltonian)
        prm = np.ascontiguousarray(parameters, dtype=np.float64)
        if len(prm) != self._n_params:
            raise ArgumentError("Binding size must match ParameterCount.")
        energy = C.c_double()
        grad = np.zeros(len(prm), dtype=np.float64)
        prof = QmProfile() if profile else None
        call("qm_vqe_program_gradient", self._h, ptr(prm), C.byref(energy), ptr(grad),
             None if prof is None else C.byref(prof))
        self.last_profile = _profile(prof)
        return VariationalEnergyResult(energy.value, [float(g) for g in grad])

    def EvaluateMany(self, circuit, hamiltonian, parameter_sets, profile: bool = False) -> np.ndarray:
        self._check(circuit, hamiltonian)
        sets = np.ascontiguousarray(parameter_sets, dtype=np.float64)
        if sets.ndim != 2 or sets.shape[1] != self._n_params:
            raise ArgumentError("Binding size must match ParameterCount.")
        out = np.zeros(sets.shape[0], dtype=np.float64)
        prof = QmProfile() if profile else None
        call("qm_vqe_program_energies", self._h, ptr(sets), sets.shape[0], ptr(out),
             None if prof is None else C.byref(prof))
        self.last_profile = _profile(prof)
        return out

    def Evaluate(self, circuit, hamiltonian, parameters) -> float:
        return float(self.EvaluateMany(circuit, hamiltonian, [parameters])[0])

    def Dispose(self) -> None:
        if self._h is not None:
            call("qm_vqe_program_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Dispose()
