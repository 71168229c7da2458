"""Variational energy and parameter-shift gradients on the HIP backend.

``HipVariationalEnergyEvaluator`` mirrors ``VariationalEnergyEvaluator``
(src/QuantumMatter.Application/Variational/VariationalEnergyEvaluator.cs:8-113):
``Evaluate`` and ``EvaluateWithGradient`` with the same shift rule (an appended
same-axis rotation of +-pi/2 after each rotation occurrence,
ParameterizedCircuit.cs:85-89), the same per-term differencing
c*(0.5*plus - 0.5*minus) and the same Neumaier accumulation per parameter.
States stay on the device; expectation values of all terms are computed from
one state read per flip mask.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .backend import C128, SCHED_FUSED, HipStateVectorBackend
from .circuit import ParameterizedCircuit, PauliHamiltonian
from .errors import ArgumentError, InvalidOperationError


@dataclass(frozen=True)
class VariationalEnergyResult:
    Energy: float
    Gradient: List[float]


class HipVariationalEnergyEvaluator:
    def __init__(self, backend: HipStateVectorBackend, precision: int = C128, schedule: int = SCHED_FUSED):
        if backend is None:
            raise ArgumentError("backend is null.")
        self._backend = backend
        self._precision = precision
        self._schedule = schedule

    def _validate(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian):
        if circuit is None or hamiltonian is None:
            raise ArgumentError("circuit and hamiltonian are required.")
        if circuit.QubitCount != hamiltonian.QubitCount:
            raise ArgumentError("Circuit and Hamiltonian sizes must match.")

    def Evaluate(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian,
                 parameters: Sequence[float]) -> float:
        self._validate(circuit, hamiltonian)
        bound = circuit.Bind(parameters)
        with self._backend.CreateSession(circuit.QubitCount, self._precision) as s:
            s.Evolve(bound, self._schedule)
            return s.ExpectationEnergy(hamiltonian)

    def EvaluateWithGradient(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian,
                             parameters: Sequence[float]) -> VariationalEnergyResult:
        self._validate(circuit, hamiltonian)
        bound = circuit.Bind(parameters)
        values = list(parameters)
        gradient = [0.0] * circuit.ParameterCount
        corrections = [0.0] * circuit.ParameterCount
        observables = [t.Observable for t in hamiltonian.Terms]
        with self._backend.CreateSession(circuit.QubitCount, self._precision) as s:
            s.Evolve(bound, self._schedule)
            energy = s.ExpectationEnergy(hamiltonian)
            for rotation, parameter in enumerate(circuit.RotationParameterIndices):
                s.Reset()
                s.Evolve(circuit.BindWithRotationShift(values, rotation, math.pi / 2), self._schedule)
                plus = self._expectations(s, hamiltonian, observables)
                s.Reset()
                s.Evolve(circuit.BindWithRotationShift(values, rotation, -math.pi / 2), self._schedule)
                minus = self._expectations(s, hamiltonian, observables)
                for term, t in enumerate(hamiltonian.Terms):
                    contribution = t.Coefficient * (0.5 * plus[term] - 0.5 * minus[term])
                    previous = gradient[parameter]
                    nxt = previous + contribution
                    corrections[parameter] += ((previous - nxt) + contribution
                                               if abs(previous) >= abs(contribution)
                                               else (contribution - nxt) + previous)
                    gradient[parameter] = nxt
                    if not (math.isfinite(nxt) and math.isfinite(corrections[parameter])):
                        raise InvalidOperationError("The gradient accumulation is nonfinite.")
        for p in range(len(gradient)):
            gradient[p] += corrections[p]
            if not math.isfinite(gradient[p]):
                raise InvalidOperationError("The gradient is nonfinite.")
        return VariationalEnergyResult(energy, gradient)

    @staticmethod
    def _expectations(s, hamiltonian, observables) -> np.ndarray:
        if not observables:
            s.ExpectationEnergy(hamiltonian)  # validates the Born mass, as the reference does
            return np.zeros(0)
        return s.ExpectationValues(observables)
