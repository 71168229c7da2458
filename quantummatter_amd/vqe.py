"""Deterministic pure-state VQE driven by the batched HIP evaluator.

``HipVqeSolver`` mirrors ``VqeSolver``
(src/QuantumMatter.Application/Variational/VqeSolver.cs:7-136) with its
options and result records (VqeOptions.cs:4-41, VqeResult.cs:4-38): steepest
descent with a bounded Armijo backtracking line search, the infinity-norm
gradient test, the fused-multiply-add trial step (VqeSolver.cs:76) and the
sufficient-decrease test compared in log space (VqeSolver.cs:103-110), with
the same evaluation accounting and termination reasons.

What runs on the device: every energy-and-gradient call is one launch of all
1 + 2R parameter-shift circuits and every line-search energy one launch, for
widths the batched kernel supports, both replayed from hipGraphs captured at
their first use (``qm_vqe_program_*``); narrower circuits use resident-state
sessions.  The solver's
own arithmetic is a handful of host FP64 operations per parameter.

Differences from the reference: the reference budgets host arrays
(``SimulationOptions.MemoryBudgetBytes``, VqeSolver.cs:120-125); states here
live on the device and are budgeted by the backend's device budget, so the
solver takes no host budget.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import enum
import math
from dataclasses import dataclass, replace
from typing import Optional, Sequence, Tuple

from .backend import C128, CancellationToken, HipStateVectorBackend
from .batched import HipVqeProgram, batch_min_qubits
from .circuit import ParameterizedCircuit, PauliHamiltonian
from .errors import ArgumentError, ArgumentOutOfRangeError, OperationCanceledError
from .variational import HipVariationalEnergyEvaluator

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.fma.restype = ctypes.c_double
_libm.fma.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]


def fused_multiply_add(a: float, b: float, c: float) -> float:
    """a*b + c with one rounding (Math.FusedMultiplyAdd; C99 fma is correctly rounded)."""
    return _libm.fma(a, b, c)


class VqeTerminationReason(enum.IntEnum):
    """VqeTerminationReason (Variational/VqeTerminationReason.cs)."""
    GradientToleranceReached = 0
    IterationLimitReached = 1
    EvaluationLimitReached = 2
    LineSearchStalled = 3


def _finite(x: float) -> bool:
    return isinstance(x, (int, float)) and math.isfinite(x)


@dataclass(frozen=True)
class VqeOptions:
    """Limits and backtracking settings (VqeOptions.cs:4-41), same defaults."""
    MaximumIterations: int = 200
    MaximumCircuitEvaluations: int = 100_000
    MaximumLineSearchIterations: int = 40
    GradientTolerance: float = 1e-6
    InitialStepSize: float = 1.0
    MinimumStepSize: float = 1e-12
    BacktrackingFactor: float = 0.5
    SufficientDecrease: float = 0.1

    def with_(self, **changes) -> "VqeOptions":
        return replace(self, **changes)

    def Validate(self) -> None:
        if self.MaximumIterations < 0:
            raise ArgumentOutOfRangeError("MaximumIterations")
        if self.MaximumCircuitEvaluations <= 0:
            raise ArgumentOutOfRangeError("MaximumCircuitEvaluations")
        if self.MaximumLineSearchIterations <= 0:
            raise ArgumentOutOfRangeError("MaximumLineSearchIterations")
        if not _finite(self.GradientTolerance) or self.GradientTolerance < 0:
            raise ArgumentOutOfRangeError("GradientTolerance")
        if not _finite(self.InitialStepSize) or self.InitialStepSize <= 0:
            raise ArgumentOutOfRangeError("InitialStepSize")
        if (not _finite(self.MinimumStepSize) or self.MinimumStepSize <= 0
                or self.MinimumStepSize > self.InitialStepSize):
            raise ArgumentOutOfRangeError("MinimumStepSize")
        if not _finite(self.BacktrackingFactor) or not 0 < self.BacktrackingFactor < 1:
            raise ArgumentOutOfRangeError("BacktrackingFactor")
        if not _finite(self.SufficientDecrease) or not 0 < self.SufficientDecrease < 1:
            raise ArgumentOutOfRangeError("SufficientDecrease")


@dataclass(frozen=True)
class VqeResult:
    """Last accepted iterate and complete evaluation accounting (VqeResult.cs:4-38)."""
    Parameters: Tuple[float, ...]
    Energy: float
    IterationCount: int
    EnergyEvaluationCount: int
    GradientEvaluationCount: int
    CircuitEvaluationCount: int
    TerminationReason: VqeTerminationReason


def _log(x: float) -> float:
    """Natural log with IEEE results at the edges (log 0 = -inf, log of a negative = NaN), as Math.Log."""
    if x == 0:
        return -math.inf
    if not x > 0:
        return math.nan
    return math.log(x)


def _check(token: Optional[CancellationToken]) -> None:
    if token is not None and token.IsCancellationRequested:
        raise OperationCanceledError("The operation was canceled.")


class HipVqeSolver:
    """VqeSolver over the HIP backend (VqeSolver.cs:7-136).

    ``evaluator`` may replace the backend's evaluators; it must provide
    ``Evaluate(circuit, hamiltonian, parameters)`` and
    ``EvaluateWithGradient(circuit, hamiltonian, parameters)`` returning an
    object with ``Energy`` and ``Gradient``.
    """

    def __init__(self, backend: Optional[HipStateVectorBackend] = None,
                 options: Optional[VqeOptions] = None, precision: int = C128, evaluator=None):
        if backend is None and evaluator is None:
            raise ArgumentError("backend is null.")
        self._options = options or VqeOptions()
        self._options.Validate()
        self._backend = backend
        self._precision = precision
        self._evaluator = evaluator

    def _evaluator_for(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian):
        """The injected evaluator; else a graphed program for batched widths (its
        gradient and line-search launches are captured once and replayed), else
        resident-state sessions.  Returns (evaluator, owned)."""
        if self._evaluator is not None:
            return self._evaluator, False
        if circuit.QubitCount >= batch_min_qubits(self._precision):
            return HipVqeProgram(self._backend, circuit, hamiltonian, self._precision), True
        return HipVariationalEnergyEvaluator(self._backend, self._precision), False

    def Solve(self, circuit: ParameterizedCircuit, hamiltonian: PauliHamiltonian,
              initial_parameters: Sequence[float],
              cancellation: Optional[CancellationToken] = None) -> VqeResult:
        """Searches from a finite parameter snapshot; returns the last accepted iterate
        (VqeSolver.cs:27-117).  Cancellation and backend or numerical failures raise
        without a partial result."""
        self._validate(circuit, hamiltonian, initial_parameters, cancellation)
        evaluator, owned = self._evaluator_for(circuit, hamiltonian)
        try:
            return self._solve(evaluator, circuit, hamiltonian, initial_parameters, cancellation)
        finally:
            if owned:
                evaluator.Dispose()

    def _solve(self, evaluator, circuit, hamiltonian, initial_parameters, cancellation) -> VqeResult:
        o = self._options
        parameters = [float(v) for v in initial_parameters]
        trial = [0.0] * len(parameters)
        energy = float(evaluator.Evaluate(circuit, hamiltonian, parameters))
        _check(cancellation)
        energies, circuits = 1, 1
        iterations, gradients = 0, 0
        gradient_cost = 1 + 2 * len(circuit.RotationParameterIndices)

        def finish(reason: VqeTerminationReason) -> VqeResult:
            _check(cancellation)
            return VqeResult(tuple(parameters), energy, iterations, energies, gradients, circuits, reason)

        while True:
            _check(cancellation)
            if iterations == o.MaximumIterations:
                return finish(VqeTerminationReason.IterationLimitReached)
            if o.MaximumCircuitEvaluations - circuits < gradient_cost:
                return finish(VqeTerminationReason.EvaluationLimitReached)
            evaluation = evaluator.EvaluateWithGradient(circuit, hamiltonian, parameters)
            _check(cancellation)
            gradients += 1
            circuits += gradient_cost
            energy = float(evaluation.Energy)
            g = [float(x) for x in evaluation.Gradient]
            norm = 0.0
            for x in g:
                norm = max(norm, abs(x))
            if norm <= o.GradientTolerance:
                return finish(VqeTerminationReason.GradientToleranceReached)
            # Dividing by the norm before squaring keeps finite gradients whose squares
            # overflow or underflow representable.
            scaled_squares = 0.0
            for x in g:
                s = x / norm
                scaled_squares += s * s
            step = o.InitialStepSize
            accepted = False
            for _ in range(o.MaximumLineSearchIterations):
                _check(cancellation)
                if circuits == o.MaximumCircuitEvaluations:
                    return finish(VqeTerminationReason.EvaluationLimitReached)
                finite, changed = True, False
                for i, p in enumerate(parameters):
                    t = fused_multiply_add(-step, g[i], p)
                    trial[i] = t
                    finite &= math.isfinite(t)
                    changed |= t != p
                if not changed:
                    return finish(VqeTerminationReason.LineSearchStalled)
                if finite:
                    candidate = float(evaluator.Evaluate(circuit, hamiltonian, trial))
                    _check(cancellation)
                    energies += 1
                    circuits += 1
                    if self._sufficient_decrease(energy, candidate, step, norm, scaled_squares):
                        parameters, trial = trial, parameters
                        energy = candidate
                        iterations += 1
                        accepted = True
                        break
                smaller = step * o.BacktrackingFactor
                if smaller < o.MinimumStepSize or smaller >= step:
                    break
                step = smaller
            if not accepted:
                return finish(VqeTerminationReason.LineSearchStalled)

    def _sufficient_decrease(self, energy: float, candidate: float, step: float, norm: float,
                             scaled_squares: float) -> bool:
        """E - E_trial >= c * step * ||g||^2, compared in log space so neither side
        overflows (VqeSolver.cs:103-110)."""
        if candidate >= energy:
            return False
        scale = max(abs(energy), abs(candidate))
        log_decrease = _log(scale) + _log(energy / scale - candidate / scale)
        log_required = (_log(self._options.SufficientDecrease) + _log(step)
                        + 2 * _log(norm) + _log(scaled_squares))
        return log_decrease >= log_required

    @staticmethod
    def _validate(circuit, hamiltonian, initial, cancellation) -> None:
        if circuit is None:
            raise ArgumentError("circuit is null.")
        if hamiltonian is None:
            raise ArgumentError("hamiltonian is null.")
        if circuit.QubitCount != hamiltonian.QubitCount:
            raise ArgumentError("Circuit and Hamiltonian sizes must match.")
        if len(initial) != circuit.ParameterCount:
            raise ArgumentError("Parameter count must match the circuit.")
        _check(cancellation)
        for v in initial:
            if not math.isfinite(float(v)):
                raise ArgumentOutOfRangeError("initialParameters", "Parameters must be finite.")
