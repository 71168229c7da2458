"""Solver cases restated from tests/QuantumMatter.Tests/VqeSolverTests.cs.

Each case takes ``make(options=None, counting=None)`` that returns a
``HipVqeSolver`` over some evaluator; ``counting`` (a ``CountingEvaluator``
wrapper factory) lets a case count circuit executions the way the reference's
RecordingBackend does.  The same cases run on the CPU with an oracle
evaluator (test_vqe_host.py) and on the device with the HIP evaluators
(test_gpu_vqe.py).
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import pytest

from quantummatter_amd.circuit import (ParameterizedCircuit, PauliHamiltonian, PauliString, PauliTerm,
                                       QuantumCircuit)
from quantummatter_amd.errors import (ArgumentError, ArgumentOutOfRangeError, InvalidOperationError,
                                      OperationCanceledError)
from quantummatter_amd.vqe import HipVqeSolver, VqeOptions, VqeTerminationReason, fused_multiply_add

R = VqeTerminationReason


def single() -> ParameterizedCircuit:
    return ParameterizedCircuit(1, 1).Ry(0, 0)


def z(coefficient: float = 1.0) -> PauliHamiltonian:
    return PauliHamiltonian(1, [PauliTerm(coefficient, PauliString("Z"))])


def near(expected, actual, tol=8e-12):
    assert abs(expected - actual) <= tol, (expected, actual)


def accounting(result, rotations):
    assert result.CircuitEvaluationCount == (result.EnergyEvaluationCount
                                             + (1 + 2 * rotations) * result.GradientEvaluationCount)


class CountingEvaluator:
    """Counts circuit executions (1 per energy, 1 + 2R per gradient), as RecordingBackend counts
    Execute calls; ``after(count_before, count_after)`` runs after each evaluator call."""

    def __init__(self, inner, after=None):
        self.inner = inner
        self.calls = 0
        self.after = after

    def _done(self, cost):
        before = self.calls
        self.calls += cost
        if self.after is not None:
            self.after(before, self.calls)

    def Evaluate(self, circuit, hamiltonian, parameters):
        e = self.inner.Evaluate(circuit, hamiltonian, parameters)
        self._done(1)
        return e

    def EvaluateWithGradient(self, circuit, hamiltonian, parameters):
        r = self.inner.EvaluateWithGradient(circuit, hamiltonian, parameters)
        self._done(1 + 2 * len(circuit.RotationParameterIndices))
        return r


def one_qubit_analytic(make):
    h = PauliHamiltonian(1, [PauliTerm(-0.2, PauliString("I")), PauliTerm(0.7, PauliString("Z")),
                             PauliTerm(0.3, PauliString("X"))])
    r = make().Solve(single(), h, [0.73])
    near(-0.2 - math.sqrt(0.58), r.Energy, 2e-11)
    theta = r.Parameters[0]
    near(-0.2 + 0.7 * math.cos(theta) + 0.3 * math.sin(theta), r.Energy)
    assert abs(-0.7 * math.sin(theta) + 0.3 * math.cos(theta)) <= 1.1e-6
    assert r.TerminationReason == R.GradientToleranceReached
    accounting(r, 1)


def two_qubit_heisenberg(make):
    h = PauliHamiltonian(2, [PauliTerm(1, PauliString("XX")), PauliTerm(1, PauliString("YY")),
                             PauliTerm(1, PauliString("ZZ"))])
    template = (ParameterizedCircuit(2, 1).Append(QuantumCircuit(2).X(0)).Ry(1, 0)
                .Append(QuantumCircuit(2).CNot(1, 0)))
    r = make().Solve(template, h, [0.2])
    # The |01>,|10> block of the dense 4x4 Hamiltonian has eigenvalues -3 and 1.
    m = [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]]
    theta = r.Parameters[0]
    st = [0, math.cos(theta / 2), math.sin(theta / 2), 0]
    energy = residual = 0.0
    for row in range(4):
        prod = sum(m[row][c] * st[c] for c in range(4))
        energy += st[row] * prod
        residual += (prod + 3 * st[row]) ** 2
    near(-3, r.Energy, 2e-11)
    near(energy, r.Energy)
    assert math.sqrt(residual) <= 1.1e-6
    assert r.TerminationReason == R.GradientToleranceReached
    accounting(r, 1)


def backtracking_counts(make):
    counter = []
    r = make(VqeOptions(MaximumIterations=1, InitialStepSize=20, SufficientDecrease=0.1), counter).Solve(single(), z(), [1])
    g = -math.sin(1)
    first = fused_multiply_add(-20, g, 1)
    accepted = fused_multiply_add(-10, g, 1)
    assert math.cos(first) > math.cos(1) - 0.1 * 20 * g * g
    assert math.cos(accepted) <= math.cos(1) - 0.1 * 10 * g * g
    near(accepted, r.Parameters[0])
    near(math.cos(accepted), r.Energy)
    assert (r.IterationCount, r.EnergyEvaluationCount, r.GradientEvaluationCount, r.CircuitEvaluationCount) == (1, 3, 1, 6)
    assert counter[0].calls == r.CircuitEvaluationCount
    assert r.TerminationReason == R.IterationLimitReached


EVALUATION_LIMITS = [(1, 1, 0, 0), (3, 1, 0, 0), (4, 4, 1, 0), (5, 5, 1, 1), (7, 5, 1, 1), (8, 8, 2, 1)]


def evaluation_limit(make, budget, calls, gradients, iterations):
    counter = []
    r = make(VqeOptions(MaximumCircuitEvaluations=budget), counter).Solve(single(), z(), [0.7])
    assert r.TerminationReason == R.EvaluationLimitReached
    assert counter[0].calls == calls == r.CircuitEvaluationCount
    assert r.GradientEvaluationCount == gradients and r.IterationCount == iterations
    near(math.cos(r.Parameters[0]), r.Energy)
    accounting(r, 1)


def iteration_limits(make):
    counter = []
    r = make(VqeOptions(MaximumIterations=0, MaximumCircuitEvaluations=2**63 - 1), counter).Solve(single(), z(), [0.7])
    assert r.TerminationReason == R.IterationLimitReached
    assert (r.IterationCount, r.GradientEvaluationCount, counter[0].calls) == (0, 0, 1)
    near(math.cos(0.7), r.Energy)
    r = make(VqeOptions(MaximumIterations=2)).Solve(single(), z(), [0.7])
    assert r.IterationCount == 2 and r.TerminationReason == R.IterationLimitReached
    assert r.CircuitEvaluationCount == 9


def stationarity(make):
    solver = make()
    maximum = solver.Solve(single(), z(), [0])
    assert maximum.TerminationReason == R.GradientToleranceReached
    assert maximum.Energy == 1 and maximum.IterationCount == 0
    constant = solver.Solve(ParameterizedCircuit(1, 0).Append(QuantumCircuit(1).X(0)), z(), [])
    assert constant.Energy == -1 and constant.Parameters == ()
    assert constant.CircuitEvaluationCount == 2
    assert constant.TerminationReason == R.GradientToleranceReached
    zero = solver.Solve(single(), PauliHamiltonian(1, []), [0.7])
    assert zero.Energy == 0 and zero.TerminationReason == R.GradientToleranceReached


def shared_parameters(make):
    counter = []
    solver = make(VqeOptions(InitialStepSize=0.25), counter)
    template = ParameterizedCircuit(1, 2).Ry(0, 0).Ry(0, 0)
    initial = [0.31, 17.0]
    r = solver.Solve(template, z(), initial)
    assert initial == [0.31, 17.0]
    assert r.Parameters[1] == 17
    near(-1, r.Energy, 2e-11)
    near(math.cos(2 * r.Parameters[0]), r.Energy)
    assert r.TerminationReason == R.GradientToleranceReached
    accounting(r, 2)
    assert counter[0].calls == r.CircuitEvaluationCount
    with pytest.raises(TypeError):
        r.Parameters[0] = 9  # the result owns an immutable copy
    assert template.OperationCount == 2


def rejected_lower_trial(make):
    r = make(VqeOptions(MaximumLineSearchIterations=1, InitialStepSize=5, SufficientDecrease=0.9)).Solve(single(), z(), [1])
    assert math.cos(1 + 5 * math.sin(1)) < math.cos(1)
    assert r.TerminationReason == R.LineSearchStalled
    assert r.IterationCount == 0 and r.Parameters[0] == 1
    near(math.cos(1), r.Energy)
    assert r.CircuitEvaluationCount == 5


def minimum_step_and_unrepresentable(make):
    m = make(VqeOptions(InitialStepSize=20, MinimumStepSize=15, SufficientDecrease=0.1)).Solve(single(), z(), [1])
    assert m.TerminationReason == R.LineSearchStalled and m.CircuitEvaluationCount == 5
    huge = make(VqeOptions(GradientTolerance=0)).Solve(single(), z(), [1e100])
    assert huge.TerminationReason == R.LineSearchStalled
    assert huge.CircuitEvaluationCount == 4  # no execution for an unchanged parameter vector
    offset = PauliHamiltonian(1, [PauliTerm(1e16, PauliString("I")), PauliTerm(1, PauliString("Z"))])
    r = make().Solve(single(), offset, [0.7])
    assert r.TerminationReason == R.LineSearchStalled and r.IterationCount == 0


SQUARE_RANGES = [(1e308, 1e-308, 1e-310), (1e-200, 1e200, 1e-12)]


def square_overflow(make, weight, step, minimum):
    r = make(VqeOptions(MaximumIterations=1, InitialStepSize=step, MinimumStepSize=minimum,
                        GradientTolerance=0)).Solve(single(), z(weight), [math.pi / 2])
    assert r.IterationCount == 1 and r.TerminationReason == R.IterationLimitReached
    assert r.Energy / weight < -0.8
    near(math.cos(r.Parameters[0]), r.Energy / weight)


def unrepresentable_candidates(make):
    counter = []
    r = make(VqeOptions(InitialStepSize=1e308, MaximumLineSearchIterations=3), counter).Solve(single(), z(1e308), [0.7])
    assert r.TerminationReason == R.LineSearchStalled
    assert counter[0].calls == 4 and r.EnergyEvaluationCount == 1 and r.Parameters[0] == 0.7


def subnormal_steps(make, rounds_back):
    x = PauliHamiltonian(1, [PauliTerm(1, PauliString("X"))])
    eps = 5e-324
    factor = math.nextafter(1.0, 0.0) if rounds_back else 0.5
    r = make(VqeOptions(InitialStepSize=eps, MinimumStepSize=eps, GradientTolerance=0,
                        BacktrackingFactor=factor)).Solve(single(), x, [0])
    assert r.TerminationReason == R.LineSearchStalled
    assert r.CircuitEvaluationCount == 5 and r.Parameters[0] == 0


def log_comparison(make):
    r = make(VqeOptions(InitialStepSize=7e-308, MinimumStepSize=1e-310, MaximumIterations=1,
                        GradientTolerance=0)).Solve(single(), z(1e308), [0.4])
    assert r.IterationCount == 1
    assert 1e308 * math.cos(0.4) - r.Energy == math.inf
    near(math.cos(r.Parameters[0]), r.Energy / 1e308)


def invalid_options_and_inputs(make):
    counter = []
    defaults = VqeOptions()
    invalid = [dict(MaximumIterations=-1), dict(MaximumCircuitEvaluations=0), dict(MaximumLineSearchIterations=0),
               dict(MinimumStepSize=2), dict(BacktrackingFactor=0), dict(BacktrackingFactor=1),
               dict(SufficientDecrease=0), dict(SufficientDecrease=1), dict(GradientTolerance=-1),
               dict(InitialStepSize=0), dict(MinimumStepSize=0)]
    for v in (math.nan, math.inf, -math.inf, -1.0):
        for name in ("GradientTolerance", "InitialStepSize", "MinimumStepSize", "BacktrackingFactor",
                     "SufficientDecrease"):
            invalid.append({name: v})
    for change in invalid:
        with pytest.raises(ArgumentOutOfRangeError):
            make(defaults.with_(**change))
    solver = make(None, counter)
    with pytest.raises(ArgumentError):
        solver.Solve(None, z(), [])
    with pytest.raises(ArgumentError):
        solver.Solve(single(), None, [1])
    with pytest.raises(ArgumentError):
        solver.Solve(single(), PauliHamiltonian(2, []), [1])
    with pytest.raises(ArgumentError):
        solver.Solve(single(), z(), [])
    for v in (math.nan, math.inf, -math.inf):
        with pytest.raises(ArgumentOutOfRangeError):
            solver.Solve(single(), z(), [v])
    assert counter[0].calls == 0


def failures_never_return_partial_results(make, cancellation_token):
    token = cancellation_token()
    token.Cancel()
    counter = []
    with pytest.raises(OperationCanceledError):
        make(None, counter).Solve(single(), z(), [0.7], token)
    assert counter[0].calls == 0
    for call in (1, 2, 5):
        during = cancellation_token()

        def cancel_at(before, after, call=call, during=during):
            if before < call <= after:
                during.Cancel()

        with pytest.raises(OperationCanceledError):
            make(None, [], cancel_at).Solve(single(), z(), [0.7], during)

        def fail_at(before, after, call=call):
            if before < call <= after:
                raise InvalidOperationError("Backend failure")

        with pytest.raises(InvalidOperationError):
            make(None, [], fail_at).Solve(single(), z(), [0.7])
    overflow = PauliHamiltonian(1, [PauliTerm(1e308, PauliString("I")), PauliTerm(1e308, PauliString("I"))])
    with pytest.raises(InvalidOperationError):
        make().Solve(single(), overflow, [0.7])


CASES = {
    "one_qubit_analytic": one_qubit_analytic,
    "two_qubit_heisenberg": two_qubit_heisenberg,
    "backtracking_counts": backtracking_counts,
    "iteration_limits": iteration_limits,
    "stationarity": stationarity,
    "shared_parameters": shared_parameters,
    "rejected_lower_trial": rejected_lower_trial,
    "minimum_step_and_unrepresentable": minimum_step_and_unrepresentable,
    "unrepresentable_candidates": unrepresentable_candidates,
    "log_comparison": log_comparison,
    "invalid_options_and_inputs": invalid_options_and_inputs,
}


def solver_factory(inner_for):
    """Builds ``make(options, counter_list, after)``: a solver over ``inner_for()`` wrapped in a
    CountingEvaluator appended to ``counter_list`` when one is given."""

    def make(options=None, counter=None, after=None):
        if options is not None:
            options.Validate()
        ev = inner_for()
        if counter is not None or after is not None:
            ev = CountingEvaluator(ev, after)
            if counter is not None:
                counter.append(ev)
        return HipVqeSolver(options=options, evaluator=ev)

    return make


def result_record(energy, gradient):
    return SimpleNamespace(Energy=energy, Gradient=list(gradient))
