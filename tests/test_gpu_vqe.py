"""HipVqeSolver over the device evaluators.

The reference's solver cases (tests/QuantumMatter.Tests/VqeSolverTests.cs)
run through the HIP session evaluator (1-2 qubits), with the same expected
energies, parameters, counts and termination reasons; a 10-qubit
hardware-efficient ansatz solved through the graphed batched
program (qm_vqe_program_*) follows the oracle-driven solve
iterate for iterate: same termination reason and counts, energies within
1e-10 and parameters within 1e-9.
"""
import numpy as np
import pytest

import test_vqe_host as host
import vqe_cases as vc
from hip_adapter import shared_backend
from quantummatter_amd import workloads
from quantummatter_amd.backend import CancellationToken
from quantummatter_amd.batched import HipVqeProgram, batch_min_qubits
from quantummatter_amd.variational import HipVariationalEnergyEvaluator
from quantummatter_amd.vqe import HipVqeSolver, VqeOptions

pytestmark = pytest.mark.gpu


def make(options=None, counter=None, after=None):
    be = shared_backend()
    if counter is None and after is None:
        if options is not None:
            options.Validate()
        return HipVqeSolver(be, options)
    return vc.solver_factory(lambda: HipVariationalEnergyEvaluator(be))(options, counter, after)


@pytest.mark.parametrize("name", sorted(vc.CASES))
def test_reference_solver_case(name):
    print(f"vqe case {name}", flush=True)
    vc.CASES[name](make)


@pytest.mark.parametrize("budget,calls,gradients,iterations", vc.EVALUATION_LIMITS)
def test_evaluation_limit_reserves_whole_gradients(budget, calls, gradients, iterations):
    vc.evaluation_limit(make, budget, calls, gradients, iterations)


@pytest.mark.parametrize("weight,step,minimum", vc.SQUARE_RANGES)
def test_sufficient_decrease_survives_square_overflow(weight, step, minimum):
    vc.square_overflow(make, weight, step, minimum)


@pytest.mark.parametrize("rounds_back", [False, True])
def test_subnormal_steps_stall(rounds_back):
    vc.subnormal_steps(make, rounds_back)


def test_failures_never_return_partial_results():
    vc.failures_never_return_partial_results(make, CancellationToken)


def test_batched_solve_follows_the_oracle_solve():
    n, layers, terms = 10, 2, 16
    assert n >= batch_min_qubits()
    circuit = workloads.hardware_efficient_ansatz(n, layers)
    ham = workloads.random_hamiltonian(n, terms, seed=11)
    theta = workloads.random_parameters(circuit.ParameterCount, seed=12)
    options = VqeOptions(MaximumIterations=4)
    print(f"vqe batched n={n} layers={layers} terms={terms}", flush=True)
    got = HipVqeSolver(shared_backend(), options).Solve(circuit, ham, theta)
    want = HipVqeSolver(options=options, evaluator=host.OracleEvaluator()).Solve(circuit, ham, theta)
    assert got.TerminationReason == want.TerminationReason
    assert (got.IterationCount, got.EnergyEvaluationCount, got.GradientEvaluationCount,
            got.CircuitEvaluationCount) == (want.IterationCount, want.EnergyEvaluationCount,
                                            want.GradientEvaluationCount, want.CircuitEvaluationCount)
    assert abs(got.Energy - want.Energy) <= 1e-10
    np.testing.assert_allclose(got.Parameters, want.Parameters, rtol=0, atol=1e-9)
    ev, owned = HipVqeSolver(shared_backend())._evaluator_for(circuit, ham)
    assert isinstance(ev, HipVqeProgram) and owned
    ev.Dispose()
