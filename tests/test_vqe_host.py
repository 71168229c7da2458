"""HipVqeSolver's host logic against the reference's solver cases, on the CPU.

The solver (quantummatter_amd/vqe.py, VqeSolver.cs:27-117) runs here over an
evaluator built on the oracle (qmo_energy_bound, qmo_energy_gradient), so
these tests pin the solver's own arithmetic and accounting: the FMA trial
step, the log-space Armijo test, the evaluation budget reserve, termination
reasons and failure propagation, with the expected values of
tests/QuantumMatter.Tests/VqeSolverTests.cs.  The device evaluators behind the
same solver are checked in test_gpu_vqe.py.
"""
import math

import pytest

import oracle
import vqe_cases as vc
from quantummatter_amd.backend import CancellationToken
from quantummatter_amd.errors import InvalidOperationError
from quantummatter_amd.vqe import fused_multiply_add


class OracleEvaluator:
    """Evaluate / EvaluateWithGradient of VariationalEnergyEvaluator.cs:23-86 on the oracle."""

    def Evaluate(self, circuit, hamiltonian, parameters):
        ops, idx = circuit.packed()
        paulis, coeffs = hamiltonian.packed()
        try:
            return oracle.energy_bound(circuit.QubitCount, ops, idx, parameters, paulis, coeffs)
        except oracle.OracleError as e:
            raise InvalidOperationError(str(e)) from e

    def EvaluateWithGradient(self, circuit, hamiltonian, parameters):
        ops, idx = circuit.packed()
        paulis, coeffs = hamiltonian.packed()
        try:
            e, g = oracle.energy_gradient(circuit.QubitCount, ops, idx, parameters, paulis, coeffs)
        except oracle.OracleError as err:
            raise InvalidOperationError(str(err)) from err
        return vc.result_record(e, g)


make = vc.solver_factory(OracleEvaluator)


@pytest.mark.parametrize("name", sorted(vc.CASES))
def test_reference_solver_case(name):
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


def test_fused_multiply_add_rounds_once():
    # (1 + 2^-30)^2 - 1 = 2^-29 + 2^-60 exactly; a separate multiply loses the 2^-60 term.
    a = 1 + 2.0 ** -30
    assert fused_multiply_add(a, a, -1.0) == 2.0 ** -29 + 2.0 ** -60
    assert a * a - 1.0 == 2.0 ** -29
    assert fused_multiply_add(1e308, 10.0, -1e308) == math.inf  # the exact sum overflows
    assert fused_multiply_add(1.5e308, 1.0, -1e308) == 1.5e308 - 1e308
    assert math.isnan(fused_multiply_add(math.inf, 0.0, 1.0))
