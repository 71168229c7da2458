"""Graph replay of variational programs (qm_vqe_program_*).

A program captures its gradient batch (and each energy batch size) as a
hipGraph on first use and replays it with new angles.  Every replay, with
eager batched calls of the same context in between, equals the one-shot
entry points (qm_energy_gradient_batch / qm_energy_batch) bit for bit, and
the oracle within 1e-11 relative (complex128).
"""
import numpy as np
import pytest

import oracle
from hip_adapter import shared_backend
from quantummatter_amd import ArgumentError, ObjectDisposedError, workloads
from quantummatter_amd.backend import C64, C128
from quantummatter_amd.batched import HipBatchedEnergyEvaluator, HipVqeProgram

pytestmark = pytest.mark.gpu


def problem(n, layers=2, terms=32, seed=3):
    circuit = workloads.hardware_efficient_ansatz(n, layers)
    ham = workloads.random_hamiltonian(n, terms, seed=seed)
    return circuit, ham


@pytest.mark.parametrize("precision", [C128, C64])
def test_gradient_replays_equal_one_shot_calls(precision):
    n = 12
    circuit, ham = problem(n)
    be = shared_backend()
    eager = HipBatchedEnergyEvaluator(be, precision)
    with HipVqeProgram(be, circuit, ham, precision) as prog:
        for k in range(6):
            theta = workloads.random_parameters(circuit.ParameterCount, seed=50 + k)
            print(f"graph replay {k} precision={precision}", flush=True)
            got = prog.EvaluateWithGradient(circuit, ham, theta)
            want = eager.EvaluateWithGradient(circuit, ham, theta)
            assert got.Energy == want.Energy
            assert got.Gradient == want.Gradient
            if precision == C128 and k < 2:
                ops, idx = circuit.packed()
                paulis, coeffs = ham.packed()
                e, g = oracle.energy_gradient(n, ops, idx, theta, paulis, coeffs)
                scale = np.abs(coeffs).sum()
                assert abs(got.Energy - e) <= 1e-11 * scale
                assert np.max(np.abs(np.array(got.Gradient) - g)) <= 1e-11 * scale


def test_energy_replays_per_set_count():
    n = 11
    circuit, ham = problem(n, seed=8)
    be = shared_backend()
    eager = HipBatchedEnergyEvaluator(be)
    with HipVqeProgram(be, circuit, ham) as prog:
        for k in range(4):
            sets = np.stack([workloads.random_parameters(circuit.ParameterCount, seed=90 + 7 * k + j)
                             for j in range(1 + (k % 2) * 2)])
            got = prog.EvaluateMany(circuit, ham, sets)
            want = eager.EvaluateMany(circuit, ham, sets)
            assert np.array_equal(got, want)
            assert prog.Evaluate(circuit, ham, sets[0]) == want[0]


def test_program_errors():
    be = shared_backend()
    circuit, ham = problem(10)
    other, _ = problem(10)
    prog = HipVqeProgram(be, circuit, ham)
    theta = workloads.random_parameters(circuit.ParameterCount, seed=1)
    with pytest.raises(ArgumentError):
        prog.EvaluateWithGradient(other, ham, theta)
    with pytest.raises(ArgumentError):
        prog.EvaluateWithGradient(circuit, ham, theta[:-1])
    prog.Dispose()
    with pytest.raises(ObjectDisposedError):
        prog.Evaluate(circuit, ham, theta)
    narrow, nham = problem(4)
    with pytest.raises(NotImplementedError):
        HipVqeProgram(be, narrow, nham)


def test_program_arenas_are_counted_and_bounded():
    """A program's graphed arenas count in the context's resident bytes while
    it lives, at most four energy batch sizes stay cached (the least recently
    used is freed first), and destroying the program returns the bytes."""
    from quantummatter_amd.backend import HipStateVectorBackend
    n = 11
    circuit, ham = problem(n, seed=9)
    with HipStateVectorBackend() as be:
        base = be.ResidentBytes()
        with HipVqeProgram(be, circuit, ham) as prog:
            theta = workloads.random_parameters(circuit.ParameterCount, seed=3)
            prog.EvaluateWithGradient(circuit, ham, theta)
            assert be.ResidentBytes() > base
            for sets in range(1, 8):
                prog.EvaluateMany(circuit, ham, np.stack([theta] * sets))
            # Sizes 4..7 are cached: evaluating them again allocates nothing.
            held = be.ResidentBytes()
            for sets in (4, 5, 6, 7):
                prog.EvaluateMany(circuit, ham, np.stack([theta] * sets))
                assert be.ResidentBytes() == held
            # Size 1 was evicted: it is rebuilt in place of the oldest (4).
            prog.EvaluateMany(circuit, ham, np.stack([theta]))
            assert be.ResidentBytes() < held
        assert be.ResidentBytes() == base
