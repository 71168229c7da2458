"""Parity of the batched entry points with the CPU oracle.

Trajectories (qm_trajectories_run): for the same per-trajectory uniform pools
and shot uniforms, every terminal sample, MeasureAll outcome and pool draw
count equals the oracle's (complex128).  Variational batches
(qm_energy_gradient_batch, qm_energy_batch): energy and parameter-shift
gradient within 1e-11 relative of the oracle in complex128 (the shifted
rotation is the host product of the two prepared rotations, so the amplitudes
differ from the reference's two-gate form by rounding only, at any finite
angle), and within 1e-4 of the coefficient norm in complex64 (FP64
reductions).
"""
import math

import numpy as np
import pytest

import oracle
from hip_adapter import shared_backend
from quantummatter_amd import (NoiseChannel, NoisyCircuit, ParameterizedCircuit, PauliHamiltonian,
                               PauliString, PauliTerm, PoolExhaustedError, QuantumCircuit,
                               workloads)
from quantummatter_amd.backend import C64, C128
from quantummatter_amd.batched import (HipBatchedEnergyEvaluator, HipTrajectoryRunner,
                                       HipVqeProgram, batch_min_qubits)

pytestmark = pytest.mark.gpu


def draws_needed(nc: NoisyCircuit) -> int:
    steps, _ = nc.packed()
    return int(np.sum(steps["is_channel"] == 1)) + 1


def run_both(nc, T, seed, shots=64, measure=True, precision=C128):
    n = nc.QubitCount
    stride = draws_needed(nc)
    pools = workloads.uniform_pools(seed, T, stride)
    su = workloads.uniforms(seed + 1, T * shots).reshape(T, shots) if shots else None
    steps, ops = nc.packed()
    c_want, o_want, s_want = oracle.trajectories(n, steps, ops, pools, shots, su, measure)
    got = HipTrajectoryRunner(shared_backend(), precision).Run(nc, pools, shots, su, measure)
    return (c_want, o_want, s_want), got


@pytest.mark.parametrize("n,layers,T", [(12, 3, 48), (13, 2, 24), (10, 3, 32), (16, 1, 8)])
def test_trajectories_match_oracle(n, layers, T):
    nc = workloads.noisy_layer_circuit(n, layers, seed=n, p_depol=0.08, gamma=0.15)
    (c_want, o_want, s_want), got = run_both(nc, T, seed=100 + n)
    assert np.array_equal(got.consumed, c_want)
    assert np.array_equal(got.outcomes, o_want)
    assert np.array_equal(got.samples, s_want)


def test_trajectories_all_channel_kinds():
    n = 12
    nc = NoisyCircuit(n)
    c = QuantumCircuit(n)
    for q in range(n):
        c.H(q)
    nc.Append(c)
    nc.ApplyNoise(NoiseChannel.BitFlip(0.3), 0).ApplyNoise(NoiseChannel.PhaseFlip(0.4), 5)
    nc.ApplyNoise(NoiseChannel.BitFlip(1.0), 3).ApplyNoise(NoiseChannel.PhaseFlip(1.0), 11)
    nc.ApplyNoise(NoiseChannel.Depolarizing(1.0), 7).ApplyNoise(NoiseChannel.Depolarizing(0.0), 2)
    nc.ApplyNoise(NoiseChannel.AmplitudeDamping(0.5), 4).ApplyNoise(NoiseChannel.AmplitudeDamping(1.0), 9)
    nc.Append(QuantumCircuit(n).CNot(4, 10).Ry(9, 0.7))
    nc.ApplyNoise(NoiseChannel.AmplitudeDamping(0.3), 10)
    (c_want, o_want, s_want), got = run_both(nc, 64, seed=7)
    assert np.array_equal(got.consumed, c_want)
    assert np.array_equal(got.outcomes, o_want)
    assert np.array_equal(got.samples, s_want)


def test_trajectories_shots_only_and_measure_only():
    nc = workloads.noisy_layer_circuit(12, 2, seed=5, p_depol=0.05, gamma=0.1)
    (c_want, _, s_want), got = run_both(nc, 16, seed=9, shots=128, measure=False)
    assert got.outcomes is None
    assert np.array_equal(got.samples, s_want)
    assert np.array_equal(got.consumed, c_want)
    (c_want, o_want, _), got = run_both(nc, 16, seed=11, shots=0, measure=True)
    assert got.samples is None
    assert np.array_equal(got.outcomes, o_want)
    assert np.array_equal(got.consumed, c_want)


def test_trajectory_pool_exhaustion():
    nc = workloads.noisy_layer_circuit(12, 1, seed=3, p_depol=0.5, gamma=0.5)
    stride = draws_needed(nc) // 3
    pools = workloads.uniform_pools(1, 4, stride)
    with pytest.raises(PoolExhaustedError):
        HipTrajectoryRunner(shared_backend()).Run(nc, pools, 0, None, True)


def test_trajectories_c64_agree_mostly():
    nc = workloads.noisy_layer_circuit(12, 2, seed=12, p_depol=0.05, gamma=0.1)
    (c_want, o_want, s_want), got = run_both(nc, 32, seed=21, precision=C64)
    assert np.mean(got.outcomes == o_want) >= 0.9
    assert np.mean(got.samples == s_want) >= 0.9


def test_batch_min_width_is_reported():
    assert batch_min_qubits(C128) <= 10
    assert batch_min_qubits(C64) <= 10


def hea_case(n, layers, terms, seed):
    circuit = workloads.hardware_efficient_ansatz(n, layers)
    theta = workloads.random_parameters(circuit.ParameterCount, seed)
    ham = workloads.random_hamiltonian(n, terms, seed + 1)
    return circuit, theta, ham


def oracle_gradient(circuit, theta, ham):
    ops, pidx = circuit.packed()
    paulis, coeffs = ham.packed()
    return oracle.energy_gradient(circuit.QubitCount, ops, pidx, theta, paulis, coeffs)


@pytest.mark.parametrize("n,layers,terms", [(12, 2, 64), (10, 3, 32), (13, 1, 48)])
def test_energy_gradient_batch_c128(n, layers, terms):
    circuit, theta, ham = hea_case(n, layers, terms, seed=n)
    e_want, g_want = oracle_gradient(circuit, theta, ham)
    res = HipBatchedEnergyEvaluator(shared_backend(), C128).EvaluateWithGradient(circuit, ham, theta)
    scale = sum(abs(t.Coefficient) for t in ham.Terms)
    assert abs(res.Energy - e_want) <= 1e-11 * scale
    assert np.max(np.abs(np.array(res.Gradient) - g_want)) <= 1e-11 * scale


def _structured_hamiltonian(n, seed):
    """Terms of every shape the after-batch Pauli reductions split on: Z-only
    and identity strings (no flip), flips inside the lowest 12 qubits only,
    flips reaching the top qubits, Y factors, and more than 8 terms on one
    flip mask (split over several groups of that mask)."""
    rng = np.random.default_rng(seed)

    def term(labels):  # labels: qubit -> 'X' | 'Y' | 'Z'
        s = "".join(labels.get(q, "I") for q in reversed(range(n)))
        return PauliTerm(float(rng.normal()), PauliString(s))

    terms = [term({}), term({0: "Z", n - 1: "Z"}), term({3: "Z", 7: "Z", 11: "Z"})]
    for lo in range(6):  # flips in the low 12 qubits
        terms.append(term({lo: "X", lo + 5: "Y"}))
        terms.append(term({lo: "Y", lo + 3: "Z"}))
    for q in (12, n - 2, n - 1):  # flips with high bits, alone and with low ones
        terms.append(term({q: "X"}))
        terms.append(term({q: "Y", 2: "X", 9: "Z"}))
    zs = [{1: "Z"}, {4: "Z"}, {n - 1: "Z"}, {6: "Z", 13: "Z"}, {}, {0: "Z"}, {8: "Z"},
          {10: "Z", 2: "Z"}, {5: "Z"}, {n - 3: "Z"}, {11: "Z"}]
    for extra in zs:  # eleven terms on one flip mask (qubits 2 and n - 1)
        terms.append(term({2: "X", n - 1: "Y", **{k: v for k, v in extra.items() if k not in (2, n - 1)}}))
    return PauliHamiltonian(n, terms)


@pytest.mark.parametrize("precision", [C128, C64])
def test_energy_gradient_after_batch_pauli_reduction(precision):
    """15 qubits (at the width where mass and numerators are reduced after
    the batch kernel) with a Hamiltonian of every group shape."""
    n = 15
    circuit = workloads.hardware_efficient_ansatz(n, 1)
    theta = workloads.random_parameters(circuit.ParameterCount, 41)
    ham = _structured_hamiltonian(n, 5)
    e_want, g_want = oracle_gradient(circuit, theta, ham)
    scale = sum(abs(t.Coefficient) for t in ham.Terms)
    tol = (1e-11 if precision == C128 else 1e-4) * scale
    res = HipBatchedEnergyEvaluator(shared_backend(), precision).EvaluateWithGradient(circuit, ham, theta)
    assert abs(res.Energy - e_want) <= tol
    assert np.max(np.abs(np.array(res.Gradient) - g_want)) <= tol


def test_energy_gradient_batch_c64():
    circuit, theta, ham = hea_case(12, 2, 64, seed=3)
    e_want, g_want = oracle_gradient(circuit, theta, ham)
    res = HipBatchedEnergyEvaluator(shared_backend(), C64).EvaluateWithGradient(circuit, ham, theta)
    scale = sum(abs(t.Coefficient) for t in ham.Terms)
    assert abs(res.Energy - e_want) <= 1e-4 * scale
    assert np.max(np.abs(np.array(res.Gradient) - g_want)) <= 1e-4 * scale


def test_energy_batch_matches_bound_energies():
    circuit, theta, ham = hea_case(12, 2, 40, seed=8)
    ops, pidx = circuit.packed()
    paulis, coeffs = ham.packed()
    sets = np.stack([theta + 0.1 * k for k in range(5)])
    want = [oracle.energy_bound(12, ops, pidx, s, paulis, coeffs) for s in sets]
    got = HipBatchedEnergyEvaluator(shared_backend()).EvaluateMany(circuit, ham, sets)
    scale = sum(abs(c) for c in coeffs)
    assert np.max(np.abs(got - np.array(want))) <= 1e-11 * scale


def test_shared_ry_parameter_shift_closed_form():
    """Two Ry rotations bound to one parameter with Z on their qubit:
    E(theta) = cos(2 theta), dE/dtheta = -2 sin(2 theta)
    (VariationalEnergyTests.cs:17-30), widened to the batch minimum width."""
    n = batch_min_qubits(C128)
    pc = ParameterizedCircuit(n, 1)
    pc.Ry(0, 0).Ry(0, 0)
    ham = PauliHamiltonian(n, [PauliTerm(1.0, PauliString("I" * (n - 1) + "Z"))])
    for theta in (0.3, 1.1, -2.4):
        res = HipBatchedEnergyEvaluator(shared_backend()).EvaluateWithGradient(pc, ham, [theta])
        assert abs(res.Energy - math.cos(2 * theta)) <= 1e-12
        assert abs(res.Gradient[0] + 2 * math.sin(2 * theta)) <= 1e-12


def _near(expected, actual):
    """VariationalEnergyTests.cs:293: |expected - actual| <= 5e-12."""
    assert abs(expected - actual) <= 5e-12, (expected, actual)


def _single_ry(n):
    return ParameterizedCircuit(n, 1).Ry(0, 0)


def _z(n, coefficient=1.0):
    return PauliHamiltonian(n, [PauliTerm(coefficient, PauliString("I" * (n - 1) + "Z"))])


def _gradient_through(path, circuit, ham, theta):
    if path == "batch":
        return HipBatchedEnergyEvaluator(shared_backend()).EvaluateWithGradient(circuit, ham, theta)
    with HipVqeProgram(shared_backend(), circuit, ham) as program:
        return program.EvaluateWithGradient(circuit, ham, theta)


@pytest.mark.parametrize("path", ["batch", "program"])
def test_large_angle_parameter_shift_cases(path):
    """VariationalEnergyTests.cs:24-44 through qm_energy_gradient_batch and
    qm_vqe_program_gradient, widened to the batch minimum width with idle
    qubits: theta = 1e100 (where 1e100 + pi/2 == 1e100, so a shift added to the
    angle is lost), a 1e16 identity offset and a 1e308 coefficient."""
    n = batch_min_qubits(C128)
    huge = 1e100
    assert huge + math.pi / 2 == huge
    res = _gradient_through(path, _single_ry(n), _z(n), [huge])
    _near(math.cos(huge), res.Energy)
    _near(-math.sin(huge), res.Gradient[0])
    angle = 0.37
    offset = PauliHamiltonian(n, [PauliTerm(1e16, PauliString("I" * n)),
                                  PauliTerm(1.0, PauliString("I" * (n - 1) + "Z"))])
    _near(-math.sin(angle), _gradient_through(path, _single_ry(n), offset, [angle]).Gradient[0])
    large = _gradient_through(path, _single_ry(n), _z(n, 1e308), [math.pi / 2])
    assert abs(large.Gradient[0] / 1e308 + 1) <= 4e-12


@pytest.mark.parametrize("path", ["batch", "program"])
def test_hea_gradient_at_large_angles_matches_oracle(path):
    """A hardware-efficient ansatz with every angle near 1e6: the oracle
    appends the shift as a second rotation; the batched shift must agree to
    the 1e-11 relative bar (an angle-added shift loses ~1e6 * 2^-53 of it)."""
    n = 11
    circuit = workloads.hardware_efficient_ansatz(n, 2)
    theta = 1e6 + workloads.random_parameters(circuit.ParameterCount, 77)
    ham = workloads.random_hamiltonian(n, 48, 78)
    e_want, g_want = oracle_gradient(circuit, theta, ham)
    res = _gradient_through(path, circuit, ham, theta)
    scale = sum(abs(t.Coefficient) for t in ham.Terms)
    assert abs(res.Energy - e_want) <= 1e-11 * scale
    assert np.max(np.abs(np.array(res.Gradient) - g_want)) <= 1e-11 * scale
