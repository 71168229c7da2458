"""Every benchmark configuration at its full size on the device.

* 30 qubits, 40 layers (16 GiB c128 / 8 GiB c64 states): the compiled
  schedule against the per-gate schedule (the reference's exact arithmetic,
  bit-identical to the oracle at 20 qubits), |<a|b>| within 1e-12 of 1 and the
  mass within 1e-12 of 1; the complex64 compiled state against the complex128
  per-gate state, infidelity at most 1e-5; dense blocks of k = 2..5 qubits on
  the matrix cores over 2 layers against the per-gate schedule.
* Variational energy and parameter-shift gradient (hardware-efficient ansatz,
  6 layers, 256 Pauli terms) at 12, 16 and 20 qubits against the oracle's
  values committed in tests/golden/c4_hea_gradients.json: 1e-11 x sum|c| in
  complex128, 1e-4 x sum|c| in complex64 (FP64 reductions).
* Noisy trajectories at 16 qubits x 30 layers with depolarizing 1e-3 after
  every gate and damping 1e-3 per qubit per layer, 1024 terminal shots each:
  every sample, MeasureAll outcome and pool draw count equal to the oracle's.
* A 33-qubit complex128 state (128 GiB) resident beside a served stream of
  widths 10..28: the stream's results equal the same requests run alone, and
  the 33-qubit state returns to |0...0> after its circuit and the inverse.
"""
import json
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from hip_adapter import shared_backend
from quantummatter_amd import QuantumCircuit, workloads
from quantummatter_amd.backend import (C64, C128, SCHED_COMPILED, SCHED_FUSED, SCHED_PER_GATE,
                                       HipStateVectorBackend)
from quantummatter_amd.batched import HipBatchedEnergyEvaluator, HipTrajectoryRunner, HipVqeProgram
from quantummatter_amd.circuit import GateKind
from quantummatter_amd.serving import AMPLITUDES, SHOTS, HipRequestServer

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def inverse_of(c: QuantumCircuit) -> QuantumCircuit:
    """The inverse of a layer circuit (rotations and CNOTs): reversed order,
    negated angles."""
    inv = QuantumCircuit(c.QubitCount)
    for op in reversed(c.Operations):
        if op.Kind == GateKind.CNot:
            inv.CNot(op.Control, op.Target)
        elif op.Kind == GateKind.Rx:
            inv.Rx(op.Target, -op.Angle)
        elif op.Kind == GateKind.Ry:
            inv.Ry(op.Target, -op.Angle)
        elif op.Kind == GateKind.Rz:
            inv.Rz(op.Target, -op.Angle)
        else:
            raise AssertionError(op)
    return inv


# ---- 30 qubits, 40 layers ------------------------------------------------------


@pytest.fixture(scope="module")
def c2_per_gate():
    """The 30-qubit, 40-layer circuit under the per-gate schedule."""
    n = 30
    c = workloads.random_layer_circuit(n, 40, seed=30)
    ops = c.to_ops()
    s = shared_backend().CreateSession(n, C128)
    s.EvolveOps(ops, SCHED_PER_GATE)
    yield n, ops, s
    s.Dispose()


def test_c2_compiled_c128_matches_per_gate(c2_per_gate):
    n, ops, ref = c2_per_gate
    with shared_backend().CreateSession(n, C128) as s:
        s.EvolveOps(ops, SCHED_COMPILED)
        assert abs(s.Mass() - 1) <= 1e-12
        assert abs(1 - abs(s.Inner(ref))) <= 1e-12
    assert abs(ref.Mass() - 1) <= 1e-12


def test_c2_compiled_c64_infidelity(c2_per_gate):
    n, ops, ref = c2_per_gate
    with shared_backend().CreateSession(n, C64) as s:
        s.EvolveOps(ops, SCHED_COMPILED)
        overlap = s.Inner(ref)
        mass = s.Mass()
    infidelity = 1 - abs(overlap) ** 2 / mass
    assert infidelity <= 1e-5


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_c2_dense_blocks_30q(k):
    n = 30
    ops = workloads.random_layer_circuit(n, 2, seed=31).to_ops()
    with shared_backend().CreateSession(n, C128) as a, shared_backend().CreateSession(n, C128) as b:
        a.EvolveOps(ops, SCHED_PER_GATE)
        b.EvolveDense(workloads.random_layer_circuit(n, 2, seed=31), k)
        assert abs(1 - abs(b.Inner(a))) <= 1e-12
        assert abs(b.Mass() - 1) <= 1e-12


def test_c2_dense_blocks_c64_30q(c2_per_gate):
    """Dense k = 5 blocks on a 30-qubit complex64 state (FP32 matrix cores),
    2 layers, against the complex128 per-gate state of the same 2 layers."""
    n = 30
    c = workloads.random_layer_circuit(n, 2, seed=31)
    with shared_backend().CreateSession(n, C128) as a, shared_backend().CreateSession(n, C64) as b:
        a.EvolveOps(c.to_ops(), SCHED_PER_GATE)
        b.EvolveDense(c, 5)
        assert 1 - abs(b.Inner(a)) ** 2 / b.Mass() <= 1e-5


# ---- variational energy and gradient ---------------------------------------------


def _c4_cases():
    with open(os.path.join(HERE, "golden", "c4_hea_gradients.json")) as f:
        return {c["qubits"]: c for c in json.load(f)["cases"]}


@pytest.mark.parametrize("precision", [C128, C64])
@pytest.mark.parametrize("n", [12, 16, 20])
def test_c4_energy_gradient_full_size(n, precision):
    case = _c4_cases()[n]
    circuit = workloads.hardware_efficient_ansatz(n, case["layers"])
    theta = workloads.random_parameters(circuit.ParameterCount, case["theta_seed"])
    ham = workloads.random_hamiltonian(n, case["terms"], case["hamiltonian_seed"])
    scale = sum(abs(t.Coefficient) for t in ham.Terms)
    tol = (1e-11 if precision == C128 else 1e-4) * scale
    g_want = np.array(case["gradient"])
    res = HipBatchedEnergyEvaluator(shared_backend(), precision).EvaluateWithGradient(circuit, ham, theta)
    assert abs(res.Energy - case["energy"]) <= tol
    assert np.max(np.abs(np.array(res.Gradient) - g_want)) <= tol
    with HipVqeProgram(shared_backend(), circuit, ham, precision) as program:
        for _ in range(2):  # the capture and one replay
            res = program.EvaluateWithGradient(circuit, ham, theta)
            assert abs(res.Energy - case["energy"]) <= tol
            assert np.max(np.abs(np.array(res.Gradient) - g_want)) <= tol


# ---- noisy trajectories ---------------------------------------------------------


def test_c3_trajectories_full_circuit():
    n, T, shots = 16, 96, 1024
    nc = workloads.noisy_layer_circuit(n, 30, seed=16, p_depol=1e-3, gamma=1e-3)
    steps, ops = nc.packed()
    stride = int(np.sum(steps["is_channel"] == 1)) + 1
    pools = workloads.uniform_pools(316, T, stride)
    su = workloads.uniforms(317, T * shots).reshape(T, shots)
    got = HipTrajectoryRunner(shared_backend(), C128).Run(nc, pools, shots, su, True)

    def chunk(lo):  # the oracle's C code runs outside the GIL
        hi = min(lo + 12, T)
        return oracle.trajectories(n, steps, ops, pools[lo:hi], shots, su[lo:hi], True)

    with ThreadPoolExecutor(max_workers=8) as ex:
        parts = list(ex.map(chunk, range(0, T, 12)))
    consumed = np.concatenate([p[0] for p in parts])
    outcomes = np.concatenate([p[1] for p in parts])
    samples = np.concatenate([p[2] for p in parts])
    assert np.array_equal(got.consumed, consumed)
    assert np.array_equal(got.outcomes, outcomes)
    assert np.array_equal(got.samples, samples)
    # Damping draws happen: not every channel is skipped.
    assert np.all(consumed > 0)


# ---- a 33-qubit resident state beside a served stream ------------------------------


def test_c5_33q_resident_state_beside_served_stream():
    n_big = 33
    big_circuit = workloads.random_layer_circuit(n_big, 3, seed=33)
    budget_left = 64 << 30  # for the served stream, beside the 128 GiB state
    with HipStateVectorBackend() as be, be.CreateSession(n_big, C128) as big:
        big.Evolve(big_circuit, SCHED_COMPILED)
        assert abs(big.Mass() - 1) <= 1e-11
        zero_before = big.MarginalProbability(list(range(n_big)), 0)

        reqs = workloads.serving_requests(40, seed=21, min_qubits=10, max_qubits=28, min_depth=10,
                                          max_depth=30)
        assert max(r["qubits"] for r in reqs) >= 25
        payloads = [workloads.serving_payload(r) for r in reqs]
        with HipRequestServer(full_lanes=3, batch_max_qubits=14, device_budget_bytes=budget_left) as sv:
            ids = [sv.Submit(c, r["kind"], C128, u, obs) for r, (c, u, obs) in zip(reqs, payloads)]
            results = [sv.Wait(i) for i in ids]
            assert sv.Stats().peak_reserved_bytes <= budget_left
        for r, (c, u, obs), res in zip(reqs, payloads, results):
            with shared_backend().CreateSession(r["qubits"], C128) as s:
                s.Evolve(c, SCHED_FUSED)
                if r["kind"] == AMPLITUDES:
                    want = s.Snapshot()
                elif r["kind"] == SHOTS:
                    want = s.Sample(u)
                else:
                    want = np.array(s.ExpectationValues(obs))
            if r["kind"] == SHOTS:
                assert np.array_equal(res.value, want), r
            else:
                assert np.max(np.abs(np.asarray(res.value) - want)) <= 1e-12, r

        # The resident state was not disturbed by the stream.
        assert big.MarginalProbability(list(range(n_big)), 0) == zero_before
        big.Evolve(inverse_of(big_circuit), SCHED_FUSED)
        assert abs(big.Mass() - 1) <= 1e-11
        assert big.MarginalProbability(list(range(n_big)), 0) >= 1 - 1e-11
