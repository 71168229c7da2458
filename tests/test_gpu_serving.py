"""The request server: results are the same alone and in the mix.

A mixed set of requests (widths 8..20, so both the batched and the
full-chip lanes and the small-width fallback are used) is submitted at once;
every result is compared with the same request run alone through a resident
session: amplitudes within 1e-12 (c128), samples identical, expectation values
within 1e-12.  Admission keeps the reserved footprint within the budget and
rejects a request that can never fit.
"""
import numpy as np
import pytest

import oracle
from hip_adapter import shared_backend
from quantummatter_amd import InvalidOperationError, workloads
from quantummatter_amd.backend import C64, C128, SCHED_FUSED
from quantummatter_amd.circuit import pauli_array
from quantummatter_amd.serving import AMPLITUDES, EXPECTATION, SHOTS, HipRequestServer

pytestmark = pytest.mark.gpu


def alone(req, circuit, u, obs, precision=C128):
    with shared_backend().CreateSession(req["qubits"], precision) as s:
        s.Evolve(circuit, SCHED_FUSED)
        if req["kind"] == AMPLITUDES:
            return s.Snapshot()
        if req["kind"] == SHOTS:
            return s.Sample(u)
        return np.array(s.ExpectationValues(obs))


def check(req, got, want):
    if req["kind"] == SHOTS:
        assert np.array_equal(got, want), req
    else:
        assert np.max(np.abs(np.asarray(got) - np.asarray(want))) <= 1e-12, req


def test_mixed_stream_same_alone_and_in_mix():
    reqs = workloads.serving_requests(60, seed=11, min_qubits=8, max_qubits=20, min_depth=2,
                                      max_depth=12)
    payloads = [workloads.serving_payload(r) for r in reqs]
    with HipRequestServer(full_lanes=2, batch_max_qubits=14) as sv:
        ids = [sv.Submit(c, r["kind"], C128, u, obs) for r, (c, u, obs) in zip(reqs, payloads)]
        results = [sv.Wait(i) for i in ids]
        stats = sv.Stats()
    lanes = {res.lane for res in results}
    assert 0 in lanes and (lanes - {0}), lanes
    assert stats.batches >= 1 and stats.batched_requests >= 2
    for r, (c, u, obs), res in zip(reqs, payloads, results):
        check(r, res.value, alone(r, c, u, obs))


def test_server_results_match_oracle_small():
    reqs = workloads.serving_requests(12, seed=5, min_qubits=10, max_qubits=13, min_depth=3,
                                      max_depth=6)
    with HipRequestServer(batch_max_qubits=16) as sv:
        pay = [workloads.serving_payload(r) for r in reqs]
        ids = [sv.Submit(c, r["kind"], C128, u, obs) for r, (c, u, obs) in zip(reqs, pay)]
        res = [sv.Wait(i) for i in ids]
    for r, (c, u, obs), got in zip(reqs, pay, res):
        psi = oracle.execute(r["qubits"], c.to_ops())
        if r["kind"] == AMPLITUDES:
            assert np.max(np.abs(got.value - psi)) <= 1e-12
        elif r["kind"] == SHOTS:
            assert np.array_equal(got.value, oracle.sample(r["qubits"], psi, u))
        else:
            pa = pauli_array(obs)
            want = [oracle.expectation(r["qubits"], psi, pa[k:k + 1]) for k in range(len(pa))]
            assert np.max(np.abs(got.value - np.array(want))) <= 1e-12


def test_c64_requests_in_the_mix():
    reqs = workloads.serving_requests(16, seed=3, min_qubits=10, max_qubits=16, min_depth=2,
                                      max_depth=6)
    with HipRequestServer(batch_max_qubits=16) as sv:
        pay = [workloads.serving_payload(r) for r in reqs]
        ids = [sv.Submit(c, AMPLITUDES, C64) for r, (c, u, obs) in zip(reqs, pay)]
        res = [sv.Wait(i) for i in ids]
    for r, (c, u, obs), got in zip(reqs, pay, res):
        psi = oracle.execute(r["qubits"], c.to_ops())
        fid = abs(np.vdot(psi, got.value.astype(np.complex128))) ** 2
        assert 1 - fid <= 1e-5


def test_admission_respects_budget_and_rejects_oversize():
    budget = 3 * (16 << 20)  # three 20-qubit c128 states
    with HipRequestServer(full_lanes=2, batch_max_qubits=12, device_budget_bytes=budget) as sv:
        ids = [sv.Submit(workloads.random_layer_circuit(20, 3, seed=s), EXPECTATION, C128, None,
                         [workloads.random_hamiltonian(20, 4, s).Terms[0].Observable])
               for s in range(6)]
        for i in ids:
            sv.Wait(i)
        assert sv.Stats().peak_reserved_bytes <= budget
        big = sv.Submit(workloads.random_layer_circuit(22, 1, seed=1), AMPLITUDES, C128)
        with pytest.raises(InvalidOperationError):
            sv.Wait(big)


def test_small_request_over_a_tiny_budget_fails_instead_of_waiting():
    """A batched-width request whose state exceeds the whole budget is failed
    at submission (its wait reports the budget error), so wait, drain and
    dispose return."""
    with HipRequestServer(full_lanes=1, batch_max_qubits=14, device_budget_bytes=4096) as sv:
        rid = sv.Submit(workloads.random_layer_circuit(10, 2, seed=1), AMPLITUDES, C128)
        with pytest.raises(InvalidOperationError):
            sv.Wait(rid)
        sv.Drain()
        assert sv.Stats().failed == 1
