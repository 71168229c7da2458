"""QM_SCHED_AUTO: interpreted on a structure's first run, compiled once its
kernels are in the cache (compiled in the background meanwhile).

The drop-in qm_execute uses this schedule: the first Execute of a circuit
structure compiles nothing on the caller's path, and after qm_jit_drain the
same structure with new angles runs through cache hits; every result stays
within 1e-12 of the oracle.  The request server takes the same schedule for
its full-chip lanes when asked.
"""
import numpy as np
import pytest

import oracle
from hip_adapter import shared_backend
from quantummatter_amd import workloads
from quantummatter_amd.backend import SCHED_AUTO, SCHED_PER_GATE, jit_drain, jit_stats
from quantummatter_amd.circuit import GateKind
from quantummatter_amd.serving import AMPLITUDES, EXPECTATION, SHOTS, HipRequestServer

pytestmark = pytest.mark.gpu


def with_angles(circuit, seed):
    """The same structure (gate kinds and qubits) with new rotation angles, as a
    variational step moves them: each angle shifted by a small amount (the
    compiled factored rotation form depends on which of |cos| and |sin| of the
    half angle is larger, so a structure's kernels are shared by angles on the
    same side of those boundaries)."""
    from quantummatter_amd import QuantumCircuit
    rng = np.random.default_rng(seed)
    out = QuantumCircuit(circuit.QubitCount)
    for op in circuit.Operations:
        if op.Kind == GateKind.CNot:
            out.CNot(op.Control, op.Target)
        else:
            getattr(out, op.Kind.name)(op.Target, op.Angle + float(rng.uniform(-1e-6, 1e-6)))
    return out


def test_execute_runs_compiled_after_the_background_compile():
    n = 17
    base = workloads.random_layer_circuit(n, 14, seed=1717)
    psi0 = np.zeros(1 << n, dtype=np.complex128)
    psi0[0] = 1
    be = shared_backend()
    jit_drain()
    compiled0, hits0, _ = jit_stats()
    a = psi0.copy()
    be.Execute(base, a)  # interpreted; its kernels are compiled in the background
    assert np.max(np.abs(a - oracle.execute(n, base.to_ops()))) <= 1e-12
    jit_drain()
    compiled1, hits1, _ = jit_stats()
    assert compiled1 > compiled0
    for seed in (1, 2):
        c = with_angles(base, seed)
        b = psi0.copy()
        be.Execute(c, b)  # every pass kernel a cache hit
        assert np.max(np.abs(b - oracle.execute(n, c.to_ops()))) <= 1e-12
    compiled2, hits2, _ = jit_stats()
    assert compiled2 == compiled1
    assert hits2 > hits1


def test_session_auto_schedule_matches_per_gate():
    n = 20
    c = workloads.random_layer_circuit(n, 10, seed=2020)
    be = shared_backend()
    with be.CreateSession(n) as ref, be.CreateSession(n) as s:
        ref.Evolve(c, SCHED_PER_GATE)
        for _ in range(2):
            s.Reset()
            s.Evolve(c, SCHED_AUTO)
            assert abs(1 - abs(s.Inner(ref))) <= 1e-12
            jit_drain()


def test_server_auto_schedule_repeated_structures():
    """A served family of one structure with new angles: the same results as the
    requests run alone."""
    n = 18
    base = workloads.random_layer_circuit(n, 12, seed=1818)
    circuits = [with_angles(base, k) for k in range(6)]
    with HipRequestServer(full_lanes=2, batch_max_qubits=12, schedule=SCHED_AUTO) as sv:
        first = sv.Submit(circuits[0], AMPLITUDES)
        sv.Wait(first)
        jit_drain()
        ids = [sv.Submit(c, AMPLITUDES) for c in circuits]
        results = [sv.Wait(i) for i in ids]
    for c, r in zip(circuits, results):
        assert np.max(np.abs(r.value - oracle.execute(n, c.to_ops()))) <= 1e-12
