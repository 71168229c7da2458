"""Dynamic programs on resident device states (qm_state_dynamic_execute).

The reference's DynamicCircuitTests.cs / NoisyDynamicCircuitTests.cs cases run
on the device with their expected registers, amplitudes and draw counts, under
the per-gate and fused schedules; random noisy dynamic programs (measurements,
conditions, resets, all channel kinds) give the oracle's classical register
and draw count exactly and its amplitudes within 1e-12 for the same pools.
"""
import numpy as np
import pytest

import dynamic_cases as dc
import oracle
from hip_adapter import shared_backend
from quantummatter_amd import ArgumentError, ArgumentOutOfRangeError, DynamicCircuit, PoolExhaustedError, QuantumCircuit
from quantummatter_amd.backend import SCHED_COMPILED, SCHED_FUSED, SCHED_PER_GATE

pytestmark = pytest.mark.gpu


def hip_run(schedule):
    def run(circuit, uniforms, state=None):
        with shared_backend().CreateSession(circuit.QubitCount) as s:
            if state is not None:
                s.Upload(np.ascontiguousarray(state, dtype=np.complex128))
            reg, used = s.EvolveDynamic(circuit, np.asarray(uniforms, dtype=np.float64), 0, schedule)
            return s.Snapshot(), reg, used
    return run


@pytest.mark.parametrize("schedule", [SCHED_PER_GATE, SCHED_FUSED])
@pytest.mark.parametrize("name", sorted(dc.CASES))
def test_reference_dynamic_case(name, schedule):
    print(f"dynamic case {name} schedule={schedule}", flush=True)
    dc.CASES[name](hip_run(schedule))


@pytest.mark.parametrize("n,seed,schedule", [(10, 1, SCHED_FUSED), (12, 2, SCHED_FUSED), (13, 3, SCHED_PER_GATE),
                                             (14, 4, SCHED_COMPILED), (16, 5, SCHED_FUSED)])
def test_random_programs_match_oracle(n, seed, schedule):
    c = dc.random_program(n, 6, 24, seed)
    u = np.random.default_rng(seed).random(64)
    steps, ops, sq, sc = c.packed()
    print(f"dynamic random n={n} seed={seed} schedule={schedule} steps={len(steps)}", flush=True)
    want, want_reg, want_used = oracle.dynamic_execute(n, steps, ops, sq, sc, u)
    got, reg, used = hip_run(schedule)(c, u)
    assert (reg, used) == (want_reg, want_used)
    assert np.max(np.abs(got - want)) <= 1e-12


def test_errors_leave_the_state_untouched():
    with shared_backend().CreateSession(2) as s:
        c = DynamicCircuit(2, 1).Append(QuantumCircuit(2).H(0)).Measure([0], [0])
        with pytest.raises(PoolExhaustedError):
            s.EvolveDynamic(c, np.zeros(0))
        with pytest.raises(ArgumentError):
            s.EvolveDynamic(DynamicCircuit(3, 1), np.zeros(1))
        # An invalid instruction (built around the builder's checks) is rejected before any work.
        steps, ops, sq, sc = DynamicCircuit(2, 1).Append(QuantumCircuit(2).X(0)).Measure([0], [0]).packed()
        sc[0] = 5
        from quantummatter_amd._lib import call, ptr
        reg = np.zeros(1, dtype=np.int32)
        cur = np.zeros(1, dtype=np.int64)
        u = np.zeros(1)
        s.Reset()
        with pytest.raises(ArgumentOutOfRangeError):
            call("qm_state_dynamic_execute", s._handle(), 1, ptr(steps), len(steps), ptr(ops), len(ops), ptr(sq),
                 ptr(sc), len(sq), SCHED_FUSED, ptr(u), 1, ptr(cur), ptr(reg), None)
        snap = s.Snapshot()
        assert snap[0] == 1 and np.all(snap[1:] == 0)
