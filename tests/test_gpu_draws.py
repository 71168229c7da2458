"""The caller's random source (qm_draw_fn) keeps its stream in step with the reference's.

CpuNoiseChannelBackend.Apply takes a System.Random and calls NextDouble() only
when a branch must be sampled: never at p == 0, not for a flip at p == 1, and
for amplitude damping only when the jump mass is positive and the stay mass
nonzero (CpuNoiseChannelBackend.cs:24-27,41-49,108-117).  A counting source
drives the resident (qm_state_noise_draw), drop-in host-buffer (qm_noise_apply),
trajectory (qm_state_noisy_execute_draw) and dynamic-program
(qm_state_dynamic_execute_draw) forms; every call count equals the draws the
oracle takes from a pool for the same case, and the states agree with the
oracle's (1e-12, complex128).  A source that fails aborts the call with the
pool status, and the drop-in form then leaves the caller's buffer unchanged.
"""
import math

import numpy as np
import pytest

import dynamic_cases as dc
import oracle
from hip_adapter import shared_backend
from quantummatter_amd import NoiseChannel, PoolExhaustedError, QuantumCircuit, workloads
from quantummatter_amd.backend import SCHED_FUSED, HipNoiseChannelBackend

pytestmark = pytest.mark.gpu


class CountingRandom:
    """System.Random's NextDouble() over a fixed uniform stream, counting calls."""

    def __init__(self, uniforms):
        self.u = np.asarray(uniforms, dtype=np.float64)
        self.calls = 0

    def NextDouble(self):
        v = self.u[self.calls]
        self.calls += 1
        return float(v)


class FailingRandom:
    def NextDouble(self):
        raise RuntimeError("source failed")


def states(n):
    """|0>, |1> on the target, a superposition, and a random state (target 1)."""
    zero = np.zeros(1 << n, dtype=np.complex128)
    zero[0] = 1
    one = np.zeros(1 << n, dtype=np.complex128)
    one[1 << 1] = 1
    sup = np.zeros(1 << n, dtype=np.complex128)
    sup[0] = math.sqrt(0.3)
    sup[1 << 1] = 1j * math.sqrt(0.7)
    rng = np.random.default_rng(n)
    rnd = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    rnd /= np.linalg.norm(rnd)
    return {"zero": zero, "one": one, "superposition": sup, "random": rnd}


CHANNELS = [("damping", NoiseChannel.AmplitudeDamping, 0.4), ("damping-full", NoiseChannel.AmplitudeDamping, 1.0),
            ("depolarizing", NoiseChannel.Depolarizing, 0.7), ("depolarizing-full", NoiseChannel.Depolarizing, 1.0),
            ("bitflip", NoiseChannel.BitFlip, 0.5), ("bitflip-full", NoiseChannel.BitFlip, 1.0),
            ("phaseflip", NoiseChannel.PhaseFlip, 0.5), ("zero", NoiseChannel.PhaseFlip, 0.0)]


@pytest.mark.parametrize("name,make,p", CHANNELS)
@pytest.mark.parametrize("state", ["zero", "one", "superposition", "random"])
def test_channel_draw_counts_and_states(name, make, p, state):
    n = 10
    psi = states(n)[state]
    ch = make(p)
    for seed in (1, 2, 3):
        u = workloads.uniforms(seed, 8)
        want, used = oracle.noise_apply(ch.Kind, p, 1, n, psi, u)
        # Resident state.
        src = CountingRandom(u)
        with shared_backend().CreateSession(n) as s:
            s.Upload(psi)
            s.ApplyNoiseRandom(ch, 1, src)
            got = s.Snapshot()
        assert src.calls == used, (name, state, seed)
        assert np.max(np.abs(got - want)) <= 1e-12
        # Drop-in host buffer.
        src = CountingRandom(u)
        host = psi.copy()
        HipNoiseChannelBackend(shared_backend()).Apply(ch, 1, host, src)
        assert src.calls == used
        assert np.max(np.abs(host - want)) <= 1e-12


def test_damping_draws_follow_the_reference_rule():
    """On |0> the jump mass is 0 (no draw); on |1> with gamma = 1 the stay mass
    is 0 (no draw, the jump is certain); otherwise one draw."""
    n = 10
    st = states(n)
    cases = [("zero", 0.4, 0), ("one", 0.4, 1), ("one", 1.0, 0), ("superposition", 0.4, 1),
             ("superposition", 1.0, 1)]
    for state, gamma, draws in cases:
        src = CountingRandom(workloads.uniforms(7, 4))
        with shared_backend().CreateSession(n) as s:
            s.Upload(st[state])
            s.ApplyNoiseRandom(NoiseChannel.AmplitudeDamping(gamma), 1, src)
        assert src.calls == draws, (state, gamma)


def test_trajectory_draws_match_oracle():
    n = 12
    nc = workloads.noisy_layer_circuit(n, 4, seed=5, p_depol=0.2, gamma=0.3)
    steps, ops = nc.packed()
    u = workloads.uniforms(11, 4096)
    want, used = oracle.noisy_execute(n, steps, ops, u)
    src = CountingRandom(u)
    with shared_backend().CreateSession(n) as s:
        s.EvolveTrajectoryRandom(nc, src)
        got = s.Snapshot()
    assert src.calls == used
    assert np.max(np.abs(got - want)) <= 1e-12


@pytest.mark.parametrize("n,seed", [(10, 1), (12, 2), (13, 3)])
def test_dynamic_program_draws_match_oracle(n, seed):
    c = dc.random_program(n, 6, 24, seed)
    u = np.random.default_rng(seed).random(64)
    steps, ops, sq, sc = c.packed()
    want, want_reg, want_used = oracle.dynamic_execute(n, steps, ops, sq, sc, u)
    src = CountingRandom(u)
    with shared_backend().CreateSession(n) as s:
        reg = s.EvolveDynamicRandom(c, src, SCHED_FUSED)
        got = s.Snapshot()
    assert (reg, src.calls) == (want_reg, want_used)
    assert np.max(np.abs(got - want)) <= 1e-12


def test_failing_source_aborts_and_drop_in_leaves_buffer_unchanged():
    n = 10
    psi = states(n)["superposition"]
    host = psi.copy()
    with pytest.raises(PoolExhaustedError):
        HipNoiseChannelBackend(shared_backend()).Apply(NoiseChannel.Depolarizing(0.5), 1, host, FailingRandom())
    assert np.array_equal(host, psi)
    with shared_backend().CreateSession(n) as s:
        s.Upload(psi)
        with pytest.raises(PoolExhaustedError):
            s.ApplyNoiseRandom(NoiseChannel.AmplitudeDamping(0.5), 1, FailingRandom())
        assert np.array_equal(s.Snapshot(), psi)
    # p == 0 never calls the source, even a failing one.
    HipNoiseChannelBackend(shared_backend()).Apply(NoiseChannel.BitFlip(0.0), 1, host, FailingRandom())
    assert np.array_equal(host, psi)


def test_drop_in_validation():
    from quantummatter_amd import ArgumentError, ArgumentOutOfRangeError
    be = HipNoiseChannelBackend(shared_backend())
    with pytest.raises(ArgumentError):
        be.Apply(NoiseChannel.BitFlip(0.5), 0, np.zeros(3, dtype=np.complex128), CountingRandom([0.1]))
    with pytest.raises(ArgumentOutOfRangeError):
        be.Apply(NoiseChannel.BitFlip(0.5), 4, np.zeros(8, dtype=np.complex128), CountingRandom([0.1]))


def test_simulator_run_counts_match_oracle_samples():
    """StateVectorSimulator.Run on the device: one draw per shot from the caller's
    source, the oracle's outcomes, counted and formatted as the reference does."""
    from quantummatter_amd.simulator import HipStateVectorSimulator, counts_of
    n = 12
    c = workloads.random_layer_circuit(n, 5, seed=44)
    u = workloads.uniforms(45, 512)
    src = CountingRandom(u)
    res = HipStateVectorSimulator(shared_backend()).Run(c, 512, src)
    assert src.calls == 512
    psi = oracle.execute(n, c.to_ops())
    assert res.Counts == counts_of(oracle.sample(n, psi, u), n)
    assert sum(res.Counts.values()) == 512
    assert np.max(np.abs(res.State - psi)) <= 1e-12
