"""Dynamic-program cases restated from tests/QuantumMatter.Tests/DynamicCircuitTests.cs
and NoisyDynamicCircuitTests.cs.

Each case takes ``run(circuit, uniforms, state=None) -> (amplitudes, classical,
draws)`` and checks the reference's expected registers, amplitudes and draw
counts.  The same cases run on the oracle (test_oracle_dynamic.py) and on the
device (test_gpu_dynamic.py).
"""
from __future__ import annotations

import cmath
import math

import numpy as np

from quantummatter_amd import DynamicCircuit, NoiseChannel, NoisyDynamicCircuit, QuantumCircuit


def teleportation(theta, phi):
    return (DynamicCircuit(3, 2)
            .Append(QuantumCircuit(3).Ry(0, theta).Rz(0, phi).H(1).CNot(1, 2).CNot(0, 1).H(0))
            .Measure([0, 1], [0, 1])
            .If(2, 2, QuantumCircuit(3).X(2))
            .If(1, 1, QuantumCircuit(3).Z(2)))


def teleportation_every_branch(run):
    """DynamicCircuitTests.cs:15-66: the draw selects each measurement branch; the
    target qubit carries (alpha, beta) afterwards."""
    for theta, phi in [(0, 0), (math.pi, 0), (math.pi / 2, 0), (math.pi / 2, math.pi / 2), (0.73, -0.61)]:
        alpha = math.cos(theta / 2) * cmath.rect(1, -phi / 2)
        beta = math.sin(theta / 2) * cmath.rect(1, phi / 2)
        circuit = teleportation(theta, phi)
        for branch in range(4):
            zero, one = abs(alpha) ** 2, abs(beta) ** 2
            lower_is_alpha = (branch & 2) == 0
            sampled = branch | (0 if (zero >= one if lower_is_alpha else one >= zero) else 4)
            draw = 0.0
            for basis in range(sampled + 1):
                mass = zero / 4 if (((basis >> 1) ^ (basis >> 2)) & 1) == 0 else one / 4
                draw += mass / 2 if basis == sampled else mass
            amps, reg, draws = run(circuit, [draw])
            assert reg == branch and draws == 1
            for basis in range(8):
                want = alpha if basis == branch else beta if basis == (branch | 4) else 0
                assert abs(amps[basis] - want) <= 3e-12


def ordered_destinations(run):
    c = (DynamicCircuit(3, 3).Append(QuantumCircuit(3).X(0).X(1)).Measure([0, 1], [0, 2])
         .Append(QuantumCircuit(3).X(0)).Measure([0], [0]).If(4, 4, QuantumCircuit(3).X(2)))
    amps, reg, draws = run(c, [0.3, 0.8])
    assert reg == 4 and draws == 2
    assert abs(abs(amps[6]) ** 2 - 1) <= 1e-15


def reordered_measurements(run):
    for basis in range(16):
        prep = QuantumCircuit(4)
        for q in range(4):
            if basis & (1 << q):
                prep.X(q)
        c = DynamicCircuit(4, 5).Append(prep).Measure([3, 0, 2], [0, 4, 2])
        expected = ((basis >> 3) & 1) | ((basis & 1) << 4) | (basis & 4)
        assert run(c, [0.5])[1] == expected


def conditions_with_zero_values(run):
    for register in range(4):
        prep = QuantumCircuit(3)
        if register & 1:
            prep.X(0)
        if register & 2:
            prep.X(1)
        c = (DynamicCircuit(3, 2).Append(prep).Measure([0, 1], [0, 1])
             .If(3, 1, QuantumCircuit(3).X(2)).If(3, 0, QuantumCircuit(3).Z(2)))
        amps, reg, _ = run(c, [0.25])
        assert reg == register
        assert abs(abs(amps[register | 4 if register == 1 else register]) ** 2 - 1) <= 1e-15


def highest_classical_bit(run):
    c = DynamicCircuit(1, 30).Append(QuantumCircuit(1).X(0)).Measure([0], [29]).If(1 << 29, 1 << 29, QuantumCircuit(1).X(0))
    amps, reg, _ = run(c, [0.9])
    assert reg == 1 << 29
    assert abs(amps[0]) ** 2 == 1


def no_implicit_measurement(run):
    amps, reg, draws = run(DynamicCircuit(1, 3).Append(QuantumCircuit(1).H(0)), [])
    assert reg == 0 and draws == 0
    assert abs(abs(amps[1]) ** 2 - 0.5) <= 1e-12


def classical_memory_starts_at_zero(run):
    c = DynamicCircuit(1, 1).If(1, 0, QuantumCircuit(1).X(0)).Measure([0], [0])
    amps, reg, _ = run(c, [0.4])
    assert reg == 1
    amps, reg, _ = run(c, [0.4], state=amps)
    assert reg == 0


def skipped_conditions_and_empty_programs(run):
    amps, reg, draws = run(DynamicCircuit(1, 1), [])
    assert (reg, draws) == (0, 0) and amps[0] == 1
    amps, reg, draws = run(DynamicCircuit(1, 1).If(1, 1, QuantumCircuit(1).X(0)), [])
    assert (reg, draws) == (0, 0) and abs(amps[0]) ** 2 == 1


def reset_preserves_spectator_phase(run):
    c = DynamicCircuit(2, 1).Append(QuantumCircuit(2).X(0).S(0).H(1).S(1)).Measure([0], [0]).Reset(0)
    amps, reg, draws = run(c, [0.6, 0.1])
    assert reg == 1 and draws == 2
    assert abs(amps[0] - 1j / math.sqrt(2)) <= 1e-12
    assert abs(amps[2] + 1 / math.sqrt(2)) <= 1e-12


def reset_and_feedback(run):
    c = (DynamicCircuit(2, 2).Append(QuantumCircuit(2).H(0).CNot(0, 1)).Measure([1], [1]).Reset(0)
         .If(2, 2, QuantumCircuit(2).X(1)).Measure([0, 1], [0, 1]))
    rng = np.random.default_rng(39)
    for _ in range(16):
        assert run(c, rng.random(3))[1] == 0


def noisy_classical_history(run):
    high = 1 << 29
    c = (NoisyDynamicCircuit(2, 30).Append(QuantumCircuit(2).X(0)).Measure([0], [29])
         .ApplyNoise(NoiseChannel.BitFlip(1), 0).If(high, high, QuantumCircuit(2).X(1)).Measure([1], [0])
         .ApplyNoise(NoiseChannel.PhaseFlip(1), 1).Reset(0).Measure([0], [29])
         .If(1, 1, QuantumCircuit(2).X(0)))
    assert c.StepCount == 9
    for draws in ([0.1, 0.5, 0.7, 0.2], [0.9, 0.9, 0.9, 0.9]):
        amps, reg, used = run(c, draws)
        assert reg == 1 and used == 4
        assert amps[3] == -1


def zero_noise_equals_clean(run):
    prep = QuantumCircuit(2).H(0).S(0).H(1)
    corr = QuantumCircuit(2).Ry(1, 0.7).CNot(1, 0)
    clean = DynamicCircuit(2, 2).Append(prep).Measure([0], [1]).If(2, 2, corr).Reset(0).Measure([1], [0])
    noisy = (NoisyDynamicCircuit(2, 2).ApplyNoise(NoiseChannel.BitFlip(0), 0).Append(prep).Measure([0], [1])
             .ApplyNoise(NoiseChannel.PhaseFlip(0), 1).If(2, 2, corr).ApplyNoise(NoiseChannel.Depolarizing(0), 0)
             .Reset(0).Measure([1], [0]).ApplyNoise(NoiseChannel.AmplitudeDamping(0), 1))
    rng = np.random.default_rng(129)
    for _ in range(8):
        u = rng.random(3)
        a1, r1, d1 = run(clean, u)
        a2, r2, d2 = run(noisy, u)
        assert (r1, d1) == (r2, d2) and d1 == 3
        assert np.array_equal(a1, a2)


def noise_feedback_and_reset_spectators(run):
    program = (NoisyDynamicCircuit(2, 2).Append(QuantumCircuit(2).X(0).H(1).S(1))
               .ApplyNoise(NoiseChannel.AmplitudeDamping(1), 0).Measure([0], [1])
               .If(2, 0, QuantumCircuit(2).X(0)).ApplyNoise(NoiseChannel.PhaseFlip(1), 1)
               .Reset(0).Measure([0], [0]))
    amps, reg, _ = run(program, [0.3, 0.6, 0.9])
    assert reg == 0
    want = [math.sqrt(0.5), 0, -1j * math.sqrt(0.5), 0]
    assert np.max(np.abs(amps - want)) <= 4e-12


def bit_flip_then_measure(run):
    c = NoisyDynamicCircuit(1, 1).ApplyNoise(NoiseChannel.BitFlip(1), 0).Measure([0], [0])
    assert run(c, [0.5])[1:] == (1, 1)
    sampled = NoisyDynamicCircuit(1, 1).ApplyNoise(NoiseChannel.BitFlip(0.5), 0).Measure([0], [0])
    assert run(sampled, [0.2, 0.5])[1:] == (1, 2)  # 0.2 < 0.5 flips
    assert run(sampled, [0.7, 0.5])[1:] == (0, 2)


CASES = {f.__name__: f for f in [
    teleportation_every_branch, ordered_destinations, reordered_measurements, conditions_with_zero_values,
    highest_classical_bit, no_implicit_measurement, classical_memory_starts_at_zero,
    skipped_conditions_and_empty_programs, reset_preserves_spectator_phase, reset_and_feedback,
    noisy_classical_history, zero_noise_equals_clean, noise_feedback_and_reset_spectators,
    bit_flip_then_measure]}


def random_program(n, cbits, steps, seed, noisy=True):
    """A random dynamic program: layers of random gates, measurements into random
    classical bits, conditioned segments, resets and channels."""
    from quantummatter_amd import workloads
    rng = np.random.default_rng(seed)
    c = NoisyDynamicCircuit(n, cbits) if noisy else DynamicCircuit(n, cbits)
    for k in range(steps):
        r = rng.random()
        if r < 0.35:
            seg = workloads.random_layer_circuit(n, 1, seed=seed * 100 + k)
            c.Append(seg)
        elif r < 0.55:
            m = int(rng.integers(1, min(n, cbits, 4) + 1))
            qs = rng.choice(n, m, replace=False).tolist()
            cs = rng.choice(cbits, m, replace=False).tolist()
            c.Measure(qs, cs)
        elif r < 0.75:
            mask = int(rng.integers(1, 1 << cbits))
            value = int(rng.integers(0, 1 << cbits)) & mask
            c.If(mask, value, workloads.random_layer_circuit(n, 1, seed=seed * 100 + k + 50))
        elif r < 0.85:
            c.Reset(int(rng.integers(0, n)))
        elif noisy:
            kind = int(rng.integers(0, 4))
            p = float(rng.choice([0.0, 0.3, 1.0]))
            ch = [NoiseChannel.BitFlip, NoiseChannel.PhaseFlip, NoiseChannel.Depolarizing,
                  NoiseChannel.AmplitudeDamping][kind](p)
            c.ApplyNoise(ch, int(rng.integers(0, n)))
    return c
