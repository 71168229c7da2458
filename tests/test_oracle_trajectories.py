"""Pins the oracle's trajectory runner and probability estimator.

qmo_trajectories (NoisyStateVectorSimulator.Run, NoisyStateVectorSimulator.cs:36-62,
plus terminal shots) and qmo_estimate_probabilities (EstimateProbabilities,
:64-106) are compositions of the pinned gate, channel and measurement
routines; these tests fix their composition rules independently: the draw
order (channel draws in step order, then the MeasureAll draw), the cursor
accounting, and the Welford mean / standard error against numpy on
trajectories whose branches are forced by the chosen uniforms.
"""
import numpy as np

import oracle
from quantummatter_amd import NoiseChannel, NoisyCircuit, QuantumCircuit


def probabilities(psi):
    p = np.abs(psi) ** 2
    return p / p.sum()


def test_noiseless_trajectory_is_measure_all_of_the_ideal_state():
    n = 5
    c = QuantumCircuit(n)
    for q in range(n):
        c.H(q).Rz(q, 0.3 * (q + 1))
    nc = NoisyCircuit(n).Append(c)
    steps, ops = nc.packed()
    pools = np.array([[0.0], [0.37], [0.999]])
    consumed, outcomes, _ = oracle.trajectories(n, steps, ops, pools)
    cdf = np.cumsum(probabilities(oracle.execute(n, c.to_ops())))
    assert list(consumed) == [1, 1, 1]
    for u, k in zip(pools[:, 0], outcomes):
        assert k == int(np.searchsorted(cdf, u, side="right"))


def test_forced_branches_and_draw_order():
    """Bit flip p=1 draws nothing; depolarizing draws once (u in [p/4, 2p/4) -> Y);
    the MeasureAll draw comes last."""
    n = 3
    nc = NoisyCircuit(n)
    nc.ApplyNoise(NoiseChannel.BitFlip(1.0), 0)
    nc.ApplyNoise(NoiseChannel.Depolarizing(0.8), 2)
    steps, ops = nc.packed()
    pools = np.array([[0.3, 0.5], [0.1, 0.5], [0.9, 0.5]])
    consumed, outcomes, _ = oracle.trajectories(n, steps, ops, pools)
    assert list(consumed) == [2, 2, 2]
    # X on 0 -> |001>; then Y on 2 (|101>), X on 2 (|101>) or I (|001>).
    assert list(outcomes) == [0b101, 0b101, 0b001]


def test_terminal_shots_precede_the_measure_draw():
    n = 2
    nc = NoisyCircuit(n).Append(QuantumCircuit(n).H(0).H(1))
    steps, ops = nc.packed()
    pools = np.array([[0.6]])
    su = np.array([[0.1, 0.3, 0.6, 0.9]])
    consumed, outcomes, samples = oracle.trajectories(n, steps, ops, pools, 4, su)
    assert list(samples[0]) == [0, 1, 2, 3]
    assert outcomes[0] == 2 and consumed[0] == 1


def test_estimate_probabilities_matches_numpy_welford():
    """Half the trajectories flip qubit 1 (u < 0.5), half do not: the estimator's
    mean and standard error equal numpy's mean and std(ddof=1)/sqrt(M)."""
    n = 3
    c = QuantumCircuit(n).H(0).Ry(2, 0.9)
    nc = NoisyCircuit(n).Append(c).ApplyNoise(NoiseChannel.BitFlip(0.5), 1)
    steps, ops = nc.packed()
    M = 6
    pools = np.array([[0.2], [0.7], [0.4], [0.9], [0.1], [0.6]])
    means, se = oracle.estimate_probabilities(n, steps, ops, pools)
    base = oracle.execute(n, c.to_ops())
    flipped = oracle.execute(n, QuantumCircuit(n).H(0).Ry(2, 0.9).X(1).to_ops())
    rows = np.array([probabilities(flipped if u < 0.5 else base) for u in pools[:, 0]])
    assert np.allclose(means, rows.mean(axis=0), atol=1e-15)
    assert np.allclose(se, rows.std(axis=0, ddof=1) / np.sqrt(M), atol=1e-15)
