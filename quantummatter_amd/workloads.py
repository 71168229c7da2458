"""Seeded synthetic workloads: the uniform stream and the circuit generators.

The uniform stream is SplitMix64 (Steele, Lea, Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014) with the top 53 bits mapped to
[0,1).  Its state is a counter, so element i of the stream seeded with s is
mix(s + (i+1)*0x9E3779B97F4A7C15), which lets host code, the oracle and the
device library all draw identical uniforms and lets every trajectory own an
independent, seekable pool.

Generators (all from one stream, draws consumed in the order written):

* ``random_layer_circuit`` -- per layer, for q = 0..n-1: axis = floor(3u) in
  {Rx, Ry, Rz}, angle = 2*pi*u; then a CNOT brick (2i -> 2i+1) on even layers and
  (2i+1 -> 2i+2) on odd layers, the lower qubit controlling.
* ``hardware_efficient_ansatz`` -- L layers of [Ry(q), Rz(q) for every q] then a
  CNOT ladder 0->1->...->n-1; 2n parameters per layer.
* ``random_hamiltonian`` -- labels I/X/Y/Z uniform per site (floor(4u)),
  coefficients N(0,1) by Box-Muller, z = sqrt(-2 ln(1-u1)) cos(2 pi u2).
* ``noisy_layer_circuit`` -- the layer circuit with depolarizing(p) on every
  qubit a gate acts on, after that gate, and amplitude damping(gamma) on every
  qubit at the end of each layer.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .circuit import (GateKind, NoiseChannel, NoisyCircuit, ParameterizedCircuit, PauliHamiltonian,
                      PauliString, PauliTerm, QuantumCircuit)

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
C1 = 0xBF58476D1CE4E5B9
C2 = 0x94D049BB133111EB


def _mix(z: int) -> int:
    z = ((z ^ (z >> 30)) * C1) & M64
    z = ((z ^ (z >> 27)) * C2) & M64
    return z ^ (z >> 31)


class SplitMix64:
    """Sequential view of the stream (scalar draws for generators)."""

    def __init__(self, seed: int):
        self.state = seed & M64

    def next_u64(self) -> int:
        self.state = (self.state + GOLDEN) & M64
        return _mix(self.state)

    def next_double(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


def uniforms(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Elements [offset, offset+count) of the stream seeded with `seed`, vectorised."""
    idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & M64) + idx * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(C1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(C2)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def pool_seed(seed: int, index: int) -> int:
    """Seed of the independent pool number `index` derived from `seed`."""
    return _mix((seed ^ ((index + 1) * 0xD1B54A32D192ED03)) & M64)


def uniform_pools(seed: int, pools: int, stride: int) -> np.ndarray:
    """A (pools, stride) array; row t is the stream seeded with pool_seed(seed, t)."""
    out = np.empty((pools, stride), dtype=np.float64)
    for t in range(pools):
        out[t] = uniforms(pool_seed(seed, t), stride)
    return out


_AXES = (GateKind.Rx, GateKind.Ry, GateKind.Rz)


def _rotation(c: QuantumCircuit, axis: int, q: int, angle: float) -> None:
    (c.Rx, c.Ry, c.Rz)[axis](q, angle)


def brick_pairs(n: int, layer: int) -> List[Tuple[int, int]]:
    if layer % 2 == 0:
        return [(2 * i, 2 * i + 1) for i in range(n // 2)]
    return [(2 * i + 1, 2 * i + 2) for i in range((n - 1) // 2)]


def random_layer_circuit(n: int, layers: int, seed: int) -> QuantumCircuit:
    rng = SplitMix64(seed)
    c = QuantumCircuit(n)
    for layer in range(layers):
        for q in range(n):
            axis = min(int(rng.next_double() * 3), 2)
            _rotation(c, axis, q, 2 * math.pi * rng.next_double())
        for a, b in brick_pairs(n, layer):
            c.CNot(a, b)
    return c


def hardware_efficient_ansatz(n: int, layers: int) -> ParameterizedCircuit:
    pc = ParameterizedCircuit(n, 2 * n * layers)
    p = 0
    for _ in range(layers):
        for q in range(n):
            pc.Ry(q, p)
            pc.Rz(q, p + 1)
            p += 2
        ladder = QuantumCircuit(n)
        for q in range(n - 1):
            ladder.CNot(q, q + 1)
        pc.Append(ladder)
    return pc


def random_parameters(count: int, seed: int) -> np.ndarray:
    return 2 * math.pi * uniforms(seed, count)


def random_hamiltonian(n: int, terms: int, seed: int) -> PauliHamiltonian:
    rng = SplitMix64(seed)
    out = []
    for _ in range(terms):
        labels = ["IXYZ"[min(int(rng.next_double() * 4), 3)] for _ in range(n)]
        u1, u2 = rng.next_double(), rng.next_double()
        coeff = math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2 * math.pi * u2)
        out.append(PauliTerm(coeff, PauliString("".join(reversed(labels)))))
    return PauliHamiltonian(n, out)


def noisy_layer_circuit(n: int, layers: int, seed: int, p_depol: float = 1e-3,
                        gamma: float = 1e-3) -> NoisyCircuit:
    rng = SplitMix64(seed)
    nc = NoisyCircuit(n)
    dep = NoiseChannel.Depolarizing(p_depol)
    damp = NoiseChannel.AmplitudeDamping(gamma)
    for layer in range(layers):
        for q in range(n):
            axis = min(int(rng.next_double() * 3), 2)
            seg = QuantumCircuit(n)
            _rotation(seg, axis, q, 2 * math.pi * rng.next_double())
            nc.Append(seg).ApplyNoise(dep, q)
        for a, b in brick_pairs(n, layer):
            nc.Append(QuantumCircuit(n).CNot(a, b)).ApplyNoise(dep, a).ApplyNoise(dep, b)
        for q in range(n):
            nc.ApplyNoise(damp, q)
    return nc


def serving_requests(count: int, seed: int, rate: float = 0.0, min_qubits: int = 10,
                     max_qubits: int = 28, min_depth: int = 10, max_depth: int = 100) -> List[dict]:
    """The mixed request stream: width log-uniform over [min_qubits, max_qubits]
    (floor of a continuous log-uniform draw), depth uniform over
    [min_depth, max_depth] layers of random_layer_circuit, kind uniform over
    {0 amplitudes, 1 shots, 2 expectation}; arrivals Poisson at `rate` per
    second (all at time 0 when rate is 0).  Each request carries its own seed
    for its circuit, its 1024 shot uniforms or its 64-term Hamiltonian."""
    rng = SplitMix64(seed)
    out = []
    t = 0.0
    lo, hi = math.log(min_qubits), math.log(max_qubits + 1)
    for i in range(count):
        n = int(math.floor(math.exp(lo + rng.next_double() * (hi - lo))))
        n = min(max(n, min_qubits), max_qubits)
        depth = min_depth + min(int(rng.next_double() * (max_depth - min_depth + 1)),
                                max_depth - min_depth)
        kind = min(int(rng.next_double() * 3), 2)
        if rate > 0:
            t += -math.log(1.0 - rng.next_double()) / rate
        out.append({"index": i, "arrival": t, "qubits": n, "depth": depth, "kind": kind,
                    "seed": pool_seed(seed, i)})
    return out


def serving_payload(request: dict):
    """(circuit, shot uniforms or None, Pauli observables or None) of one request."""
    n, seed = request["qubits"], request["seed"]
    circuit = random_layer_circuit(n, request["depth"], seed=seed)
    u = uniforms(seed ^ 0x5A5A, 1024) if request["kind"] == 1 else None
    obs = ([t.Observable for t in random_hamiltonian(n, 64, seed ^ 0xA5A5).Terms]
           if request["kind"] == 2 else None)
    return circuit, u, obs
