"""StateVectorSimulator (src/QuantumMatter.Application/StateVectorSimulator.cs) on the device.

``Run`` evolves the circuit on a resident state, samples terminal shots with
the reference's strict-upper-bound CDF rule on the device, one draw per shot
in shot order (StateVectorSimulator.cs:44-90), and formats the counts as the
reference does (:92-98): each outcome's binary string padded to the qubit
count, in ordinal key order.  The random source is any object with
NextDouble() (System.Random's shape); the reference draws nothing else in
Run, so the shot draws are taken up front in shot order.
"""
from __future__ import annotations

from collections import OrderedDict
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from .backend import C128, SCHED_FUSED, HipStateVectorBackend
from .circuit import QuantumCircuit
from .errors import ArgumentError, ArgumentOutOfRangeError


def counts_of(outcomes, qubits: int) -> "OrderedDict[str, int]":
    """StateVectorSimulator.cs:92-98: bit strings of the outcomes, padded to
    `qubits` characters (qubit 0 rightmost), with their counts, in ordinal
    (byte-wise) key order."""
    tally: Dict[str, int] = {}
    for o in np.asarray(outcomes, dtype=np.int64).tolist():
        key = format(int(o), "b").rjust(qubits, "0")
        tally[key] = tally.get(key, 0) + 1
    return OrderedDict(sorted(tally.items(), key=lambda kv: kv[0].encode("ascii")))


@dataclass
class SimulationResult:
    """SimulationResult (StateVectorSimulator.cs:100): the final state and the counts."""
    State: np.ndarray
    Counts: "OrderedDict[str, int]"


class HipStateVectorSimulator:
    def __init__(self, backend: HipStateVectorBackend, schedule: int = SCHED_FUSED):
        if backend is None:
            raise ArgumentError("backend is null.")
        self._backend = backend
        self._schedule = schedule

    def Simulate(self, circuit: QuantumCircuit) -> np.ndarray:
        if circuit is None:
            raise ArgumentError("circuit is null.")
        with self._backend.CreateSession(circuit.QubitCount, C128) as s:
            s.Evolve(circuit, self._schedule)
            return s.Snapshot()

    def Run(self, circuit: QuantumCircuit, shots: int = 1024, random=None,
            uniforms: Optional[np.ndarray] = None) -> SimulationResult:
        """One evolution, `shots` terminal samples without collapse.  Draws come
        from `random.NextDouble()` (one per shot, in order) or, when given,
        from `uniforms[:shots]`."""
        if shots < 1:
            raise ArgumentOutOfRangeError("shots", "shots must be at least 1.")
        if circuit is None:
            raise ArgumentError("circuit is null.")
        if uniforms is None:
            if random is None:
                raise ArgumentError("a random source or uniforms are required.")
            uniforms = np.array([random.NextDouble() for _ in range(shots)], dtype=np.float64)
        u = np.ascontiguousarray(uniforms[:shots], dtype=np.float64)
        if len(u) != shots:
            raise ArgumentError("one uniform per shot is required.")
        with self._backend.CreateSession(circuit.QubitCount, C128) as s:
            s.Evolve(circuit, self._schedule)
            outcomes = s.Sample(u)
            state = s.Snapshot()
        return SimulationResult(state, counts_of(outcomes, circuit.QubitCount))
