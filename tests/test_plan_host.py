"""CPU verification of the fused-pass planner and the tile kernel's indexing.

qm_plan_apply_host runs a fused schedule on the host with the tile kernel's
exact arithmetic and index mapping (loads, LDS transposes, register bits,
lane/outside-tile controls, diagonal placement).  Checked against the oracle
on random circuits covering every gate kind, for several tile shapes.
"""
import math

import numpy as np
import pytest

import oracle
from quantummatter_amd import GateKind, QuantumCircuit, SingleQubitUnitary, workloads
from quantummatter_amd.backend import plan_apply_host, plan_fused

SHAPES = [(12, 8, 3), (13, 8, 4), (10, 6, 3), (8, 5, 2), (9, 4, 3), (6, 3, 1)]


def random_all_kinds(n, gates, seed):
    rng = np.random.default_rng(seed)
    c = QuantumCircuit(n)
    for _ in range(gates):
        k = int(rng.integers(17))
        q = [int(x) for x in rng.permutation(n)[:3]]
        a = float(rng.normal() * 3)
        kind = GateKind(k)
        if kind in (GateKind.Rx, GateKind.Ry, GateKind.Rz, GateKind.Phase):
            getattr(c, kind.name)(q[0], a)
        elif kind == GateKind.CNot:
            c.CNot(q[1], q[0])
        elif kind == GateKind.Swap:
            c.Swap(q[0], q[1])
        elif kind == GateKind.Toffoli:
            c.Toffoli(q[1], q[2], q[0])
        elif kind == GateKind.Unitary:
            t = rng.random(3) * 2 * math.pi
            u = SingleQubitUnitary.Create(
                complex(math.cos(t[2]) * math.cos(t[0]), math.cos(t[2]) * math.sin(t[0])),
                -complex(math.sin(t[2]) * math.cos(t[1]), math.sin(t[2]) * math.sin(t[1])),
                complex(math.sin(t[2]) * math.cos(t[1]), -math.sin(t[2]) * math.sin(t[1])),
                complex(math.cos(t[2]) * math.cos(t[0]), -math.cos(t[2]) * math.sin(t[0])))
            if rng.random() < 0.5:
                c.ControlledU(q[1], q[0], u)
            else:
                c.U(q[0], u)
        else:
            if rng.random() < 0.3 and kind != GateKind.I:
                c.Controlled(kind, q[1], q[0], a)
            else:
                getattr(c, kind.name)(q[0])
    return c


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("n", [10, 13, 15])
def test_fused_plan_matches_oracle_all_kinds(shape, n):
    if n < shape[0]:
        pytest.skip("register narrower than the tile")
    c = random_all_kinds(n, 300, seed=n * 31 + shape[0])
    ops = c.to_ops()
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    exp = oracle.execute(n, ops, psi)
    got = plan_apply_host(n, ops, psi, *shape)
    assert np.max(np.abs(got - exp)) <= 1e-13


# (tile bits, lane bits, contiguous, remap): passes move qubits into the low
# index bits (1: with a transpose before the store when needed, 2: only where
# the store stays direct).
REMAP_SHAPES = [(12, 8, 6, 1), (13, 8, 4, 1), (13, 8, 5, 2), (10, 6, 4, 1), (8, 5, 3, 2), (9, 4, 4, 1),
                (6, 3, 3, 1), (13, 8, 4, 1, True), (12, 8, 3, 0, True), (10, 6, 4, 2, True), (8, 5, 3, 1, True)]


@pytest.mark.parametrize("shape", REMAP_SHAPES)
@pytest.mark.parametrize("n", [10, 13, 15])
def test_remapped_plan_matches_oracle_all_kinds(shape, n):
    """Every gate kind, controls on every kind of bit, with the low index bits
    reassigned between passes: the result is back in natural order and equals
    the oracle's."""
    if n < shape[0]:
        pytest.skip("register narrower than the tile")
    c = random_all_kinds(n, 300, seed=n * 37 + shape[0] + shape[2])
    ops = c.to_ops()
    rng = np.random.default_rng(n + 5)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    exp = oracle.execute(n, ops, psi)
    got = plan_apply_host(n, ops, psi, *shape)
    assert np.max(np.abs(got - exp)) <= 1e-13


@pytest.mark.parametrize("n", [12, 16, 20])
def test_remapped_plan_layer_circuit(n):
    c = workloads.random_layer_circuit(n, 12, seed=78)
    ops = c.to_ops()
    exp = oracle.execute(n, ops)
    for shape in [(12, 8, 6, 1), (13, 8, 6, 1), (13, 8, 4, 2), (12, 7, 5, 1), (13, 8, 4, 1, True),
                  (12, 8, 4, 0, True)]:
        if n < shape[0]:
            continue
        got = plan_apply_host(n, ops, oracle.zero_state(n), *shape)
        assert np.max(np.abs(got - exp)) <= 1e-13


@pytest.mark.parametrize("kind,n", [("all", 13), ("all", 15), ("layer", 16), ("layer", 20)])
@pytest.mark.parametrize("shape", [(13, 8, 3, 1, True), (14, 9, 3, 1, True), (12, 8, 4, 2, True)])
def test_replayed_plan_equals_fresh_plan(kind, n, shape):
    """A plan replayed from the decisions recorded on the same structure with
    other coefficients equals the fresh plan in every table and op (checked in
    the library) and its state equals the oracle's."""
    if n < shape[0]:
        pytest.skip("register narrower than the tile")
    c = random_all_kinds(n, 300, seed=n * 41 + shape[0]) if kind == "all" else \
        workloads.random_layer_circuit(n, 16, seed=n)
    ops = c.to_ops()
    exp = oracle.execute(n, ops)
    got = plan_apply_host(n, ops, oracle.zero_state(n), *shape, replay=True)
    assert np.max(np.abs(got - exp)) <= 1e-13


@pytest.mark.parametrize("kind,n", [("all", 13), ("all", 15), ("layer", 16), ("layer", 20)])
@pytest.mark.parametrize("shape", [(13, 8, 3, 1, True), (12, 8, 4, 2, True)])
def test_searched_plan_matches_oracle(kind, n, shape):
    """The compiled schedule's searched plan (perturbed restarts and iterated
    greedy moves over the tiles) replays from its recorded trace to the same
    plan (checked in the library) and its state equals the oracle's."""
    if n < shape[0]:
        pytest.skip("register narrower than the tile")
    c = random_all_kinds(n, 300, seed=n * 43 + shape[0]) if kind == "all" else \
        workloads.random_layer_circuit(n, 16, seed=n + 5)
    ops = c.to_ops()
    exp = oracle.execute(n, ops)
    got = plan_apply_host(n, ops, oracle.zero_state(n), *shape, searched=True)
    assert np.max(np.abs(got - exp)) <= 1e-13


@pytest.mark.parametrize("kind,n", [("all", 13), ("all", 15), ("layer", 18)])
@pytest.mark.parametrize("shape", [(13, 8, 3, 1, True), (12, 7, 4, 2, True)])
@pytest.mark.parametrize("how", ["plain", "replay", "searched"])
def test_lane_layout_plan_matches_oracle(kind, n, shape, how):
    """Lane layout: each pass stores under the index-bit permutation that puts
    the next pass's load lanes right above the contiguous run (tiles then
    store to other index bits than they load, so passes alternate between
    the state and a second buffer).  The state after the last pass, stored in
    natural order, equals the oracle's; replayed and searched plans are
    checked in the library as without the layout."""
    if n < shape[0]:
        pytest.skip("register narrower than the tile")
    c = random_all_kinds(n, 300, seed=n * 47 + shape[0]) if kind == "all" else \
        workloads.random_layer_circuit(n, 16, seed=n + 9)
    ops = c.to_ops()
    exp = oracle.execute(n, ops)
    got = plan_apply_host(n, ops, oracle.zero_state(n), *shape, lane_layout=True,
                          replay=how == "replay", searched=how == "searched")
    assert np.max(np.abs(got - exp)) <= 1e-13


def test_remapped_single_gates_are_exact_permutations():
    """A permutation gate under remapping moves amplitudes exactly."""
    n = 12
    psi = np.exp(1j * np.arange(1 << n) * 0.37) / math.sqrt(1 << n)
    for build in (lambda c: c.CNot(2, 9), lambda c: c.Swap(1, 8), lambda c: c.Toffoli(0, 5, 11),
                  lambda c: (c.X(7), c.Swap(0, 11), c.CNot(11, 3), c.Y(5))):
        c = QuantumCircuit(n)
        build(c)
        ops = c.to_ops()
        got, want = plan_apply_host(n, ops, psi, 12, 8, 6, True), oracle.execute(n, ops, psi)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [12, 16])
def test_fused_plan_layer_circuit(n):
    c = workloads.random_layer_circuit(n, 12, seed=77)
    ops = c.to_ops()
    exp = oracle.execute(n, ops)
    for shape in [(12, 8, 3), (13, 8, 4)]:
        if n < shape[0]:
            continue
        got = plan_apply_host(n, ops, oracle.zero_state(n), *shape)
        assert np.max(np.abs(got - exp)) <= 1e-13


def test_single_gates_in_one_tile():
    """Permutations are exact; arithmetic gates differ from the reference's
    rounding by at most the one rounding a fused multiply-add saves."""
    n = 12
    psi = np.exp(1j * np.arange(1 << n) * 0.37) / math.sqrt(1 << n)
    for build, exact in ((lambda c: c.Rx(3, 0.7), False), (lambda c: c.Ry(11, -2.1), False),
                         (lambda c: c.Rz(0, 0.3), False), (lambda c: c.H(7), False),
                         (lambda c: c.T(10), False), (lambda c: c.CNot(2, 9), True),
                         (lambda c: c.Y(4), True), (lambda c: c.Swap(1, 8), True),
                         (lambda c: c.Toffoli(0, 5, 11), True)):
        c = QuantumCircuit(n)
        build(c)
        ops = c.to_ops()
        got, want = plan_apply_host(n, ops, psi, 12, 8, 3), oracle.execute(n, ops, psi)
        if exact:
            assert np.array_equal(got, want)
        else:
            assert np.max(np.abs(got - want)) <= 2.0 ** -50 * np.max(np.abs(psi))


def test_schedule_shape_for_layer_circuits():
    """Passes per layer of the 30-qubit random layer circuit stay bounded."""
    c = workloads.random_layer_circuit(30, 40, seed=1)
    ops = c.to_ops()
    passes, transposes, gates = plan_fused(30, ops, 0)
    assert gates == 1780
    assert passes / 40 <= 4.0, passes
    passes64, _, _ = plan_fused(30, ops, 1)
    assert passes64 <= passes
