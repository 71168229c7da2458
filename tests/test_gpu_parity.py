"""Parity of the HIP library with the CPU oracle on the benchmark workloads.

Tolerances (from the backend contract): complex128 amplitudes within 1e-12
of the oracle (the per-gate schedule is bit-identical); sampled, measured and
Kraus-branch outcomes identical for the same uniforms; complex64 infidelity
1 - |<psi64|psi128>|^2 at most 1e-5.
"""
import math

import numpy as np
import pytest

import oracle
from hip_adapter import shared_backend
from quantummatter_amd import (InvalidOperationError, NoiseChannel, ObjectDisposedError,
                               OperationCanceledError, PauliString, QuantumCircuit, workloads)
from quantummatter_amd.backend import (C64, C128, SCHED_FUSED, SCHED_PER_GATE, CancellationToken,
                                       HipBackendOptions, HipStateVectorBackend)
from test_plan_host import random_all_kinds

pytestmark = pytest.mark.gpu


def session(n, precision=C128, state=None):
    s = shared_backend().CreateSession(n, precision)
    if state is not None:
        s.Upload(state)
    return s


@pytest.fixture(scope="module")
def c1():
    """20-qubit, 20-layer seeded random circuit and its oracle state."""
    c = workloads.random_layer_circuit(20, 20, seed=1)
    ops = c.to_ops()
    return c, ops, oracle.execute(20, ops)


def test_c1_per_gate_is_bit_identical(c1):
    c, ops, expected = c1
    with session(20) as s:
        s.Evolve(c, SCHED_PER_GATE)
        got = s.Snapshot()
    assert np.array_equal(got, expected)


def test_c1_fused_within_1e12(c1):
    c, ops, expected = c1
    with session(20) as s:
        s.Evolve(c, SCHED_FUSED)
        got = s.Snapshot()
    assert np.max(np.abs(got - expected)) <= 1e-12


@pytest.mark.parametrize("schedule", [SCHED_PER_GATE, SCHED_FUSED])
def test_c1_shots_identical(c1, schedule):
    c, ops, expected = c1
    u = workloads.uniforms(4242, 1024)
    want = oracle.sample(20, expected, u)
    with session(20) as s:
        s.Evolve(c, schedule)
        got = s.Sample(u)
        print("boundary fallbacks:", s.BoundaryFallbacks())
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [12, 13, 16])
@pytest.mark.parametrize("schedule", [SCHED_PER_GATE, SCHED_FUSED])
def test_all_gate_kinds_match_oracle(n, schedule):
    c = random_all_kinds(n, 400, seed=1000 + n)
    ops = c.to_ops()
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    expected = oracle.execute(n, ops, psi)
    with session(n, state=psi) as s:
        s.Evolve(c, schedule)
        got = s.Snapshot()
    if schedule == SCHED_PER_GATE:
        assert np.array_equal(got, expected)
    else:
        assert np.max(np.abs(got - expected)) <= 1e-12


def test_c64_infidelity(c1):
    c, ops, expected = c1
    with session(20, C64) as s:
        s.Evolve(c, SCHED_FUSED)
        got = s.Snapshot()
    overlap = np.vdot(expected, got)
    infidelity = 1 - abs(overlap) ** 2 / (np.vdot(got, got).real * np.vdot(expected, expected).real)
    assert infidelity <= 1e-5


def test_c64_device_inner_product_matches_host(c1):
    c, ops, expected = c1
    with session(20) as a, session(20, C64) as b:
        a.Evolve(c)
        b.Evolve(c)
        ip = a.Inner(b)
        host = np.vdot(a.Snapshot(), b.Snapshot())
    assert abs(ip - host) <= 1e-12


@pytest.mark.parametrize("schedule", [SCHED_PER_GATE, SCHED_FUSED])
def test_noisy_trajectories_match_oracle(schedule):
    n = 12
    nc = workloads.noisy_layer_circuit(n, 4, seed=9, p_depol=0.05, gamma=0.08)
    steps, ops = nc.packed()
    for t in range(6):
        pool = workloads.uniforms(workloads.pool_seed(5, t), 600)
        exp_state, exp_cursor = oracle.noisy_execute(n, steps, ops, pool)
        with session(n) as s:
            cursor = s.EvolveTrajectory(nc, pool)
            got = s.Snapshot()
            assert cursor == exp_cursor
            assert np.max(np.abs(got - exp_state)) <= 1e-12
            u = pool[cursor]
            want, _ = oracle.measure_all(n, exp_state, u)
            assert s.MeasureAll(u) == want


def test_measure_subset_and_marginal_match_oracle(c1):
    c, ops, expected = c1
    sel = [3, 17, 0]
    for u in (0.1, 0.5, 0.93):
        want, want_state = oracle.measure(20, expected, sel, u)
        with session(20, state=expected) as s:
            p = s.MarginalProbability(sel, 5)
            assert abs(p - oracle.marginal(20, expected, sel, 5)) <= 1e-12
            assert s.Measure(sel, u) == want
            assert np.max(np.abs(s.Snapshot() - want_state)) <= 1e-12


def test_expectations_and_energy_match_oracle(c1):
    c, ops, expected = c1
    ham = workloads.random_hamiltonian(20, 64, seed=3)
    paulis, coeffs = ham.packed()
    with session(20, state=expected) as s:
        vals = s.ExpectationValues([t.Observable for t in ham.Terms])
        e = s.ExpectationEnergy(ham)
    for k in range(len(coeffs)):
        assert abs(vals[k] - oracle.expectation(20, expected, paulis[k:k + 1])) <= 1e-12
    assert abs(e - oracle.energy(20, expected, paulis, coeffs)) <= 1e-11


def test_large_register_properties():
    """26 qubits: fused equals per-gate, the norm is kept, and the inverse returns |0>."""
    n = 26
    c = workloads.random_layer_circuit(n, 6, seed=11)
    inv = QuantumCircuit(n)
    for op in reversed(c.Operations):
        if op.Kind.name == "CNot":
            inv.CNot(op.Control, op.Target)
        else:
            getattr(inv, op.Kind.name)(op.Target, -op.Angle)
    with session(n) as a, session(n) as b:
        a.Evolve(c, SCHED_FUSED)
        b.Evolve(c, SCHED_PER_GATE)
        assert abs(a.Inner(b) - 1) <= 1e-12
        assert abs(a.Mass() - 1) <= 1e-12
        a.Evolve(inv, SCHED_FUSED)
        assert abs(a.MarginalProbability(list(range(n)), 0) - 1) <= 1e-12


def test_drop_in_execute_contract():
    n = 14
    c = workloads.random_layer_circuit(n, 4, seed=2)
    ops = c.to_ops()
    psi = np.exp(1j * 0.17 * np.arange(1 << n)) / math.sqrt(1 << n)
    expected = oracle.execute(n, ops, psi)
    with HipStateVectorBackend(HipBackendOptions(EnableProfiling=True)) as be:
        buf = psi.copy()
        prof = be.ExecuteProfiled(c, buf)
        assert np.max(np.abs(buf - expected)) <= 1e-12
        assert prof.KernelLaunches >= 1 and prof.StateBytes == 16 << n
        # Cancellation before work leaves the input unchanged.
        tok = CancellationToken()
        tok.Cancel()
        buf = psi.copy()
        with pytest.raises(OperationCanceledError):
            be.Execute(c, buf, tok)
        assert np.array_equal(buf, psi)
    # Budget rejection happens before allocation; input unchanged; backend reusable.
    with HipStateVectorBackend(HipBackendOptions(DeviceMemoryBudgetBytes=1024)) as small:
        buf = psi.copy()
        with pytest.raises(InvalidOperationError):
            small.Execute(c, buf)
        assert np.array_equal(buf, psi)
        tiny = QuantumCircuit(4).H(0)
        t = np.zeros(16, complex)
        t[0] = 1
        small.Execute(tiny, t)
        assert abs(t[1] - math.sqrt(0.5)) <= 1e-15
    # Profiling must be enabled at construction; use after dispose fails.
    be = HipStateVectorBackend()
    with pytest.raises(InvalidOperationError):
        be.ExecuteProfiled(c, psi.copy())
    be.Dispose()
    be.Dispose()
    with pytest.raises(ObjectDisposedError):
        be.Execute(c, psi.copy())


def test_depolarizing_and_damping_branches_on_device():
    n = 13
    rng = np.random.default_rng(4)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    for kind in range(4):
        for u in (0.01, 0.3, 0.6, 0.9):
            ch = NoiseChannel(kind, 0.37)
            want, used = oracle.noise_apply(kind, 0.37, 5, n, psi, np.array([u]))
            with session(n, state=psi) as s:
                cur = s.ApplyNoise(ch, 5, np.array([u]))
                assert cur == used
                assert np.max(np.abs(s.Snapshot() - want)) <= 1e-12


@pytest.mark.parametrize("n,layers", [(22, 3), (24, 1)])
def test_wide_sampling_is_exact(n, layers):
    """At these widths the sequential-sum rounding exceeds single probabilities,
    so only an exact emulation of the reference's cumulative gives its samples."""
    c = workloads.random_layer_circuit(n, layers, seed=n)
    psi = oracle.execute(n, c.to_ops())
    u = workloads.uniforms(77 + n, 1024)
    want = oracle.sample(n, psi, u)
    with session(n, state=psi) as s:
        got = s.Sample(u)
        assert s.BoundaryFallbacks() == 0
    assert np.array_equal(got, want)


def test_uniform_superposition_sampling_is_exact():
    """Equal probabilities 2^-n: every cumulative is exact, draws land on boundaries."""
    n = 20
    c = QuantumCircuit(n)
    for q in range(n):
        c.H(q)
    psi = oracle.execute(n, c.to_ops())
    u = np.concatenate([np.arange(0, 1024) / 1024.0, workloads.uniforms(5, 1024)])
    want = oracle.sample(n, psi, u)
    with session(n, state=psi) as s:
        got = s.Sample(u)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [12, 13, 14, 16])
def test_compiled_schedule_all_kinds(n):
    from quantummatter_amd.backend import SCHED_COMPILED
    c = random_all_kinds(n, 240, seed=100 + n)
    ops = c.to_ops()
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = oracle.execute(n, ops, psi)
    with session(n, state=psi) as s:
        s.EvolveOps(ops, SCHED_COMPILED)
        got = s.Snapshot()
    assert np.max(np.abs(got - want)) <= 1e-12


def test_compiled_schedule_c1_and_cache(c1):
    from quantummatter_amd.backend import SCHED_COMPILED
    c, ops, expected = c1
    with session(20) as s:
        s.Evolve(c, SCHED_COMPILED)
        got = s.Snapshot()
        s.Reset()
        s.Evolve(c, SCHED_COMPILED)  # second run: every pass kernel from the cache
        again = s.Snapshot()
    assert np.max(np.abs(got - expected)) <= 1e-12
    assert np.array_equal(got, again)


def test_compiled_schedule_c64_infidelity():
    from quantummatter_amd.backend import SCHED_COMPILED
    n = 16
    c = workloads.random_layer_circuit(n, 8, seed=4)
    want = oracle.execute(n, c.to_ops())
    with session(n, C64) as s:
        s.Evolve(c, SCHED_COMPILED)
        got = s.Snapshot()
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-5


def test_compiled_schedule_c64_all_kinds():
    """Every gate kind through the complex64 generated kernels at the default
    16384-amplitude tiles (folded LDS offsets), against the oracle."""
    from quantummatter_amd.backend import SCHED_COMPILED
    n = 15
    c = random_all_kinds(n, 240, seed=515)
    want = oracle.execute(n, c.to_ops())
    with session(n, C64) as s:
        s.Evolve(c, SCHED_COMPILED)
        got = s.Snapshot()
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-5


@pytest.mark.parametrize("precision", [C128, C64])
def test_compiled_schedule_explicit_8192_tiles(precision):
    """The previous default shape (8192-amplitude tiles of 256 threads, two
    workgroups per CU), selected explicitly on a private context."""
    from quantummatter_amd.backend import SCHED_COMPILED, HipStateVectorBackend
    n = 17
    c = workloads.random_layer_circuit(n, 12, seed=17)
    want = oracle.execute(n, c.to_ops())
    with HipStateVectorBackend() as be:
        be.SetFusedShape(precision, 13, 8)
        assert be.GetFusedShape(precision) == (13, 8)
        with be.CreateSession(n, precision) as s:
            s.Evolve(c, SCHED_COMPILED)
            got = s.Snapshot()
    assert 1 - abs(np.vdot(want, got)) ** 2 <= (1e-5 if precision == C64 else 1e-12)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("n", [10, 13])
def test_dense_blocks_on_matrix_cores(n, k):
    c = random_all_kinds(n, 160, seed=200 + 7 * n + k)
    ops = c.to_ops()
    rng = np.random.default_rng(k)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = oracle.execute(n, ops, psi)
    with session(n, state=psi) as s:
        s.EvolveDense(c, k)
        got = s.Snapshot()
    assert np.max(np.abs(got - want)) <= 1e-12


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("n", [10, 13])
def test_dense_blocks_c64_on_fp32_matrix_cores(n, k):
    """The complex64 dense variant (FP32 MFMA, block matrices rounded to FP32):
    infidelity against the oracle's complex128 state at most 1e-5, every block
    width (the FP32 output layout is checked here: a wrong row order would
    scramble the state)."""
    c = random_all_kinds(n, 160, seed=300 + 7 * n + k)
    ops = c.to_ops()
    rng = np.random.default_rng(k)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = oracle.execute(n, ops, psi)
    with session(n, C64, state=psi) as s:
        s.EvolveDense(c, k)
        got = s.Snapshot()
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-5
    assert np.max(np.abs(got - want)) <= 1e-5


def test_dense_blocks_layer_circuit_16q():
    n = 16
    c = workloads.random_layer_circuit(n, 10, seed=8)
    want = oracle.execute(n, c.to_ops())
    with session(n) as s:
        s.EvolveDense(c, 5)
        got = s.Snapshot()
    assert np.max(np.abs(got - want)) <= 1e-12


def test_compiled_schedule_plans_with_both_low_bit_counts():
    """The complex128 compiled schedule plans with 3 and with 4 low bits kept
    in the tiles (full and 9-phase budgets) and runs the plan of lowest
    modelled time; on this circuit that is the 3-bit plan at the default
    8192-amplitude tiles (14 passes against 16 with 4 bits; qm_jit_compile_check
    plans with 4), and the state matches the per-gate schedule, also when the
    plan is replayed from the cached decisions (second call, new angles)."""
    import ctypes as C
    from quantummatter_amd._lib import lib, ptr
    from quantummatter_amd.backend import (SCHED_COMPILED, SCHED_PER_GATE, HipBackendOptions,
                                           HipStateVectorBackend)
    n = 26
    ops = workloads.random_layer_circuit(n, 40, seed=2).to_ops()
    four = C.c_int32()
    with HipStateVectorBackend(HipBackendOptions(EnableProfiling=True)) as be:
        K, L = be.GetFusedShape(C128)
        assert lib().qm_jit_compile_check(n, 0, ptr(ops), len(ops), K, L, C.byref(four)) == 0
        with be.CreateSession(n, C128) as a, be.CreateSession(n, C128) as b:
            a.EvolveOps(ops, SCHED_PER_GATE)
            p = b.EvolveOps(ops, SCHED_COMPILED, profile=True)
            assert p.Passes <= four.value
            assert abs(1 - abs(b.Inner(a))) <= 1e-12
            ops2 = ops.copy()
            ops2["angle"] = ops2["angle"] * 0.7 + 0.3
            a.Reset()
            b.Reset()
            a.EvolveOps(ops2, SCHED_PER_GATE)
            p2 = b.EvolveOps(ops2, SCHED_COMPILED, profile=True)
            assert p2.Passes == p.Passes
            assert abs(1 - abs(b.Inner(a))) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("n,layers,seed", [(25, 40, 2), (28, 40, 1), (29, 12, 1)])
def test_compiled_schedule_wide_layer_circuits(n, layers, seed):
    """Searched, remapped compiled passes at widths where every tile holds a
    minority of the qubits, against the per-gate schedule (exact arithmetic of
    the reference): the states agree to rounding."""
    from quantummatter_amd.backend import SCHED_COMPILED, SCHED_PER_GATE
    ops = workloads.random_layer_circuit(n, layers, seed=seed).to_ops()
    with session(n) as a, session(n) as b:
        a.EvolveOps(ops, SCHED_PER_GATE)
        b.EvolveOps(ops, SCHED_COMPILED)
        assert abs(1 - abs(b.Inner(a))) <= 1e-12
        assert abs(b.Mass() - 1) <= 1e-12


def test_known_wrong_code_variant_is_refused_on_the_device():
    """Factored forms without per-tile coefficient loads (QM_JIT_VARIANT=17)
    computed wrong device states for some passes; the library refuses to load
    such a variant (NotSupported) before compiling anything, while the default
    runs.  Checked in a child process: the variant is read once per process."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "from quantummatter_amd import workloads\n"
        "from quantummatter_amd.backend import HipStateVectorBackend, SCHED_COMPILED\n"
        "c = workloads.random_layer_circuit(16, 4, seed=3)\n"
        "with HipStateVectorBackend() as be, be.CreateSession(16) as s:\n"
        "    try:\n"
        "        s.Evolve(c, SCHED_COMPILED)\n"
        "        print('ran')\n"
        "    except NotImplementedError as e:\n"
        "        print('refused', e)\n")
    for variant, want in (("17", "refused"), ("11665", "ran")):
        env = dict(os.environ, QM_JIT_VARIANT=variant, PYTHONPATH=root)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        assert r.stdout.split()[0] == want, r.stdout
