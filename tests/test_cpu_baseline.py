"""The parallel CPU baseline (oracle/cpu_parallel.cpp) equals the scalar oracle bit for
bit under every dispatch mode of ParallelCpuStateVectorBackend.cs:246-276: AVX-512
(simd + avx512), 256-bit AVX with the SIMD Hadamard (simd only), and scalar, with
and without blocking."""
import numpy as np
import pytest

import oracle
from quantummatter_amd import workloads
from test_plan_host import random_all_kinds


@pytest.mark.parametrize("simd,avx512,blocking", [(True, True, True), (True, True, False),
                                                  (True, False, True), (True, False, False),
                                                  (False, False, True), (False, False, False)])
@pytest.mark.parametrize("workers", [1, 4])
def test_parallel_baseline_matches_scalar(simd, avx512, blocking, workers):
    for n, c in [(17, random_all_kinds(17, 120, seed=5)), (16, workloads.random_layer_circuit(16, 6, 3))]:
        ops = c.to_ops()
        exp = oracle.execute(n, ops)
        got = oracle.parallel_execute(n, ops, workers=workers, simd=simd, avx512=avx512, blocking=blocking)
        assert np.array_equal(got, exp)


def test_blocking_on_high_targets():
    from quantummatter_amd import QuantumCircuit
    n = 14
    c = QuantumCircuit(n)
    for k in range(20):
        c.H(13).Rz(13, 0.1 * k).Rx(13, 0.3).Y(13) if k % 2 else c.Ry(12, 0.2 * k).T(12)
    ops = c.to_ops()
    rng = np.random.default_rng(1)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    exp = oracle.execute(n, ops, psi)
    got = oracle.parallel_execute(n, ops, psi, workers=2)
    assert np.array_equal(got, exp)


def test_avx_paths_on_q0_q1_and_unit_halves():
    """The AVX kernel's q0 lane-permute path, H on q0 and q1 (SIMD Hadamard above
    q0), diagonals with identity halves (Z, S, T on q1..q3, skipped), odd ranges."""
    from quantummatter_amd import QuantumCircuit
    n = 9
    c = QuantumCircuit(n)
    for k in range(6):
        c.H(0).X(0).Rx(0, 0.3 + k).Ry(0, -0.7).Rz(0, 1.1).T(0).S(0).Z(0)
        c.H(1).Rz(1, 0.2 * k).T(1).S(2).Z(3).Phase(2, 0.4).Y(1).H(4).X(2)
    ops = c.to_ops()
    rng = np.random.default_rng(4)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    exp = oracle.execute(n, ops, psi)
    for avx512 in (False, True):
        got = oracle.parallel_execute(n, ops, psi, workers=1, simd=True, avx512=avx512, blocking=False)
        assert np.array_equal(got, exp)
