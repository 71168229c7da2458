"""Dense-block schedule on the host: the blocked matrices reproduce the oracle.

plan_dense multiplies runs of gates into 2^k x 2^k matrices (k <= 5) and the
host reference applies them with the kernel's index mapping; against the
oracle at 1e-12 for every gate kind and block width, and the block count
falls as the width grows.
"""
import numpy as np
import pytest

import oracle
from quantummatter_amd import workloads
from quantummatter_amd.backend import dense_apply_host
from test_plan_host import random_all_kinds


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_blocks_match_oracle_all_kinds(k):
    n = 9
    ops = random_all_kinds(n, 150, seed=k).to_ops()
    rng = np.random.default_rng(k)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    got, blocks = dense_apply_host(n, ops, psi, k)
    assert np.max(np.abs(got - oracle.execute(n, ops, psi))) <= 1e-12
    assert blocks >= 1


def test_wider_blocks_mean_fewer_passes():
    n = 12
    ops = workloads.random_layer_circuit(n, 6, seed=2).to_ops()
    want = oracle.execute(n, ops)
    counts = []
    for k in (2, 3, 4, 5):
        got, blocks = dense_apply_host(n, ops, oracle.zero_state(n), k)
        assert np.max(np.abs(got - want)) <= 1e-12
        counts.append(blocks)
    assert counts == sorted(counts, reverse=True) and counts[-1] < counts[0]
