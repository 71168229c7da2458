"""The autotuner (qm_autotune) on the device.

It times every tile shape of a schedule on a prefix of the caller's circuit,
keeps the fastest for that width, and later executions of the width use it:
their pass count is the winning row's, and their amplitudes stay within 1e-12
of the oracle (complex128) or 1e-5 infidelity (complex64).  An explicit
qm_ctx_set_fused_shape replaces the tuned shape.
"""
import numpy as np
import pytest

import oracle
from quantummatter_amd import ArgumentError
from quantummatter_amd.backend import (C64, C128, SCHED_COMPILED, SCHED_FUSED, SCHED_PER_GATE,
                                       HipBackendOptions, HipStateVectorBackend)
from test_plan_host import random_all_kinds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    be = HipStateVectorBackend(HipBackendOptions(EnableProfiling=True))
    yield be
    be.Dispose()


def evolve(be, n, ops, precision, schedule):
    with be.CreateSession(n, precision) as s:
        prof = s.EvolveOps(ops, schedule, profile=True)
        return s.Snapshot(), prof


@pytest.mark.parametrize("schedule,precision", [(SCHED_FUSED, C128), (SCHED_COMPILED, C128),
                                                (SCHED_FUSED, C64), (SCHED_COMPILED, C64)])
def test_tuned_shape_is_used_and_exact(backend, schedule, precision):
    n = 14
    c = random_all_kinds(n, 300, seed=5 + schedule)
    ops = c.to_ops()
    print(f"autotune n={n} schedule={schedule} precision={precision}", flush=True)
    best, rows = backend.Autotune(c, precision, schedule)
    assert len(rows) >= 2
    assert best["ms"] == min(r["ms"] for r in rows)
    assert all(r["passes"] >= 1 and r["ms"] > 0 for r in rows)
    got, prof = evolve(backend, n, ops, precision, schedule)
    assert prof.Passes == best["passes"]  # the whole circuit is the tuning prefix here
    want = oracle.execute(n, ops)
    if precision == C128:
        assert np.max(np.abs(got - want)) <= 1e-12
    else:
        fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
        assert 1 - fid <= 1e-5


def test_explicit_shape_replaces_the_tuned_one(backend):
    n = 13
    c = random_all_kinds(n, 200, seed=41)
    ops = c.to_ops()
    best, rows = backend.Autotune(c, C128, SCHED_FUSED)
    other = next(r for r in rows if (r["tile_bits"], r["lane_bits"]) != (best["tile_bits"], best["lane_bits"])
                 and r["passes"] != best["passes"])
    assert evolve(backend, n, ops, C128, SCHED_FUSED)[1].Passes == best["passes"]
    backend.SetFusedShape(C128, other["tile_bits"], other["lane_bits"])
    got, prof = evolve(backend, n, ops, C128, SCHED_FUSED)
    assert prof.Passes == other["passes"]
    assert np.max(np.abs(got - oracle.execute(n, ops))) <= 1e-12
    backend.SetFusedShape(C128, 13, 8)


def test_autotune_rejects_what_has_no_shape(backend):
    c = random_all_kinds(14, 50, seed=3)
    with pytest.raises(ArgumentError):
        backend.Autotune(c, C128, SCHED_PER_GATE)
    with pytest.raises(NotImplementedError):
        backend.Autotune(random_all_kinds(6, 50, seed=3), C128, SCHED_FUSED)
