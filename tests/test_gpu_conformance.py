"""The reference's backend conformance suite against the HIP library (both schedules)."""
import pytest

from conformance import ConformanceSuite
from hip_adapter import HipBackend
from quantummatter_amd.backend import SCHED_FUSED, SCHED_PER_GATE

pytestmark = pytest.mark.gpu


class TestHipConformancePerGate(ConformanceSuite):
    backend = HipBackend(SCHED_PER_GATE)


class TestHipConformanceFused(ConformanceSuite):
    backend = HipBackend(SCHED_FUSED)
