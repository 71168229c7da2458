"""The HIP library behind the conformance suite's backend interface (tests/conformance.py)."""
import math

import numpy as np

from quantummatter_amd.backend import C128, SCHED_FUSED, HipStateVectorBackend
from quantummatter_amd.circuit import PauliString

_BACKEND = None


def shared_backend():
    global _BACKEND
    if _BACKEND is None:
        _BACKEND = HipStateVectorBackend()
    return _BACKEND


class HipBackend:
    def __init__(self, schedule=SCHED_FUSED, precision=C128):
        self.schedule = schedule
        self.precision = precision

    def session(self, n, state=None):
        s = shared_backend().CreateSession(n, self.precision)
        if state is not None:
            s.Upload(np.asarray(state, dtype=np.complex128))
        return s

    def simulate(self, c, state=None):
        with self.session(c.QubitCount, state) as s:
            s.Evolve(c, self.schedule)
            return s.Snapshot()

    def measure_all(self, n, state, u):
        with self.session(n, state) as s:
            out = s.MeasureAll(u)
            return out, s.Snapshot()

    def measure(self, n, state, sel, u):
        with self.session(n, state) as s:
            out = s.Measure(sel, u)
            return out, s.Snapshot()

    def marginal(self, n, state, sel, outcome):
        with self.session(n, state) as s:
            return s.MarginalProbability(sel, outcome)

    def expectation(self, n, state, p):
        with self.session(n, state) as s:
            return s.ExpectationValue(p)

    def energy(self, n, state, ham):
        with self.session(n, state) as s:
            return s.ExpectationEnergy(ham)

    def sample(self, n, state, u):
        with self.session(n, state) as s:
            return s.Sample(u)

    def noise_apply(self, ch, target, n, state, u):
        with self.session(n, state) as s:
            cur = s.ApplyNoise(ch, target, u)
            return s.Snapshot(), cur

    def energy_gradient(self, pc, ham, params):
        from quantummatter_amd.variational import HipVariationalEnergyEvaluator
        ev = HipVariationalEnergyEvaluator(shared_backend(), precision=self.precision)
        r = ev.EvaluateWithGradient(pc, ham, params)
        return r.Energy, np.asarray(r.Gradient)
