"""Pins the CPU oracle against the reference's known-answer tests and the
independent dense operators (tests/conformance.py)."""
from conformance import ConformanceSuite, OracleBackend


class TestOracleConformance(ConformanceSuite):
    backend = OracleBackend()
