"""Request server on one MI355X (``qm_server_*`` in include/qm_hip.h).

``HipRequestServer`` queues circuits with what to return (amplitudes, terminal
shots from caller uniforms, or Pauli expectation values), admits them by HBM
footprint, batches small widths into one launch per tile shape and runs wider
ones on concurrent full-chip lanes.  ``Submit`` returns at once; ``Wait``
blocks for one request; results land in the arrays the request was given.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np

from ._lib import call, lib, ptr
from .backend import C64, C128
from .circuit import PauliString, QuantumCircuit, pauli_array
from .errors import ArgumentError, raise_for_status

AMPLITUDES, SHOTS, EXPECTATION = 0, 1, 2


class QmServerOptions(C.Structure):
    _fields_ = [("device_index", C.c_int32), ("full_lanes", C.c_int32),
                ("batch_max_qubits", C.c_int32), ("max_batch", C.c_int32),
                ("device_budget_bytes", C.c_int64), ("schedule", C.c_int32), ("pad", C.c_int32)]


class QmRequest(C.Structure):
    _fields_ = [("qubits", C.c_int32), ("precision", C.c_int32), ("kind", C.c_int32),
                ("pad", C.c_int32), ("ops", C.c_void_p), ("n_ops", C.c_int64),
                ("shots", C.c_int64), ("uniforms", C.c_void_p), ("terms", C.c_void_p),
                ("n_terms", C.c_int64), ("out", C.c_void_p)]


class QmRequestStatus(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("lane", C.c_int32),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("fallbacks", C.c_int64)]


class QmServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("pending", C.c_int64), ("reserved_bytes", C.c_int64),
                ("peak_reserved_bytes", C.c_int64), ("budget_bytes", C.c_int64),
                ("batches", C.c_int64), ("batched_requests", C.c_int64), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("span_ms", C.c_double)]


@dataclass
class RequestResult:
    id: int
    status: int
    lane: int
    submit_ms: float
    start_ms: float
    done_ms: float
    value: np.ndarray  # amplitudes, samples or expectation values

    @property
    def latency_ms(self) -> float:
        return self.done_ms - self.submit_ms


class HipRequestServer:
    def __init__(self, device_index: int = 0, full_lanes: int = 3, batch_max_qubits: int = 14,
                 max_batch: int = 1024, device_budget_bytes: int = 0, schedule: int = 0):
        o = QmServerOptions(device_index, full_lanes, batch_max_qubits, max_batch, device_budget_bytes,
                            schedule, 0)
        h = C.c_void_p()
        call("qm_server_create", C.byref(o), C.byref(h))
        self._h = h
        self._live: Dict[int, tuple] = {}

    def Submit(self, circuit: QuantumCircuit, kind: int, precision: int = C128,
               uniforms: Optional[np.ndarray] = None,
               observables: Optional[Sequence[PauliString]] = None) -> int:
        if circuit is None:
            raise ArgumentError("circuit is null.")
        n = circuit.QubitCount
        ops = circuit.to_ops()
        u = terms = None
        if kind == AMPLITUDES:
            out = np.zeros(1 << n, dtype=np.complex128 if precision == C128 else np.complex64)
        elif kind == SHOTS:
            u = np.ascontiguousarray(uniforms, dtype=np.float64)
            out = np.zeros(len(u), dtype=np.int64)
        elif kind == EXPECTATION:
            terms = pauli_array(list(observables))
            out = np.zeros(len(terms), dtype=np.float64)
        else:
            raise ArgumentError("unknown request kind")
        rq = QmRequest(n, precision, kind, 0, ptr(ops), len(ops), 0 if u is None else len(u), ptr(u),
                       ptr(terms), 0 if terms is None else len(terms), ptr(out))
        rid = C.c_int64()
        call("qm_server_submit", self._h, C.byref(rq), C.byref(rid))
        self._live[rid.value] = (ops, u, terms, out)
        return rid.value

    def Wait(self, rid: int) -> RequestResult:
        st = QmRequestStatus()
        rc = lib().qm_server_wait(self._h, rid, C.byref(st))
        keep = self._live.pop(rid, None)
        if rc:
            msg = lib().qm_last_error()
            raise_for_status(rc, msg.decode() if msg else "")
        return RequestResult(st.id, st.status, st.lane, st.submit_ms, st.start_ms, st.done_ms,
                             keep[3] if keep else None)

    def Drain(self) -> None:
        call("qm_server_drain", self._h)

    def Stats(self) -> QmServerStats:
        s = QmServerStats()
        call("qm_server_get_stats", self._h, C.byref(s))
        return s

    def ResetStats(self) -> None:
        call("qm_server_reset_stats", self._h)

    def Dispose(self) -> None:
        if self._h:
            call("qm_server_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Dispose()
