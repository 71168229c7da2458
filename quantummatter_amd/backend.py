"""Host-side mirror of the reference's GPU backend API over libqmhip.so.

``HipStateVectorBackend`` has the public shape of ``CudaStateVectorBackend``
(src/QuantumMatter.Infrastructure/CudaStateVectorBackend.cs:10-237 and
CudaBackendOptions.cs): options record, device info, ``GetDevices``,
``Execute``, ``ExecuteProfiled``, ``Dispose``.  ``HipStateSession`` is the
additive device-resident form of ``StateVector`` (StateVector.cs,
StateVector.Observations.cs, StateVector.Energy.cs): the state stays in HBM
between evolution, noise, measurement and observation calls.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import QmCtxOptions, QmDeviceInfo, QmProfile, call, ptr
from .circuit import (NoiseChannel, NoisyCircuit, PauliHamiltonian, PauliString, QuantumCircuit,
                      pauli_array)
from .errors import ArgumentError, ObjectDisposedError, OperationCanceledError

C128 = 0
C64 = 1
SCHED_PER_GATE = 0
SCHED_FUSED = 1
SCHED_COMPILED = 2  # fused passes compiled to straight-line kernels, cached by structure
SCHED_AUTO = 3  # compiled when the structure's kernels are cached, else interpreted + background compile


def jit_drain() -> None:
    """Waits for the background compilations SCHED_AUTO started (qm_jit_drain)."""
    call("qm_jit_drain")


def jit_stats():
    """(kernels compiled, cache hits, compile seconds) so far in this process."""
    a, b, c = C.c_int64(), C.c_int64(), C.c_double()
    call("qm_jit_stats", C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


@dataclass(frozen=True)
class HipBackendOptions:
    """CudaBackendOptions (CudaBackendOptions.cs:4-20); 0 bytes = no budget beyond the device."""
    DeviceIndex: int = 0
    DeviceMemoryBudgetBytes: int = 0
    StagingMemoryBudgetBytes: int = 0
    EnableProfiling: bool = False


@dataclass(frozen=True)
class HipDeviceInfo:
    """CudaDeviceInfo (CudaBackendOptions.cs:29)."""
    Index: int
    Name: str
    Architecture: str
    DriverVersion: str
    MemoryBytes: int
    ComputeUnits: int


@dataclass(frozen=True)
class HipExecutionProfile:
    """CudaExecutionProfile (CudaBackendOptions.cs:35) plus the schedule's shape."""
    DeviceEvolutionMs: float
    StateBytes: int
    KernelLaunches: int
    Passes: int
    AlgorithmicBytes: float


class CancellationToken:
    """A flag polled by the library between launches (stands in for .NET's token)."""

    def __init__(self):
        self._flag = np.zeros(1, dtype=np.int32)

    def Cancel(self):
        self._flag[0] = 1

    @property
    def IsCancellationRequested(self):
        return bool(self._flag[0])

    def _ptr(self):
        return ptr(self._flag)


def _cancel_ptr(token: Optional[CancellationToken]):
    return None if token is None else token._ptr()


class HipStateVectorBackend:
    """Complex128 AMD execution; owns a HIP context.  Dispose after use."""

    def __init__(self, options: Optional[HipBackendOptions] = None):
        self._options = options or HipBackendOptions()
        o = QmCtxOptions(self._options.DeviceIndex, 1 if self._options.EnableProfiling else 0,
                         self._options.DeviceMemoryBudgetBytes, self._options.StagingMemoryBudgetBytes)
        h = C.c_void_p()
        call("qm_ctx_create", C.byref(o), C.byref(h))
        self._ctx = h
        self._lock = threading.Lock()
        self._disposed = False
        self.Device = _device_info(self._options.DeviceIndex)

    @staticmethod
    def GetDevices() -> List[HipDeviceInfo]:
        n = C.c_int32()
        call("qm_device_count", C.byref(n))
        return [_device_info(i) for i in range(n.value)]

    def _handle(self):
        if self._disposed:
            raise ObjectDisposedError("HipStateVectorBackend")
        return self._ctx

    def Execute(self, circuit: QuantumCircuit, amplitudes: np.ndarray,
                cancellation: Optional[CancellationToken] = None) -> None:
        self._execute(circuit, amplitudes, None, cancellation)

    def ExecuteProfiled(self, circuit: QuantumCircuit, amplitudes: np.ndarray,
                        cancellation: Optional[CancellationToken] = None) -> HipExecutionProfile:
        prof = QmProfile()
        self._execute(circuit, amplitudes, prof, cancellation)
        return HipExecutionProfile(prof.device_ms, prof.state_bytes, prof.kernel_launches, prof.passes,
                                   prof.algorithmic_bytes)

    def _execute(self, circuit, amplitudes, prof, cancellation):
        if circuit is None:
            raise ArgumentError("circuit is null.")
        if not (isinstance(amplitudes, np.ndarray) and amplitudes.dtype == np.complex128
                and amplitudes.flags.c_contiguous):
            raise ArgumentError("amplitudes must be a contiguous complex128 array.")
        ops = circuit.to_ops()
        with self._lock:
            call("qm_execute", self._handle(), circuit.QubitCount, ptr(ops), len(ops), ptr(amplitudes),
                 amplitudes.size, _cancel_ptr(cancellation), None if prof is None else C.byref(prof))

    def SetFusedShape(self, precision: int, tile_bits: int, lane_bits: int) -> None:
        """Selects a compiled tile shape for the fused schedule (a tuning knob)."""
        call("qm_ctx_set_fused_shape", self._handle(), precision, tile_bits, lane_bits)

    def GetFusedShape(self, precision: int):
        k, l = C.c_int32(), C.c_int32()
        call("qm_ctx_get_fused_shape", self._handle(), precision, C.byref(k), C.byref(l))
        return k.value, l.value

    def Autotune(self, circuit: QuantumCircuit, precision: int = C128, schedule: int = SCHED_FUSED):
        """Times every tile shape of `schedule` on a prefix of `circuit` and keeps the
        fastest for this width and precision (qm_autotune).  Returns (best, rows), each
        row a dict of tile_bits, lane_bits, variant, passes, ms, ms_per_gate."""
        from ._lib import QmTuneEntry
        ops = circuit.to_ops()
        cap = 64
        rows = (QmTuneEntry * cap)()
        count, best = C.c_int32(), C.c_int32()
        with self._lock:
            call("qm_autotune", self._handle(), circuit.QubitCount, precision, schedule, ptr(ops), len(ops),
                 rows, cap, C.byref(count), C.byref(best))
        out = [{f: getattr(rows[i], f) for f, _ in QmTuneEntry._fields_} for i in range(min(count.value, cap))]
        return out[best.value], out

    def ResidentBytes(self) -> int:
        b = C.c_int64()
        call("qm_ctx_resident_bytes", self._handle(), C.byref(b))
        return b.value

    def CreateSession(self, qubits: int, precision: int = C128) -> "HipStateSession":
        return HipStateSession(self, qubits, precision)

    def Dispose(self) -> None:
        with self._lock:
            if self._disposed:
                return
            self._disposed = True
            call("qm_ctx_destroy", self._ctx)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Dispose()


class HipNoiseChannelBackend:
    """INoiseChannelBackend (src/QuantumMatter.Core/Noise/INoiseChannelBackend.cs:13-14)
    on the device: Apply(channel, target, amplitudes, random) validates like
    CpuNoiseChannelBackend.Apply (CpuNoiseChannelBackend.cs:13-22), uploads the
    host amplitudes, samples the branch with the caller's random source
    (NextDouble(), called exactly when the reference calls it) and writes the
    conditional state back only after success."""

    def __init__(self, backend: HipStateVectorBackend):
        if backend is None:
            raise ArgumentError("backend is null.")
        self._backend = backend

    def Apply(self, channel: NoiseChannel, target: int, amplitudes: np.ndarray, random,
              cancellation: Optional[CancellationToken] = None) -> None:
        if channel is None or random is None:
            raise ArgumentError("channel and random are required.")
        if not (isinstance(amplitudes, np.ndarray) and amplitudes.dtype == np.complex128
                and amplitudes.flags.c_contiguous):
            raise ArgumentError("amplitudes must be a contiguous complex128 array.")
        n = amplitudes.size
        if n < 2 or n & (n - 1):
            raise ArgumentError("State length must be a power of two, at least two.")
        cb = _lib.draw_callback(random)
        call("qm_noise_apply", self._backend._handle(), n.bit_length() - 1, int(channel.Kind),
             channel.Probability, target, ptr(amplitudes), n, C.cast(cb, C.c_void_p), None,
             _cancel_ptr(cancellation))


def _device_info(i: int) -> HipDeviceInfo:
    info = QmDeviceInfo()
    call("qm_device_get_info", i, C.byref(info))
    return HipDeviceInfo(info.index, info.name.decode(), info.arch.decode(), info.driver.decode(),
                         info.memory_bytes, info.compute_units)


class HipStateSession:
    """A device-resident pure state with StateVector's operations.

    Uniform draws replace System.Random: every drawing call takes the uniform
    (or a pool and a cursor) explicitly, consumed under the reference's rules.
    """

    def __init__(self, backend: HipStateVectorBackend, qubits: int, precision: int = C128):
        self._backend = backend
        h = C.c_void_p()
        call("qm_state_create", backend._handle(), qubits, precision, C.byref(h))
        self._h = h
        self.QubitCount = qubits
        self.Precision = precision
        self.Length = 1 << qubits

    # ---- lifetime ------------------------------------------------------------
    def Dispose(self):
        if self._h is not None:
            call("qm_state_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Dispose()

    def _handle(self):
        if self._h is None:
            raise ObjectDisposedError("HipStateSession")
        self._backend._handle()
        return self._h

    # ---- state ---------------------------------------------------------------
    def Reset(self):
        call("qm_state_reset", self._handle())

    def Upload(self, amplitudes: np.ndarray):
        a = np.ascontiguousarray(amplitudes, dtype=np.complex128)
        call("qm_state_upload", self._handle(), ptr(a), a.size)

    def Snapshot(self) -> np.ndarray:
        out = np.empty(self.Length, dtype=np.complex128)
        call("qm_state_download", self._handle(), ptr(out), out.size)
        return out

    def CopyFrom(self, other: "HipStateSession"):
        call("qm_state_copy", self._handle(), other._handle())

    # ---- evolution -----------------------------------------------------------
    def Evolve(self, circuit: QuantumCircuit, schedule: int = SCHED_FUSED,
               cancellation: Optional[CancellationToken] = None) -> Optional[HipExecutionProfile]:
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit and state sizes must match.")
        return self.EvolveOps(circuit.to_ops(), schedule, cancellation)

    def EvolveOps(self, ops: np.ndarray, schedule: int = SCHED_FUSED,
                  cancellation: Optional[CancellationToken] = None, profile: bool = False):
        prof = QmProfile() if profile else None
        call("qm_state_execute", self._handle(), ptr(ops), len(ops), schedule, _cancel_ptr(cancellation),
             None if prof is None else C.byref(prof))
        if prof is None:
            return None
        return HipExecutionProfile(prof.device_ms, prof.state_bytes, prof.kernel_launches, prof.passes,
                                   prof.algorithmic_bytes)

    def EvolveDynamic(self, circuit, uniforms: np.ndarray, cursor: int = 0, schedule: int = SCHED_FUSED,
                      cancellation: Optional[CancellationToken] = None):
        """StateVector.EvolveDynamic / EvolveNoisyDynamic (StateVector.Dynamic.cs:11-25,
        StateVector.NoisyDynamic.cs:11-27) for a DynamicCircuit or NoisyDynamicCircuit:
        draws come from uniforms[cursor...]; returns (classical register, advanced cursor)."""
        if circuit is None:
            raise ArgumentError("circuit is null.")
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit and state sizes must match.")
        steps, ops, sel_q, sel_c = circuit.packed()
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        cur = np.array([cursor], dtype=np.int64)
        reg = np.zeros(1, dtype=np.int32)
        call("qm_state_dynamic_execute", self._handle(), circuit.ClassicalBitCount, ptr(steps), len(steps),
             ptr(ops), len(ops), ptr(sel_q), ptr(sel_c), len(sel_q), schedule, ptr(u), len(u), ptr(cur),
             ptr(reg), _cancel_ptr(cancellation))
        return int(reg[0]), int(cur[0])

    def ApplyNoise(self, channel: NoiseChannel, target: int, uniforms: np.ndarray, cursor: int = 0) -> int:
        """INoiseChannelBackend.Apply; returns the advanced cursor."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        cur = np.array([cursor], dtype=np.int64)
        call("qm_state_noise", self._handle(), int(channel.Kind), channel.Probability, target, ptr(u),
             len(u), ptr(cur))
        return int(cur[0])

    def ApplyNoiseRandom(self, channel: NoiseChannel, target: int, random) -> None:
        """INoiseChannelBackend.Apply (Noise/INoiseChannelBackend.cs:13-14) with the
        caller's random source (NextDouble()), drawn exactly when
        CpuNoiseChannelBackend.Apply draws."""
        if channel is None or random is None:
            raise ArgumentError("channel and random are required.")
        cb = _lib.draw_callback(random)
        call("qm_state_noise_draw", self._handle(), int(channel.Kind), channel.Probability, target,
             C.cast(cb, C.c_void_p), None)

    def EvolveTrajectoryRandom(self, circuit: NoisyCircuit, random) -> None:
        """StateVector.EvolveTrajectory (StateVector.cs:74-88) with the caller's random source."""
        if circuit is None or random is None:
            raise ArgumentError("circuit and random are required.")
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit and state sizes must match.")
        steps, ops = circuit.packed()
        cb = _lib.draw_callback(random)
        call("qm_state_noisy_execute_draw", self._handle(), ptr(steps), len(steps), ptr(ops),
             C.cast(cb, C.c_void_p), None)

    def EvolveDynamicRandom(self, circuit, random, schedule: int = SCHED_FUSED,
                            cancellation: Optional[CancellationToken] = None) -> int:
        """EvolveDynamic / EvolveNoisyDynamic with the caller's random source;
        returns the classical register."""
        if circuit is None or random is None:
            raise ArgumentError("circuit and random are required.")
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit and state sizes must match.")
        steps, ops, sel_q, sel_c = circuit.packed()
        reg = np.zeros(1, dtype=np.int32)
        cb = _lib.draw_callback(random)
        call("qm_state_dynamic_execute_draw", self._handle(), circuit.ClassicalBitCount, ptr(steps), len(steps),
             ptr(ops), len(ops), ptr(sel_q), ptr(sel_c), len(sel_q), schedule, C.cast(cb, C.c_void_p), None,
             ptr(reg), _cancel_ptr(cancellation))
        return int(reg[0])

    def EvolveDense(self, circuit: QuantumCircuit, max_block_qubits: int = 4, profile: bool = False):
        """Gates fused into dense blocks of at most max_block_qubits qubits, one
        matrix-core pass per block (complex128).  Returns (profile, flops)."""
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit and state sizes must match.")
        ops = circuit.to_ops()
        prof = QmProfile() if profile else None
        call("qm_state_execute_dense", self._handle(), ptr(ops), len(ops), max_block_qubits,
             None if prof is None else C.byref(prof))
        flops = C.c_double()
        call("qm_dense_flops", self._backend._handle(), C.byref(flops))
        if prof is None:
            return None, flops.value
        return HipExecutionProfile(prof.device_ms, prof.state_bytes, prof.kernel_launches, prof.passes,
                                   prof.algorithmic_bytes), flops.value

    def EvolveTrajectory(self, circuit: NoisyCircuit, uniforms: np.ndarray, cursor: int = 0) -> int:
        if circuit.QubitCount != self.QubitCount:
            raise ArgumentError("Circuit and state sizes must match.")
        steps, ops = circuit.packed()
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        cur = np.array([cursor], dtype=np.int64)
        call("qm_state_noisy_execute", self._handle(), ptr(steps), len(steps), ptr(ops), ptr(u), len(u),
             ptr(cur))
        return int(cur[0])

    # ---- observations --------------------------------------------------------
    def Mass(self) -> float:
        m = C.c_double()
        call("qm_state_mass", self._handle(), C.byref(m))
        return m.value

    def MeasureAll(self, u: float) -> int:
        out = C.c_int64()
        call("qm_state_measure_all", self._handle(), float(u), C.byref(out))
        return out.value

    def Measure(self, qubits: Sequence[int], u: float) -> int:
        sel = np.asarray(qubits, dtype=np.int32)
        out = C.c_int64()
        call("qm_state_measure", self._handle(), ptr(sel), len(sel), float(u), C.byref(out))
        return out.value

    def MarginalProbability(self, qubits: Sequence[int], outcome: int) -> float:
        sel = np.asarray(qubits, dtype=np.int32)
        p = C.c_double()
        call("qm_state_marginal", self._handle(), ptr(sel), len(sel), int(outcome), C.byref(p))
        return p.value

    def Sample(self, uniforms: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.empty(len(u), dtype=np.int64)
        call("qm_state_sample", self._handle(), ptr(u), len(u), ptr(out))
        return out

    def ExpectationValues(self, observables: Sequence[PauliString]) -> np.ndarray:
        terms = pauli_array(observables)
        out = np.empty(len(terms), dtype=np.float64)
        call("qm_state_expectations", self._handle(), ptr(terms), len(terms), ptr(out))
        return out

    def ExpectationValue(self, observable: PauliString) -> float:
        if observable.QubitCount != self.QubitCount:
            raise ArgumentError("Observable and state sizes must match.")
        return float(self.ExpectationValues([observable])[0])

    def ExpectationEnergy(self, hamiltonian: PauliHamiltonian) -> float:
        if hamiltonian.QubitCount != self.QubitCount:
            raise ArgumentError("Hamiltonian and state sizes must match.")
        terms, coeffs = hamiltonian.packed()
        e = C.c_double()
        call("qm_state_energy", self._handle(), ptr(terms), ptr(coeffs), len(coeffs), C.byref(e))
        return e.value

    def Inner(self, other: "HipStateSession") -> complex:
        re, im = C.c_double(), C.c_double()
        call("qm_state_inner", self._handle(), other._handle(), C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def BoundaryFallbacks(self) -> int:
        c = C.c_int64()
        call("qm_state_boundary_fallbacks", self._handle(), C.byref(c))
        return c.value


def plan_fused(qubits: int, ops: np.ndarray, precision: int = C128):
    """(passes, transposes, gates) of the fused schedule; host only."""
    p, t, g = C.c_int32(), C.c_int32(), C.c_int64()
    call("qm_plan_fused", qubits, precision, ptr(ops), len(ops), C.byref(p), C.byref(t), C.byref(g))
    return p.value, t.value, g.value


def dense_apply_host(n: int, ops: np.ndarray, state: np.ndarray, max_block_qubits: int):
    """Host reference of the dense-block schedule: (new state, blocks)."""
    out = np.array(state, dtype=np.complex128, copy=True)
    blocks = C.c_int32()
    call("qm_plan_dense_apply_host", n, ptr(ops), len(ops), max_block_qubits, ptr(out), C.byref(blocks))
    return out, blocks.value


def plan_apply_host(qubits: int, ops: np.ndarray, state: np.ndarray, tile_bits: int, lane_bits: int,
                    contiguous: int, remap: int = 0, search: bool = False, replay: bool = False,
                    searched: bool = False, lane_layout: bool = False) -> np.ndarray:
    """Runs the fused schedule on the host with the kernels' arithmetic (CPU verification).
    remap: remapping mode of the low index bits (0, 1 or 2); search: tile search;
    replay: the plan is replayed from decisions recorded on the same gates with
    other coefficients (and checked against a fresh plan); searched: the
    compiled schedule's searched planning (restarts and iterated greedy moves),
    replayed from its trace (and checked against the searched plan);
    lane_layout: every pass stores under the index-bit permutation that puts
    the next pass's load lanes above the contiguous run (passes out of place)."""
    s = np.array(state, dtype=np.complex128, copy=True)
    call("qm_plan_apply_host", qubits, ptr(ops), len(ops), tile_bits, lane_bits, contiguous,
         int(remap) | (4 if search else 0) | (8 if replay else 0) | (16 if searched else 0)
         | (32 if lane_layout else 0), ptr(s))
    return s
