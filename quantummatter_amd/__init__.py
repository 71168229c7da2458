"""quantummatter_amd -- MI355X (gfx950) state-vector execution backend for QuantumMatter.

The product is the C-ABI library ``libqmhip.so`` (HIP kernels + C++ runtime,
declared in include/qm_hip.h).  This package is its host-side mirror of the
reference's circuit/observable/noise API (``circuit``), the seeded synthetic
workloads (``workloads``) and the ctypes binding of the library (``backend``).
"""
from .errors import (ArgumentError, ArgumentOutOfRangeError, DeviceError, InvalidOperationError,
                     ObjectDisposedError, OperationCanceledError, PlatformNotSupportedError,
                     PoolExhaustedError)
from .circuit import (DynamicCircuit, GateKind, GateOperation, NoiseChannel, NoiseKind, NoisyCircuit,
                      NoisyDynamicCircuit, ParameterizedCircuit, PauliHamiltonian, PauliString, PauliTerm,
                      QuantumCircuit, SingleQubitUnitary, OP_DTYPE, PAULI_DTYPE, STEP_DTYPE)

__all__ = [
    "ArgumentError", "ArgumentOutOfRangeError", "DeviceError", "InvalidOperationError",
    "ObjectDisposedError", "OperationCanceledError", "PlatformNotSupportedError", "PoolExhaustedError",
    "DynamicCircuit", "NoisyDynamicCircuit", "GateKind", "GateOperation", "NoiseChannel", "NoiseKind",
    "NoisyCircuit", "ParameterizedCircuit",
    "PauliHamiltonian", "PauliString", "PauliTerm", "QuantumCircuit", "SingleQubitUnitary",
    "OP_DTYPE", "PAULI_DTYPE", "STEP_DTYPE",
]
