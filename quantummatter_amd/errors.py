"""Exception kinds mirroring the .NET exceptions of the reference API.

The C ABI returns integer status codes (include/qm_hip.h, QM_ERR_*); each maps
to the exception the reference raises for the same condition
(CudaStateVectorBackend.cs, StateVector.cs, CpuNoiseChannelBackend.cs).
"""


class ArgumentError(ValueError):
    """System.ArgumentException."""


class ArgumentOutOfRangeError(ArgumentError):
    """System.ArgumentOutOfRangeException."""

    def __init__(self, param: str = "", message: str = ""):
        super().__init__(f"{param}: {message}" if message else param)
        self.param = param


class InvalidOperationError(RuntimeError):
    """System.InvalidOperationException (budget, mass, profiling misuse)."""


class OperationCanceledError(RuntimeError):
    """System.OperationCanceledException."""


class PlatformNotSupportedError(RuntimeError):
    """System.PlatformNotSupportedException (no HIP device)."""


class ObjectDisposedError(RuntimeError):
    """System.ObjectDisposedException."""


class DeviceError(RuntimeError):
    """A HIP runtime failure reported by the device library."""


class PoolExhaustedError(InvalidOperationError):
    """A uniform pool ran out of draws (stands in for System.Random)."""


# Status codes of include/qm_hip.h.
QM_OK = 0
QM_ERR_ARG = 1
QM_ERR_RANGE = 2
QM_ERR_MASS = 3
QM_ERR_NONFINITE = 4
QM_ERR_POOL = 5
QM_ERR_UNSUPPORTED = 6
QM_ERR_BUDGET = 7
QM_ERR_CANCELLED = 8
QM_ERR_NODEVICE = 9
QM_ERR_DISPOSED = 10
QM_ERR_DEVICE = 11
QM_ERR_STATE = 12

_MAP = {
    QM_ERR_ARG: ArgumentError,
    QM_ERR_RANGE: ArgumentOutOfRangeError,
    QM_ERR_MASS: InvalidOperationError,
    QM_ERR_NONFINITE: InvalidOperationError,
    QM_ERR_POOL: PoolExhaustedError,
    QM_ERR_UNSUPPORTED: NotImplementedError,
    QM_ERR_BUDGET: InvalidOperationError,
    QM_ERR_CANCELLED: OperationCanceledError,
    QM_ERR_NODEVICE: PlatformNotSupportedError,
    QM_ERR_DISPOSED: ObjectDisposedError,
    QM_ERR_DEVICE: DeviceError,
    QM_ERR_STATE: InvalidOperationError,
}


def raise_for_status(code: int, message: str = "") -> None:
    if code == QM_OK:
        return
    exc = _MAP.get(code, RuntimeError)
    if exc is ArgumentOutOfRangeError:
        raise exc("", message)
    raise exc(message or f"status {code}")
