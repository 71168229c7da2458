"""ctypes binding of libqmhip.so (declarations in include/qm_hip.h).

The library is built in-tree (``quantummatter_amd/libqmhip.so``) by
``quantummatter_amd/build.py``.  Nothing here falls back to another
implementation: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os

from .errors import raise_for_status

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqmhip.so")

P = C.c_void_p
I32, I64, F64 = C.c_int32, C.c_int64, C.c_double


class QmDeviceInfo(C.Structure):
    _fields_ = [("index", C.c_int32), ("compute_units", C.c_int32), ("memory_bytes", C.c_int64),
                ("name", C.c_char * 128), ("arch", C.c_char * 64), ("driver", C.c_char * 64)]


class QmTuneEntry(C.Structure):
    _fields_ = [("tile_bits", C.c_int32), ("lane_bits", C.c_int32), ("variant", C.c_int32),
                ("passes", C.c_int32), ("ms", C.c_double), ("ms_per_gate", C.c_double),
                ("contiguous", C.c_int32), ("remap", C.c_int32)]


class QmCtxOptions(C.Structure):
    _fields_ = [("device_index", C.c_int32), ("enable_profiling", C.c_int32),
                ("device_budget_bytes", C.c_int64), ("staging_budget_bytes", C.c_int64)]


class QmProfile(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("state_bytes", C.c_int64), ("kernel_launches", C.c_int32),
                ("passes", C.c_int32), ("algorithmic_bytes", C.c_double)]


# qm_draw_fn: int (*)(void *user, double *value); returns 0 with one uniform in *value.
QM_DRAW_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double))


def draw_callback(random):
    """A qm_draw_fn over a random source with NextDouble() (System.Random's
    shape); any exception in the source aborts the library call (QM_ERR_POOL).
    Keep the returned object alive for the duration of the call."""
    def fn(_user, out):
        try:
            out[0] = float(random.NextDouble())
            return 0
        except Exception:  # noqa: BLE001 - reported to the library as a failed draw
            return 1
    return QM_DRAW_FN(fn)


# name -> argtypes; every function returns int status unless listed in _RESTYPE.
SIGNATURES = {
    "qm_last_error": [],
    "qm_abi_version": [],
    "qm_device_count": [P],
    "qm_device_get_info": [I32, P],
    "qm_ctx_create": [P, P],
    "qm_ctx_destroy": [P],
    "qm_ctx_resident_bytes": [P, P],
    "qm_ctx_set_fused_shape": [P, I32, I32, I32],
    "qm_autotune": [P, I32, I32, I32, P, I64, P, I32, P, P],
    "qm_ctx_get_fused_shape": [P, I32, P, P],
    "qm_execute": [P, I32, P, I64, P, I64, P, P],
    "qm_state_create": [P, I32, I32, P],
    "qm_state_destroy": [P],
    "qm_state_info": [P, P, P],
    "qm_state_reset": [P],
    "qm_state_upload": [P, P, I64],
    "qm_state_download": [P, P, I64],
    "qm_state_copy": [P, P],
    "qm_state_execute": [P, P, I64, I32, P, P],
    "qm_state_noise": [P, I32, F64, I32, P, I64, P],
    "qm_state_noisy_execute": [P, P, I64, P, P, I64, P],
    "qm_state_noise_draw": [P, I32, F64, I32, P, P],
    "qm_noise_apply": [P, I32, I32, F64, I32, P, I64, P, P, P],
    "qm_state_noisy_execute_draw": [P, P, I64, P, P, P],
    "qm_state_dynamic_execute_draw": [P, I32, P, I64, P, I64, P, P, I64, I32, P, P, P, P],
    "qm_state_mass": [P, P],
    "qm_state_measure_all": [P, F64, P],
    "qm_state_measure": [P, P, I32, F64, P],
    "qm_state_marginal": [P, P, I32, I64, P],
    "qm_state_sample": [P, P, I64, P],
    "qm_state_expectations": [P, P, I64, P],
    "qm_state_energy": [P, P, P, I64, P],
    "qm_state_energy_noalloc": [P, P, P, I64, P],
    "qm_state_inner": [P, P, P, P],
    "qm_state_boundary_fallbacks": [P, P],
    "qm_plan_fused": [I32, I32, P, I64, P, P, P],
    "qm_plan_apply_host": [I32, P, I64, I32, I32, I32, I32, P],
    "qm_batch_min_qubits": [I32, P],
    "qm_trajectories_run": [P, I32, I32, P, I64, P, I64, I64, P, I64, I64, P, I32, P, P, P, P, P],
    "qm_trajectories_estimate": [P, I32, I32, P, I64, P, I64, I64, P, I64, P, P, P, P, P],
    "qm_state_dynamic_execute": [P, I32, P, I64, P, I64, P, P, I64, I32, P, I64, P, P, P],
    "qm_energy_gradient_batch": [P, I32, I32, P, P, I64, P, I64, P, P, I64, P, P, P],
    "qm_energy_batch": [P, I32, I32, P, P, I64, P, I64, I64, P, P, I64, P, P],
    "qm_vqe_program_create": [P, I32, I32, P, P, I64, I64, P, P, I64, P],
    "qm_vqe_program_gradient": [P, P, P, P, P],
    "qm_vqe_program_energies": [P, P, I64, P, P],
    "qm_vqe_program_destroy": [P],
    "qm_jit_stats": [P, P, P],
    "qm_jit_drain": [],
    "qm_state_execute_dense": [P, P, I64, I32, P],
    "qm_dense_flops": [P, P],
    "qm_plan_dense_apply_host": [I32, P, I64, I32, P, P],
    "qm_jit_compile_check": [I32, I32, P, I64, I32, I32, P],
    "qm_server_create": [P, P],
    "qm_server_submit": [P, P, P],
    "qm_server_wait": [P, I64, P],
    "qm_server_drain": [P],
    "qm_server_get_stats": [P, P],
    "qm_server_reset_stats": [P],
    "qm_server_destroy": [P],
}
_RESTYPE = {"qm_last_error": C.c_char_p, "qm_abi_version": C.c_int32}

_LIB = None


def lib():
    """Loads libqmhip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run quantummatter_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _LIB = L
    return _LIB


def call(name: str, *args) -> None:
    """Calls an entry point and raises the mapped exception on a nonzero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().qm_last_error()
        raise_for_status(rc, msg.decode() if msg else "")


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
