"""The C ABI library loads on the CPU and exports every symbol the header declares."""
import os
import re

import pytest

from quantummatter_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "qm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(qm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qm_execute" in names and "qm_state_sample" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    missing = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_abi_version_and_error_text():
    assert _lib.lib().qm_abi_version() == 1
    assert _lib.lib().qm_last_error() is not None


def test_no_device_is_reported_not_faked():
    import ctypes as C
    n = C.c_int32(-1)
    assert _lib.lib().qm_device_count(C.byref(n)) == 0
    assert n.value >= 0
    if n.value == 0:
        h = C.c_void_p()
        assert _lib.lib().qm_ctx_create(None, C.byref(h)) == 9  # QM_ERR_NODEVICE


def test_packed_record_layouts():
    import ctypes as C
    from quantummatter_amd.circuit import OP_DTYPE, PAULI_DTYPE, STEP_DTYPE
    assert OP_DTYPE.itemsize == 88 and PAULI_DTYPE.itemsize == 24 and STEP_DTYPE.itemsize == 40
    assert C.sizeof(_lib.QmProfile) == 32
    from quantummatter_amd.serving import QmServerOptions
    assert C.sizeof(QmServerOptions) == 32
