"""qm_state_energy_noalloc: the energy with no allocation, against qm_state_energy.

The entry point keeps every device buffer in the library's static scratch
array (its device address found once per device when a context is created),
so its result must equal qm_state_energy's bit for bit: the same kernels, block
counts and fold order.  Two calls in a row, the second with a smaller state
and fewer terms (the array keeps the first call's contents), then the term
limit and the errors, all through the C ABI.
"""
import ctypes as C

import numpy as np
import pytest

from hip_adapter import shared_backend
from quantummatter_amd import workloads
from quantummatter_amd._lib import lib, ptr
from quantummatter_amd.backend import C64, C128, SCHED_FUSED

pytestmark = pytest.mark.gpu

QM_NOALLOC_MAX_TERMS = 512


def energies(session, ham):
    paulis, coeffs = ham.packed()
    want, got = C.c_double(), C.c_double()
    rc = lib().qm_state_energy(session._handle(), ptr(paulis), ptr(coeffs), len(coeffs), C.byref(want))
    assert rc == 0, lib().qm_last_error()
    rc = lib().qm_state_energy_noalloc(session._handle(), ptr(paulis), ptr(coeffs), len(coeffs), C.byref(got))
    assert rc == 0, lib().qm_last_error()
    return want.value, got.value


@pytest.mark.parametrize("precision", [C128, C64])
def test_noalloc_energy_equals_energy_on_two_calls(precision):
    be = shared_backend()
    cases = [(20, 256, 5), (12, 40, 6)]  # the second call is smaller in state and terms
    for n, terms, seed in cases:
        print(f"noalloc energy: {n} qubits, {terms} terms, precision {precision}", flush=True)
        ham = workloads.random_hamiltonian(n, terms, seed)
        with be.CreateSession(n, precision) as s:
            s.Evolve(workloads.random_layer_circuit(n, 6, seed=seed), SCHED_FUSED)
            want, got = energies(s, ham)
        assert got == want, (n, terms, got, want)


def test_noalloc_energy_limits_and_errors():
    be = shared_backend()
    n = 10
    with be.CreateSession(n, C128) as s:
        s.Evolve(workloads.random_layer_circuit(n, 3, seed=1), SCHED_FUSED)
        # Exactly the limit of nonzero terms, plus zero-coefficient terms (not counted).
        ham = workloads.random_hamiltonian(n, QM_NOALLOC_MAX_TERMS, 9)
        paulis, coeffs = ham.packed()
        paulis2 = np.concatenate([paulis, paulis[:7]])
        coeffs2 = np.concatenate([coeffs, np.zeros(7)])
        want, got = C.c_double(), C.c_double()
        assert lib().qm_state_energy(s._handle(), ptr(paulis2), ptr(coeffs2), len(coeffs2), C.byref(want)) == 0
        assert lib().qm_state_energy_noalloc(s._handle(), ptr(paulis2), ptr(coeffs2), len(coeffs2),
                                             C.byref(got)) == 0
        assert got.value == want.value
        # One nonzero term beyond the limit: QM_ERR_RANGE, nothing written.
        coeffs3 = np.concatenate([coeffs, [0.5]])
        paulis3 = np.concatenate([paulis, paulis[:1]])
        out = C.c_double(123.0)
        assert lib().qm_state_energy_noalloc(s._handle(), ptr(paulis3), ptr(coeffs3), len(coeffs3),
                                             C.byref(out)) == 2
        assert out.value == 123.0
        # Empty Hamiltonian: 0, as qm_state_energy.
        assert lib().qm_state_energy_noalloc(s._handle(), None, None, 0, C.byref(out)) == 0
        assert out.value == 0.0
        # A zero state: no finite positive mass (QM_ERR_MASS), as qm_state_energy.
        s.Upload(np.zeros(1 << n, dtype=np.complex128))
        assert lib().qm_state_energy_noalloc(s._handle(), ptr(paulis), ptr(coeffs), 4, C.byref(out)) == 3
        assert lib().qm_state_energy(s._handle(), ptr(paulis), ptr(coeffs), 4, C.byref(out)) == 3
