"""The compiled-pass generator emits kernels that hipRTC compiles for gfx950.

qm_jit_compile_check plans a circuit, generates every pass kernel of the
compiled schedule (straight-line register code) and compiles it without a
device; covers every gate kind, controls on register, lane and outside-tile
bits, merged diagonals, both precisions and two tile shapes.
"""
import ctypes as C

import pytest

from quantummatter_amd import workloads
from quantummatter_amd._lib import lib, ptr
from test_plan_host import random_all_kinds


@pytest.mark.parametrize("n,precision,K,L,kind", [(14, 0, 12, 8, "layer"), (13, 0, 12, 8, "all"),
                                                   (13, 1, 13, 8, "all"), (12, 0, 12, 9, "all"),
                                                   (15, 0, 13, 8, "all"), (16, 0, 13, 8, "layer")])
def test_generated_pass_kernels_compile(n, precision, K, L, kind):
    circuit = workloads.random_layer_circuit(n, 3, 1) if kind == "layer" else random_all_kinds(n, 90, n)
    ops = circuit.to_ops()
    passes = C.c_int32()
    rc = lib().qm_jit_compile_check(n, precision, ptr(ops), len(ops), K, L, C.byref(passes))
    assert rc == 0, lib().qm_last_error()
    assert passes.value >= 1


def test_compile_helper_writes_code_objects_and_errors(tmp_path):
    """qm_jitc (the child-process compiler) writes kI.co for a good source and
    kI.err with the compiler's message for a bad one, exiting 1."""
    import os
    import subprocess
    helper = os.path.join(os.path.dirname(lib()._name), "qm_jitc")
    assert os.access(helper, os.X_OK)
    (tmp_path / "k0.hip").write_text('extern "C" __global__ void qm_pass(double *p) { p[threadIdx.x] *= 2.0; }\n')
    (tmp_path / "k1.hip").write_text('extern "C" __global__ void qm_pass(double *p) { p[threadIdx.x] = q; }\n')
    r = subprocess.run([helper, str(tmp_path), "0", "1"], capture_output=True, timeout=120)
    assert r.returncode == 1
    assert (tmp_path / "k0.co").stat().st_size > 0
    assert not (tmp_path / "k1.co").exists()
    assert "undeclared identifier" in (tmp_path / "k1.err").read_text()
