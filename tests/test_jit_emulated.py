"""The compiled schedule's generated pass kernels, run on the host.

qm_jit_compile_check writes the source and coefficients of every pass kernel
(QM_JIT_DUMP).  This test compiles those exact sources with g++ and runs each
kernel with one host thread per GPU thread (std::barrier for __syncthreads, a
host array for the LDS), then compares the state with the oracle.  It checks
the generator's code (register updates, factored rotations, exchange flags,
LDS transposes, remapped stores) without a device.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from quantummatter_amd import workloads
from quantummatter_amd._lib import lib, ptr
from test_plan_host import random_all_kinds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS_HEAD = r'''
#include <barrier>
#include <cmath>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(a, b)
#define __shared__
#define QM_OCC 2
#define QJ_CONST
struct Dim { unsigned x; };
thread_local Dim threadIdx, blockIdx;
Dim gridDim{1};
thread_local std::barrier<> *g_bar;
inline void __syncthreads() { g_bar->arrive_and_wait(); }
thread_local double *g_lds;
#define lds g_lds
'''

HARNESS_MAIN = r'''
int main(int argc, char **argv) {
  const int n = std::atoi(argv[1]), K = std::atoi(argv[2]), threads = std::atoi(argv[3]);
  const int blocks = argc > 7 ? std::atoi(argv[7]) : 1, stop = argc > 8 ? std::atoi(argv[8]) : NPASS;
  gridDim.x = (unsigned)blocks;
  std::vector<std::vector<double>> shared(blocks, std::vector<double>(1 << 15));
  const size_t amps = (size_t)1 << n;
  std::vector<qj_c128> s(amps);
  FILE *f = std::fopen(argv[4], "rb");
  if (std::fread(s.data(), sizeof(qj_c128), amps, f) != amps) return 2;
  std::fclose(f);
  for (int i = 0; i < stop; i++) {
    std::vector<double> cf;
    char path[4096];
    std::snprintf(path, sizeof path, "%s/pass%d.coef", argv[5], i);
    FILE *c = std::fopen(path, "r");
    double v;
    while (c && std::fscanf(c, "%lf", &v) == 1) cf.push_back(v);
    if (c) std::fclose(c);
    cf.push_back(0.0);
    std::vector<std::unique_ptr<std::barrier<>>> bars;
    for (int b = 0; b < blocks; b++) bars.emplace_back(new std::barrier<>(threads));
    std::vector<std::thread> th;
    for (int b = 0; b < blocks; b++)
      for (int t = 0; t < threads; t++)
        th.emplace_back([&, b, t] {
          threadIdx.x = (unsigned)t;
          blockIdx.x = (unsigned)b;
          g_bar = bars[b].get();
          g_lds = shared[b].data();
          KERNELS[i](s.data(), cf.data(), (unsigned long long)(amps >> K));
        });
    for (auto &x : th) x.join();
  }
  f = std::fopen(argv[6], "wb");
  std::fwrite(s.data(), sizeof(qj_c128), amps, f);
  std::fclose(f);
  return 0;
}
'''


def _prelude():
    src = open(os.path.join(ROOT, "quantummatter_amd", "csrc", "jit.hip")).read()
    return src.split('R"QMJIT(')[1].split(')QMJIT"')[0]


def run_generated(tmp_path, n, ops, psi, K=13, L=8, blocks=1, stop=None):
    dump = tmp_path / "dump"
    dump.mkdir()
    os.environ["QM_JIT_DUMP"] = str(dump)
    try:
        passes = C.c_int32()
        rc = lib().qm_jit_compile_check(n, 0, ptr(ops), len(ops), K, L, C.byref(passes))
        assert rc == 0, lib().qm_last_error()
    finally:
        del os.environ["QM_JIT_DUMP"]
    npass = passes.value
    parts = [HARNESS_HEAD, _prelude()]
    for i in range(npass):
        k = (dump / f"pass{i}.hip").read_text()
        k = k.replace('extern "C" ', "").replace("void qm_pass(", f"void qm_pass_{i}(")
        k = re.sub(r"extern\s+__shared__\s+F\s+lds\[\];", "", k)
        k = re.sub(r'"\+[sv]"\(', '"+r"(', k)  # empty asm operands: host register class
        parts.append(k)
    parts.append(f"#define NPASS {npass}\ntypedef void (*Kernel)(qj_c128 *, const double *, unsigned long long);\n"
                 "Kernel KERNELS[] = {" +
                 ", ".join(f"qm_pass_{i}" for i in range(npass)) + "};\n")
    parts.append(HARNESS_MAIN)
    cpp = tmp_path / "emu.cpp"
    cpp.write_text("\n".join(parts))
    exe = tmp_path / "emu"
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", "-ffp-contract=off", str(cpp), "-o", str(exe)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(psi, dtype=np.complex128).tofile(inp)
    args = [str(exe), str(n), str(K), str(1 << L), str(inp), str(dump), str(out), str(blocks)]
    if stop is not None:
        args.append(str(stop))
    r = subprocess.run(args,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return np.fromfile(out, dtype=np.complex128), npass


pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.mark.parametrize("kind,n,seed,K,L", [("all", 14, 3, 13, 8), ("all", 15, 8, 13, 8), ("layer", 16, 5, 13, 8),
                                             ("all", 15, 13, 14, 9), ("layer", 16, 14, 14, 9)])
def test_generated_kernels_match_oracle(tmp_path, kind, n, seed, K, L):
    """Every gate kind (controls on register, thread and outside bits, swaps,
    Toffolis, diagonals) or the random layer workload, through the generated
    kernels of the default code variant, for 8192-amplitude tiles of 256
    threads and 16384-amplitude tiles of 512 threads."""
    c = random_all_kinds(n, 200, seed) if kind == "all" else workloads.random_layer_circuit(n, 10, seed)
    ops = c.to_ops()
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = oracle.execute(n, ops, psi)
    got, npass = run_generated(tmp_path, n, ops, psi, K=K, L=L)
    assert npass >= 1
    assert np.max(np.abs(got - want)) <= 1e-12


def _variant_case(variant, kind, n, seed, tmp_path):
    """Runs one emulated case in a fresh process with QM_JIT_VARIANT set (the
    library reads the variant once per process); returns the max error."""
    import sys
    code = (
        "import pathlib, sys\n"
        "import numpy as np\n"
        "import oracle\n"
        "from quantummatter_amd import workloads\n"
        "from test_plan_host import random_all_kinds\n"
        "from test_jit_emulated import run_generated\n"
        f"n, seed, kind = {n}, {seed}, {kind!r}\n"
        "c = random_all_kinds(n, 200, seed) if kind == 'all' else workloads.random_layer_circuit(n, 10, seed)\n"
        "ops = c.to_ops()\n"
        "rng = np.random.default_rng(seed)\n"
        "psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)\n"
        "psi /= np.linalg.norm(psi)\n"
        "got, _ = run_generated(pathlib.Path(sys.argv[1]), n, ops, psi)\n"
        "print(float(np.max(np.abs(got - oracle.execute(n, ops, psi)))))\n")
    env = dict(os.environ, QM_JIT_VARIANT=str(variant),
               PYTHONPATH=os.pathsep.join([ROOT, os.path.join(ROOT, "tests")]))
    r = subprocess.run([sys.executable, "-c", code, str(tmp_path)], env=env, capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    return float(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("variant,kind,n,seed", [(49, "all", 14, 3), (49, "all", 15, 1003), (49, "layer", 16, 6),
                                                 (17, "all", 14, 4), (17, "layer", 16, 7), (145, "all", 15, 9),
                                                 (129, "layer", 16, 8), (401, "all", 15, 10),
                                                 (401, "layer", 16, 11), (913, "all", 15, 12),
                                                 (913, "layer", 16, 13), (1425, "all", 15, 14),
                                                 (1425, "layer", 16, 15), (1937, "all", 15, 16),
                                                 (3473, "all", 15, 17), (3985, "layer", 16, 18),
                                                 (7601, "all", 15, 19), (7601, "layer", 16, 20),
                                                 (11665, "all", 15, 21), (12177, "layer", 16, 22)])
def test_generated_kernel_variants_match_oracle(tmp_path, variant, kind, n, seed):
    """Optional code variants: factored forms (bit 4), exchange flags (bit 5),
    per-tile coefficient loads (bit 7), per-exchange lane addresses (bit 8),
    folded LDS offsets (bit 9), scalar coefficient loads (bit 11) and flips
    carried through register CNOTs and folded into exchange and store
    addresses (bit 12) and merged register diagonals (bit 13)."""
    assert _variant_case(variant, kind, n, seed, tmp_path) <= 1e-12
