"""Timing of the CPU baseline beside the GPU.  TEST/BENCH INFRASTRUCTURE ONLY.

The reference's parallel CPU path (restated in oracle/cpu_parallel.cpp:
ParallelCpuStateVectorBackend.cs with its AVX-512, AVX and SIMD-Hadamard
kernels and optional blocking) is timed the way the reference's benchmark
program times it (benchmarks/QuantumMatter.Benchmarks/Program.cs:161-209: warmup
runs, then timed repetitions, median reported with min and max): on a state
allocated and touched before the timed region, for several worker counts and
with blocking on and off.  Results are checked against the scalar oracle.

The scalar restatement (oracle/qm_oracle.c, one thread) times the other
workloads: terminal sampling, noisy trajectories and parameter-shift circuits.
Only bench.py's cpu_baseline leg and the tests call this module.
"""
from __future__ import annotations

import os
import statistics
import subprocess
import time
from typing import Callable, Dict, List

import numpy as np

from . import execute, lib, parallel_execute, zero_state


def host_description() -> Dict:
    """CPU model, logical CPUs, the CPUs this process may run on, physical cores."""
    info = {"model": "unknown", "logical_cpus": os.cpu_count() or 1,
            "allowed_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count(),
            "physical_cores": None, "avx512f": False}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        fields = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        fields = {k.strip(): v.strip() for k, v in fields.items()}
        info["model"] = fields.get("Model name", info["model"])
        cores = fields.get("Core(s) per socket")
        sockets = fields.get("Socket(s)")
        if cores and sockets and cores.isdigit() and sockets.isdigit():
            info["physical_cores"] = int(cores) * int(sockets)
        info["avx512f"] = "avx512f" in fields.get("Flags", "")
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    return info


def timed(fn: Callable[[], None], warmup: int, reps: int) -> Dict:
    """Runs fn warmup times untimed, then reps times timed: median, min, max seconds."""
    for _ in range(warmup):
        fn()
    runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        runs.append(time.perf_counter() - t0)
    return {"median_s": statistics.median(runs), "min_s": min(runs), "max_s": max(runs), "reps": reps,
            "warmup": warmup}


def layer_circuit_rows(n: int, layers: int, workers_list: List[int], modes, warmup: int = 2, reps: int = 5,
                       seed: int = 1) -> List[Dict]:
    """The parallel CPU path on the seeded layer circuit (n qubits, `layers`
    layers), one row per (workers, simd, avx512, blocking).  The state is
    allocated and written once before the first warmup and evolved in place by
    every run (the work per run is the same), so page faults stay out of the
    timed runs.  Each row's first warmup output is compared with the scalar
    oracle's state when n is small enough to run it."""
    from quantummatter_amd import workloads
    import ctypes as C
    ops = np.ascontiguousarray(workloads.random_layer_circuit(n, layers, seed=seed).to_ops())
    state = np.empty(1 << n, dtype=np.complex128)
    P = C.c_void_p
    rows = []
    for workers in workers_list:
        for simd, avx512, blocking in modes:
            state[:] = 0
            state[0] = 1

            def run():
                rc = lib().qmo_parallel_execute(n, ops.ctypes.data_as(P), len(ops), state.ctypes.data_as(P),
                                                workers, int(simd), int(avx512), int(blocking))
                if rc != 0:
                    raise RuntimeError(f"qmo_parallel_execute: status {rc}")

            row = {"workers": workers, "simd": simd, "avx512": avx512, "blocking": blocking}
            row.update(timed(run, warmup, reps))
            rows.append(row)
    return rows


def check_parallel_against_scalar(n: int = 18, layers: int = 4, workers: int = 8) -> bool:
    """The timed path's state equals the scalar oracle's bit for bit on a smaller instance."""
    from quantummatter_amd import workloads
    ops = workloads.random_layer_circuit(n, layers, seed=1).to_ops()
    want = execute(n, ops)
    for simd, avx512, blocking in ((True, True, True), (True, True, False), (True, False, False)):
        if not np.array_equal(parallel_execute(n, ops, workers=workers, simd=simd, avx512=avx512,
                                               blocking=blocking), want):
            return False
    return True


def sampling_row(n: int = 20, layers: int = 20, shots: int = 1024, workers: int = 8, warmup: int = 2,
                 reps: int = 5) -> Dict:
    """StateVectorSimulator.Run on the CPU: the parallel path evolves the
    20-qubit 20-layer circuit, then the scalar restatement of Run's CDF and
    strict-upper-bound search draws the shots (StateVectorSimulator.cs:37-101)."""
    from quantummatter_amd import workloads
    from . import sample
    ops = workloads.random_layer_circuit(n, layers, seed=1).to_ops()
    u = workloads.uniforms(2, shots)

    def run():
        s = parallel_execute(n, ops, workers=workers)
        sample(n, s, u)

    row = timed(run, warmup, reps)
    row.update({"qubits": n, "layers": layers, "shots": shots, "workers": workers,
                "circuits_per_s": 1.0 / row["median_s"], "gate_layers_per_s": layers / row["median_s"]})
    return row


def trajectories_row(n: int = 16, layers: int = 30, trajectories: int = 16, shots: int = 1024) -> Dict:
    """NoisyStateVectorSimulator.Run on the CPU (scalar restatement, one thread,
    trajectories one after another): a sample of the 4096-trajectory workload."""
    from quantummatter_amd import workloads
    from . import trajectories as run_trajectories
    nc = workloads.noisy_layer_circuit(n, layers, seed=3)
    steps, ops = nc.packed()
    stride = int(np.sum(steps["is_channel"] == 1)) + 1
    pools = workloads.uniform_pools(7, trajectories, stride)
    su = workloads.uniforms(8, trajectories * shots).reshape(trajectories, shots)
    row = timed(lambda: run_trajectories(n, steps, ops, pools, shots, su, True), 0, 1)
    row.update({"qubits": n, "layers": layers, "trajectories": trajectories, "shots": shots, "threads": 1,
                "trajectories_per_s": trajectories / row["median_s"]})
    return row


def variational_rows(widths=(12, 16, 20), circuits=(16, 4, 1), layers: int = 6, terms: int = 256) -> List[Dict]:
    """Parameter-shift evaluation on the CPU (scalar restatement, one thread):
    each of the 2P+1 circuits is one bound-circuit evolution plus the 256-term
    energy (VariationalEnergyEvaluator.cs:34-86); a sample of `circuits`
    bound circuits per width gives circuits/s."""
    from quantummatter_amd import workloads
    from . import energy_bound
    rows = []
    for n, count in zip(widths, circuits):
        circ = workloads.hardware_efficient_ansatz(n, layers)
        theta = workloads.random_parameters(circ.ParameterCount, 1)
        ham = workloads.random_hamiltonian(n, terms, 2)
        ops, pidx = circ.packed()
        paulis, coeffs = ham.packed()

        def run():
            for k in range(count):
                energy_bound(n, ops, pidx, theta + 1e-3 * k, paulis, coeffs)

        row = timed(run, 0, 1)
        row.update({"qubits": n, "circuits_timed": count, "threads": 1,
                    "circuits_per_s": count / row["median_s"],
                    "gradient_circuits": 1 + 2 * circ.ParameterCount})
        rows.append(row)
    return rows
