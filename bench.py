"""Benchmark of the state-vector hot path on one MI355X.

Workload (the metric's configuration): a seeded 30-qubit random layer circuit,
40 layers (Rx/Ry/Rz with angles U[0,2pi) on every qubit, then a CNOT brick),
on a resident complex128 state (16 GiB), fused schedule.  A "step" is one
execution of the full 40-layer circuit on the resident state.  `value` is
gate-layers per second over the timed steps (wall clock around K steps, each
synchronised), with the state already in HBM.

Also reported: the dominant kernel's roofline (the fused tile pass: 2 x 16 B x
2^30 algorithmic bytes per launch, average duration from HIP events recorded
by the library on its stream around the timed sequence), and the CPU baseline
(the oracle's scalar restatement of the reference path, timed on a bounded
sample on this host and scaled to the same unit).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    return ap.parse_args()


def pmc_traffic(path):
    """HBM bytes per launch of the fused pass from a committed rocprofv3 --pmc summary."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def cpu_baseline(n_target, layers_target):
    """Scalar oracle on a bounded sample: 22 qubits, 2 layers, scaled by 2^(30-22)."""
    import oracle
    from quantummatter_amd import workloads
    n_s, l_s = 22, 2
    c = workloads.random_layer_circuit(n_s, l_s, seed=1)
    ops = c.to_ops()
    oracle.lib()
    t0 = time.perf_counter()
    oracle.execute(n_s, ops)
    dt = time.perf_counter() - t0
    per_layer_at_target = dt / l_s * (2 ** (n_target - n_s))
    return {
        "value": 1.0 / per_layer_at_target,
        "unit": "gate-layers/s",
        "cores": 1,
        "kind": "port",
        "sample": f"scalar oracle, {n_s}q x {l_s} layers of the same generator ({len(ops)} gates), "
                  f"{dt:.2f} s, scaled linearly in amplitudes to {n_target}q",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note "
                                   "(DESIGN.md), --gpus must be 1"}))
        sys.exit(2)
    from quantummatter_amd import workloads
    from quantummatter_amd.backend import (C64, C128, SCHED_FUSED, SCHED_PER_GATE, HipBackendOptions,
                                           HipStateVectorBackend, plan_fused)
    prec = C128 if args.precision == "c128" else C64
    n, L = args.qubits, args.layers
    circuit = workloads.random_layer_circuit(n, L, seed=1)
    ops = circuit.to_ops()
    passes, transposes, gates = plan_fused(n, ops, prec)
    amp_bytes = 16 if prec == C128 else 8
    with HipStateVectorBackend(HipBackendOptions(EnableProfiling=True)) as be:
        with be.CreateSession(n, prec) as s:
            for _ in range(args.warmup):
                s.Reset()
                s.EvolveOps(ops, SCHED_FUSED)
            device_ms = []
            t_total = 0.0
            for _ in range(args.steps):
                s.Reset()
                t0 = time.perf_counter()
                prof = s.EvolveOps(ops, SCHED_FUSED, profile=True)
                t_total += time.perf_counter() - t0
                device_ms.append(prof.DeviceEvolutionMs)
            ms_per_step = t_total / args.steps * 1e3
            dev_ms = sum(device_ms) / len(device_ms)
            launch_ms = dev_ms / prof.Passes
            bytes_per_launch = 2.0 * amp_bytes * (1 << n)
            achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
            extras = {}
            if not args.no_extras:
                # The reference's GPU shape: one launch per gate, on two layers.
                two = workloads.random_layer_circuit(n, 2, seed=1).to_ops()
                s.Reset()
                p2 = s.EvolveOps(two, SCHED_PER_GATE, profile=True)
                extras["per_gate_schedule"] = {
                    "gate_layers_per_s": 2 / (p2.DeviceEvolutionMs * 1e-3),
                    "launches_per_layer": p2.KernelLaunches / 2,
                    "algorithmic_GBps": p2.AlgorithmicBytes / (p2.DeviceEvolutionMs * 1e-3) / 1e9,
                }
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", f"fused_pass_{args.precision}_{n}q_pmc.txt"))
    line = {
        "metric": "30q c128 gate-layers/s + % HBM peak; served circuits/s with p50/p99 latency",
        "value": L * args.steps / t_total,
        "unit": "gate-layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == C128 else "f32",
        "data": "synthetic (seeded SplitMix64 random layer circuit, seed 1)",
        "config": {"workload": f"{n}q {args.precision} random layer circuit, {L} layers "
                               f"({gates} gates), fused tile passes",
                   "qubits": n, "layers": L, "gates": gates, "passes": passes,
                   "passes_per_layer": passes / L, "lds_transposes": transposes,
                   "device_ms_per_step": dev_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_fused (tile pass)", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": launch_ms},
    }
    line.update(extras)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, L)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
