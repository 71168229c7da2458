"""Benchmark of the state-vector hot path on one MI355X.

Workload (the metric's configuration): a seeded 30-qubit random layer circuit,
40 layers (Rx/Ry/Rz with angles U[0,2pi) on every qubit, then a CNOT brick),
on a resident complex128 state (16 GiB), fused schedule.  A "step" is one
execution of the full 40-layer circuit on the resident state.  `value` is
gate-layers per second over the timed steps (wall clock around K steps, each
synchronised), with the state already in HBM.

The passes (tiles chosen by the planner's tile search, the lowest index bits
remapped between passes) run as compiled straight-line kernels (one per
planned pass, compiled during the warmup and cached by circuit structure;
compile time is reported in config), or with --schedule fused through the
interpreting kernel.  The same circuit on a complex64 state is reported under
"c64_compiled".
Also reported: the dominant kernel's roofline (the tile pass: 2 x 16 B x
2^30 algorithmic bytes per launch, average duration from HIP events recorded
by the library on its stream around the timed sequence), and the CPU baseline
(the reference's parallel AVX-512 CPU path, restated in C++ in oracle/, timed
on a bounded sample on this host's cores and scaled to the same unit).  The
batched forms (noisy trajectories, parameter-shift VQE gradients) are reported
beside it under "batched".
"""
import argparse
import json
import os
import sys
import time

import numpy as np

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
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-form extras")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream extras")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense-block extras")
    ap.add_argument("--report", metavar="PATH", default=None,
                    help="also write the accelerator comparison rows in the reference's benchmark report "
                         "schema (PATH as CSV and PATH.metadata.json), then exit")
    ap.add_argument("--schedule", choices=["compiled", "fused"], default="compiled",
                    help="compiled: fused passes as straight-line kernels (compiled during warmup, "
                         "cached by circuit structure); fused: the interpreting tile-pass kernel")
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


def _cpu_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_target):
    """The reference's parallel CPU path (restated in C++: workers over
    contiguous pair ranges, the AVX-512 / AVX / SIMD-Hadamard dispatch of
    ParallelCpuStateVectorBackend.cs:246-276, optional 32-gate/512-pair
    blocking) timed as the reference's benchmark program times it: 2 warmups,
    5 timed runs, median/min/max, on a state written before the timed region;
    all available physical cores and 8 workers, blocking on and off, AVX-512
    on, plus the AVX path without AVX-512.  Sample: 28 qubits x 1 layer of the
    same generator (a 4 GiB state, DRAM-resident like the 30-qubit one), scaled
    by 2^(n_target - 28) to the target width's gate-layers/s.  Also the other
    workloads on the CPU: terminal sampling (20q x 20 layers + 1024 shots),
    noisy trajectories and parameter-shift circuits (scalar restatement, one
    thread, bounded samples)."""
    from oracle import baseline as B
    host = B.host_description()
    all_cores = max(1, min(host["allowed_cpus"] or 1, host["physical_cores"] or host["allowed_cpus"] or 1))
    workers_list = sorted({all_cores, min(8, all_cores)}, reverse=True)
    n_s, l_s = 28, 1
    rows = B.layer_circuit_rows(n_s, l_s, workers_list, [(True, True, True), (True, True, False)])
    rows += B.layer_circuit_rows(n_s, l_s, [all_cores], [(True, False, False)])
    scale = 2 ** (n_target - n_s) / l_s
    for r in rows:
        r["gate_layers_per_s"] = 1.0 / (r["median_s"] * scale)
    best = max(rows, key=lambda r: r["gate_layers_per_s"])
    checked = B.check_parallel_against_scalar(workers=min(8, all_cores))
    return {
        "value": best["gate_layers_per_s"],
        "unit": "gate-layers/s",
        "cores": best["workers"],
        "kind": "port",
        "sample": f"parallel CPU path, {n_s}q x {l_s} layer of the seeded layer circuit, median of 5 after 2 "
                  f"warmups, best of {len(rows)} settings (workers {workers_list}, blocking on/off, AVX-512 "
                  f"on; AVX only), scaled by 2^{n_target - n_s} to {n_target}q; host {host['model']}, "
                  f"{host['physical_cores']} physical cores, {host['allowed_cpus']} CPUs allowed",
        "method": "Program.cs:161-209 (warmups, timed repetitions, median/min/max); state touched before timing",
        "host": host,
        "equal_to_scalar_oracle": checked,
        "rows": rows,
        "sampling_20q_20l_1024_shots": B.sampling_row(workers=all_cores),
        "trajectories_16q_30l": B.trajectories_row(),
        "vqe_circuits_hea6_256terms": B.variational_rows(),
    }


REPORT_COLUMNS = ("backend,qubits,gates,mix,max_workers,avx512_requested,gate_blocking,median_ms,min_ms,max_ms,"
                  "median_allocated_bytes,state_bytes,device_state_bytes,host_staging_bytes,"
                  "profiled_device_median_ms,process_peak_working_set_bytes,max_amplitude_error,max_norm_error")


def _report_circuit(n, depth, mix):
    """The accelerator benchmark circuits (AcceleratorBenchmarks.cs:170-195,
    CudaBenchmarks.cs:118-146): repeated H, Rx or Phase on the top qubit, an
    8-gate mixed cycle on it, or a mixed cycle over the top three qubits."""
    import math
    from quantummatter_amd import QuantumCircuit
    from quantummatter_amd.circuit import SingleQubitUnitary
    c = QuantumCircuit(n)
    if mix == "varying-high":
        for i in range(depth):
            t = n - 1 - i % 3
            ctl = (t + 1) % n
            k = i % 8
            if k == 0:
                c.H(t)
            elif k == 1:
                c.Rx(t, 0.37)
            elif k == 2:
                c.CNot(ctl, t)
            elif k == 3:
                c.Ry(t, -0.71)
            elif k == 4:
                c.CRz(ctl, t, 1.19)
            elif k == 5:
                c.Swap(t, ctl)
            elif k == 6:
                c.Phase(t, -0.29)
            else:
                c.Toffoli(ctl, (t + 2) % n, t)
        return c
    t, ctl = n - 1, n // 2
    h = math.sqrt(0.5)
    u = SingleQubitUnitary.Create(h, 1j * h, 1j * h, h)
    for i in range(depth):
        if mix == "hadamard-high":
            c.H(t)
        elif mix == "rx-high":
            c.Rx(t, 0.37)
        elif mix == "phase-high":
            c.Phase(t, -0.29)
        else:
            k = i % 8
            if k == 0:
                c.H(t)
            elif k == 1:
                c.Rx(t, 0.37)
            elif k == 2:
                c.Ry(t, -0.71)
            elif k == 3:
                c.Rz(t, 1.19)
            elif k == 4:
                c.CNot(ctl, t)
            elif k == 5:
                c.Swap(t, (t + 1) % n)
            elif k == 6:
                c.CS(ctl, t)
            else:
                c.U(t, u)
    return c


def _report_check(expected, state):
    """CudaBenchmarks.Check (CudaBenchmarks.cs:148-170): max amplitude error and
    norm error, at most 3e-12 and 1e-11."""
    err = float(np.max(np.abs(expected - state)))
    # Pairwise summation (numpy) instead of the reference's Kahan loop: both
    # bound the summation error far below the 1e-11 tolerance.
    nerr = abs(float(np.sum(state.real * state.real + state.imag * state.imag)) - 1)
    if not (err <= 3e-12 and nerr <= 1e-11):
        raise RuntimeError(f"numerical validation failed: amplitude {err!r}, norm {nerr!r}")
    return err, nerr


def accelerator_report(be, path, repetitions=5, warmups=2, sizes=(12, 20, 25)):
    """The reference's accelerator comparison (AcceleratorBenchmarks.cs:10-167)
    in its report schema, with this backend's drop-in Execute as the "hip" row
    (host buffer in and out: device allocation, upload, fused passes, download)
    and its profiled device time, beside the parallel CPU path restated in
    oracle/ (cpu-{256,512}-w{1,4,8}-block-{false,true}, every width forced
    parallel).  Every output is checked against the scalar oracle."""
    import ctypes as C
    import resource
    import oracle
    from oracle import baseline
    from quantummatter_amd.backend import jit_drain
    sizes = list(sizes)
    configs = []
    for workers in (1, 4, 8):
        for blocking in (False, True):
            for wide in (False, True):
                configs.append((f"cpu-{512 if wide else 256}-w{workers}-block-{str(blocking).lower()}",
                                workers, wide, blocking))
    configs.append(("hip", 0, False, False))
    meta = {
        "timestampUtc": time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime()),
        "privacy": "Synthetic benchmark; host CPU model recorded, nothing else about the installation.",
        "precision": "double real/imaginary; the cpu rows round every product and sum (no FMA); the hip rows "
                     "use fused multiply-adds in the fused passes (amplitudes within 3e-12 of the oracle)",
        "sizes": sizes,
        "depths": "64 gates for 12/20 qubits; 8 gates for 25 qubits",
        "mixes": "hadamard-high, mixed-high, varying-high; plus rx-high and phase-high at 20 qubits",
        "repetitions": repetitions, "warmupsPerBackendAndCase": warmups,
        "configurations": [{"Name": c[0], "Workers": c[1], "Wide": c[2], "Blocking": c[3]} for c in configs],
        "forcedParallelThresholdQubits": 1,
        "totalTiming": "Ordinary synchronous Execute; hip includes device allocation/free, upload, the tile "
                       "passes (compiled kernels of the structure, built in the background during the warmups), "
                       "download and commit into the host buffer",
        "deviceTiming": "Median of five ExecuteProfiled calls after the timed ones: HIP events around the "
                        "schedule on the library's stream, excluding transfers and allocation",
        "excluded": "Context creation, input reset, warmup and numerical validation",
        "ordering": "Rotate all CPU/HIP variants each repetition; reverse on odd repetitions",
        "allocations": "median_allocated_bytes is 0: no managed runtime; the hip row allocates one device "
                       "state per call (device_state_bytes)",
        "workingSet": "Process peak resident set (ru_maxrss), cumulative",
        "cpuMemory": "16N state; blocked CPU adds up to 16 KiB stack per active worker",
        "hipMemory": "16N device state per call; no host staging copy (the host buffer is copied directly)",
        "validation": "Independent scalar oracle (oracle/qm_oracle.c); amplitude <=3e-12, norm error "
                      "(pairwise sum) <=1e-11; warmups and every timed/profiled output checked",
        "host": baseline.host_description(),
        "device": {"name": be.Device.Name, "architecture": be.Device.Architecture,
                   "memory_bytes": be.Device.MemoryBytes},
    }
    with open(path + ".metadata.json", "w") as f:
        json.dump(meta, f, indent=2)
    lines = [REPORT_COLUMNS]
    P = C.c_void_p
    for n in sizes:
        mixes = (["hadamard-high", "mixed-high", "varying-high", "rx-high", "phase-high"] if n == 20
                 else ["hadamard-high", "mixed-high", "varying-high"])
        depth = 8 if n == 25 else 64
        for mix in mixes:
            circuit = _report_circuit(n, depth, mix)
            ops = np.ascontiguousarray(circuit.to_ops())
            N = 1 << n
            inp = (1 / np.sqrt(N)) * np.exp(1j * 0.17 * np.arange(N))  # Complex.FromPolarCoordinates(scale, 0.17 j)
            expected = oracle.execute(n, ops, inp)
            state = np.empty(N, dtype=np.complex128)
            samples = [[] for _ in configs]
            errors = [0.0] * len(configs)
            norms = [0.0] * len(configs)

            def run(k):
                name, workers, wide, blocking = configs[k]
                state[:] = inp
                t0 = time.perf_counter()
                if name == "hip":
                    be.Execute(circuit, state)
                else:
                    rc = oracle.lib().qmo_parallel_execute_ex(n, ops.ctypes.data_as(P), len(ops),
                                                              state.ctypes.data_as(P), workers, 1, int(wide),
                                                              int(blocking), 1)
                    if rc:
                        raise RuntimeError(f"qmo_parallel_execute_ex: {rc}")
                dt = (time.perf_counter() - t0) * 1e3
                e, ne = _report_check(expected, state)
                errors[k] = max(errors[k], e)
                norms[k] = max(norms[k], ne)
                return dt

            for k in range(len(configs)):
                for _ in range(warmups):
                    run(k)
            # The drop-in Execute compiles a structure's pass kernels in the
            # background on its first run; wait for them and warm them up.
            jit_drain()
            run(len(configs) - 1)
            for it in range(repetitions):
                for off in range(len(configs)):
                    k = (it + off) % len(configs)
                    if it & 1:
                        k = len(configs) - 1 - k
                    samples[k].append(run(k))
            dev = []
            for _ in range(repetitions):
                state[:] = inp
                dev.append(be.ExecuteProfiled(circuit, state).DeviceEvolutionMs)
                e, ne = _report_check(expected, state)
                errors[-1] = max(errors[-1], e)
                norms[-1] = max(norms[-1], ne)
            dev.sort()
            for k, (name, workers, wide, blocking) in enumerate(configs):
                t = sorted(samples[k])
                hip = name == "hip"
                sb = 16 * N
                peak = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss * 1024
                row = [name, n, depth, mix, workers, str(wide), str(blocking), repr(t[len(t) // 2]), repr(t[0]),
                       repr(t[-1]), 0, sb, sb if hip else 0, 0, repr(dev[len(dev) // 2]) if hip else 0, peak,
                       repr(errors[k]), repr(norms[k])]
                lines.append(",".join(str(v) for v in row))
            hip_t = sorted(samples[-1])
            print(f"report: {n}q {depth} gates {mix}: hip {hip_t[len(hip_t) // 2]:.4f} ms "
                  f"(device {dev[len(dev) // 2]:.4f} ms)", file=sys.stderr, flush=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def batched_extras(be):
    """Batched forms: 4096 noisy 16q trajectories x 30 layers with 1024 shots each,
    and parameter-shift VQE energy+gradient (HEA, 6 layers, 256 Pauli terms)."""
    import numpy as np
    from quantummatter_amd import workloads
    from quantummatter_amd.backend import C64, C128
    from quantummatter_amd.batched import HipBatchedEnergyEvaluator, HipTrajectoryRunner, HipVqeProgram
    out = {}
    n, L, T, shots = 16, 30, 4096, 1024
    nc = workloads.noisy_layer_circuit(n, L, seed=3)
    steps, _ = nc.packed()
    stride = int(np.sum(steps["is_channel"] == 1)) + 1
    pools = workloads.uniform_pools(7, T, stride)
    su = workloads.uniforms(8, T * shots).reshape(T, shots)
    ref = None
    for name, prec in (("c128", C128), ("c64", C64)):
        r = HipTrajectoryRunner(be, prec)
        r.Run(nc, pools[:64], shots, su[:64], True)
        t0 = time.perf_counter()
        res = r.Run(nc, pools, shots, su, True, profile=True)
        wall = time.perf_counter() - t0
        p = res.profile
        gbps = p.AlgorithmicBytes / (p.DeviceEvolutionMs * 1e-3) / 1e9
        row = {
            "trajectories_per_s": T / wall, "wall_s": wall, "device_ms": p.DeviceEvolutionMs,
            "passes_per_trajectory": p.Passes, "algorithmic_GBps": gbps, "hbm_frac": gbps / HBM_PEAK_GBS,
            "shots_per_trajectory": shots, "boundary_fallbacks": res.boundary_fallbacks}
        if ref is None:
            ref = res
        else:
            # complex64 against complex128 on the same pools and shot draws.
            row["vs_c128"] = {
                "trajectories_same_draw_count": float(np.mean(res.consumed == ref.consumed)),
                "measure_all_outcomes_equal": float(np.mean(res.outcomes == ref.outcomes)),
                "samples_equal": float(np.mean(res.samples == ref.samples))}
        out[f"trajectories_16q_30l_{name}"] = row
    for q in (12, 16, 20):
        circ = workloads.hardware_efficient_ansatz(q, 6)
        theta = workloads.random_parameters(circ.ParameterCount, 1)
        ham = workloads.random_hamiltonian(q, 256, 2)
        ref = None
        for name, prec in (("c128", C128), ("c64", C64)):
            ev = HipBatchedEnergyEvaluator(be, prec)
            ev.EvaluateWithGradient(circ, ham, theta)
            t0 = time.perf_counter()
            got = ev.EvaluateWithGradient(circ, ham, theta, profile=True)
            wall = time.perf_counter() - t0
            p = ev.last_profile
            circuits = 1 + 2 * circ.ParameterCount
            gbps = p.AlgorithmicBytes / (p.DeviceEvolutionMs * 1e-3) / 1e9
            # AlgorithmicBytes counts one read of each state per Pauli flip group;
            # those reads are served by the last-level cache, so the HBM figure
            # takes each state's passes plus one final read.
            hbm_bytes = circuits * p.StateBytes * (2.0 * p.Passes + 1.0)
            hbm_gbps = hbm_bytes / (p.DeviceEvolutionMs * 1e-3) / 1e9
            # The same gradient replayed from the captured graph of a VQE program
            # (arena kept, angles uploaded from pinned memory), mean of 5 calls.
            with HipVqeProgram(be, circ, ham, prec) as prog:
                prog.EvaluateWithGradient(circ, ham, theta)
                t1 = time.perf_counter()
                for _ in range(5):
                    prog.EvaluateWithGradient(circ, ham, theta)
                replay = (time.perf_counter() - t1) / 5
            row = {
                "circuits_per_s": circuits / wall, "circuits": circuits, "wall_s": wall,
                "device_ms": p.DeviceEvolutionMs, "algorithmic_GBps": gbps,
                "hbm_GBps": hbm_gbps, "hbm_frac": hbm_gbps / HBM_PEAK_GBS,
                "graph_replay_circuits_per_s": circuits / replay, "graph_replay_wall_s": replay}
            if ref is None:
                ref = got
            else:
                scale = float(np.sum(np.abs(ham.packed()[1])))
                row["vs_c128"] = {
                    "energy_abs_error": abs(got.Energy - ref.Energy),
                    "gradient_max_abs_error": float(np.max(np.abs(np.array(got.Gradient) - np.array(ref.Gradient)))),
                    "relative_to_sum_abs_coefficients": scale}
            out[f"vqe_gradient_{q}q_{name}"] = row
    return out


def dense_extras(be, n=30, layers=4):
    """The dense-block variant: gates fused into blocks of k <= 5 qubits, one
    matrix-core pass per block (FP64 MFMA in complex128, FP32 MFMA in
    complex64), on the same generator; the complex64 rows carry their
    infidelity against the complex128 dense state of the same circuit."""
    from quantummatter_amd import workloads
    from quantummatter_amd.backend import C64, C128
    c = workloads.random_layer_circuit(n, layers, seed=1)
    out = {}
    peaks = {C128: 78.6e12, C64: 157.3e12}
    with be.CreateSession(n, C128) as s, be.CreateSession(n, C64) as s64:
        for k in (2, 3, 4, 5):
            for name, st in (("c128", s), ("c64", s64)):
                prec = C128 if name == "c128" else C64
                st.Reset()
                p, flops = st.EvolveDense(c, k, profile=True)
                t = p.DeviceEvolutionMs * 1e-3
                row = {"gate_layers_per_s": layers / t, "passes": p.Passes,
                       "ms_per_pass": p.DeviceEvolutionMs / p.Passes,
                       "algorithmic_GBps": p.AlgorithmicBytes / t / 1e9,
                       "mfma_TFLOPs": flops / t / 1e12, "mfma_frac_peak": flops / t / peaks[prec],
                       "mfma_peak_TFLOPs": peaks[prec] / 1e12}
                if name == "c64":
                    row["infidelity_vs_c128"] = 1.0 - abs(st.Inner(s)) ** 2 / (st.Mass() * s.Mass())
                out[f"k{k}_{name}"] = row
    return out


def _percentiles(lat_ms):
    import numpy as np
    return float(np.percentile(lat_ms, 50)), float(np.percentile(lat_ms, 99))


def served_extras(be, count=48, full_lanes=3, batch_max_qubits=14, rate_fractions=(0.25, 0.5, 1.0)):
    """The mixed request stream (widths log-uniform 10..28, depth 10..100 layers,
    amplitudes / 1024 shots / 64-term expectation) while a 33-qubit complex128
    state (128 GiB, evolved through 4 layers) stays resident on the device: the
    stream is admitted against the HBM left beside it.  Plain: one request
    after another through resident sessions (each request's service time is
    measured).  Served: the request server (admission by HBM footprint, batched
    small widths, concurrent full-chip lanes, pinned staging of amplitude
    results).  First with every request arriving at t = 0 (saturation), then
    with Poisson arrivals at fractions of the served saturation rate: requests
    are submitted at their arrival times and latency is completion minus
    arrival; the plain figures at the same arrivals come from its measured
    service times in a first-come-first-served queue."""
    import math
    import numpy as np
    from quantummatter_amd import workloads
    from quantummatter_amd.backend import C128, SCHED_FUSED
    from quantummatter_amd.serving import AMPLITUDES, SHOTS, HipRequestServer
    reqs = workloads.serving_requests(count, seed=2024)
    pay = [workloads.serving_payload(r) for r in reqs]
    resident = be.CreateSession(33, C128)
    resident.Evolve(workloads.random_layer_circuit(33, 4, seed=99), SCHED_FUSED)
    budget = be.Device.MemoryBytes - (16 << 33) - (8 << 30)
    sweep = []
    try:
        service = []
        t0 = time.perf_counter()
        for r, (c, u, obs) in zip(reqs, pay):
            t1 = time.perf_counter()
            with be.CreateSession(r["qubits"], C128) as s:
                s.Evolve(c, SCHED_FUSED)
                if r["kind"] == AMPLITUDES:
                    s.Snapshot()
                elif r["kind"] == SHOTS:
                    s.Sample(u)
                else:
                    s.ExpectationValues(obs)
            service.append(time.perf_counter() - t1)
        plain_span = time.perf_counter() - t0
        plain_done = np.cumsum(service) * 1e3
        with HipRequestServer(full_lanes=full_lanes, batch_max_qubits=batch_max_qubits,
                              device_budget_bytes=budget) as sv:
            t0 = time.perf_counter()
            ids = [sv.Submit(c, r["kind"], C128, u, obs) for r, (c, u, obs) in zip(reqs, pay)]
            for i in ids:
                sv.Wait(i)
            span = time.perf_counter() - t0
            st = sv.Stats()
            sat_rate = count / span
            gaps = workloads.uniforms(77, count)
            for frac in rate_fractions:
                lam = frac * sat_rate
                arrivals = np.cumsum([-math.log(1.0 - g) / lam for g in gaps])
                arrivals -= arrivals[0]
                sv.ResetStats()
                t0 = time.perf_counter()
                ids = []
                for a, r, (c, u, obs) in zip(arrivals, reqs, pay):
                    delay = a - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                    ids.append(sv.Submit(c, r["kind"], C128, u, obs))
                results = [sv.Wait(i) for i in ids]
                wall = time.perf_counter() - t0
                lat = [res.latency_ms for res in results]
                # The plain path at the same arrivals: a first-come-first-served queue.
                done, plain_lat = 0.0, []
                for a, sv_s in zip(arrivals, service):
                    done = max(done, a) + sv_s
                    plain_lat.append((done - a) * 1e3)
                p50, p99 = _percentiles(lat)
                q50, q99 = _percentiles(plain_lat)
                sweep.append({"rate_per_s": lam, "fraction_of_saturation": frac,
                              "circuits_per_s": count / wall, "p50_ms": p50, "p99_ms": p99,
                              "plain_circuits_per_s": count / done, "plain_p50_ms": q50, "plain_p99_ms": q99,
                              "p99_vs_plain": q99 / p99 if p99 > 0 else None})
    finally:
        resident.Dispose()
    widths = [r["qubits"] for r in reqs]
    q50, q99 = _percentiles(plain_done)
    return {
        "requests": len(reqs), "widths": [min(widths), max(widths)],
        "resident_33q_c128_GiB": 128, "budget_GB": budget / 1e9,
        "circuits_per_s": len(reqs) / span, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
        "plain_circuits_per_s": len(reqs) / plain_span,
        "plain_p50_ms": q50, "plain_p99_ms": q99,
        "speedup_vs_plain": plain_span / span, "batched_launches": st.batches,
        "batched_requests": st.batched_requests, "peak_hbm_GB": st.peak_reserved_bytes / 1e9,
        "arrivals": "all at t=0 (saturation); Poisson sweep below", "full_lanes": full_lanes,
        "batch_max_qubits": batch_max_qubits, "poisson": sweep}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU (--gpus 1); sharding a state or "
                                   "a batch across GPUs is not built"}))
        sys.exit(2)
    from quantummatter_amd import workloads
    import ctypes as C
    from quantummatter_amd._lib import lib
    from quantummatter_amd.backend import (C64, C128, SCHED_COMPILED, SCHED_FUSED, SCHED_PER_GATE,
                                           HipBackendOptions, HipStateVectorBackend, plan_fused)
    sched = SCHED_COMPILED if args.schedule == "compiled" else SCHED_FUSED
    prec = C128 if args.precision == "c128" else C64
    n, L = args.qubits, args.layers
    circuit = workloads.random_layer_circuit(n, L, seed=1)
    ops = circuit.to_ops()
    gates = plan_fused(n, ops, prec)[2]
    amp_bytes = 16 if prec == C128 else 8
    if args.report:
        with HipStateVectorBackend(HipBackendOptions(EnableProfiling=True)) as be:
            accelerator_report(be, args.report)
        return
    with HipStateVectorBackend(HipBackendOptions(EnableProfiling=True)) as be:
        with be.CreateSession(n, prec) as s:
            t_warm = time.perf_counter()
            for _ in range(max(args.warmup, 1 if sched == SCHED_COMPILED else 0)):
                s.Reset()
                s.EvolveOps(ops, sched)
            t_warm = time.perf_counter() - t_warm
            compiled, hits, compile_s = C.c_int64(), C.c_int64(), C.c_double()
            lib().qm_jit_stats(C.byref(compiled), C.byref(hits), C.byref(compile_s))
            device_ms = []
            t_total = 0.0
            for _ in range(args.steps):
                s.Reset()
                t0 = time.perf_counter()
                prof = s.EvolveOps(ops, sched, profile=True)
                t_total += time.perf_counter() - t0
                device_ms.append(prof.DeviceEvolutionMs)
            ms_per_step = t_total / args.steps * 1e3
            dev_ms = sum(device_ms) / len(device_ms)
            launch_ms = dev_ms / prof.Passes
            bytes_per_launch = 2.0 * amp_bytes * (1 << n)
            achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
            extras = {}
            if not args.no_extras and sched == SCHED_COMPILED:
                # The same passes through the interpreting kernel (no compilation).
                s.Reset()
                pf = s.EvolveOps(ops, SCHED_FUSED, profile=True)
                extras["interpreted_fused_schedule"] = {
                    "gate_layers_per_s": L / (pf.DeviceEvolutionMs * 1e-3),
                    "ms_per_pass": pf.DeviceEvolutionMs / pf.Passes,
                    "algorithmic_GBps": pf.AlgorithmicBytes / (pf.DeviceEvolutionMs * 1e-3) / 1e9,
                }
            if not args.no_extras and sched == SCHED_COMPILED and prec == C128:
                # The same circuit on a complex64 state (8 GiB), compiled passes
                # (FP32 arithmetic), and its infidelity against the complex128
                # state of the same circuit (FP64 inner product on the device).
                with be.CreateSession(n, C64) as s64:
                    s64.EvolveOps(ops, sched)  # compiles the c64 kernels
                    s64.Reset()
                    p64 = s64.EvolveOps(ops, sched, profile=True)
                    ov = s64.Inner(s)
                    infid = 1.0 - abs(ov) ** 2 / (s64.Mass() * s.Mass())
                    extras["c64_compiled"] = {
                        "gate_layers_per_s": L / (p64.DeviceEvolutionMs * 1e-3),
                        "passes": p64.Passes, "ms_per_pass": p64.DeviceEvolutionMs / p64.Passes,
                        "algorithmic_GBps": p64.AlgorithmicBytes / (p64.DeviceEvolutionMs * 1e-3) / 1e9,
                        "infidelity_vs_c128": infid,
                    }
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
        if not args.no_extras and not args.no_batched:
            extras["batched"] = batched_extras(be)
        if not args.no_extras and not args.no_served:
            extras["served"] = served_extras(be)
        if not args.no_extras and not args.no_dense:
            extras["dense_blocks_mfma"] = dense_extras(be, n)
    traffic = pmc_traffic(os.path.join(ROOT, "profiles",
                                       f"{args.schedule}_pass_{args.precision}_{n}q_pmc.txt"))
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
                               f"({gates} gates), {args.schedule} tile passes",
                   "qubits": n, "layers": L, "gates": gates, "passes": prof.Passes,
                   "passes_per_layer": prof.Passes / L,
                   "device_ms_per_step": dev_ms, "schedule": args.schedule,
                   "pass_kernels_compiled": compiled.value, "compile_s": compile_s.value,
                   "warmup_s": t_warm},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "qm_pass (compiled tile pass)" if sched == SCHED_COMPILED
                     else "k_fused (tile pass)", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": launch_ms},
    }
    line.update(extras)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
