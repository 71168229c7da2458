"""Writes c4_hea_gradients.json: the oracle's energy and parameter-shift
gradient for the variational workload (hardware-efficient ansatz of 6 layers,
Ry and Rz on every qubit then a CNOT ladder; 256 random Pauli strings with
N(0,1) coefficients) at 12, 16 and 20 qubits.

The inputs are regenerated from their seeds by quantummatter_amd.workloads, so
the fixture holds only the seeds and the oracle's outputs.  The oracle
(oracle/qm_oracle.c, qmo_energy_gradient) is single-threaded; the 20-qubit case
takes tens of minutes, so the cases run in parallel processes.

Usage: python tests/golden/make_c4_gradients.py [widths...]
"""
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

LAYERS, TERMS = 6, 256


def case_inputs(n):
    from quantummatter_amd import workloads
    circuit = workloads.hardware_efficient_ansatz(n, LAYERS)
    theta = workloads.random_parameters(circuit.ParameterCount, 1000 + n)
    ham = workloads.random_hamiltonian(n, TERMS, 2000 + n)
    return circuit, theta, ham


def one(n):
    import oracle
    circuit, theta, ham = case_inputs(n)
    ops, pidx = circuit.packed()
    paulis, coeffs = ham.packed()
    t0 = time.time()
    e, g = oracle.energy_gradient(n, ops, pidx, theta, paulis, coeffs)
    return {"qubits": n, "layers": LAYERS, "terms": TERMS, "theta_seed": 1000 + n,
            "hamiltonian_seed": 2000 + n, "energy": e, "gradient": [float(x) for x in g],
            "oracle_seconds": round(time.time() - t0, 1)}


def main(widths):
    path = os.path.join(HERE, "c4_hea_gradients.json")
    cases = {}
    if os.path.exists(path):
        cases = {str(c["qubits"]): c for c in json.load(open(path))["cases"]}
    with ProcessPoolExecutor(max_workers=len(widths)) as ex:
        for res in ex.map(one, widths):
            cases[str(res["qubits"])] = res
            print(f"{res['qubits']} qubits: energy {res['energy']!r}, {res['oracle_seconds']} s", flush=True)
            out = {"generator": "tests/golden/make_c4_gradients.py",
                   "cases": [cases[k] for k in sorted(cases, key=int)]}
            with open(path, "w") as f:
                json.dump(out, f, indent=1)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [12, 16, 20])
