"""Writes the workload fixtures of this directory (run from the repository root:
``python tests/golden/make_fixtures.py``).

* ``c1_random_layers_20q_seed0.qasm`` -- the 20-qubit, 20-layer random-layer
  circuit (600 gates, seed 0) in the reference's input dialect
  (src/main.cpp:68-231), angles written with 17 significant digits.
* ``workload_hashes.json`` -- SHA-256 of the same qasm text for the 30-qubit
  (30 layers), 33-qubit and 34-qubit (24 layers) instances, with gate counts.
* ``c1_state_digest.json`` -- the FP64 oracle's final state of the 20-qubit
  circuit, digested: amplitudes 0..63, sum |a|^4, the squared norm and
  <Z_q> for every qubit (hex floats, exact).
* ``c2_state_digest.json`` (``--c2``: about 15 minutes and 24 GiB of memory on
  8 cores) -- the same for the 30-qubit, 30-layer circuit (1350 gates, seed
  0) run on the oracle's plain per-gate path: 4096 amplitudes (indices 0..63
  and 4032 indices drawn by numpy PCG64 seed 30), sum |a|^4, the squared
  norm, <Z_q> for every qubit and <Z_q Z_(q+1)> for every neighbouring pair.

The generator is the test-side restatement (helpers.random_layers), which
tests/test_host.py ties bit for bit to the library's add_random_layers."""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from helpers import random_layers, to_oracle, to_qasm  # noqa: E402

WORKLOADS = {"C1": (20, 20, 0), "C2": (30, 30, 0), "C3": (33, 24, 0), "C3_stretch": (34, 24, 0)}


def qasm_of(n, layers, seed):
    return to_qasm(n, random_layers(n, layers, seed))


def state_digest(psi):
    p = np.abs(psi) ** 2
    n = int(round(np.log2(len(psi))))
    idx = np.arange(len(psi))
    z = [float(np.sum(p * (1 - 2 * ((idx >> q) & 1)))) for q in range(n)]
    return {"amps": [[float(a.real).hex(), float(a.imag).hex()] for a in psi[:64]],
            "sum_p2": float(np.sum(p * p)).hex(), "norm2": float(np.sum(p)).hex(),
            "z": [x.hex() for x in z]}


def sample_indices(n, count=4096, seed=30):
    """0..63 and count - 64 more distinct indices in [64, 2^n), ascending."""
    rs = np.random.default_rng(seed)
    extra = set()
    while len(extra) < count - 64:
        extra.update(int(x) for x in rs.integers(64, 1 << n, size=count))
    return list(range(64)) + sorted(extra)[: count - 64]


def c2_digest():
    """FP64 oracle run of the 30-qubit circuit, in place (16 GiB), digested."""
    n, layers, seed = WORKLOADS["C2"]
    psi = oracle.zero_state(n)
    psi = oracle.run_plain(to_oracle(random_layers(n, layers, seed)), n, psi, copy=False)
    idx = sample_indices(n)
    amps = psi[np.array(idx)]
    p = np.empty(1 << n)
    np.abs(psi, out=p)
    del psi
    np.square(p, out=p)
    z, zz = [], []
    for q in range(n):
        v = p.reshape(-1, 2, 1 << q).sum(axis=(0, 2))
        z.append(float(v[0] - v[1]))
    for q in range(n - 1):
        v = p.reshape(-1, 2, 2, 1 << q).sum(axis=(0, 3))
        zz.append(float(v[0, 0] + v[1, 1] - v[0, 1] - v[1, 0]))
    return {"workload": "random layers, 30 qubits, 30 layers, seed 0", "indices": idx,
            "amps": [[float(a.real).hex(), float(a.imag).hex()] for a in amps],
            "sum_p2": float(np.dot(p, p)).hex(), "norm2": float(np.sum(p)).hex(),
            "z": [x.hex() for x in z], "zz": [x.hex() for x in zz]}


def main():
    if "--c2" in sys.argv[1:]:
        d = c2_digest()
        with open(os.path.join(HERE, "c2_state_digest.json"), "w") as f:
            json.dump(d, f)
            f.write("\n")
        return
    n, layers, seed = WORKLOADS["C1"]
    with open(os.path.join(HERE, "c1_random_layers_20q_seed0.qasm"), "w") as f:
        f.write(qasm_of(n, layers, seed))
    hashes = {}
    for name, (n, layers, seed) in WORKLOADS.items():
        text = qasm_of(n, layers, seed)
        hashes[name] = {"qubits": n, "layers": layers, "seed": seed, "gates": layers * (n + n // 2),
                        "qasm_sha256": hashlib.sha256(text.encode()).hexdigest()}
    with open(os.path.join(HERE, "workload_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1)
        f.write("\n")
    n, layers, seed = WORKLOADS["C1"]
    psi = oracle.run_plain(to_oracle(random_layers(n, layers, seed)), n)
    with open(os.path.join(HERE, "c1_state_digest.json"), "w") as f:
        json.dump({"workload": "random layers, 20 qubits, 20 layers, seed 0", **state_digest(psi)}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
