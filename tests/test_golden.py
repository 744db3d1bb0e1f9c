"""Workload fixtures (tests/golden, written by make_fixtures.py): the 20-qubit
random-layer circuit as qasm, SHA-256 of the 30/33/34-qubit instances, and a
digest of the 20-qubit circuit's final state from the FP64 oracle."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from helpers import random_layers, to_oracle, to_qasm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f) if name.endswith(".json") else f.read()


def _digest_arrays(d):
    amps = np.array([complex(float.fromhex(r), float.fromhex(i)) for r, i in d["amps"]])
    return amps, float.fromhex(d["sum_p2"]), float.fromhex(d["norm2"]), np.array([float.fromhex(x) for x in d["z"]])


def test_workload_hashes():
    for name, w in _load("workload_hashes.json").items():
        text = to_qasm(w["qubits"], random_layers(w["qubits"], w["layers"], w["seed"]))
        assert text.count(";") - 3 == w["gates"], name
        assert hashlib.sha256(text.encode()).hexdigest() == w["qasm_sha256"], name


def test_c1_qasm_is_the_library_workload(hq):
    text = _load("c1_random_layers_20q_seed0.qasm")
    assert hashlib.sha256(text.encode()).hexdigest() == _load("workload_hashes.json")["C1"]["qasm_sha256"]
    parsed = hq.Circuit.from_qasm(text).gate_array()
    built = hq.Circuit(20).add_random_layers(20, 0).gate_array()
    assert len(parsed) == len(built) == 600
    for f in ("type", "target", "control", "control2"):
        assert np.array_equal(parsed[f], built[f])
    # 17 significant digits round-trip every angle exactly
    assert np.array_equal(parsed["mat"], built["mat"])


def test_oracle_reproduces_c1_digest():
    amps, p2, norm2, z = _digest_arrays(_load("c1_state_digest.json"))
    psi = oracle.run_plain(to_oracle(random_layers(20, 20, 0)), 20)
    p = np.abs(psi) ** 2
    idx = np.arange(len(psi))
    assert np.max(np.abs(psi[:64] - amps)) <= 1e-15
    assert abs(np.sum(p * p) - p2) <= 1e-15 and abs(np.sum(p) - norm2) <= 1e-13
    zz = np.array([np.sum(p * (1 - 2 * ((idx >> q) & 1))) for q in range(20)])
    assert np.max(np.abs(zz - z)) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [None, 1, 3])
def test_gpu_c1_matches_digest(gpu, mode):
    amps, p2, norm2, z = _digest_arrays(_load("c1_state_digest.json"))
    c = gpu.Circuit.from_qasm(_load("c1_random_layers_20q_seed0.qasm"))
    if mode is not None:
        c.set_option("mode", mode)
    c.compile()
    c.run(copy_back=False, destroy=False)
    psi = c.logical_state()
    p = np.abs(psi) ** 2
    assert np.max(np.abs(psi[:64] - amps)) <= 1e-12
    assert abs(np.sum(p * p) - p2) <= 1e-9 * p2 and abs(np.sum(p) - norm2) <= 1e-12
    idx = np.arange(len(psi))
    zz = np.array([np.sum(p * (1 - 2 * ((idx >> q) & 1))) for q in range(20)])
    assert np.max(np.abs(zz - z)) <= 1e-11
    c.close()
