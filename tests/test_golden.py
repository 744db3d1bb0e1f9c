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


def _load_c2():
    d = _load("c2_state_digest.json")
    amps = np.array([complex(float.fromhex(r), float.fromhex(i)) for r, i in d["amps"]])
    return (np.array(d["indices"], dtype=np.int64), amps, float.fromhex(d["sum_p2"]), float.fromhex(d["norm2"]),
            np.array([float.fromhex(x) for x in d["z"]]), np.array([float.fromhex(x) for x in d["zz"]]))


def test_c2_digest_is_consistent():
    """The 30-qubit digest (FP64 oracle, plain per-gate path): the sampled
    indices are the documented draw, the norm is 1, and the sampled
    amplitudes carry about their share of the probability."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_fixtures import sample_indices
    idx, amps, p2, norm2, z, zz = _load_c2()
    assert list(idx) == sample_indices(30)
    assert abs(norm2 - 1) <= 1e-12
    assert len(z) == 30 and len(zz) == 29 and np.all(np.abs(z) < 1) and np.all(np.abs(zz) < 1)
    # a scrambled 30-qubit state: sum |a|^4 near 2 / 2^30 (Porter-Thomas)
    assert 1.5 < p2 * 2 ** 30 < 2.5
    share = np.sum(np.abs(amps) ** 2) * 2 ** 30 / len(amps)
    assert 0.8 < share < 1.2


@pytest.mark.gpu
def test_gpu_c2_matches_digest(gpu):
    """The 30-qubit, 30-layer circuit with the benchmark's default compile
    (hybrid partitioner, specialised kernels, per-group tuning, zero-tile
    skipping) against the FP64 oracle's digest: 4096 amplitudes within 1e-10
    relative (2-norm), sum |a|^4, <Z_q>, <Z_q Z_q+1>.  Then complex64 at the
    same size against the complex128 state: 1 - fidelity <= 1e-5."""
    idx, amps, p2, norm2, z, zz = _load_c2()
    n = 30
    c = gpu.Circuit(n).add_random_layers(30, 0)
    c.compile()
    c.run(copy_back=False, destroy=False)
    got = np.array([c.ampAtGPU(int(i)) for i in idx])
    err = np.linalg.norm(got - amps) / np.linalg.norm(amps)
    print(f"c128 30q: sample rel err {err:.2e}, max |d amp| {np.max(np.abs(got - amps)):.2e}", flush=True)
    assert err <= 1e-10
    pos = c.final_pos()
    st = c.state_view(gpu.HQ_C128)
    gz, gzz = st.expect_z()
    # expect_z reads physical qubits; map them to logical ones
    gz = np.array([gz[pos[q]] for q in range(n)])
    assert np.max(np.abs(gz - z)) <= 1e-10
    if pos == list(range(n)):
        assert np.max(np.abs(gzz - zz)) <= 1e-10
    assert abs(st.norm2() - norm2) <= 1e-12
    psi = st.download()
    p = np.empty(len(psi))
    np.abs(psi, out=p)
    del psi
    np.square(p, out=p)
    assert abs(float(np.dot(p, p)) - p2) <= 1e-9 * p2
    del p
    # complex64, same circuit and size
    c64 = gpu.Circuit(n).add_random_layers(30, 0)
    c64.set_option("precision", gpu.HQ_C64)
    c64.compile()
    c64.run(copy_back=False, destroy=False)
    s64 = c64.state_view(gpu.HQ_C64)
    ov = st.inner(s64) if c64.final_pos() == pos else None
    got64 = np.array([c64.ampAtGPU(int(i)) for i in idx])
    d64 = float(np.max(np.abs(got64 - amps)))
    print(f"c64 30q: max |d amp| {d64:.2e}, 1 - F {1 - abs(ov) ** 2 if ov is not None else None}", flush=True)
    assert d64 <= 1e-6
    if ov is not None:
        assert 1 - abs(ov) ** 2 <= 1e-5
    c64.close()
    c.close()
