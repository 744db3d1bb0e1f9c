"""Pins of the CPU oracle (oracle/hq_oracle.c) against checks that do not
restate its formulas: textbook matrix definitions, a brute-force dense
simulator built from the Kronecker definition, known-answer circuits,
invariants, hand-traced partitions and hand-written text.  CPU only."""
import math

import numpy as np
import pytest

import oracle
from helpers import (CCX, CNOT, CRX, CRY, CRZ, CU1, CY, CZ, H, RX, RY, RZ, S, SDG, T, TDG, U1, U2, U3, X, Y, Z,
                     random_mixed, rel_err, to_oracle)

I2 = np.eye(2)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]])
PZ = np.diag([1.0, -1.0]).astype(complex)


def expm_pauli(P, theta):
    # exp(-i theta/2 P) for a Pauli P (P^2 = I)
    return math.cos(theta / 2) * I2 - 1j * math.sin(theta / 2) * P


def textbook(t, p):
    e = np.exp
    if t in (X, CNOT, CCX):
        return PX
    if t in (Y, CY):
        return PY
    if t in (Z, CZ):
        return PZ
    if t == H:
        return (PX + PZ) / math.sqrt(2)
    if t == S:
        return np.diag([1, 1j])
    if t == SDG:
        return np.diag([1, -1j])
    if t == T:
        return np.diag([1, e(1j * math.pi / 4)])
    if t == TDG:
        return np.diag([1, e(-1j * math.pi / 4)])
    if t in (RX, CRX):
        return expm_pauli(PX, p[0])
    if t in (RY, CRY):
        return expm_pauli(PY, p[0])
    if t in (RZ, CRZ):
        return expm_pauli(PZ, p[0])
    if t in (U1, CU1):
        return np.diag([1, e(1j * p[0])])
    if t == U3:
        th, ph, lam = p
        return np.array([[math.cos(th / 2), -e(1j * lam) * math.sin(th / 2)],
                         [e(1j * ph) * math.sin(th / 2), e(1j * (ph + lam)) * math.cos(th / 2)]])
    if t == U2:
        return textbook(U3, (math.pi / 2, p[0], p[1]))
    raise ValueError(t)


def mat_of(rec):
    m = np.asarray(rec["mat"])
    return (m[0::2] + 1j * m[1::2]).reshape(2, 2)


NPAR = {RX: 1, RY: 1, RZ: 1, U1: 1, CRX: 1, CRY: 1, CRZ: 1, CU1: 1, U2: 2, U3: 3}


@pytest.mark.parametrize("t", [H, X, Y, Z, S, SDG, T, TDG, RX, RY, RZ, U1, U2, U3, CNOT, CY, CZ, CRX, CRY, CRZ,
                               CU1, CCX])
def test_gate_matrices_closed_form(ora, t):
    rs = np.random.default_rng(t)
    for _ in range(5):
        p = tuple(rs.uniform(-7, 7, NPAR.get(t, 0)))
        ctl = 1 if t in (CNOT, CY, CZ, CRX, CRY, CRZ, CU1, CCX) else -1
        c2 = 2 if t == CCX else -1
        rec = oracle.gate(t, 0, ctl, c2, p)
        m = mat_of(rec)
        assert np.allclose(m, textbook(t, p), atol=1e-15, rtol=0)
        assert np.allclose(m.conj().T @ m, I2, atol=1e-15)


def test_gate_make_rejects_bad_arity(ora):
    with pytest.raises(ValueError):
        oracle.gate(RX, 0)            # missing angle
    with pytest.raises(ValueError):
        oracle.gate(CNOT, 0)          # missing control
    with pytest.raises(ValueError):
        oracle.gate(H, 0, 1)          # control on a single-qubit gate


@pytest.mark.parametrize("n,seed", [(1, 0), (3, 1), (4, 2), (6, 3)])
def test_plain_equals_dense_definition(ora, n, seed):
    spec = random_mixed(n, 40, seed, ccx=n >= 3)
    g = to_oracle(spec)
    U = oracle.dense_unitary(g, n)
    assert np.allclose(U.conj().T @ U, np.eye(1 << n), atol=1e-12)
    for col in (0, (1 << n) - 1):
        psi0 = np.zeros(1 << n, dtype=complex)
        psi0[col] = 1
        assert rel_err(oracle.run_plain(g, n, psi0), U[:, col]) <= 1e-13


def test_dense_definition_single_gates(ora):
    # 2 qubits: the embedding convention (qubit 0 = least significant bit)
    U = oracle.dense_unitary(to_oracle([(X, 0, -1, ())]), 2)
    assert np.allclose(U, np.kron(I2, PX))
    U = oracle.dense_unitary(to_oracle([(CNOT, 1, 0, ())]), 2)   # control 0, target 1
    want = np.zeros((4, 4))
    for i in range(4):
        j = i ^ 2 if i & 1 else i
        want[j, i] = 1
    assert np.allclose(U, want)


def test_known_answer_hadamard_ghz(ora):
    n = 10
    psi = oracle.run_plain(to_oracle([(H, q, -1, ()) for q in range(n)]), n)
    assert np.allclose(psi, 2 ** (-n / 2), atol=1e-15)
    ghz = [(H, 0, -1, ())] + [(CNOT, q + 1, q, ()) for q in range(n - 1)]
    psi = oracle.run_plain(to_oracle(ghz), n)
    want = np.zeros(1 << n, dtype=complex)
    want[0] = want[-1] = 1 / math.sqrt(2)
    assert np.allclose(psi, want, atol=1e-15)


def test_known_answer_qft(ora):
    n, x = 7, 0b1011001
    N = 1 << n
    spec = [(X, q, -1, ()) for q in range(n) if (x >> q) & 1]
    for j in range(n - 1, -1, -1):
        spec.append((H, j, -1, ()))
        for k in range(j - 1, -1, -1):
            spec.append((CU1, j, k, (math.pi / 2 ** (j - k),)))
    psi = oracle.run_plain(to_oracle(spec), n)
    rev = [int(format(i, f"0{n}b")[::-1], 2) for i in range(N)]
    want = np.exp(2j * math.pi * x * np.arange(N) / N) / math.sqrt(N)
    assert np.allclose(psi[rev], want, atol=1e-14)


def test_known_answer_bernstein_vazirani(ora):
    n, secret = 9, 0b10110101
    anc = n - 1
    spec = [(X, anc, -1, ()), (H, anc, -1, ())] + [(H, q, -1, ()) for q in range(n - 1)]
    spec += [(CNOT, anc, q, ()) for q in range(n - 1) if (secret >> q) & 1]
    spec += [(H, q, -1, ()) for q in range(n - 1)]
    p = np.abs(oracle.run_plain(to_oracle(spec), n)) ** 2
    data = np.arange(1 << n) & ((1 << (n - 1)) - 1)
    assert abs(p[data == secret].sum() - 1) <= 1e-13


@pytest.mark.parametrize("n,mask", [(8, 0b11110111), (10, 0b1100111011), (9, 0b111111111), (11, 0b00111111100)])
def test_sharemem_semantics_equal_plain(ora, n, mask):
    from helpers import group_for_mask, rand_state
    g = group_for_mask(n, mask, 150, n)
    psi0 = rand_state(n, 1)
    assert rel_err(oracle.sharemem_group(psi0, n, g, mask), oracle.run_plain(g, n, psi0)) <= 1e-13


def test_fuse_equals_embedded_product(ora):
    rs = np.random.default_rng(5)
    for mq in (1, 2, 4):
        n = mq
        spec = random_mixed(n, 25, mq, ccx=n >= 3) if n > 1 else [(RX, 0, -1, (0.3,)), (H, 0, -1, ())]
        g = to_oracle(spec)
        pos = list(range(n))
        m = oracle.fuse_group(g, pos, mq)
        K = 1 << mq
        U = m.reshape(K, K).T
        assert np.allclose(U, oracle.dense_unitary(g, n), atol=1e-13)
        assert np.allclose(U.conj().T @ U, np.eye(K), atol=1e-13)
    # positions: gates on logical qubits 5, 9 mapped to matrix bits 0, 1
    g = to_oracle([(H, 5, -1, ()), (CNOT, 9, 5, ())])
    pos = [0] * 10
    pos[5], pos[9] = 0, 1
    U = oracle.fuse_group(g, pos, 2).reshape(4, 4).T
    assert np.allclose(U, oracle.dense_unitary(to_oracle([(H, 0, -1, ()), (CNOT, 1, 0, ())]), 2))


def test_transpose_properties(ora):
    n = 9
    rs = np.random.default_rng(0)
    psi = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    assert np.array_equal(oracle.transpose(psi, n, list(range(n))), psi)
    perm = [int(x) for x in rs.permutation(n)]
    inv = [0] * n
    for i, p in enumerate(perm):
        inv[p] = i
    assert np.array_equal(oracle.transpose(oracle.transpose(psi, n, perm), n, inv), psi)
    # swapping bit positions 0 and 3 maps index i to i with those bits exchanged
    sw = list(range(n))
    sw[0], sw[3] = 3, 0
    out = oracle.transpose(psi, n, sw)
    idx = np.arange(1 << n)
    src = (idx & ~0b1001) | ((idx & 1) << 3) | ((idx >> 3) & 1)
    assert np.array_equal(out, psi[src])


def test_dense_runs_matmul(ora):
    rs = np.random.default_rng(1)
    n, mq = 8, 3
    K = 1 << mq
    U = rs.normal(size=(K, K)) + 1j * rs.normal(size=(K, K))
    psi = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    out = oracle.dense_runs(psi, n, U.T.reshape(-1), mq)
    assert np.allclose(out.reshape(-1, K), (U @ psi.reshape(-1, K).T).T, atol=1e-12)


# ---- partitioner ------------------------------------------------------------------

def _check_partition_invariants(spec, n, groups, local_size, mat_lim):
    seen = sorted(i for g in groups for i in g["gates"])
    assert seen == list(range(len(spec)))
    order = [i for g in groups for i in g["gates"]]
    # per-qubit order preserved
    last = {}
    for i in order:
        s = spec[i]
        qs = [s[1]] + ([s[2]] if s[2] >= 0 else []) + ([s[4]] if len(s) > 4 else [])
        for q in qs:
            assert last.get(q, -1) < i
            last[q] = i
    for g in groups:
        if g["backend"] == oracle.PERGATE:
            assert bin(g["related"]).count("1") <= local_size
        else:
            assert bin(g["related"]).count("1") <= mat_lim


def _execute_partition(spec, n, groups):
    """Run a partition on the oracle: PerGate groups gate by gate on the
    current layout, BLAS groups as transpose + fused matrix."""
    g = to_oracle(spec)
    psi = oracle.zero_state(n)
    pos = list(range(n))
    for grp in groups:
        sub = g[grp["gates"]]
        if grp["backend"] == oracle.PERGATE:
            phys = sub.copy()
            for r in phys:
                r["target"] = pos[r["target"]]
                if r["control"] >= 0:
                    r["control"] = pos[r["control"]]
                if r["control2"] >= 0:
                    r["control2"] = pos[r["control2"]]
            psi = oracle.run_plain(phys, n, psi)
        else:
            perm = grp["perm"]
            psi = oracle.transpose(psi, n, perm)
            newpos = [0] * n
            for q in range(n):
                newpos[q] = perm.index(pos[q])
            pos = newpos
            m = oracle.fuse_group(sub, pos, grp["mat_qubit"])
            psi = oracle.dense_runs(psi, n, m, grp["mat_qubit"])
    return psi, pos


@pytest.mark.parametrize("mode", [oracle.MODE_GROUP, oracle.MODE_BLAS, oracle.MODE_MIX])
@pytest.mark.parametrize("n,seed", [(12, 0), (14, 1)])
def test_partition_invariants_and_amplitudes(ora, mode, n, seed):
    spec = random_mixed(n, 250, seed, ccx=True)
    p = oracle.partition_params(n, mode)
    groups, fpos, _ = oracle.partition(to_oracle(spec), n, p)
    _check_partition_invariants(spec, n, groups, p.local_size, p.blas_mat_limit if mode == oracle.MODE_BLAS else 7)
    psi, pos = _execute_partition(spec, n, groups)
    assert pos == fpos
    want = oracle.run_plain(to_oracle(spec), n)
    idx = np.arange(1 << n)
    phys = np.zeros_like(idx)
    for q in range(n):
        phys |= ((idx >> q) & 1) << fpos[q]
    assert rel_err(psi[phys], want) <= 1e-12


def test_partition_hand_traced_group_mode(ora):
    """12 qubits, H on every qubit then CNOT(0 -> 11), tile 10, coalesce 3:
    the greedy pass takes the chain of qubit 0 (with 11 and the CNOT), the
    coalesced qubits 1, 2, then single chains 3..8 until the tile is full;
    H(9), H(10) form the second group."""
    n = 12
    spec = [(H, q, -1, ()) for q in range(n)] + [(CNOT, 11, 0, ())]
    groups, fpos, _ = oracle.partition(to_oracle(spec), n, oracle.partition_params(n, oracle.MODE_GROUP))
    assert [g["gates"] for g in groups] == [[0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 12], [9, 10]]
    assert groups[0]["related"] == 0b100111111111
    assert groups[1]["related"] == 0b011000000111
    assert fpos == list(range(n))


def test_partition_hand_traced_blas_mode(ora):
    """8 qubits, H on each then CNOT(0 -> 1), matrices of at most 6 qubits:
    group 1 = H0..H5 + CNOT on qubits 0..5 (already low, identity
    permutation); group 2 = H6, H7, moved to positions 0, 1 by swapping with
    qubits 0, 1 (MIN_MAT_SIZE pads the matrix to 4 qubits)."""
    n = 8
    spec = [(H, q, -1, ()) for q in range(n)] + [(CNOT, 1, 0, ())]
    groups, fpos, flay = oracle.partition(to_oracle(spec), n, oracle.partition_params(n, oracle.MODE_BLAS))
    assert [g["gates"] for g in groups] == [[0, 1, 2, 3, 4, 5, 8], [6, 7]]
    assert [g["backend"] for g in groups] == [oracle.BLAS, oracle.BLAS]
    assert [g["related"] for g in groups] == [0x3F, 0xC0]
    assert [g["mat_qubit"] for g in groups] == [6, 4]
    assert groups[0]["perm"] == list(range(8))
    assert groups[1]["perm"] == [6, 7, 2, 3, 4, 5, 0, 1]
    assert fpos == [6, 7, 2, 3, 4, 5, 0, 1]
    assert flay == [6, 7, 2, 3, 4, 5, 0, 1]


# ---- MIX mode: the per-group backend choice (compiler.cpp:248-279, evaluator.cpp:168-229) ----
#
# Circuit (12 qubits): H on 8, 9, 10, 11, CNOT(8 -> 9), CNOT(10 -> 11), then H
# on 0..7 (gates 6..13).  Tile 10, coalesce 3, matrices of 4..7 qubits.
#
# Tables "flat": every gate costs 512 (512 / 1000 / 512 = 0.001 ms per gate),
# overhead 2^(12-27) ms per group, BLAS_perf[4] = blas4, BLAS_perf[5] = blas5,
# BLAS_perf[6] = BLAS_perf[7] = 1, cutt 0.
#
# Group 1, OShareMem candidate (related starts at {0, 1, 2}): chains q8 = q9 =
# {0, 1, 4}, q10 = q11 = {2, 3, 5}, q0..q7 one H each.  The merge takes q8
# (3 gates; q0..q2 are absorbed: their related sets lie inside {0,1,2,8,9}),
# then q10 (q11 absorbed), then q3, q4, q5 (10 qubits); q6, q7 do not fit:
# gates 0..11, eff = (12 * 0.001 + 2^-15) / 12 = 1.0025e-3.
# BLAS candidates (nothing preset): ms = 4 takes q8 + q10 chains, gates 0..5
# (eff blas4 / 6); ms = 5 adds H(0) (blas5 / 7); ms = 6, 7 add H(1), H(2)
# (1 / 8, 1 / 9).
PG, BL = oracle.PERGATE, oracle.BLAS
MIX_SPEC = ([(H, q, -1, ()) for q in (8, 9, 10, 11)] + [(CNOT, 9, 8, ()), (CNOT, 11, 10, ())] +
            [(H, q, -1, ()) for q in range(8)])
MIX_CASES = {
    # blas4 / 6 = 8.3e-4 wins group 1 (ms = 4, qubits 8..11 swap with 0..3:
    # perm = pos = [8,9,10,11,4,5,6,7,0,1,2,3]).  Group 2: tile preset to the
    # logical qubits at physical 0..2 = {8, 9, 10}; H(0)..H(6) fill it (eff
    # (7e-3 + 2^-15) / 7 = 1.0044e-3) against BLAS ms = 4 on H(0)..H(3):
    # 0.005 / 4 = 1.25e-3 -> OShareMem.  Group 3: H(7) alone (1.03e-3 < 0.005).
    "blas4": dict(blas={4: 0.005, 5: 1.0},
                  groups=[(BL, [0, 1, 2, 3, 4, 5], 0xF00, 4), (PG, [6, 7, 8, 9, 10, 11, 12], 0x77F, 0),
                          (PG, [13], 0x780, 0)],
                  perm=[8, 9, 10, 11, 4, 5, 6, 7, 0, 1, 2, 3]),
    # blas5 / 7 = 7.86e-4 beats blas4 / 6 = 8.3e-4: ms = 5, gates 0..6 on
    # qubits {0, 8, 9, 10, 11}; qubit 0 is already low, 1..4 swap with 8..11:
    # perm = [0,8,9,10,11,5,6,7,1,2,3,4].  Group 2: preset {0, 8, 9}; H(1)..H(7)
    # fill the tile exactly (1.0044e-3) against ms = 4: 1.25e-3, ms = 5:
    # 0.0055 / 5 = 1.1e-3 -> OShareMem.
    "blas5": dict(blas={4: 0.005, 5: 0.0055},
                  groups=[(BL, [0, 1, 2, 3, 4, 5, 6], 0xF01, 5), (PG, [7, 8, 9, 10, 11, 12, 13], 0x3FF, 0)],
                  perm=[0, 8, 9, 10, 11, 5, 6, 7, 1, 2, 3, 4]),
    # the reference's built-in V100 tables: group 1 costs (10 * 352 + 2 * 195)
    # / 512000 + 2^-15 = 7.67e-3 ms for 12 gates (6.4e-4 per gate) against
    # (BLAS_perf[4] = 0) + 11.37 ms for 6 gates; group 2 = H(6), H(7)
    "v100": dict(blas=None, groups=[(PG, list(range(12)), 0xF3F, 0), (PG, [12, 13], 0x0C7, 0)],
                 perm=list(range(12))),
}
MIX_FLAT_COST = 512.0


def mix_params(case):
    p = oracle.partition_params(12, oracle.MODE_MIX)
    blas = MIX_CASES[case]["blas"]
    if blas is not None:
        for t in range(22):
            p.single_cost[t] = MIX_FLAT_COST
            p.ctr_cost[t] = MIX_FLAT_COST
        for i in range(64):
            p.blas_cost[i] = 1.0
        for k, v in blas.items():
            p.blas_cost[k] = v
        p.cutt_cost = 0.0
    return p


@pytest.mark.parametrize("case", sorted(MIX_CASES))
def test_partition_hand_traced_mix_mode(ora, case):
    want = MIX_CASES[case]
    groups, fpos, flay = oracle.partition(to_oracle(MIX_SPEC), 12, mix_params(case))
    assert [(g["backend"], g["gates"], g["related"], g["mat_qubit"] if g["backend"] == oracle.BLAS else 0)
            for g in groups] == want["groups"]
    blas = [g for g in groups if g["backend"] == oracle.BLAS]
    if blas:
        assert blas[0]["perm"] == want["perm"]
    assert fpos == want["perm"] and flay == want["perm"]
    # the mixed schedule, executed on the oracle, is the circuit
    psi, pos = _execute_partition(MIX_SPEC, 12, groups)
    assert pos == fpos
    ref = oracle.run_plain(to_oracle(MIX_SPEC), 12)
    idx = np.arange(1 << 12)
    phys = np.zeros_like(idx)
    for q in range(12):
        phys |= ((idx >> q) & 1) << fpos[q]
    assert rel_err(psi[phys], ref) <= 1e-12


# ---- readout and parser ---------------------------------------------------------------

def test_print_state_text(ora):
    n = 8
    psi = oracle.run_plain(to_oracle([(H, 0, -1, ()), (X, 7, -1, ())]), n)
    text = oracle.print_state(psi, n, list(range(n)))
    lines = text.splitlines()
    assert len(lines) == 128 + 2
    assert lines[0] == "0 0.000000000000: 0.000000000000 0.000000000000"
    assert lines[127] == "127 0.000000000000: 0.000000000000 0.000000000000"
    assert lines[128] == "128 0.500000000000: 0.707106781187 0.000000000000"
    assert lines[129] == "129 0.500000000000: 0.707106781187 0.000000000000"
    # a layout that swaps qubits 0 and 7: logical indices are printed
    phys = oracle.run_plain(to_oracle([(H, 7, -1, ()), (X, 0, -1, ())]), n)
    pos = list(range(n))
    pos[0], pos[7] = 7, 0
    assert oracle.print_state(phys, n, pos) == text
    # values below 1e-14 print as zero
    assert oracle.lib().ora_zero_wrapper(-5e-15) == 0.0
    assert oracle.lib().ora_zero_wrapper(2e-14) == 2e-14


QASM = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[5];
// a comment line
h q[0];
cx q[0],q[1];
rz(pi*0.5) q[2];
ry(pi/4) q[3];
u3(0.1,0.2,0.3) q[1];
ccx q[0],q[1],q[4];
cu1(0.25) q[3],q[4];
tdg q[2];
"""


def test_parse_qasm_dialect(ora):
    n, g = oracle.parse_qasm(QASM)
    assert n == 5
    got = [(int(r["type"]), int(r["target"]), int(r["control"]), int(r["control2"])) for r in g]
    assert got == [(H, 0, -1, -1), (CNOT, 1, 0, -1), (RZ, 2, -1, -1), (RY, 3, -1, -1), (U3, 1, -1, -1),
                   (CCX, 4, 0, 1), (CU1, 4, 3, -1), (TDG, 2, -1, -1)]
    assert np.allclose(mat_of(g[2]), textbook(RZ, (math.pi * 0.5,)))
    assert np.allclose(mat_of(g[3]), textbook(RY, (math.pi / 4,)))
    assert np.allclose(mat_of(g[4]), textbook(U3, (0.1, 0.2, 0.3)))
    assert np.allclose(mat_of(g[6]), textbook(U1, (0.25,)))


@pytest.mark.parametrize("bad", ["qreg q[2];\nmeasure q[0];\n", "qreg q[2];\nrx q[0];\n", "h q[0];\n",
                                 "qreg q[2];\ncx q[0];\n", "qreg q[2];\nu3(0.1,0.2 q[0];\n"])
def test_parse_qasm_rejects(ora, bad):
    with pytest.raises(ValueError):
        oracle.parse_qasm(bad)


def test_inverse_helper_is_the_adjoint():
    # U then U^dagger is the identity on a random state: pins the adjoint rules
    # of helpers.inverse used by the full-size round-trip GPU tests
    from helpers import inverse, random_layers
    n = 8
    extra = [(S, 1, -1, ()), (SDG, 2, -1, ()), (TDG, 3, -1, ()), (X, 0, -1, ()), (Y, 4, -1, ()), (Z, 5, -1, ()),
             (CZ, 1, 0, ()), (CRX, 2, 1, (0.3,)), (CRY, 3, 2, (1.1,)), (CRZ, 4, 3, (-0.7,)), (CU1, 5, 4, (2.2,)),
             (CCX, 6, 0, (), 1)]
    spec = random_layers(n, 6, 3) + extra + random_layers(n, 2, 9)
    rng = np.random.default_rng(1)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    out = oracle.run_plain(to_oracle(spec + inverse(spec)), n, psi0)
    assert rel_err(out, psi0) <= 1e-13
    half = oracle.run_plain(to_oracle(spec), n, psi0)
    assert rel_err(half, psi0) > 0.1


@pytest.mark.parametrize("n,positions", [(8, [5, 1, 6, 3]), (7, [0, 6, 2, 4, 1]), (9, [8, 7, 6, 5, 0, 1])])
def test_dense_at_reference_matches_brute_force(n, positions):
    """The reference for an in-place TransMM group (U on qubits where they
    are: transpose, dense runs, transpose back) equals U scattered amplitude
    by amplitude: out[x with matrix bits r] += U[r, k] psi[x with bits k]."""
    from helpers import dense_at_reference, rand_state
    m = len(positions)
    K = 1 << m
    rs = np.random.default_rng(n)
    U = rs.normal(size=(K, K)) + 1j * rs.normal(size=(K, K))
    psi = rand_state(n, n + 1)
    got = dense_at_reference(psi, n, positions, U)
    want = np.zeros_like(psi)
    mask = sum(1 << p for p in positions)
    for x in range(1 << n):
        k = sum(((x >> p) & 1) << j for j, p in enumerate(positions))
        base = x & ~mask
        for r in range(K):
            y = base | sum(((r >> j) & 1) << p for j, p in enumerate(positions))
            want[y] += U[r, k] * psi[x]
    assert np.max(np.abs(got - want)) <= 1e-12
