"""Test-side helpers: an independent restatement of the synthetic workload
generators (to check the library's) and circuit builders for the oracle."""
from __future__ import annotations

import math

import numpy as np

import oracle

MASK64 = (1 << 64) - 1

# gate type codes (src/gate.h:7-9)
CCX, CNOT, CY, CZ, CRX, CRY, CU1, CRZ, U1, U2, U3, H, X, Y, Z, S, SDG, T, TDG, RX, RY, RZ = range(22)
ID = 23


class Rng:
    """xoshiro256** seeded by SplitMix64 (same definition as hq::Rng)."""

    def __init__(self, seed: int):
        x = seed & MASK64
        self.s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & MASK64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & MASK64

    def next(self):
        s = self.s
        result = (self._rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def below(self, m):
        return ((self.next() >> 11) * m) >> 53


def random_layers(n, layers, seed):
    """(type, target, control, params) tuples of the random-layer workload."""
    rng = Rng(seed)
    two_pi = 2 * math.acos(-1.0)
    out = []
    for _ in range(layers):
        for q in range(n):
            k = rng.below(5)
            if k == 0:
                out.append((H, q, -1, ()))
            elif k == 1:
                out.append((RX, q, -1, (two_pi * rng.uniform(),)))
            elif k == 2:
                out.append((RY, q, -1, (two_pi * rng.uniform(),)))
            elif k == 3:
                out.append((RZ, q, -1, (two_pi * rng.uniform(),)))
            else:
                out.append((T, q, -1, ()))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for i in range(0, n - 1, 2):
            out.append((CNOT, perm[i + 1], perm[i], ()))
    return out


def hea_layers(n, layers, seed):
    rng = Rng(seed)
    two_pi = 2 * math.acos(-1.0)
    out = []
    for _ in range(layers):
        for q in range(n):
            out.append((RY, q, -1, (two_pi * rng.uniform(),)))
            out.append((RZ, q, -1, (two_pi * rng.uniform(),)))
        for q in range(n - 1):
            out.append((CZ, q + 1, q, ()))
    return out


_DAGGER = {S: SDG, SDG: S, T: TDG, TDG: T}
_NEGATE = {RX, RY, RZ, U1, CRX, CRY, CRZ, CU1}


def inverse(spec):
    """The inverse circuit: gates in reverse order, each replaced by its
    adjoint (self-inverse gates kept, S/T swapped with their daggers,
    rotation angles negated)."""
    out = []
    for s in reversed(spec):
        t = s[0]
        if t in _DAGGER:
            out.append((_DAGGER[t],) + tuple(s[1:]))
        elif t in _NEGATE:
            out.append((t, s[1], s[2], tuple(-x for x in s[3])) + tuple(s[4:]))
        elif t in (H, X, Y, Z, CNOT, CY, CZ, CCX, ID):
            out.append(s)
        else:
            raise ValueError(f"no inverse rule for gate type {t}")
    return out


def to_oracle(spec):
    """Tuples -> oracle gate array."""
    arr = np.zeros(len(spec), dtype=oracle.GATE_DTYPE)
    for i, s in enumerate(spec):
        t, tgt, c = s[0], s[1], s[2]
        params = s[3] if len(s) > 3 else ()
        c2 = s[4] if len(s) > 4 else -1
        arr[i] = oracle.gate(t, tgt, c, c2, params)
    return arr


ONE_Q = [(H, 0), (X, 0), (Y, 0), (Z, 0), (S, 0), (SDG, 0), (T, 0), (TDG, 0), (RX, 1), (RY, 1), (RZ, 1), (U1, 1),
         (U2, 2), (U3, 3)]
TWO_Q = [(CNOT, 0), (CY, 0), (CZ, 0), (CRX, 1), (CRY, 1), (CU1, 1), (CRZ, 1)]


def random_mixed(n, m, seed, ccx=True):
    """Random circuit over every gate type (tuples)."""
    rs = np.random.default_rng(seed)
    out = []
    for _ in range(m):
        r = rs.random()
        if r < 0.55 or n < 2:
            t, npar = ONE_Q[rs.integers(len(ONE_Q))]
            q = int(rs.integers(n))
            out.append((t, q, -1, tuple(rs.uniform(0, 2 * np.pi, npar))))
        elif r < 0.93 or not ccx or n < 3:
            t, npar = TWO_Q[rs.integers(len(TWO_Q))]
            a, b = rs.choice(n, 2, replace=False)
            out.append((t, int(b), int(a), tuple(rs.uniform(0, 2 * np.pi, npar))))
        else:
            a, b, c = rs.choice(n, 3, replace=False)
            out.append((CCX, int(c), int(a), (), int(b)))
    return out


def rel_err(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rand_state(n, seed, dtype=np.complex128):
    rs = np.random.default_rng(seed)
    v = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(dtype)


def group_spec_for_mask(n, mask, m, seed):
    """Random gates (all types, as tuples) whose non-diagonal targets lie in the tile `mask`."""
    rs = np.random.default_rng(seed)
    tile = [q for q in range(n) if (mask >> q) & 1]
    spec = []
    for s in random_mixed(n, m, seed):
        t, tgt, c = s[0], s[1], s[2]
        if not oracle.lib().ora_is_diagonal(t) and tgt not in tile:
            new_t = int(rs.choice([q for q in tile if q != c and (len(s) < 5 or q != s[4])]))
            s = (t, new_t) + tuple(s[2:])
        spec.append(s)
    return spec


def group_for_mask(n, mask, m, seed):
    """Random gates (all types) whose non-diagonal targets lie in the tile `mask`."""
    return to_oracle(group_spec_for_mask(n, mask, m, seed))


def reangle(spec, seed):
    """The same gates with every angle drawn again (same structure)."""
    rs = np.random.default_rng(seed)
    out = []
    for s in spec:
        params = s[3] if len(s) > 3 else ()
        s = (s[0], s[1], s[2], tuple(float(x) for x in rs.uniform(-math.pi, math.pi, len(params)))) + tuple(s[4:])
        out.append(s)
    return out


_QASM_NAMES = {CCX: "ccx", CNOT: "cx", CY: "cy", CZ: "cz", CRX: "crx", CRY: "cry", CRZ: "crz", CU1: "cu1",
               U1: "u1", U3: "u3", H: "h", X: "x", Y: "y", Z: "z", S: "s", SDG: "sdg", T: "t", TDG: "tdg",
               RX: "rx", RY: "ry", RZ: "rz"}


def to_qasm(n, spec):
    """The circuit as text in the reference's input dialect (src/main.cpp:68-231):
    angles as plain decimals written to round-trip (%.17g), qubits ordered
    controls first."""
    lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', f"qreg q[{n}];"]   # no creg: the dialect rejects it
    for s in spec:
        t, tgt, ctl = s[0], s[1], s[2]
        params = s[3] if len(s) > 3 else ()
        c2 = s[4] if len(s) > 4 else -1
        name = _QASM_NAMES[t]
        if params:
            name += "(" + ",".join(f"{p:.17g}" for p in params) + ")"
        qs = ([ctl] if ctl >= 0 else []) + ([c2] if c2 >= 0 else []) + [tgt]
        if t == CCX:
            qs = [ctl, c2, tgt]
        lines.append(name + " " + ",".join(f"q[{q}]" for q in qs) + ";")
    return "\n".join(lines) + "\n"


def qft_spec(n, x):
    """X gates preparing |x>, then the textbook QFT without the final swaps."""
    spec = [(X, q, -1, ()) for q in range(n) if (x >> q) & 1]
    for j in range(n - 1, -1, -1):
        spec.append((H, j, -1, ()))
        for k in range(j - 1, -1, -1):
            spec.append((CU1, j, k, (math.pi / 2 ** (j - k),)))
    return spec


def bv_spec(n, secret):
    """Bernstein-Vazirani on n-1 data qubits and ancilla n-1."""
    anc = n - 1
    spec = [(X, anc, -1, ()), (H, anc, -1, ())] + [(H, q, -1, ()) for q in range(n - 1)]
    spec += [(CNOT, anc, q, ()) for q in range(n - 1) if (secret >> q) & 1]
    spec += [(H, q, -1, ()) for q in range(n - 1)]
    return spec


def ghz_spec(n):
    return [(H, 0, -1, ())] + [(CNOT, q + 1, q, ()) for q in range(n - 1)]


def dense_block(qubits, m, seed):
    """m gates on a few qubits: 30 % CNOT between two of them, else H, RX,
    RY, RZ or T on one (numpy PCG64 draws)."""
    rs = np.random.default_rng(seed)
    out = []
    for _ in range(m):
        if rs.random() < 0.3:
            a, b = rs.choice(qubits, 2, replace=False)
            out.append((CNOT, int(b), int(a), ()))
        else:
            t = int(rs.choice([H, RX, RY, RZ, T]))
            q = int(rs.choice(qubits))
            out.append((t, q, -1, (float(rs.uniform(0, 2 * math.pi)),) if t in (RX, RY, RZ) else ()))
    return out


def blocks_and_layers(n, m=100):
    """Random layers over all n qubits alternating with deep blocks on six
    qubits: a circuit whose gates crowd onto few qubits in places, where a
    dense TransMM group pays off beside OShareMem groups."""
    assert n >= 20
    return (random_layers(n, 1, 1) + dense_block([2, 6, 9, 12, 15, 17], m, 2) + random_layers(n, 2, 3) +
            dense_block([0, 4, 7, 11, 16, 19], m, 4) + random_layers(n, 2, 5) +
            dense_block([1, 3, 8, 13, 14, 18], m, 6) + random_layers(n, 1, 7))


def flat_cost_tables(blas_ms, gate_cost=512.0, cutt=0.0):
    """A parameter file in the reference's <n>qubits.out format
    (src/evaluator.cpp:60-103, param_type 1): every gate costs `gate_cost`,
    BLAS_perf[log2 K] = blas_ms for every K, transpose cost `cutt`."""
    rows = ["1"] + [str(gate_cost)] * 14 + [""] + [str(gate_cost)] * 7 + [""]
    rows += [f"{1 << i} {blas_ms}" for i in range(10)] + ["", str(cutt)]
    return "\n".join(rows) + "\n"


def dense_at_reference(psi, n, positions, U):
    """U (K x K, matrix index bit j = qubit positions[j]) applied to psi,
    composed from the oracle: transpose those qubits to the lowest bits in
    that order, apply U to every run of K amplitudes, transpose back."""
    m = len(positions)
    perm = list(positions) + [q for q in range(n) if q not in positions]   # new bit i = old bit perm[i]
    inv = [0] * n
    for i, q in enumerate(perm):
        inv[q] = i
    x = oracle.transpose(psi, n, perm)
    x = oracle.dense_runs(x, n, np.asarray(U).T.reshape(-1), m)
    return oracle.transpose(x, n, inv)
