"""GPU parity of the device layer (C ABI) against the CPU oracle.

Tolerances: complex128 paths <= 1e-12 relative (2-norm) against the FP64
oracle, transposes bit-exact, complex64 paths <= 2e-6 relative.
"""
import numpy as np
import pytest

import oracle
from helpers import dense_at_reference, H, group_for_mask as _group_for_mask, rand_state as _rand_state, rel_err, random_mixed, \
    to_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [3, 8, 13])
def test_plain_all_gate_types(gpu, n):
    spec = random_mixed(n, 200, seed=n)
    g = to_oracle(spec)
    psi0 = _rand_state(n, 1)
    want = oracle.run_plain(g, n, psi0)
    s = gpu.State(n)
    s.upload(psi0)
    s.apply_plain(g)
    got = s.download()
    assert rel_err(got, want) <= 1e-12


def test_plain_c64(gpu):
    n = 12
    g = to_oracle(random_mixed(n, 150, seed=7))
    psi0 = _rand_state(n, 2)
    want = oracle.run_plain(g, n, psi0)
    s = gpu.State(n, gpu.HQ_C64)
    s.upload(psi0.astype(np.complex64))
    s.apply_plain(g)
    assert rel_err(s.download(), want) <= 2e-6


@pytest.mark.parametrize("n,k,seed", [(7, 7, 0), (10, 9, 1), (12, 10, 2), (14, 12, 3), (15, 12, 4), (16, 13, 5)])
def test_sharemem_matches_oracle(gpu, n, k, seed):
    rs = np.random.default_rng(seed)
    low = list(range(min(3, k)))
    rest = rs.choice(np.arange(3, n), k - len(low), replace=False)
    mask = 0
    for q in low + [int(x) for x in rest]:
        mask |= 1 << q
    g = _group_for_mask(n, mask, 120, seed)
    psi0 = _rand_state(n, seed + 10)
    want = oracle.run_plain(g, n, psi0)
    # the oracle's restatement of the OShareMem semantics agrees too
    assert rel_err(oracle.sharemem_group(psi0, n, g, mask), want) <= 1e-13
    s = gpu.State(n)
    s.upload(psi0)
    p = s.program_sharemem(g, mask)
    tile, stages, ops = p.info()
    # one op per gate, plus the final scale of the factored scalars
    assert tile == k and stages >= 1 and ops in (len(g), len(g) + 1)
    s.launch(p)
    assert rel_err(s.download(), want) <= 1e-12


@pytest.mark.parametrize("n,k,seed,prec", [(12, 10, 21, 0), (14, 12, 22, 0), (16, 13, 23, 0), (15, 12, 24, 1),
                                            (16, 13, 25, 1)])
def test_sharemem_specialized_matches_oracle(gpu, n, k, seed, prec):
    """The hiprtc-specialised kernel of a group gives the interpreter's result."""
    rs = np.random.default_rng(seed)
    low = list(range(3))
    rest = rs.choice(np.arange(3, n), k - 3, replace=False)
    mask = 0
    for q in low + [int(x) for x in rest]:
        mask |= 1 << q
    g = _group_for_mask(n, mask, 150, seed)
    psi0 = _rand_state(n, seed)
    want = oracle.run_plain(g, n, psi0)
    s = gpu.State(n, prec)
    s.upload(psi0.astype(s.dtype))
    p = s.program_sharemem(g, mask).specialize()
    assert p.is_specialized()
    s.launch(p)
    assert rel_err(s.download(), want) <= (1e-12 if prec == 0 else 2e-6)


def test_sharemem_c64(gpu):
    n, mask = 15, (1 << 13) - 1
    g = _group_for_mask(n, mask, 100, 9)
    psi0 = _rand_state(n, 3)
    want = oracle.run_plain(g, n, psi0)
    s = gpu.State(n, gpu.HQ_C64)
    s.upload(psi0.astype(np.complex64))
    s.apply_sharemem(g, mask)
    assert rel_err(s.download(), want) <= 2e-6


def test_sharemem_rejects_non_diagonal_outside_tile(gpu):
    n = 12
    s = gpu.State(n)
    g = to_oracle([(H, 11, -1, ())])
    with pytest.raises(gpu.HQError):
        s.program_sharemem(g, (1 << 10) - 1)


@pytest.mark.parametrize("n", [6, 10, 13])
def test_transpose_bit_exact(gpu, n):
    rs = np.random.default_rng(n)
    for trial in range(3):
        perm = [int(x) for x in rs.permutation(n)]
        psi0 = _rand_state(n, trial)
        want = oracle.transpose(psi0, n, perm)
        s = gpu.State(n, with_buffer=True)
        s.upload(psi0)
        s.apply_transpose(perm)
        assert np.array_equal(s.download(), want)


def _rand_unitary(K, seed):
    rs = np.random.default_rng(seed)
    a = rs.normal(size=(K, K)) + 1j * rs.normal(size=(K, K))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("mq", [1, 3, 4, 5, 6, 7])
def test_dense_runs(gpu, mq):
    n = 13
    K = 1 << mq
    U = _rand_unitary(K, mq)
    colmajor = U.T.reshape(-1)          # colmajor[col*K + row] = U[row, col]
    psi0 = _rand_state(n, mq)
    want = oracle.dense_runs(psi0, n, colmajor, mq)
    assert rel_err(want, (U @ psi0.reshape(-1, K).T).T.reshape(-1)) <= 1e-13
    s = gpu.State(n, with_buffer=True)
    s.upload(psi0)
    s.apply_dense(colmajor, mq)
    assert rel_err(s.download(), want) <= 1e-12


@pytest.mark.parametrize("prec,mq", [("c128", 4), ("c128", 6), ("c128", 7), ("c64", 6), ("c64", 7)])
def test_dense_runs_many_tiles_per_block(gpu, prec, mq):
    """24 qubits: every block of the GEMM runs several column tiles, so the
    pipelined path (next tile's B loaded during the current tile's MFMAs,
    two LDS buffers) is exercised past its first tile."""
    n = 24
    K = 1 << mq
    U = _rand_unitary(K, 200 + mq)
    colmajor = U.T.reshape(-1)
    psi0 = _rand_state(n, 300 + mq)
    want = oracle.dense_runs(psi0, n, colmajor, mq)
    c64 = prec == "c64"
    s = gpu.State(n, gpu.HQ_C64 if c64 else gpu.HQ_C128, with_buffer=True)
    s.upload(psi0.astype(np.complex64) if c64 else psi0)
    s.apply_dense(colmajor, mq)
    assert rel_err(s.download(), want) <= (2e-6 if c64 else 1e-12)
    s.close()


@pytest.mark.parametrize("prec,n,mq,seed", [
    ("c128", 13, 4, 1), ("c128", 13, 5, 2), ("c128", 14, 6, 3), ("c128", 13, 7, 4),
    ("c128", 24, 6, 5), ("c128", 24, 7, 6), ("c128", 22, 4, 7),
    ("c64", 14, 4, 8), ("c64", 14, 6, 9), ("c64", 13, 7, 10), ("c64", 24, 6, 11), ("c64", 24, 7, 12)])
def test_dense_in_place_matches_oracle(gpu, prec, n, mq, seed):
    """hq_program_create_dense: a TransMM group's matrix on qubits at
    arbitrary (unsorted) positions, in place and without a work buffer,
    against the oracle's transpose + dense + inverse transpose; 24 qubits
    give several tiles per block (the pipelined path), 13-14 the one-tile
    minimum; positions include and exclude the lowest qubits."""
    rs = np.random.default_rng(seed)
    positions = [int(x) for x in rs.choice(n, mq, replace=False)]
    if seed % 3 == 0 and 0 not in positions:
        positions[0] = 0
    K = 1 << mq
    U = _rand_unitary(K, 400 + seed)
    psi0 = _rand_state(n, 500 + seed)
    want = dense_at_reference(psi0, n, positions, U)
    c64 = prec == "c64"
    s = gpu.State(n, gpu.HQ_C64 if c64 else gpu.HQ_C128)
    s.upload(psi0.astype(np.complex64) if c64 else psi0)
    p = s.program_dense(positions, U.T.reshape(-1))
    s.launch(p)
    s.sync()
    got = s.download()
    print(prec, n, mq, positions, rel_err(got, want), flush=True)
    assert rel_err(got, want) <= (2e-6 if c64 else 1e-12)
    # a second launch applies U again
    s.launch(p)
    s.sync()
    want2 = dense_at_reference(want, n, positions, U)
    assert rel_err(s.download(), want2) <= (4e-6 if c64 else 1e-12)
    p.close()
    s.close()


def test_dense_in_place_rejects_bad_arguments(gpu):
    s = gpu.State(13)
    U = np.eye(16)
    with pytest.raises(RuntimeError):
        s.program_dense([0, 1, 2, 2], U)         # repeated qubit
    with pytest.raises(RuntimeError):
        s.program_dense([0, 1, 2, 13], U)        # outside the state
    with pytest.raises(RuntimeError):
        s.program_dense([0, 1, 2], np.eye(8))    # below 4 matrix qubits
    small = gpu.State(8)
    with pytest.raises(RuntimeError):
        small.program_dense([0, 1, 2, 3], U)     # smaller than one tile
    small.close()
    s.close()


@pytest.mark.parametrize("mq", [4, 6, 7])
def test_dense_runs_c64(gpu, mq):
    n = 14
    K = 1 << mq
    U = _rand_unitary(K, 100 + mq)
    colmajor = U.T.reshape(-1)
    psi0 = _rand_state(n, mq)
    want = oracle.dense_runs(psi0, n, colmajor, mq)
    s = gpu.State(n, gpu.HQ_C64, with_buffer=True)
    s.upload(psi0.astype(np.complex64))
    s.apply_dense(colmajor, mq)
    assert rel_err(s.download(), want) <= 2e-6


def test_readouts(gpu):
    n = 14
    psi0 = _rand_state(n, 5)
    s = gpu.State(n)
    s.upload(psi0)
    p = np.abs(psi0) ** 2
    assert abs(s.norm2() - p.sum()) <= 1e-13
    idx = np.arange(1 << n)
    for t in (0, 5, 13):
        assert abs(s.measure_zero(t) - p[((idx >> t) & 1) == 0].sum()) <= 1e-13
    z, zz = s.expect_z()
    for q in range(n):
        b = (idx >> q) & 1
        assert abs(z[q] - (p * (1 - 2 * b)).sum()) <= 1e-12
        if q + 1 < n:
            b2 = (idx >> (q + 1)) & 1
            assert abs(zz[q] - (p * (1 - 2 * (b ^ b2))).sum()) <= 1e-12
    s2 = gpu.State(n)
    psi1 = _rand_state(n, 6)
    s2.upload(psi1)
    assert abs(s.inner(s2) - np.vdot(psi0, psi1)) <= 1e-13
    assert s.get_amp(12345) == psi0[12345]
    s.reset()
    d = s.download()
    assert d[0] == 1 and np.count_nonzero(d) == 1
