"""CPU checks of the batching / serving layer: the readout definitions the
device kernels implement, and the engine's error path without a GPU."""
import numpy as np
import pytest


def readouts(psi, n):
    """norm, sum_q <Z_q>, sum_q <Z_q Z_{q+1}> of psi by per-qubit marginals."""
    p = np.abs(psi) ** 2
    idx = np.arange(1 << n, dtype=np.uint64)
    bits = [((idx >> np.uint64(q)) & np.uint64(1)).astype(np.int64) for q in range(n)]
    z = sum(float(np.dot(p, 1 - 2 * b)) for b in bits)
    zz = sum(float(np.dot(p, (1 - 2 * bits[q]) * (1 - 2 * bits[q + 1]))) for q in range(n - 1))
    return float(p.sum()), z, zz


def test_readout_popcount_form_matches_marginals():
    # the device evaluates Z = n - 2 popcount(i), ZZ = (n-1) - 2 popcount((i ^ i>>1) & (2^(n-1)-1))
    rng = np.random.default_rng(3)
    for n in (2, 5, 9):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi)
        p = np.abs(psi) ** 2
        idx = np.arange(1 << n)
        pc = np.vectorize(lambda x: bin(int(x)).count("1"))
        z = float(np.dot(p, n - 2 * pc(idx)))
        zz = float(np.dot(p, (n - 1) - 2 * pc((idx ^ (idx >> 1)) & ((1 << (n - 1)) - 1))))
        _, z2, zz2 = readouts(psi, n)
        assert abs(z - z2) < 1e-12 and abs(zz - zz2) < 1e-12


def test_readouts_known_states():
    # |0..0>: every Z = +1; uniform superposition: every <Z> = <ZZ> = 0; GHZ: ZZ = n-1, Z = 0
    n = 6
    e0 = np.zeros(1 << n, complex); e0[0] = 1
    assert readouts(e0, n) == (1.0, n, n - 1)
    u = np.full(1 << n, 2 ** (-n / 2), complex)
    _, z, zz = readouts(u, n)
    assert abs(z) < 1e-12 and abs(zz) < 1e-12
    g = np.zeros(1 << n, complex); g[0] = g[-1] = 2 ** -0.5
    _, z, zz = readouts(g, n)
    assert abs(z) < 1e-12 and abs(zz - (n - 1)) < 1e-12


def test_engine_without_device(hq):
    if hq.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(hq.HQError) as ei:
        hq.Engine(hq.HQ_C128, 1 << 30)
    assert ei.value.status == hq.HQ_ERR_NO_DEVICE
