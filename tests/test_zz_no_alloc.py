"""hq_expect_z_noalloc (the allocation-free variant of hq_expect_z, partials
in the library's fixed device array) against hq_expect_z, through the C ABI:
bit-identical <Z_q> and <Z_q Z_q+1> on two calls in a row, the second on a
smaller state, in both precisions; complex128 also against the FP64 oracle."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import random_layers, to_oracle

pytestmark = pytest.mark.gpu


def _call(L, fn, st, n):
    z = (ctypes.c_double * n)()
    zz = (ctypes.c_double * max(1, n - 1))()
    rc = getattr(L, fn)(st.handle, z, zz)
    assert rc == 0, (fn, L.hq_last_error())
    return np.array(z[:]), np.array(zz[: n - 1])


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_equals_expect_z(gpu, prec):
    L = gpu.lib()
    p = gpu.HQ_C128 if prec == 0 else gpu.HQ_C64
    for n, seed in ((22, 1), (14, 2)):
        print(f"noalloc: n={n} prec={p}", flush=True)
        c = gpu.Circuit(n).add_random_layers(6, seed)
        c.set_option("precision", p)
        c.set_option("mode", gpu.HQ_MODE_GROUP)
        c.run(copy_back=False, destroy=False)
        st = c.state_view(p)
        z0, zz0 = _call(L, "hq_expect_z_noalloc", st, n)
        z1, zz1 = _call(L, "hq_expect_z", st, n)
        assert np.array_equal(z0, z1) and np.array_equal(zz0, zz1)
        # and again: the array keeps the last call's contents, which must not leak
        z2, zz2 = _call(L, "hq_expect_z_noalloc", st, n)
        assert np.array_equal(z2, z1) and np.array_equal(zz2, zz1)
        if p == gpu.HQ_C128:
            psi = oracle.run_plain(to_oracle(random_layers(n, 6, seed)), n)
            pr = np.abs(psi) ** 2
            idx = np.arange(1 << n)
            pos = c.final_pos()
            bit = [(idx >> q) & 1 for q in range(n)]
            zw = np.array([np.sum(pr * (1 - 2 * bit[q])) for q in range(n)])
            assert np.max(np.abs(np.array([z0[pos[q]] for q in range(n)]) - zw)) <= 1e-12
        c.close()
