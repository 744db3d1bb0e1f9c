"""GPU parity of the batching / serving engine (hq_engine_*, hq_batch_*) and
the top-k readout against the FP64 oracle, plus the properties the forms
promise: a result is bit-identical alone and in a mix, graph replays repeat
bit-for-bit, admission never exceeds the HBM budget, complex64 stays within
1e-5 of complex128."""
import numpy as np
import pytest

import oracle
from helpers import hea_layers, random_layers, to_oracle
from test_engine_cpu import readouts as _readouts

pytestmark = pytest.mark.gpu


def _topk_ref(psi, k=16):
    p = np.abs(psi) ** 2
    order = np.lexsort((np.arange(p.size), -p))[:k]
    return [int(i) for i in order], [float(p[i]) for i in order]


def _hea(hq, n, depth, seed):
    return hq.Circuit(n).add_hea_layers(depth, seed)


def _rl(hq, n, depth, seed):
    return hq.Circuit(n).add_random_layers(depth, seed)


@pytest.mark.parametrize("n,depth", [(10, 6), (13, 4), (14, 5), (17, 3)])
def test_engine_hea_matches_oracle(gpu, n, depth):
    hq = gpu
    e = hq.Engine(hq.HQ_C128, 1 << 30)
    c = _hea(hq, n, depth, seed=n)
    j = e.submit(c, hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
    e.drain()
    r = e.result(j, topk=True)
    psi = oracle.run_plain(to_oracle(hea_layers(n, depth, n)), n)
    norm, z, zz = _readouts(psi, n)
    assert abs(r["norm"] - 1) < 1e-12
    assert abs(r["z"] - z) < 1e-10 * n and abs(r["zz"] - zz) < 1e-10 * n
    idx, prob = _topk_ref(psi)
    assert r["topk"][0] == idx
    np.testing.assert_allclose(r["topk"][1], prob, rtol=1e-10, atol=1e-15)
    e.close()


@pytest.mark.parametrize("n,prec", [(12, "c128"), (16, "c128"), (20, "c128"), (24, "c128"),
                                    (16, "c64"), (22, "c64"), (24, "c64")])
def test_engine_random_layers_matches_oracle(gpu, n, prec):
    """A random-layer request (the serving stream's other circuit kind) in
    either precision: readouts and top-k against the FP64 oracle; complex64
    within its rounding (1e-5 per qubit on the sums, 1e-6 on probabilities)."""
    hq = gpu
    c64 = prec == "c64"
    e = hq.Engine(hq.HQ_C64 if c64 else hq.HQ_C128, 1 << 31)
    depth = 6
    j = e.submit(_rl(hq, n, depth, seed=5), hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
    e.drain()
    r = e.result(j, topk=True)
    psi = oracle.run_plain(to_oracle(random_layers(n, depth, 5)), n)
    nrm, z, zz = _readouts(psi, n)
    tol = 1e-5 if c64 else 1e-10
    assert abs(r["norm"] - nrm) < tol
    assert abs(r["z"] - z) < tol * n and abs(r["zz"] - zz) < tol * n
    idx, prob = _topk_ref(psi)
    if c64:
        # near-equal probabilities may swap places under complex64 rounding
        got = np.array(r["topk"][0])
        np.testing.assert_allclose(r["topk"][1], prob, rtol=0, atol=1e-6)
        np.testing.assert_allclose(np.abs(psi[got]) ** 2, prob, rtol=0, atol=1e-6)
    else:
        assert r["topk"][0] == idx
        np.testing.assert_allclose(r["topk"][1], prob, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_alone_equals_in_mix(gpu, prec):
    hq = gpu
    p = hq.HQ_C128 if prec == "c128" else hq.HQ_C64
    probe = [(11, 7), (15, 5), (19, 4)]
    alone = {}
    for n, d in probe:
        e = hq.Engine(p, 1 << 30)
        j = e.submit(_hea(hq, n, d, 100 + n), hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
        e.drain()
        alone[n] = e.result(j, topk=True)
        e.close()
    # the same circuits interleaved with others, submitted while rounds run
    e = hq.Engine(p, 1 << 30)
    ids = {}
    others = [(10, 9), (21, 3), (13, 12), (16, 6), (12, 3)]
    for k, (n, d) in enumerate(others[:2]):
        e.submit(_rl(hq, n, d, k), hq.HQ_OUT_EXPECT)
    e.step()
    for n, d in probe:
        ids[n] = e.submit(_hea(hq, n, d, 100 + n), hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
        e.step()
        for k, (m, dd) in enumerate(others[2:]):
            e.submit(_hea(hq, m, dd, 7 * k + n), hq.HQ_OUT_EXPECT)
    e.drain()
    for n, _ in probe:
        r = e.result(ids[n], topk=True)
        a = alone[n]
        assert (r["norm"], r["z"], r["zz"]) == (a["norm"], a["z"], a["zz"])
        assert r["topk"] == a["topk"]


def test_batch_graph_matches_engine_and_replays(gpu):
    hq = gpu
    rng = np.random.default_rng(11)
    specs = [(int(rng.integers(10, 19)), int(rng.integers(2, 9))) for _ in range(24)]
    circuits = [_hea(hq, n, d, 1000 + i) for i, (n, d) in enumerate(specs)]
    e = hq.Engine(hq.HQ_C128, 1 << 31)
    b = e.batch(circuits, hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
    b.run()
    r1, t1 = b.results()
    e2 = hq.Engine(hq.HQ_C128, 1 << 31)
    jobs = [e2.submit(c, hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK) for c in circuits]
    e2.drain()
    for i, j in enumerate(jobs):
        r = e2.result(j, topk=True)
        assert (r["norm"], r["z"], r["zz"]) == tuple(r1[i])
        assert r["topk"][0] == [int(x) for x in t1[0][i]]
    for _ in range(3):
        b.run()
        r2, t2 = b.results()
        assert np.array_equal(r1, r2) and np.array_equal(t1[0], t2[0]) and np.array_equal(t1[1], t2[1])
    # spot-check two circuits against the oracle
    for i in (0, len(specs) - 1):
        n, d = specs[i]
        psi = oracle.run_plain(to_oracle(hea_layers(n, d, 1000 + i)), n)
        _, z, zz = _readouts(psi, n)
        assert abs(r1[i][1] - z) < 1e-10 * n and abs(r1[i][2] - zz) < 1e-10 * n
    b.close()


def test_batch_graph_c64_large_circuits(gpu):
    """The batching form as the bench runs it: complex64, one hipGraph, 19- to
    22-qubit ansatz circuits next to small ones; two circuits against the FP64
    oracle (complex64 rounding: 1e-5 per qubit), replays bit for bit."""
    hq = gpu
    specs = [(22, 5), (19, 8), (21, 4), (12, 6), (20, 7), (22, 3), (10, 9), (19, 4)]
    circuits = [_hea(hq, n, d, 2000 + i) for i, (n, d) in enumerate(specs)]
    e = hq.Engine(hq.HQ_C64, 1 << 31)
    b = e.batch(circuits, hq.HQ_OUT_EXPECT)
    b.run()
    r1, _ = b.results()
    for _ in range(2):
        b.run()
        r2, _ = b.results()
        assert np.array_equal(r1, r2)
    for i in (0, 1):
        n, d = specs[i]
        psi = oracle.run_plain(to_oracle(hea_layers(n, d, 2000 + i)), n)
        nrm, z, zz = _readouts(psi, n)
        assert abs(r1[i][0] - nrm) < 1e-5
        assert abs(r1[i][1] - z) < 1e-5 * n and abs(r1[i][2] - zz) < 1e-5 * n
    assert np.all(np.abs(r1[:, 0] - 1) < 1e-5)
    b.close()
    e.close()


@pytest.mark.parametrize("jit", [0, 1])
def test_side_streams_give_identical_results(gpu, jit):
    """Option "streams": a round's group launches spread over side streams
    (fork and join by events, captured into the batch graph) give the same
    results bit for bit as one stream, in a batch and in continuous batching."""
    hq = gpu
    specs = [(16, 6, 1), (17, 5, 2), (19, 4, 3), (14, 8, 4), (16, 6, 5), (20, 3, 6), (12, 5, 7), (18, 4, 8)]
    out = {}
    for streams in (1, 4):
        e = hq.Engine(hq.HQ_C64, 1 << 31)
        e.set_option("streams", streams)
        e.set_option("jit", jit)
        b = e.batch([_hea(hq, n, d, s) for n, d, s in specs], hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
        b.run()
        r, t = b.results()
        b.run()
        r2, t2 = b.results()
        assert np.array_equal(r, r2) and np.array_equal(t[1], t2[1])
        b.close()
        jobs = [e.submit(_rl(hq, n, d, s), hq.HQ_OUT_EXPECT) for n, d, s in specs]
        e.drain()
        cont = [e.result(j)["z"] for j in jobs]
        e.close()
        out[streams] = (r, t, cont)
    assert np.array_equal(out[1][0], out[4][0])
    assert np.array_equal(out[1][1][0], out[4][1][0]) and np.array_equal(out[1][1][1], out[4][1][1])
    assert out[1][2] == out[4][2]


def test_batch_pass_bytes(gpu):
    """hq_batch_pass_bytes: a whole-state circuit is one write-only pass; a
    partitioned one writes once, then reads and writes at most the whole
    state per later pass (tiles still zero are skipped)."""
    hq = gpu
    e = hq.Engine(hq.HQ_C64, 1 << 30)
    b = e.batch([_hea(hq, 10, 3, 1)], hq.HQ_OUT_EXPECT)
    assert b.info()["pass_bytes"] == 8 * (1 << 10)
    b.close()
    b = e.batch([_hea(hq, 16, 6, 2)], hq.HQ_OUT_EXPECT)
    info = b.info()
    state = 8 * (1 << 16)
    assert info["rounds"] >= 2
    assert state < info["pass_bytes"] <= state + 2 * state * (info["rounds"] - 1)
    b.close()
    e.close()


def test_budget_admission(gpu):
    hq = gpu
    budget = 1 << 27  # 128 MiB: about three 20-qubit complex128 states at a time
    e = hq.Engine(hq.HQ_C128, budget)
    jobs = [e.submit(_rl(hq, 20, 4, s), hq.HQ_OUT_EXPECT) for s in range(8)]
    big = e.submit(_rl(hq, 26, 2, 0), hq.HQ_OUT_EXPECT)  # 1 GiB: can never fit
    e.drain()
    st = e.stats()
    assert st["bytes_peak"] <= budget and st["bytes_reserved"] <= budget
    assert e.status(big) == hq.HQ_JOB_FAILED
    with pytest.raises(hq.HQError):
        e.result(big)
    ref = oracle.run_plain(to_oracle(random_layers(20, 4, 3)), 20)
    _, z, zz = _readouts(ref, 20)
    r = e.result(jobs[3])
    assert abs(r["z"] - z) < 1e-9 and abs(r["zz"] - zz) < 1e-9
    assert all(e.status(j) == hq.HQ_JOB_DONE for j in jobs)
    assert st["bytes_used"] == 0
    # blocks that each fit a chunk but not together: refused, not waited for forever
    e2 = hq.Engine(hq.HQ_C128, 1 << 26)             # one 64 MiB chunk
    j = e2.submit(_rl(hq, 22, 2, 0), hq.HQ_OUT_EXPECT)   # 64 MiB state + programs + partials
    e2.drain()
    assert e2.status(j) == hq.HQ_JOB_FAILED
    e2.close()


def test_failed_batch_keeps_admission_accounting(gpu):
    """A batch whose create fails partway (its second state no longer fits the
    budget) must leave the engine's count of bytes held by live batches as it
    was: a later job that fits the budget but not beside the live batch, with
    nothing else running, is then failed instead of waiting for memory that
    only the live batch holds."""
    import time
    hq = gpu
    e = hq.Engine(hq.HQ_C128, 1 << 27)                        # one 128 MiB chunk
    live = e.batch([_rl(hq, 21, 2, s) for s in range(2)])    # 2 x 32 MiB, stays live
    with pytest.raises(hq.HQError):
        e.batch([_rl(hq, 21, 2, s) for s in range(3)])
    j = e.submit(_rl(hq, 22, 2, 0), hq.HQ_OUT_EXPECT)        # 64 MiB
    t0 = time.perf_counter()
    while e.step() > 0 and time.perf_counter() - t0 < 20:
        time.sleep(0.01)
    assert e.status(j) == hq.HQ_JOB_FAILED
    live.run()
    ex, _ = live.results()
    assert np.all(np.abs(ex[:, 0] - 1) < 1e-12)
    live.close()
    # with the batch gone the same job fits
    j2 = e.submit(_rl(hq, 22, 2, 0), hq.HQ_OUT_EXPECT)
    e.drain()
    assert e.status(j2) == hq.HQ_JOB_DONE
    e.close()


def test_c64_close_to_c128(gpu):
    hq = gpu
    res = {}
    for prec in (hq.HQ_C128, hq.HQ_C64):
        e = hq.Engine(prec, 1 << 30)
        js = [e.submit(_hea(hq, n, 10, n), hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK) for n in (11, 14, 18)]
        e.drain()
        res[prec] = [e.result(j, topk=True) for j in js]
    for a, b in zip(res[hq.HQ_C128], res[hq.HQ_C64]):
        assert abs(a["z"] - b["z"]) < 1e-4 and abs(a["zz"] - b["zz"]) < 1e-4
        assert abs(b["norm"] - 1) < 1e-5
        np.testing.assert_allclose(a["topk"][1], b["topk"][1], atol=1e-6)


def test_state_topk(gpu):
    hq = gpu
    for n in (7, 12, 21):
        s = hq.State(n)
        psi = np.random.default_rng(n).normal(size=1 << n) + 0j
        psi /= np.linalg.norm(psi)
        s.upload(psi)
        idx, prob = s.topk(16)
        ridx, rprob = _topk_ref(psi)
        assert idx == ridx
        np.testing.assert_allclose(prob, rprob, rtol=1e-14)
        s.close()


def test_state_topk_ties(gpu):
    hq = gpu
    s = hq.State(13)
    psi = np.full(1 << 13, 1 / np.sqrt(1 << 13), dtype=np.complex128)
    s.upload(psi)
    idx, prob = s.topk(16)
    assert idx == list(range(16))
    s.close()


@pytest.mark.parametrize("n,prec", [(9, "c128"), (20, "c128"), (20, "c64"), (13, "c64")])
def test_state_sample_matches_inverse_cdf(gpu, n, prec):
    from helpers import Rng
    hq = gpu
    p = hq.HQ_C128 if prec == "c128" else hq.HQ_C64
    s = hq.State(n, p)
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi = (psi / np.linalg.norm(psi)).astype(np.complex128 if p == hq.HQ_C128 else np.complex64)
    s.upload(psi)
    shots = 3000
    got = s.sample(shots, seed=77)
    prob = np.abs(psi.astype(np.complex128)) ** 2
    cdf = np.cumsum(prob)
    r = Rng(77)
    u = np.array([r.uniform() for _ in range(shots)]) * cdf[-1]
    want = np.searchsorted(cdf, u, side="right")
    assert np.array_equal(got.astype(np.int64), want)
    s.close()


def test_state_sample_basis_state(gpu):
    hq = gpu
    s = hq.State(16)
    psi = np.zeros(1 << 16, complex)
    psi[12345] = 1j
    s.upload(psi)
    assert set(s.sample(100, 3).tolist()) == {12345}
    s.close()


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_specialised_batch_kernels(gpu, prec):
    """Option "jit": the engine's groups run as specialised batched kernels,
    shared by groups of the same structure (angles in per-circuit
    coefficients).  A batch of ansatz circuits with repeated (qubits, depth)
    and random-layer circuits, with readout and top-k, against the oracle;
    the same results alone and in a mix; graph replays repeat bit for bit."""
    hq = gpu
    p = hq.HQ_C128 if prec == "c128" else hq.HQ_C64
    tol = 1e-10 if prec == "c128" else 2e-5
    specs = [("hea", 16, 6, 1), ("hea", 16, 6, 2), ("hea", 16, 6, 3), ("hea", 19, 4, 4), ("hea", 11, 8, 5),
             ("hea", 19, 4, 6), ("rl", 17, 5, 7), ("rl", 13, 9, 8), ("hea", 12, 3, 9), ("rl", 20, 3, 10)]

    def build(kind, n, d, s):
        return _hea(hq, n, d, s) if kind == "hea" else _rl(hq, n, d, s)

    e = hq.Engine(p, 1 << 31)
    e.set_option("jit", 1)
    e.set_option("compile_threads", 8)
    b = e.batch([build(*s) for s in specs], hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
    b.run()
    r1, t1 = b.results()
    for i, (kind, n, d, s) in enumerate(specs):
        spec = hea_layers(n, d, s) if kind == "hea" else random_layers(n, d, s)
        psi = oracle.run_plain(to_oracle(spec), n)
        nrm, z, zz = _readouts(psi, n)
        print(i, kind, n, d, r1[i], (nrm, z, zz), flush=True)
        assert abs(r1[i][0] - nrm) <= tol and abs(r1[i][1] - z) <= tol * n and abs(r1[i][2] - zz) <= tol * n
        idx, prob = _topk_ref(psi)
        np.testing.assert_allclose(t1[1][i], prob, rtol=0, atol=tol)
        # the indices carry those probabilities (ties may come in either order)
        got_idx = np.array([int(x) for x in t1[0][i]])
        np.testing.assert_allclose(np.abs(psi[got_idx]) ** 2, prob, rtol=0, atol=tol)
    for _ in range(2):
        b.run()
        r2, t2 = b.results()
        assert np.array_equal(r1, r2) and np.array_equal(t1[1], t2[1])
    b.close()
    # continuous batching with the same kernels: alone and in a mix
    alone = {}
    for i in (0, 3, 6):
        e1 = hq.Engine(p, 1 << 31)
        e1.set_option("jit", 1)
        j = e1.submit(build(*specs[i]), hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
        e1.drain()
        alone[i] = e1.result(j, topk=True)
        e1.close()
    jobs = {}
    for i, s in enumerate(specs):
        jobs[i] = e.submit(build(*s), hq.HQ_OUT_EXPECT | hq.HQ_OUT_TOPK)
        if i % 3 == 0:
            e.step()
    e.drain()
    for i in (0, 3, 6):
        r = e.result(jobs[i], topk=True)
        a = alone[i]
        assert (r["norm"], r["z"], r["zz"]) == (a["norm"], a["z"], a["zz"])
        assert r["topk"] == a["topk"]
    e.close()
