"""Batching and serving forms on one GPU, over the engine (``hq_engine_*``).

Workloads (portable: numpy's PCG64 generator seeded with the given seed):

* ``hea_batch(count, seed)``: hardware-efficient-ansatz circuits, qubit count
  ``round(exp(U[ln 10, ln 22]))``, depth uniform in [4, 40], each with its own
  generator seed; the output is sum_i <Z_i Z_{i+1}>.
* ``request_stream(count, rate, seed)``: a Poisson stream (exponential gaps of
  mean 1/rate seconds) of random-layer or ansatz circuits (even odds), qubits
  ``round(exp(U[ln 12, ln 28]))``, depth uniform in [8, 60], asking for
  sum_i <Z_i> or the 16 most probable basis states (even odds).

Drivers:

* ``serve(engine, requests)``: real-time loop; a request is submitted when its
  arrival time has passed, the engine advances every admitted circuit one
  group per round, new arrivals join at the next round (continuous batching).
* ``serve_hybrid(hq, requests)``: the baseline, one circuit at a time in
  arrival order with the hybrid partitioner (OShareMem / TransMM groups) in
  complex128, each run on a resident state (one per qubit count, allocated
  before the stream starts).
* ``serve_plain(hq, requests)``: a second baseline, one circuit at a time,
  plain per-gate kernels, complex128, a fresh state per request.

Both return one record per request: arrival, done, latency (seconds from
arrival to the result), and the result.
"""
from __future__ import annotations

import math
import time

import numpy as np

_LN = math.log


def _log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(_LN(lo), _LN(hi)))))


def hea_batch(count: int, seed: int, nmin: int = 10, nmax: int = 22, dmin: int = 4, dmax: int = 40):
    """[(num_qubits, depth, circuit_seed)] of the batching workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = _log_uniform_int(rng, nmin, nmax)
        d = int(rng.integers(dmin, dmax + 1))
        out.append((n, d, int(rng.integers(0, 2 ** 63))))
    return out


def request_stream(count: int, rate: float, seed: int, nmin: int = 12, nmax: int = 28, dmin: int = 8,
                   dmax: int = 60):
    """Requests of the serving workload: dict(arrival, kind, n, depth, seed, output)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        out.append(dict(arrival=t,
                        kind="rl" if rng.random() < 0.5 else "hea",
                        n=_log_uniform_int(rng, nmin, nmax),
                        depth=int(rng.integers(dmin, dmax + 1)),
                        seed=int(rng.integers(0, 2 ** 63)),
                        output="z" if rng.random() < 0.5 else "topk"))
    return out


def build_circuit(hq, kind: str, n: int, depth: int, seed: int):
    c = hq.Circuit(n)
    if kind == "hea":
        c.add_hea_layers(depth, seed)
    else:
        c.add_random_layers(depth, seed)
    return c


def autotune_tiles(hq, engine, specs, circuits, per_n: int = 12, reps: int = 3):
    """Autotuner of the batching form: for every qubit count in `specs`, time
    each candidate tile (the whole state when it fits, else 10..13 qubits) on up
    to `per_n` of the batch's circuits of that count as a captured batch, keep
    the fastest in `engine` (hq_engine_set_tile) and return {n: tile}."""
    by_n = {}
    for i, (n, _, _) in enumerate(specs):
        by_n.setdefault(n, []).append(i)
    chosen = {}
    for n, idx in sorted(by_n.items()):
        sample = [circuits[i] for i in idx[:per_n]]
        cands = sorted({k for k in (n, 13, 12, 11, 10) if 7 <= k <= min(n, 13)})
        best = None
        for k in cands:
            engine.set_tile(n, k)
            b = engine.batch(sample)
            b.run()
            t = min(b.run() for _ in range(reps))
            b.close()
            if best is None or t < best[0]:
                best = (t, k)
        engine.set_tile(n, best[1])
        chosen[n] = best[1]
    return chosen


def percentile(values, q):
    if not values:
        return None
    return float(np.percentile(np.asarray(values), q))


def serve(hq, engine, requests, circuits=None, idle_sleep: float = 2e-4):
    """Run `requests` through `engine` in real time; returns (records, wall seconds)."""
    circuits = circuits or [build_circuit(hq, r["kind"], r["n"], r["depth"], r["seed"]) for r in requests]
    n = len(requests)
    job_of = {}
    req_of = {}
    done = {}
    nxt = 0
    t0 = engine.now()
    while len(done) < n:
        now = engine.now() - t0
        while nxt < n and requests[nxt]["arrival"] <= now:
            r = requests[nxt]
            out = hq.HQ_OUT_EXPECT | (hq.HQ_OUT_TOPK if r["output"] == "topk" else 0)
            j = engine.submit(circuits[nxt], out)
            job_of[nxt] = j
            req_of[j] = nxt
            nxt += 1
        pending = engine.step()
        for j in engine.poll():
            i = req_of[j]
            want_topk = requests[i]["output"] == "topk"
            if engine.status(j) == hq.HQ_JOB_FAILED:
                done[i] = dict(error=True, done=engine.now() - t0)
            else:
                res = engine.result(j, topk=want_topk)
                done[i] = dict(error=False, done=res["times"][2] - t0, result=res)
            engine.release(j)
        if pending == 0 and nxt < n:
            wait = requests[nxt]["arrival"] - (engine.now() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.01))
        elif pending and idle_sleep:
            pass
    recs = []
    for i, r in enumerate(requests):
        d = done[i]
        recs.append(dict(arrival=r["arrival"], done=d["done"], latency=d["done"] - r["arrival"],
                         error=d["error"], result=d.get("result")))
    return recs, max(x["done"] for x in recs)


def _logical_topk(idx, prob, pos):
    """Top-k of a permuted state: physical basis indices to logical ones
    (bit pos[q] of a physical index is logical qubit q)."""
    out = []
    for i in idx:
        v = 0
        for q, p in enumerate(pos):
            v |= ((int(i) >> p) & 1) << q
        out.append(v)
    return out, list(prob)


# From this many qubits every TransMM group runs in place (matrices of 4..7
# qubits, either precision), so resident states need no work buffer.
INPLACE_MIN_QUBITS = 12


def resident_states(hq, requests):
    """One complex128 state per qubit count of `requests` (with a work buffer
    below INPLACE_MIN_QUBITS, where TransMM groups may run out of place)."""
    return {n: hq.State(n, hq.HQ_C128, with_buffer=n < INPLACE_MIN_QUBITS)
            for n in sorted({r["n"] for r in requests})}


def serve_hybrid(hq, requests, circuits=None, states=None):
    """Baseline: one circuit at a time in arrival order, hybrid partitioner
    (OShareMem / TransMM groups as the cost model picks, interpreter kernels:
    no per-request kernel build), complex128, on resident states."""
    circuits = circuits or [build_circuit(hq, r["kind"], r["n"], r["depth"], r["seed"]) for r in requests]
    own = states is None
    states = states or resident_states(hq, requests)
    t0 = time.perf_counter()
    recs = []
    for i, r in enumerate(requests):
        wait = r["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        c = circuits[i]
        c.set_option("mode", hq.HQ_MODE_MIX)
        c.set_option("precision", hq.HQ_C128)
        c.set_option("jit", 0)
        c.use_state(states[r["n"]])
        c.compile()
        c.run(copy_back=False, destroy=False)
        st = c.state_view(hq.HQ_C128)
        z, zz = st.expect_z()
        pos = c.final_pos()
        # sum <Z_i> does not depend on the layout; neighbour pairs do (known
        # for the identity layout of an OShareMem-only schedule)
        res = dict(norm=st.norm2(), z=float(z.sum()),
                   zz=float(zz.sum()) if pos == list(range(len(pos))) else None)
        if r["output"] == "topk":
            idx, prob = st.topk(16)
            res["topk"] = _logical_topk(idx, prob, pos)
        c.use_state(None)
        c.close()
        t = time.perf_counter() - t0
        recs.append(dict(arrival=r["arrival"], done=t, latency=t - r["arrival"], error=False, result=res))
    if own:
        for st in states.values():
            st.close()
    return recs, max(x["done"] for x in recs)


def serve_plain(hq, requests, circuits=None):
    """Baseline: one circuit at a time, arrival order, plain per-gate kernels, complex128."""
    circuits = circuits or [build_circuit(hq, r["kind"], r["n"], r["depth"], r["seed"]) for r in requests]
    t0 = time.perf_counter()
    recs = []
    for i, r in enumerate(requests):
        wait = r["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        c = circuits[i]
        c.set_option("mode", hq.HQ_MODE_SERIAL)
        c.set_option("precision", hq.HQ_C128)
        c.run(copy_back=False, destroy=False)
        st = c.state_view(hq.HQ_C128)
        z, zz = st.expect_z()
        res = dict(norm=st.norm2(), z=float(z.sum()), zz=float(zz.sum()))
        if r["output"] == "topk":
            res["topk"] = st.topk(16)
        c.close()
        t = time.perf_counter() - t0
        recs.append(dict(arrival=r["arrival"], done=t, latency=t - r["arrival"], error=False, result=res))
    return recs, max(x["done"] for x in recs)


def summarize(recs, wall):
    lat = [x["latency"] for x in recs if not x["error"]]
    ok = len(lat)
    return dict(requests=len(recs), completed=ok, failed=len(recs) - ok,
                circuits_per_s=ok / wall if wall > 0 else None,
                p50_ms=None if not lat else percentile(lat, 50) * 1e3,
                p99_ms=None if not lat else percentile(lat, 99) * 1e3,
                makespan_s=wall)
