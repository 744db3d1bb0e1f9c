"""Cost-model calibration on the GPU (the step the reference performs with
evaluator-preprocess/process.cpp).

    python -m hyquas_amd.calibrate [--qubits 30] [--out DIR] [--tables 22,24,26,28]

Writes
  DIR/mi355x.params      this build's model: OShareMem pass bandwidth by tile
                         run length (complex128 and complex64), VALU rates,
                         MFMA rates of the TransMM GEMM, transpose bandwidth,
                         launch overhead ("key value" lines);
  DIR/<n>qubits.out      the reference's table format for each n in --tables
                         (param_type 1: 14 single-gate rows, 7 controlled-gate
                         rows, "K ms" rows for K = 1..512, transpose cost; the
                         rows are times of 512-gate circuits in microseconds,
                         like process.cpp:20-82), extended to complex64 with
                         --c64; the dense rows K = 2..128 are measured, K = 1
                         repeats K = 2 and K = 256, 512 are the K = 128 time
                         scaled by K / 128 (the dense kernel stops at 7
                         qubits, and the partitioner reads rows 4..7 only).

Either file loads with ``Circuit.load_cost_model(path)``
(``hq_circuit_load_cost_model``).  Prints one JSON summary line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

import hyquas_amd as hq

TILE = 12        # VALU-rate group tile
RUN_TILE = 13    # tile of the run-length passes: the specialised kernels' default


def _time_program(st, prog, reps=5):
    st.launch(prog)
    st.sync()
    st.profile_read()
    for _ in range(reps):
        st.launch(prog)
    st.sync()
    t = sorted(ms for _, ms in st.profile_read())
    return t[len(t) // 2]


def run_length_table(n, precision, samples=4, seed=0, tile=RUN_TILE):
    """Bandwidth (B/s) of an OShareMem pass by the run length L of its tile,
    L = 0..tile (capped at 12): the tile is the lowest L qubits plus tile - L
    qubits drawn at random above qubit L (median of `samples` draws).  The pass
    is a specialised kernel holding one diagonal gate (a kernel without gates
    loads and stores the same values, which the compiler removes)."""
    tile = min(tile, n - 1)
    st = hq.State(n, precision)
    st.profile(True)
    b = 16 if precision == hq.HQ_C128 else 8
    rng = np.random.default_rng(seed)
    out = []
    for L in range(13):
        ts = []
        for _ in range(samples if L < tile else 1):
            rest = rng.choice(np.arange(L + 1, n), tile - L, replace=False) if L < tile else []
            mask = (1 << L) - 1
            for q in rest:
                mask |= 1 << int(q)
            low = min(q for q in range(n) if (mask >> q) & 1)
            p = st.program_sharemem(hq.gates_to_array([hq.Gate.RZ(low, 0.25)]), mask)
            p.specialize()
            ts.append(_time_program(st, p, reps=3))
            p.close()
        ms = float(np.median(ts))
        out.append(2.0 * (1 << n) * b / (ms * 1e-3))
    st.close()
    return out


def valu_rate(n, precision):
    """FP64 / FP32 lane-ops per second of RX updates on register qubits (compute-bound group)."""
    st = hq.State(n, precision)
    st.profile(True)
    mask = (1 << TILE) - 1
    gates = [hq.Gate.RX(q % 4 + 8, 0.1 + 0.01 * i) for i, q in enumerate(range(256))]
    p = st.program_sharemem(hq.gates_to_array(gates), mask)
    p.specialize()
    ms = _time_program(st, p)
    p.close()
    st.close()
    # factored RX: 4 FMA per pair = 2 per amplitude
    return 2.0 * len(gates) * (1 << n) / (ms * 1e-3)


def dense_and_transpose(n, precision):
    st = hq.State(n, precision, with_buffer=True)
    st.profile(True)
    b = 16 if precision == hq.HQ_C128 else 8
    rates, times = {}, {}
    rng = np.random.default_rng(0)
    for mq in range(1, 8):
        K = 1 << mq
        m = np.linalg.qr(rng.normal(size=(K, K)) + 1j * rng.normal(size=(K, K)))[0]
        colmajor = np.empty(2 * K * K)
        colmajor[0::2] = m.T.reshape(-1).real
        colmajor[1::2] = m.T.reshape(-1).imag
        p = st.program_blas(None, colmajor, mq)
        ms = _time_program(st, p)
        times[K] = ms
        rates[K] = 8.0 * K * (1 << n) / (ms * 1e-3)
        p.close()
    ttimes = []
    for k in range(5):
        perm = list(range(n))
        for _ in range(4):
            i, j = rng.integers(0, n, 2)
            perm[i], perm[j] = perm[j], perm[i]
        if perm == list(range(n)):
            perm[0], perm[n - 1] = perm[n - 1], perm[0]
        p = st.program_blas(perm, None, 0)
        ttimes.append(_time_program(st, p))
        p.close()
    st.close()
    tms = float(np.median(ttimes))
    return rates, times, tms, 2.0 * (1 << n) * b / (tms * 1e-3)


def launch_overhead():
    st = hq.State(10)
    g = hq.gates_to_array([hq.Gate.H(0)] * 2000)
    st.apply_plain(g[:10])
    st.sync()
    t0 = time.perf_counter()
    st.apply_plain(g)
    st.sync()
    dt = (time.perf_counter() - t0) / len(g)
    st.close()
    return dt


def reference_tables(n, precision, path):
    """The reference's <n>qubits.out: 512-gate circuits per gate type (us),
    K = 1..512 dense times (ms), transpose cost (ms)."""
    singles = [hq.GateType.U1, hq.GateType.U2, hq.GateType.U3, hq.GateType.H, hq.GateType.X, hq.GateType.Y,
               hq.GateType.Z, hq.GateType.S, hq.GateType.SDG, hq.GateType.T, hq.GateType.TDG, hq.GateType.RX,
               hq.GateType.RY, hq.GateType.RZ]
    ctrls = [hq.GateType.CNOT, hq.GateType.CY, hq.GateType.CZ, hq.GateType.CRX, hq.GateType.CRY,
             hq.GateType.CU1, hq.GateType.CRZ]
    npar = {hq.GateType.U1: 1, hq.GateType.U2: 2, hq.GateType.U3: 3, hq.GateType.RX: 1, hq.GateType.RY: 1,
            hq.GateType.RZ: 1, hq.GateType.CRX: 1, hq.GateType.CRY: 1, hq.GateType.CU1: 1, hq.GateType.CRZ: 1}
    rng = np.random.default_rng(n)

    def circuit_us(gates):
        c = hq.Circuit(n)
        for g in gates:
            c.addGate(g)
        c.set_option("mode", hq.HQ_MODE_GROUP)
        c.set_option("precision", precision)
        c.compile()
        c.run(copy_back=False, destroy=False)
        us = min(c.run(copy_back=False, destroy=False) for _ in range(3))
        c.close()
        return us

    su = [circuit_us([hq.Gate.make(t, -1, -1, 1, tuple(rng.uniform(0, 2 * math.pi, npar.get(t, 0))))
                      for _ in range(512)]) for t in singles]
    cu = [circuit_us([hq.Gate.make(t, -1, 0, 2, tuple(rng.uniform(0, 2 * math.pi, npar.get(t, 0))))
                      for _ in range(512)]) for t in ctrls]
    _, times, tms, _ = dense_and_transpose(n, precision)
    with open(path, "w") as f:
        f.write(reference_table_text(su, cu, times, tms))


def reference_table_text(singles_us, ctrls_us, dense_ms, transpose_ms):
    """The reference's <n>qubits.out text (param_type 1, read by
    src/evaluator.cpp:60-103): the 14 single-gate rows (U1, U2, U3, H, X, Y, Z,
    S, SDG, T, TDG, RX, RY, RZ), a blank line, the 7 controlled-gate rows
    (CNOT, CY, CZ, CRX, CRY, CU1, CRZ), the "K ms" rows for K = 1..512 and the
    transpose cost in ms.  dense_ms maps K = 2..128 to measured times."""
    assert len(singles_us) == 14 and len(ctrls_us) == 7
    lines = ["1"] + [str(v) for v in singles_us] + [""] + [str(v) for v in ctrls_us] + [""]
    for i in range(10):
        K = 1 << i
        ms = dense_ms[K] if K in dense_ms else dense_ms[2] if K == 1 else dense_ms[128] * K / 128
        lines.append(f"{K} {ms:.6f}")
    lines += ["", f"{transpose_ms:.6f}"]
    return "\n".join(lines) + "\n"


def params_text(n, hbm_bw, valu64, valu32, mfma64, mfma32, launch, bw64, bw32):
    """This build's model file ("key value" lines, read by CostModel::load)."""
    out = [f"# MI355X cost model, calibrated at {n} qubits by hyquas_amd.calibrate",
           f"hbm_bw {hbm_bw:.6e}", f"valu_rate {valu64:.6e}", f"valu_rate_f32 {valu32:.6e}",
           f"mfma_f64 {mfma64:.6e}", f"mfma_f32 {mfma32:.6e}", f"launch_overhead {launch:.6e}"]
    out += [f"shm_bw_run_{L} {b:.6e}" for L, b in enumerate(bw64)]
    out += [f"shm_bw_run_f32_{L} {b:.6e}" for L, b in enumerate(bw32)]
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--out", default="calibration")
    ap.add_argument("--tables", default="", help="qubit counts for the reference-format tables, e.g. 22,24,26,28")
    ap.add_argument("--c64", action="store_true", help="reference-format tables in complex64")
    ap.add_argument("--samples", type=int, default=4, help="random tiles per run length")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    n = a.qubits
    t0 = time.perf_counter()

    def done(what):
        print(f"calibrate: {what} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    bw64 = run_length_table(n, hq.HQ_C128, a.samples)
    done("complex128 run-length table")
    bw32 = run_length_table(n, hq.HQ_C64, a.samples)
    done("complex64 run-length table")
    v64 = valu_rate(n, hq.HQ_C128)
    v32 = valu_rate(n, hq.HQ_C64)
    done("VALU rates")
    r64, _, t64, tbw = dense_and_transpose(n - 1, hq.HQ_C128)
    r32, _, _, _ = dense_and_transpose(n - 1, hq.HQ_C64)
    done("dense GEMM and transpose")
    lo = launch_overhead()
    params = os.path.join(a.out, "mi355x.params")
    with open(params, "w") as f:
        f.write(params_text(n, tbw, v64, v32, r64[64], r32[128], lo, bw64, bw32))
    tables = []
    for m in [int(x) for x in a.tables.split(",") if x]:
        path = os.path.join(a.out, f"{m}qubits.out")
        reference_tables(m, hq.HQ_C64 if a.c64 else hq.HQ_C128, path)
        tables.append(path)
        done(f"reference-format table for {m} qubits")
    print(json.dumps({"params": params, "tables": tables, "qubits": n, "seconds": round(time.perf_counter() - t0, 1),
                      "shm_bw_run_c128": [round(b / 1e12, 3) for b in bw64],
                      "shm_bw_run_c64": [round(b / 1e12, 3) for b in bw32],
                      "valu_rate_c128": v64, "valu_rate_c64": v32,
                      "dense_flops_c128": {str(k): round(v / 1e12, 2) for k, v in r64.items()},
                      "dense_flops_c64": {str(k): round(v / 1e12, 2) for k, v in r32.items()},
                      "transpose_GBps": round(tbw / 1e9, 1), "launch_overhead_us": round(lo * 1e6, 2)}),
          flush=True)


if __name__ == "__main__":
    sys.exit(main())
