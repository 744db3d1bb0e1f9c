"""Benchmark: 30-qubit random-layer circuit, complex128, hybrid partitioner,
one MI355X.

A step is one full simulation of the circuit from |0...0> on the GPU:
state reset + every group kernel (OShareMem / TransMM as the partitioner
chose) + stream sync.  The compiled schedule and every group program are
resident on the device before timing starts.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30]
                    [--layers 30] [--seed 0] [--cpu-gates 45] [--no-cpu]

roofline: the dominant kernel is the OShareMem group kernel; its algorithmic
traffic per launch is one read + one write of the state, 2 * 2^n * 16 bytes
(the first group of a run builds |0...0> in registers and only writes), averaged
over the launches of a step and divided by the average launch duration (HIP
events recorded on the state's stream around every launch in the timed steps).

cpu_baseline: the FP64 OpenMP per-gate CPU path (oracle/, kernelSimple.cu
semantics) timed on the first --cpu-gates gates of the same 30-qubit
circuit on this host, scaled to the whole circuit.

The default line also carries the metric's second half (--no-forms skips
it): "batch", 4096 hardware-efficient-ansatz circuits batched in complex64
(specialised batched kernels, one per distinct group structure; their build
time is reported, not timed) against the FP64 hybrid path one circuit at a
time on resident states and
the plain per-gate path (both on 128 evenly spaced circuits of the batch);
"serve", a Poisson stream of 500 requests at each of --form-rates served by
the complex64 engine against the FP64 hybrid path one circuit at a time
(circuits/s, p50, p99 latency), and the plain per-gate path on the first
100 requests of each stream.  --form batch|serve|precision run one form
alone with more detail.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK = 78.6e12


def _traffic_from_profiles(config, alg_bytes):
    """HBM bytes per OShareMem launch measured by rocprofv3 PMC counters
    (FETCH_SIZE, WRITE_SIZE), if a summary of exactly this workload is
    committed under profiles/: same circuit, precision and mode, and the same
    algorithmic bytes per launch (so the same schedule); None otherwise."""
    path = os.path.join(ROOT, "profiles", f"sharemem_traffic_{config['qubits']}q.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("config") != config or abs(d.get("algorithmic_bytes_per_launch", 0) - alg_bytes) > 1e-3 * alg_bytes:
        return None
    return d.get("bytes_per_launch")


def _batch_valu_from_profiles(name, config, launches):
    """VALU wave-instructions of one batch replay measured by rocprofv3
    (SQ_INSTS_VALU summed over the replay's dispatches), if a summary of
    exactly this batch is committed under profiles/: same circuits, kernels
    and tiles, and the same launches per replay; (count, peak rate) or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "batch_valu_hea4096.json")) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    rec = d.get(name)
    if d.get("config") != config or not rec or rec.get("launches_per_replay") != launches:
        return None
    return rec["valu_wave_instr_per_replay"], d["peak_wave_instr_per_s"]


def cpu_baseline(n, layers, seed, gates_sample):
    import numpy as np

    import oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import random_layers, to_oracle

    spec = random_layers(n, layers, seed)
    total = len(spec)
    sample = to_oracle(spec[:gates_sample])
    threads = oracle.max_threads()
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    t0 = time.perf_counter()
    oracle.run_plain(sample, n, psi, threads)
    dt = time.perf_counter() - t0
    per_gate = dt / len(sample)
    return {
        "value": per_gate * total,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {len(sample)} of {total} gates of the {n}-qubit circuit ({dt:.1f} s), "
                  f"per-gate time scaled to the whole circuit",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--cpu-gates", type=int, default=45)
    ap.add_argument("--share-qubits", type=int, default=0, help="OShareMem tile qubits (0: default)")
    ap.add_argument("--reg-qubits", type=int, default=0, help="register qubits per thread (0: default)")
    ap.add_argument("--coalesce", type=int, default=0, help="low qubits always in the tile (0: default)")
    ap.add_argument("--beam", type=int, default=0, help="partitions kept by the beam search (0: default)")
    ap.add_argument("--starts", type=int, default=0, help="merge starts per beam step (0: default)")
    ap.add_argument("--tune", type=int, choices=[-1, 0, 1, 2], default=-1,
                    help="per-group kernel variant timing at compile: -1 auto, 0 off, 1 on, "
                         "2 whole-line variant everywhere")
    ap.add_argument("--mode", choices=["mix", "group", "blas", "serial"], default="mix")
    ap.add_argument("--blas-mat", type=int, default=0, help="TransMM matrix qubits in blas mode (0: default)")
    ap.add_argument("--transmm-inplace", type=int, choices=[-1, 0, 1], default=-1,
                    help="TransMM groups: 1 one in-place dense pass, 0 transpose + GEMM (-1: library default, 1)")
    ap.add_argument("--jit", type=int, choices=[-1, 0, 1], default=-1,
                    help="specialised group kernels: -1 auto (from 22 qubits), 0 interpreter, 1 always")
    ap.add_argument("--cost-model", default="", help="parameter file for the partitioner's cost model")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["sim", "batch", "serve", "precision"], default="sim",
                    help="sim: one 30-qubit circuit (default); batch: 4096 ansatz circuits; "
                         "serve: Poisson request stream")
    ap.add_argument("--count", type=int, default=4096, help="batch: circuits")
    ap.add_argument("--plain-sample", type=int, default=256,
                    help="batch: circuits of the plain FP64 one-at-a-time baseline (evenly spaced)")
    ap.add_argument("--no-autotune", action="store_true", help="batch: default tiles instead of autotuned ones")
    ap.add_argument("--batch-jit", type=int, choices=[0, 1], default=1,
                    help="batch: 1 specialised batched kernels shared by same-structure groups (default), "
                         "0 the interpreter kernel with autotuned tiles")
    ap.add_argument("--batch-tile", type=int, default=0,
                    help="batch/serve: engine tile qubits of partitioned circuits (0: engine default)")
    ap.add_argument("--streams", type=int, default=0,
                    help="batch/serve: streams a round's group launches spread over (0: engine default, 4)")
    ap.add_argument("--requests", type=int, default=500, help="serve: requests per rate")
    ap.add_argument("--rates", default="20,100", help="serve: arrival rates (requests/s), comma separated")
    ap.add_argument("--no-plain", action="store_true", help="serve: skip the plain per-gate baseline")
    ap.add_argument("--no-forms", action="store_true",
                    help="default form: skip the batching and serving summaries")
    ap.add_argument("--form-count", type=int, default=4096,
                    help="default form: circuits of the batching summary")
    ap.add_argument("--form-rates", default="100,400",
                    help="default form: arrival rates (requests/s) of the serving summary")
    ap.add_argument("--copy-pass", action="store_true",
                    help="before timing, run one zero-gate OShareMem pass (a copy of known size, "
                         "2 * 2^n * amplitude bytes) to calibrate HBM counters")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the state-vector path shards by global qubits, which is not part of this build)",
              file=sys.stderr)
        sys.exit(2)

    import hyquas_amd as hq

    if a.form == "batch":
        return main_batch(a, hq)
    if a.form == "serve":
        return main_serve(a, hq)
    if a.form == "precision":
        return main_precision(a, hq)

    n, layers = a.qubits, a.layers
    prec = hq.HQ_C128 if a.precision == "c128" else hq.HQ_C64
    amp_bytes = 16 if prec == hq.HQ_C128 else 8
    c = hq.Circuit(n).add_random_layers(layers, a.seed)
    c.set_option("precision", prec)
    c.set_option("profile_events", 1)
    c.set_option("mode", dict(serial=0, group=1, blas=3, mix=4)[a.mode])
    if a.share_qubits:
        c.set_option("share_qubits", a.share_qubits)
    if a.reg_qubits:
        c.set_option("reg_qubits", a.reg_qubits)
    if a.coalesce:
        c.set_option("coalesce", a.coalesce)
    if a.blas_mat:
        c.set_option("blas_mat", a.blas_mat)
    c.set_option("jit", a.jit)
    c.set_option("tune", a.tune)
    if a.transmm_inplace >= 0:
        c.set_option("transmm_inplace", a.transmm_inplace)
    if a.beam:
        c.set_option("beam", a.beam)
    if a.starts:
        c.set_option("starts", a.starts)
    jit_on = a.jit == 1 or (a.jit == -1 and n >= 22)
    if a.cost_model:
        c.load_cost_model(a.cost_model)
    t0 = time.perf_counter()
    c.compile()
    compile_s = time.perf_counter() - t0
    sched = c.schedule()
    ngates = sum(g["num_gates"] for g in sched)

    for _ in range(a.warmup):
        c.run(copy_back=False, destroy=False)

    if a.copy_pass:
        st = c.state_view(prec)
        mask = (1 << min(12, n)) - 1
        prog = st.program_sharemem(hq.gates_to_array([]), mask)
        st.launch(prog)
        st.sync()
        prog.close()

    launches = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        c.run(copy_back=False, destroy=False)
        launches.extend(c.launch_times())
    elapsed = time.perf_counter() - t0
    sec_per_step = elapsed / a.steps

    # sanity: the state stays normalised
    st = c.state_view(prec)
    norm_err = abs(st.norm2() - 1.0)

    shm = [ms for k, ms in launches if k == hq.KIND_SHAREMEM]
    bytes_per_pass = 2.0 * (1 << n) * amp_bytes
    # algorithmic bytes of each OShareMem pass: the share of a full read +
    # write of the state it moves (the library reports it: 0.5 for the first
    # pass, which builds |0...0> in registers and only writes; the live tiles'
    # share for the early passes that skip tiles known to be zero)
    pass_units = [g["live_fraction"] for g in sched if g["backend"] == hq.HQ_BACKEND_PERGATE]
    roof = None
    if shm:
        avg_ms = sum(shm) / len(shm)
        per_step = len(shm) // a.steps
        alg_bytes = bytes_per_pass * (sum(pass_units) if len(pass_units) == per_step else per_step) / per_step
        achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
        traffic = _traffic_from_profiles({"qubits": n, "layers": layers, "seed": a.seed,
                                          "precision": a.precision, "mode": a.mode}, alg_bytes)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "hq_group" if jit_on else "k_sharemem", "avg_launch_ms": round(avg_ms, 4),
                "launches_per_step": len(shm) // a.steps,
                "pass_units_per_step": round(sum(pass_units), 4)}
    kinds = {}
    for k, ms in launches:
        kinds[k] = kinds.get(k, 0.0) + ms
    groups_pg = sum(1 for g in sched if g["backend"] == hq.HQ_BACKEND_PERGATE)
    groups_blas = sum(1 for g in sched if g["backend"] == hq.HQ_BACKEND_BLAS)
    # whole-step roofline: every group needs one HBM pass (TransMM in the
    # transpose form two)
    dense = [ms for k, ms in launches if k == hq.KIND_DENSE]
    trans = [ms for k, ms in launches if k == hq.KIND_TRANSPOSE]
    blas_passes = (len(dense) + len(trans)) / a.steps
    t_ideal = (sum(pass_units) + blas_passes) * bytes_per_pass / (HBM_PEAK_GBS * 1e9)

    # TransMM groups: GEMM flops (8 K 2^n per group) over the dense kernels'
    # time, transposes as bytes moved (read + write of the state)
    transmm = None
    if dense:
        flops = sum(8.0 * (1 << g["mat_qubit"]) * (1 << n) for g in sched
                    if g["backend"] == hq.HQ_BACKEND_BLAS) * a.steps
        mfma_peak = 78.6 if prec == hq.HQ_C128 else 157.3
        tflops = flops / (sum(dense) * 1e-3) / 1e12
        davg = sum(dense) / len(dense)
        transmm = {"form": "transpose + GEMM" if trans else "in place",
                   "dense_launches": len(dense), "dense_avg_ms": round(davg, 3),
                   "dense_GBps": round(bytes_per_pass / (davg * 1e-3) / 1e9, 1),
                   "dense_TFLOPs": round(tflops, 2), "mfma_peak_TFLOPs": mfma_peak,
                   "mfma_frac": round(tflops / mfma_peak, 4),
                   "mat_qubits": sorted({g["mat_qubit"] for g in sched if g["backend"] == hq.HQ_BACKEND_BLAS})}
        if trans:
            tavg = sum(trans) / len(trans)
            transmm.update({"transpose_launches": len(trans), "transpose_avg_ms": round(tavg, 3),
                            "transpose_GBps": round(bytes_per_pass / (tavg * 1e-3) / 1e9, 1)})

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(n, layers, a.seed, a.cpu_gates)
        except Exception as e:  # reported, not fatal
            cpu = {"error": str(e)}

    forms = None
    if not a.no_forms and n == 30 and prec == hq.HQ_C128 and a.mode == "mix":
        # the metric's second half: the batched and served complex64 forms
        # against FP64 one circuit at a time on the same GPU
        c.close()
        rates = [float(x) for x in a.form_rates.split(",") if x]
        forms = {"batch": batch_form(a, hq, a.form_count, 128, precisions=("c64",))[0],
                 "serve": {"workload": SERVE_WORKLOAD.format(requests=a.requests, seed=a.seed),
                           "sweep": [serve_form(a, hq, r, a.requests, a.seed + k, plain=True, plain_sample=100)
                                     for k, r in enumerate(rates)]}}

    out = {
        "metric": "30q fp64 sim time + % HBM roofline; served circuits/s + p99 vs plain",
        "value": round(sec_per_step, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(sec_per_step * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == hq.HQ_C128 else "f32",
        "data": "synthetic",
        "config": {"workload": f"random-layer circuit, {n} qubits, {layers} layers (H/RX/RY/RZ/T + CNOT "
                               f"matching), seed {a.seed}, {a.precision} state, hybrid partitioner",
                   "qubits": n, "layers": layers, "gates": ngates, "groups_sharemem": groups_pg,
                   "groups_transmm": groups_blas, "share_qubits": a.share_qubits or None,
                   "reg_qubits": a.reg_qubits or None, "mode": a.mode, "parallelism": "single GPU"},
        "roofline": roof,
        "transmm": transmm,
        "step_hbm_roofline_frac": round(t_ideal / sec_per_step, 4),
        "device_ms_per_step": {str(k): round(v / a.steps, 3) for k, v in sorted(kinds.items())},
        "compile_s": round(compile_s, 3),
        "norm_error": norm_err,
        "cpu_baseline": cpu,
    }
    if forms:
        out.update(forms)
    print(json.dumps(out), flush=True)


def _progress(msg):
    # one line per finished step on stderr: long forms stay visibly alive
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _fidelity(hq, spec, kinds="hea"):
    """1 - |<psi64|psi128>|^2 for one circuit (OShareMem path, both precisions)."""
    from hyquas_amd.serving import build_circuit
    n, d, seed = spec
    cs = []
    for prec in (hq.HQ_C128, hq.HQ_C64):
        c = build_circuit(hq, kinds, n, d, seed)
        c.set_option("mode", hq.HQ_MODE_GROUP)
        c.set_option("precision", prec)
        c.run(copy_back=False, destroy=False)
        cs.append(c)
    ov = cs[0].state_view(hq.HQ_C128).inner(cs[1].state_view(hq.HQ_C64))
    for c in cs:
        c.close()
    return 1.0 - abs(ov) ** 2


def _build_batch(hq, e, circuits, name):
    """e.batch(circuits) on a worker thread; while it compiles, report how many
    specialised kernels the engine has built (the compile is the long part of
    a cold batch)."""
    import threading
    box = {}

    def work():
        try:
            box["b"] = e.batch(circuits, hq.HQ_OUT_EXPECT)
        except Exception as ex:  # re-raised below
            box["err"] = ex

    th = threading.Thread(target=work)
    t0 = time.perf_counter()
    th.start()
    while th.is_alive():
        th.join(10.0)
        if th.is_alive():
            _progress(f"batch: {name} building, {e.kernels()} specialised kernels compiled "
                      f"({time.perf_counter() - t0:.0f} s)")
    if "err" in box:
        raise box["err"]
    return box["b"]


def _one_at_a_time(hq, specs, kind, mode):
    """Circuits of `specs` one at a time on the same GPU, complex128: `mode`
    "hybrid" = the hybrid partitioner (interpreter kernels, no per-circuit
    kernel build) on resident states, one per qubit count allocated up front;
    "plain" = per-gate kernels, a fresh state per circuit.  Returns (circuits/s,
    [sum Z_i Z_(i+1)])."""
    from hyquas_amd.serving import INPLACE_MIN_QUBITS, build_circuit
    states = ({n: hq.State(n, hq.HQ_C128, with_buffer=n < INPLACE_MIN_QUBITS) for n in sorted({s[0] for s in specs})}
              if mode == "hybrid" else {})
    out = []
    t0 = time.perf_counter()
    for n, d, seed in specs:
        c = build_circuit(hq, kind, n, d, seed)
        c.set_option("precision", hq.HQ_C128)
        if mode == "hybrid":
            c.set_option("mode", hq.HQ_MODE_MIX)
            c.set_option("jit", 0)
            c.use_state(states[n])
        else:
            c.set_option("mode", hq.HQ_MODE_SERIAL)
        c.run(copy_back=False, destroy=False)
        pos = c.final_pos()
        _, zzv = c.state_view(hq.HQ_C128).expect_z()
        out.append(float(zzv.sum()) if pos == list(range(n)) else float("nan"))
        c.use_state(None) if mode == "hybrid" else None
        c.close()
    dt = time.perf_counter() - t0
    for st in states.values():
        st.close()
    return len(specs) / dt, out


def batch_form(a, hq, count, plain_sample, precisions=("c64", "c128")):
    """Batching form: `count` HEA circuits, n log-uniform in [10, 22], depth in
    [4, 40]; output sum_i <Z_i Z_(i+1)> per circuit; circuits/s of the batch (one
    hipGraph replay of every round) in each precision, against the FP64 hybrid
    path one circuit at a time on resident states and the plain per-gate path,
    both on an evenly spaced sample of the same circuits."""
    import numpy as np
    from hyquas_amd.serving import autotune_tiles, build_circuit, hea_batch
    specs = hea_batch(count, a.seed)
    circuits = [build_circuit(hq, "hea", n, d, s) for n, d, s in specs]
    res = {}
    zz = {}
    for name in precisions:
        prec = hq.HQ_C64 if name == "c64" else hq.HQ_C128
        e = hq.Engine(prec, 280 * 10 ** 9)
        if a.streams:
            e.set_option("streams", a.streams)
        if a.batch_tile:
            e.set_option("share_qubits", a.batch_tile)
        t0 = time.perf_counter()
        if a.batch_jit:
            # specialised batched kernels (one per group structure) on the
            # default tiles: the whole state up to 13 qubits, else 12 (13 in complex128)
            e.set_option("compile_threads", 16)
            e.set_option("jit", 1)
            tiles = {}
        else:
            tiles = {} if a.no_autotune else autotune_tiles(hq, e, specs, circuits)
        tune_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        b = _build_batch(hq, e, circuits, name)
        build_s = time.perf_counter() - t0
        for _ in range(max(1, a.warmup)):
            b.run()
        steps = max(1, min(a.steps, 10))
        t0 = time.perf_counter()
        dev = [b.run() for _ in range(steps)]
        wall = (time.perf_counter() - t0) / steps
        ex, _ = b.results()
        zz[name] = ex[:, 2].copy()
        info = b.info()
        dev_ms = float(np.mean(dev))
        # the replay's tile passes against HBM: algorithmic bytes (each pass's
        # live tiles read and written once) over the replay's device time
        gbps = info["pass_bytes"] / (dev_ms * 1e-3) / 1e9
        res[name] = {"kernels": "specialised" if a.batch_jit else "interpreter",
                     "circuits_per_s": round(count / wall, 1), "s_per_batch": round(wall, 6),
                     "device_ms": round(min(dev), 3), "rounds": info["rounds"], "launches": info["launches"],
                     "device_bytes": info["device_bytes"], "build_s": round(build_s, 3),
                     "autotune_s": round(tune_s, 3), "tiles": tiles,
                     "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBS, 4),
                                  "pass_bytes": info["pass_bytes"], "device_ms_mean": round(dev_ms, 3)},
                     "max_norm_error": float(np.max(np.abs(ex[:, 0] - 1)))}
        # the same replay against the VALU issue rate (instruction count from
        # the committed counter profile when this is the profiled batch)
        valu = _batch_valu_from_profiles(name, {"count": count, "seed": a.seed, "batch_jit": a.batch_jit,
                                                "batch_tile": a.batch_tile}, info["launches"])
        if valu:
            rate = valu[0] / (dev_ms * 1e-3)
            res[name]["valu_roofline"] = {"bound": "valu issue", "achieved": round(rate / 1e9, 1),
                                          "peak": round(valu[1] / 1e9, 1), "unit": "G wave-instr/s",
                                          "frac": round(rate / valu[1], 4), "wave_instr_per_replay": valu[0]}
        b.close()
        e.close()
        _progress(f"batch: {name} {res[name]['circuits_per_s']} circuits/s")
    idx = list(range(0, count, max(1, count // max(1, plain_sample))))[:plain_sample]
    sample = [specs[i] for i in idx]
    hyb_cps, hz = _one_at_a_time(hq, sample, "hea", "hybrid")
    _progress(f"batch: hybrid FP64 one at a time {hyb_cps:.1f} circuits/s")
    plain_cps, pz = _one_at_a_time(hq, sample, "hea", "plain")
    _progress(f"batch: plain FP64 one at a time {plain_cps:.1f} circuits/s")
    ref = zz["c128"] if "c128" in zz else zz["c64"]
    out = {"workload": f"{count} hardware-efficient-ansatz circuits (RY, RZ per qubit, CZ ladder), qubits "
                       f"round(exp(U[ln10, ln22])), depth U[4, 40], seed {a.seed}, output sum <Z_i Z_(i+1)>",
           "batch": res,
           "hybrid_fp64_one_at_a_time": {"circuits_per_s": round(hyb_cps, 2),
                                         "sample": f"{len(sample)} evenly spaced circuits"},
           "plain_fp64_one_at_a_time": {"circuits_per_s": round(plain_cps, 2),
                                        "sample": f"{len(sample)} evenly spaced circuits"},
           "max_abs_dzz_hybrid_vs_batch": float(np.nanmax(np.abs(np.array(hz) - ref[idx]))),
           "max_abs_dzz_plain_vs_batch": float(np.max(np.abs(np.array(pz) - ref[idx])))}
    for name in res:
        out[f"speedup_{name}_vs_hybrid"] = round(res[name]["circuits_per_s"] / hyb_cps, 2)
        out[f"speedup_{name}_vs_plain"] = round(res[name]["circuits_per_s"] / plain_cps, 2)
    if "c64" in zz and "c128" in zz:
        out["max_abs_dzz_c64_vs_c128"] = float(np.max(np.abs(zz["c64"] - zz["c128"])))
    return out, specs, idx


def main_batch(a, hq):
    """Batching form alone: value = circuits/s of the complex64 batch."""
    res, specs, idx = batch_form(a, hq, a.count, a.plain_sample)
    fid_idx = sorted(set([max(range(len(specs)), key=lambda i: specs[i][0])] + idx[:4]))
    fids = [_fidelity(hq, specs[i]) for i in fid_idx]
    c64 = res["batch"]["c64"]
    out = {
        "metric": "batched HEA circuits/s (complex64) vs FP64 one-at-a-time",
        "value": c64["circuits_per_s"],
        "unit": "circuits/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(c64["s_per_batch"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": res.pop("workload"), "parallelism": "single GPU"},
        **res,
        "one_minus_fidelity_c64": {"max": max(fids), "circuits": fid_idx},
    }
    print(json.dumps(out), flush=True)


def serve_form(a, hq, rate, requests, seed, plain=True, plain_sample=0):
    """Serving form at one arrival rate: the complex64 engine (continuous
    batching, HBM admission under 280 GB) against the FP64 hybrid path one
    circuit at a time on resident states (and the plain per-gate path, on the
    first `plain_sample` requests of the same stream when that is > 0: the
    plain path runs a few circuits per second)."""
    from hyquas_amd.serving import build_circuit, request_stream, resident_states, serve, serve_hybrid, \
        serve_plain, summarize
    reqs = request_stream(requests, rate, seed)
    e = hq.Engine(hq.HQ_C64, 280 * 10 ** 9)
    if a.streams:
        e.set_option("streams", a.streams)
    if a.batch_tile:
        e.set_option("share_qubits", a.batch_tile)
    circuits = [build_circuit(hq, r["kind"], r["n"], r["depth"], r["seed"]) for r in reqs]
    recs, wall = serve(hq, e, reqs, circuits)
    st = e.stats()
    e.close()
    _progress(f"serve: rate {rate} engine done in {wall:.1f} s")
    states = resident_states(hq, reqs)
    circuits = [build_circuit(hq, r["kind"], r["n"], r["depth"], r["seed"]) for r in reqs]
    hrecs, hwall = serve_hybrid(hq, reqs, circuits, states)
    for s in states.values():
        s.close()
    _progress(f"serve: rate {rate} hybrid FP64 done in {hwall:.1f} s")
    dz = [abs(x["result"]["z"] - y["result"]["z"]) for x, y in zip(recs, hrecs) if not x["error"]]
    topk_same = sum(1 for x, y, r in zip(recs, hrecs, reqs) if r["output"] == "topk" and not x["error"]
                    and x["result"]["topk"][0][:4] == list(y["result"]["topk"][0][:4]))
    out = {"rate": rate, "requests": requests, "engine_c64": summarize(recs, wall),
           "hybrid_fp64": summarize(hrecs, hwall), "peak_bytes": st["bytes_peak"], "rounds": st["rounds"],
           "max_abs_dz_c64_vs_hybrid_fp64": max(dz) if dz else None,
           "topk_top4_agree": f"{topk_same}/{sum(1 for r in reqs if r['output'] == 'topk')}"}
    hc = out["hybrid_fp64"]["circuits_per_s"]
    out["speedup_vs_hybrid_fp64"] = round(out["engine_c64"]["circuits_per_s"] / hc, 2) if hc else None
    if plain:
        preqs = reqs[:plain_sample] if plain_sample > 0 else reqs
        circuits = [build_circuit(hq, r["kind"], r["n"], r["depth"], r["seed"]) for r in preqs]
        precs, pwall = serve_plain(hq, preqs, circuits)
        _progress(f"serve: rate {rate} plain done in {pwall:.1f} s")
        out["plain_fp64"] = summarize(precs, pwall)
        if len(preqs) < len(reqs):
            out["plain_fp64"]["sample"] = f"first {len(preqs)} of the {len(reqs)} requests"
        pc = out["plain_fp64"]["circuits_per_s"]
        out["speedup_vs_plain_fp64"] = round(out["engine_c64"]["circuits_per_s"] / pc, 2) if pc else None
    return out


def main_serve(a, hq):
    """Serving form alone: Poisson stream per rate."""
    rates = [float(x) for x in a.rates.split(",") if x]
    sweep = [serve_form(a, hq, rate, a.requests, a.seed + k, plain=not a.no_plain) for k, rate in enumerate(rates)]
    top = sweep[-1]
    out = {
        "metric": "served circuits/s + p99 vs FP64 one-at-a-time",
        "value": top["engine_c64"]["circuits_per_s"],
        "unit": "circuits/s",
        "n_gpus": 1, "steps": len(rates), "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": SERVE_WORKLOAD.format(requests=a.requests, seed=a.seed), "rates": rates,
                   "parallelism": "single GPU"},
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


SERVE_WORKLOAD = ("Poisson stream, {requests} requests per rate, random-layer or ansatz (even odds), qubits "
                  "round(exp(U[ln12, ln28])), depth U[8, 60], output sum <Z_i> or top-16 (even odds), seed {seed} + k "
                  "for the k-th rate, HBM cap 280 GB")


def _timed_run(hq, c, steps, warmup):
    for _ in range(warmup):
        c.run(copy_back=False, destroy=False)
    t0 = time.perf_counter()
    launches = []
    for _ in range(steps):
        c.run(copy_back=False, destroy=False)
        launches.extend(c.launch_times())
    return (time.perf_counter() - t0) / steps, launches


def main_precision(a, hq):
    """Precision form: random layers in complex64 at 33 qubits (64 GiB) and 34
    qubits (128 GiB), 24 layers; the TransMM path (FP64 fusion rounded to FP32,
    FP32 MFMA) at 30 qubits; max |d amp| and 1 - fidelity of complex64 against a
    complex128 run of the 30-qubit instance of the same circuit."""
    import numpy as np
    layers = 24
    out_runs = {}
    # 33 and 34 qubits, complex64, OShareMem groups
    for n, steps, warm in ((33, a.steps if a.steps < 5 else 3, 1), (34, 1, 0)):
        c = hq.Circuit(n).add_random_layers(layers, a.seed)
        c.set_option("precision", hq.HQ_C64)
        c.set_option("mode", hq.HQ_MODE_GROUP)
        c.set_option("profile_events", 1)
        t0 = time.perf_counter()
        c.compile()
        comp = time.perf_counter() - t0
        _progress(f"precision: {n}q complex64 compiled in {comp:.1f} s")
        sec, launches = _timed_run(hq, c, steps, warm)
        _progress(f"precision: {n}q complex64 {sec:.4f} s per simulation")
        shm = [ms for k, ms in launches if k == hq.KIND_SHAREMEM]
        per_step = len(shm) // steps
        # live share of each pass (first: write only; early passes skip zero tiles)
        live = [g["live_fraction"] for g in c.schedule() if g["backend"] == hq.HQ_BACKEND_PERGATE]
        alg = 2.0 * (1 << n) * 8 * (sum(live) if len(live) == per_step else per_step) / per_step
        avg = sum(shm) / len(shm)
        norm = c.state_view(hq.HQ_C64).norm2()
        out_runs[f"{n}q_c64"] = {"s_per_sim": round(sec, 4), "groups": per_step, "compile_s": round(comp, 2),
                                 "avg_launch_ms": round(avg, 3), "achieved_GBps": round(alg / avg / 1e6, 1),
                                 "hbm_frac": round(alg / avg / 1e6 / HBM_PEAK_GBS, 4), "norm_error": abs(norm - 1),
                                 "state_GiB": (1 << n) * 8 / 2 ** 30}
        c.close()
    # 30 qubits: complex128 reference, complex64 OShareMem, complex64 TransMM (FP32 MFMA)
    n = 30
    ref = hq.Circuit(n).add_random_layers(layers, a.seed)
    ref.set_option("precision", hq.HQ_C128)
    ref.set_option("mode", hq.HQ_MODE_GROUP)
    ref.run(copy_back=False, destroy=False)
    ref_state = ref.state_view(hq.HQ_C128)
    psi_ref = ref_state.download()
    _progress("precision: 30q complex128 reference done")
    errs = {}
    for name, mode, mat in (("c64_group", hq.HQ_MODE_GROUP, 0), ("c64_transmm_k64", hq.HQ_MODE_BLAS, 6),
                            ("c64_transmm_k128", hq.HQ_MODE_BLAS, 7)):
        c = hq.Circuit(n).add_random_layers(layers, a.seed)
        c.set_option("precision", hq.HQ_C64)
        c.set_option("mode", mode)
        c.set_option("profile_events", 1)
        if mat:
            c.set_option("blas_mat", mat)
        c.compile()
        sec, launches = _timed_run(hq, c, 1, 1 if mode == hq.HQ_MODE_GROUP else 0)
        _progress(f"precision: 30q {name} {sec:.4f} s per simulation")
        # squared norm from the device's FP64 (Kahan) reduction
        if mode == hq.HQ_MODE_GROUP:
            ov = ref_state.inner(c.state_view(hq.HQ_C64))
            n64 = c.state_view(hq.HQ_C64).norm2()
            psi = c.state_view(hq.HQ_C64).download()
        else:
            psi = c.logical_state(hq.HQ_C64)
            st = hq.State(n, hq.HQ_C64)
            st.upload(psi)
            ov = ref_state.inner(st)
            n64 = st.norm2()
            st.close()
        d = float(np.max(np.abs(psi.astype(np.complex128) - psi_ref)))
        dense = [ms for k, ms in launches if k == hq.KIND_DENSE]
        tr = [ms for k, ms in launches if k == hq.KIND_TRANSPOSE]
        # raw: |<ref|psi>|^2 with psi as computed; normalized: psi scaled to unit norm first
        rec = {"s_per_sim": round(sec, 4), "max_abs_damp": d, "one_minus_fidelity": 1.0 - abs(ov) ** 2,
               "one_minus_fidelity_normalized": 1.0 - abs(ov) ** 2 / n64, "norm_error": abs(n64 - 1.0)}
        if dense:
            K = 1 << mat
            fl = 8.0 * K * (1 << n)
            rec.update({"dense_launches": len(dense), "dense_avg_ms": round(sum(dense) / len(dense), 3),
                        "dense_TFLOPs": round(fl / (sum(dense) / len(dense)) / 1e9, 2),
                        "mfma_frac_fp32": round(fl / (sum(dense) / len(dense)) / 1e9 / 157.3, 4),
                        "transpose_launches": len(tr),
                        "transpose_avg_ms": round(sum(tr) / len(tr), 3) if tr else None})
        errs[name] = rec
        del psi
        c.close()
    ref.close()
    top = out_runs["33q_c64"]
    out = {
        "metric": "33q complex64 sim time",
        "value": top["s_per_sim"], "unit": "s", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": round(top["s_per_sim"] * 1e3, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"random-layer circuit, {layers} layers, seed {a.seed}, complex64; errors of the "
                               f"30-qubit instance against complex128", "parallelism": "single GPU"},
        "runs": out_runs,
        "errors_vs_c128_30q": errs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
