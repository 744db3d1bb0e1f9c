"""The calibration tool (hyquas_amd/calibrate.py, the step of
evaluator-preprocess/process.cpp): its two writers on the CPU, loaded back
by the partitioner; the whole tool at a small size on the GPU."""
import json
import os

import pytest

import hyquas_amd.calibrate as cal
from helpers import blocks_and_layers
from test_oracle import MIX_CASES, MIX_FLAT_COST, MIX_SPEC


def _partition(hq, n, spec, cost_file, **opts):
    c = hq.Circuit(n)
    for s in spec:
        c.addGate(hq.Gate.make(s[0], s[4] if len(s) > 4 else -1, s[2], s[1], s[3]))
    for k, v in opts.items():
        c.set_option(k, v)
    c.set_option("mode", hq.HQ_MODE_MIX)
    c.load_cost_model(str(cost_file))
    c.partition()
    order = c.schedule_order()
    out, k = [], 0
    for g in c.schedule():
        out.append((g["backend"], order[k:k + g["num_gates"]], g["related"],
                    g["mat_qubit"] if g["backend"] == hq.HQ_BACKEND_BLAS else 0))
        k += g["num_gates"]
    return out, c.final_pos()


def test_reference_table_writer_round_trip(hq, tmp_path):
    """Rows in the order src/evaluator.cpp:60-103 reads them: a table written
    with the hand-traced MIX case's numbers (BLAS_perf[4] = 0.005 ms,
    BLAS_perf[5] = 0.0055 ms) reproduces that case's partition."""
    dense = {1 << i: 1.0 for i in range(1, 8)}
    dense[16], dense[32] = 0.005, 0.0055
    text = cal.reference_table_text([MIX_FLAT_COST] * 14, [MIX_FLAT_COST] * 7, dense, 0.0)
    rows = [r for r in text.splitlines() if r.strip()]
    assert rows[0] == "1" and len(rows) == 1 + 14 + 7 + 10 + 1
    assert [int(r.split()[0]) for r in rows[22:32]] == [1 << i for i in range(10)]
    assert float(rows[22].split()[1]) == 1.0 and float(rows[31].split()[1]) == 4.0   # K = 1, 512
    f = tmp_path / "12qubits.out"
    f.write_text(text)
    got, pos = _partition(hq, 12, MIX_SPEC, f, profile=1)
    assert got == MIX_CASES["blas5"]["groups"] and pos == MIX_CASES["blas5"]["perm"]


def test_params_writer_round_trip(hq, tmp_path):
    """The model file loads, and its numbers reach the partitioner: a slow VALU
    rate makes dense TransMM groups pay off on a circuit whose gates crowd onto
    six qubits in places; the calibrated defaults keep OShareMem everywhere."""
    bw = [5e12] * 13
    spec = blocks_and_layers(20)
    slow = tmp_path / "slow.params"
    slow.write_text(cal.params_text(20, 5.4e12, 1e12, 1e12, 3.9e13, 1.1e14, 2.6e-6, bw, bw))
    fast = tmp_path / "fast.params"
    fast.write_text(cal.params_text(20, 5.4e12, 3.8e13, 3.8e13, 3.9e13, 1.1e14, 2.6e-6, bw, bw))
    g_slow, _ = _partition(hq, 20, spec, slow, jit=1)
    g_fast, _ = _partition(hq, 20, spec, fast, jit=1)
    assert any(g[0] == hq.HQ_BACKEND_BLAS for g in g_slow)
    assert all(g[0] == hq.HQ_BACKEND_PERGATE for g in g_fast)
    lines = [l for l in slow.read_text().splitlines() if not l.startswith("#")]
    assert len(lines) == 6 + 26


@pytest.mark.gpu
def test_calibrate_small(gpu, tmp_path):
    """The whole tool at 16 qubits (reference tables at 14): both files load
    and the measured numbers are positive."""
    out = tmp_path / "cal"
    cal.main(["--qubits", "16", "--out", str(out), "--tables", "14", "--samples", "1"])
    params = out / "mi355x.params"
    table = out / "14qubits.out"
    vals = {l.split()[0]: float(l.split()[1]) for l in params.read_text().splitlines() if not l.startswith("#")}
    assert all(v > 0 for v in vals.values()) and len(vals) == 32
    rows = [r for r in table.read_text().splitlines() if r.strip()]
    assert len(rows) == 33 and all(float(x) > 0 for r in rows[1:] for x in r.split())
    for path, n in ((params, 16), (table, 14)):
        c = gpu.Circuit(n).add_random_layers(4, 1)
        c.load_cost_model(str(path))
        c.compile()
        c.run(copy_back=False, destroy=False)
        assert abs(c.state_view().norm2() - 1) < 1e-12
        c.close()
