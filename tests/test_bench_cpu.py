"""CPU checks of bench.py's lookups of committed counter profiles: a figure
measured on one workload is reported only for exactly that workload."""
import importlib.util
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_lookup_matches_only_the_profiled_workload():
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "sharemem_traffic_30q.json")) as f:
        d = json.load(f)
    cfg, alg = d["config"], d["algorithmic_bytes_per_launch"]
    assert b._traffic_from_profiles(dict(cfg), alg) == d["bytes_per_launch"]
    assert b._traffic_from_profiles(dict(cfg, precision="c64"), alg) is None
    assert b._traffic_from_profiles(dict(cfg, mode="blas"), alg) is None
    # another schedule of the same circuit moves other bytes per launch
    assert b._traffic_from_profiles(dict(cfg), alg * 1.05) is None


def test_batch_valu_lookup_matches_only_the_profiled_batch():
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "batch_valu_hea4096.json")) as f:
        d = json.load(f)
    cfg = d["config"]
    for name in ("c64", "c128"):
        rec = d[name]
        got = b._batch_valu_from_profiles(name, dict(cfg), rec["launches_per_replay"])
        assert got == (rec["valu_wave_instr_per_replay"], d["peak_wave_instr_per_s"])
        assert b._batch_valu_from_profiles(name, dict(cfg), rec["launches_per_replay"] + 1) is None
        assert b._batch_valu_from_profiles(name, dict(cfg, batch_tile=13), rec["launches_per_replay"]) is None
        assert b._batch_valu_from_profiles(name, dict(cfg, batch_jit=0), rec["launches_per_replay"]) is None
        assert b._batch_valu_from_profiles(name, dict(cfg, count=1024), rec["launches_per_replay"]) is None
    # the nominal issue rate: one wave64 VALU instruction per 4 cycles per
    # SIMD, 4 SIMDs per CU, 256 CUs at 2.4 GHz
    assert d["peak_wave_instr_per_s"] == 256 * 4 * 2.4e9 / 4
