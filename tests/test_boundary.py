"""The boundary end to end: the CLI in the reference harness's shape
(main.cpp:233-251 -> circuit.cpp:250-309, diffed as scripts/check_wrapper.sh
does), and a driver written against the reference's C++ API compiled with
this library's host headers.  Compiling runs on the CPU; running needs the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from helpers import CCX, CNOT, CU1, H, RY, TDG, U3, to_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "hyquas_amd", "csrc", "host")
PKG = os.path.join(ROOT, "hyquas_amd")
QASM_C1 = os.path.join(ROOT, "tests", "golden", "c1_random_layers_20q_seed0.qasm")


def _build_driver(out):
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
           os.path.join(ROOT, "tests", "cpp", "ref_style_driver.cpp"), "-o", str(out), "-L" + PKG, "-lhyquas_amd",
           "-Wl,-rpath," + PKG]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_reference_style_driver_compiles(hq, tmp_path):
    exe = tmp_path / "driver"
    _build_driver(exe)
    assert exe.exists()


def driver_spec(n):
    """The circuit tests/cpp/ref_style_driver.cpp builds."""
    import math
    spec = [(H, i, -1, ()) for i in range(n)] + [(CNOT, i + 1, i, ()) for i in range(n - 1)]
    spec += [(RY, i, -1, (0.1 * (i + 1),)) for i in range(n)]
    spec += [(CCX, i + 2, i, (), i + 1) for i in range(0, n - 2, 2)]
    spec += [(CU1, n - 1, 0, (math.pi / 3,)), (U3, 1, -1, (0.3, 0.2, 0.1)), (TDG, 2, -1, ())]
    return spec


def _state_lines(text):
    """Lines of printState output: Logger/CLUSTER lines dropped (the filter of
    scripts/check_wrapper.sh:18)."""
    return [l for l in text.splitlines() if l.strip() and not re.search("Logger|CLUSTER", l)]


def _compare_text(got_lines, want_lines):
    """Same lines in the same order; values within print precision
    (scripts/compare.py:1-32 compares the printed numbers); returns the share
    of byte-identical lines."""
    assert len(got_lines) == len(want_lines)
    same = 0
    for g, w in zip(got_lines, want_lines):
        gi, gv = g.split(" ", 1)
        wi, wv = w.split(" ", 1)
        assert gi == wi
        a = [float(x.rstrip(":")) for x in gv.split()]
        b = [float(x.rstrip(":")) for x in wv.split()]
        assert np.max(np.abs(np.array(a) - np.array(b))) <= 2e-12, (g, w)
        same += g == w
    return same / len(want_lines)


@pytest.mark.gpu
@pytest.mark.parametrize("args", [[], ["--mode", "group"], ["--mode", "blas"], ["--reference"]])
def test_cli_prints_the_oracle_state(gpu, args):
    """hyquas_main on the committed 20-qubit qasm against the oracle's
    printState text of the same circuit (the check_wrapper.sh diff)."""
    exe = os.path.join(PKG, "hyquas_main")
    print("cli:", args, flush=True)
    out = subprocess.run([exe] + args + [QASM_C1], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    n, g = oracle.parse_qasm(open(QASM_C1).read())
    psi = oracle.run_plain(g, n)
    want = oracle.print_state(psi, n, list(range(n))).splitlines()
    share = _compare_text(_state_lines(out.stdout), want)
    assert share >= 0.95


@pytest.mark.gpu
def test_reference_style_driver_runs(gpu, tmp_path):
    exe = tmp_path / "driver"
    _build_driver(exe)
    n = 12
    out = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    psi = oracle.run_plain(to_oracle(driver_spec(n)), n)
    want = oracle.print_state(psi, n, list(range(n))).splitlines()
    assert _compare_text(_state_lines(out.stdout), want) >= 0.95
    m = re.search(r"CN CCX amp\[3\] = ([^\s,]+) ([^\s,]+)", out.stderr)
    assert m, out.stderr[-500:]
    assert abs(complex(float(m.group(1)), float(m.group(2))) - psi[3]) <= 1e-12
