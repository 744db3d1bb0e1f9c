"""GPU parity of the full Circuit path (partition -> lower -> run) against the
FP64 oracle: amplitudes within 1e-10 relative (2-norm) for complex128,
1 - fidelity <= 1e-5 for complex64."""
import numpy as np
import pytest

import oracle
from helpers import (blocks_and_layers, flat_cost_tables, inverse, random_layers, random_mixed, rel_err,
                     to_oracle)
from test_oracle import MIX_CASES, MIX_FLAT_COST, MIX_SPEC

pytestmark = pytest.mark.gpu


def _circuit(hq, n, spec):
    c = hq.Circuit(n)
    for s in spec:
        t, tgt, ctl = s[0], s[1], s[2]
        params = s[3] if len(s) > 3 else ()
        c2 = s[4] if len(s) > 4 else -1
        c.addGate(hq.Gate.make(t, c2, ctl, tgt, params))
    return c


def _logical_state(c, n):
    return np.array([c.ampAt(i) for i in range(1 << n)])


@pytest.mark.parametrize("mode", ["serial", "group", "blas", "mix"])
@pytest.mark.parametrize("n", [5, 9, 14])
def test_modes_match_oracle(gpu, n, mode):
    spec = random_mixed(n, 160, seed=n * 7)
    c = _circuit(gpu, n, spec)
    c.set_option("mode", dict(serial=0, group=1, blas=3, mix=4)[mode])
    if mode == "blas":
        c.set_option("blas_mat", min(6, n))
    c.compile()
    c.run()
    got = _logical_state(c, n)
    want = oracle.run_plain(to_oracle(spec), n)
    assert rel_err(got, want) <= 1e-10


@pytest.mark.parametrize("n,jit,tune,inplace", [(20, 1, 0, 0), (20, 1, 1, 0), (20, 0, 0, 0), (22, 1, 0, 0),
                                                (20, 1, 0, 1), (20, 0, 0, 1), (22, 1, 1, 1)])
def test_mixed_schedule_interleaves_transmm_and_sharemem(gpu, tmp_path, n, jit, tune, inplace):
    """A schedule whose groups alternate between OShareMem and TransMM, chosen
    per group by cost per gate (src/compiler.cpp:248-279) under tables that
    make a dense group pay off where gates crowd onto six qubits: OShareMem
    groups lowered through the layouts the TransMM permutations leave (or,
    with TransMM in place, through the unchanged layout), zero-tile tracking
    stopped at the first TransMM group, the logical readout after the
    sequence; complex128 against the per-gate oracle at 1e-10."""
    spec = blocks_and_layers(n)
    f = tmp_path / "cost.out"
    f.write_text(flat_cost_tables(0.03))
    c = _circuit(gpu, n, spec)
    c.set_option("mode", gpu.HQ_MODE_MIX)
    c.set_option("jit", jit)
    c.set_option("tune", tune)
    c.set_option("transmm_inplace", inplace)
    c.load_cost_model(str(f))
    c.compile()
    sched = c.schedule()
    kinds = "".join("B" if g["backend"] == gpu.HQ_BACKEND_BLAS else "P" for g in sched)
    print(n, jit, tune, inplace, kinds, [round(g["live_fraction"], 4) for g in sched], flush=True)
    assert kinds.count("B") >= 2 and "BP" in kinds and "PB" in kinds
    first_b = kinds.index("B")
    if inplace:
        # no transpose moved a qubit; zero tiles are still tracked past the
        # TransMM groups (their qubits count as touched)
        assert c.final_pos() == list(range(n))
        assert all(0 < g["live_fraction"] <= 1.0 for g in sched)
    else:
        assert all(g["live_fraction"] == 1.0 for g in sched[first_b + 1:])
    if first_b > 0:
        assert sched[0]["live_fraction"] == 0.5
    assert sum(g["num_gates"] for g in sched) == len(spec)
    want = oracle.run_plain(to_oracle(spec), n)
    c.run(copy_back=False, destroy=False)
    assert rel_err(c.logical_state(), want) <= 1e-10
    # a second run starts over from |0...0> (the layout of the first run's end
    # does not leak into it)
    c.run(copy_back=False, destroy=False)
    assert rel_err(c.logical_state(), want) <= 1e-10


@pytest.mark.parametrize("case", sorted(MIX_CASES))
def test_hand_traced_mix_schedules_run(gpu, tmp_path, case):
    """The hand-traced MIX-mode partitions of the oracle's tests, run on the
    GPU ("profile" = 1: the reference's transposes): a TransMM group, then
    OShareMem groups on the permuted layout."""
    n = 12
    c = _circuit(gpu, n, MIX_SPEC)
    c.set_option("profile", 1)
    c.set_option("mode", gpu.HQ_MODE_MIX)
    if MIX_CASES[case]["blas"] is not None:
        f = tmp_path / "12qubits.out"
        blas = MIX_CASES[case]["blas"]
        rows = ["1"] + [str(MIX_FLAT_COST)] * 14 + [""] + [str(MIX_FLAT_COST)] * 7 + [""]
        rows += [f"{1 << i} {blas.get(i, 1.0)}" for i in range(10)] + ["", "0"]
        f.write_text("\n".join(rows) + "\n")
        c.load_cost_model(str(f))
    c.compile()
    assert [g["backend"] for g in c.schedule()] == [g[0] for g in MIX_CASES[case]["groups"]]
    c.run(copy_back=False, destroy=False)
    assert c.final_pos() == MIX_CASES[case]["perm"]
    want = oracle.run_plain(to_oracle(MIX_SPEC), n)
    assert rel_err(c.logical_state(), want) <= 1e-12


def test_random_layers_20q_full_compare(gpu):
    """20 qubits, 20 random layers, seed 0: all 2^20 amplitudes (hybrid mode)."""
    n = 20
    c = gpu.Circuit(n).add_random_layers(20, 0)
    spec = random_layers(n, 20, 0)
    assert len(c.gates()) == len(spec) == 600
    c.compile()
    c.run(copy_back=True, destroy=False)
    got = c.logical_state()
    want = oracle.run_plain(to_oracle(spec), n)
    assert rel_err(got, want) <= 1e-10
    assert float(np.max(np.abs(got - want))) <= 1e-12
    assert sum(g["num_gates"] for g in c.schedule()) == 600
    # ampAt agrees with the device readout
    for i in (0, 1, 12345, (1 << n) - 1):
        assert c.ampAt(i) == got[i]


def test_random_layers_serial_plain_20q(gpu):
    n = 20
    c = gpu.Circuit(n).add_random_layers(20, 0)
    c.set_option("mode", gpu.HQ_MODE_SERIAL)
    c.compile()
    c.run(copy_back=False, destroy=False)
    want = oracle.run_plain(to_oracle(random_layers(n, 20, 0)), n)
    assert rel_err(c.logical_state(), want) <= 1e-12


@pytest.mark.parametrize("jit", [0, 1])
def test_random_layers_jit_and_interpreter(gpu, jit):
    n = 17
    c = gpu.Circuit(n).add_random_layers(10, 5)
    c.set_option("jit", jit)
    c.compile()
    c.run(copy_back=False, destroy=False)
    want = oracle.run_plain(to_oracle(random_layers(n, 10, 5)), n)
    assert rel_err(c.logical_state(), want) <= 1e-10


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("mode", [1, 2])
def test_tuned_kernels_equal_plain(gpu, prec, mode):
    # mode 1: the tuned schedule mixes plain and whole-line non-temporal
    # kernels; mode 2: whole-line kernels everywhere.  Both kernels apply the
    # same operations in the same order (the compiler may still contract
    # multiply-adds differently in the two: rounding level)
    n = 22
    states = []
    for tune in (0, mode):
        c = gpu.Circuit(n).add_random_layers(12, 9)
        c.set_option("jit", 1)
        c.set_option("precision", prec)
        c.set_option("tune", tune)
        c.compile()
        c.run(copy_back=False, destroy=False)
        states.append(c.logical_state(prec))
        c.close()
    assert rel_err(states[1], states[0]) <= (1e-14 if prec == 0 else 1e-6)
    if prec == 0:
        want = oracle.run_plain(to_oracle(random_layers(n, 12, 9)), n)
        assert rel_err(states[1], want) <= 1e-10


@pytest.mark.parametrize("jit", [0, 1])
def test_zero_tiles_skipped(gpu, jit):
    # early passes skip the tiles that are still zero (the run starts from
    # |0...0>); the state must equal the per-gate oracle's
    n = 22
    c = gpu.Circuit(n).add_random_layers(10, 21)
    c.set_option("jit", jit)
    c.set_option("mode", gpu.HQ_MODE_GROUP)
    c.compile()
    live = [g["live_fraction"] for g in c.schedule()]
    assert live[0] == 0.5 and live[-1] == 1.0   # the first pass only writes
    assert 0 < live[1] < 1.0
    assert all(0 < f <= 1.0 for f in live)
    c.run(copy_back=False, destroy=False)
    want = oracle.run_plain(to_oracle(random_layers(n, 10, 21)), n)
    assert rel_err(c.logical_state(), want) <= 1e-10
    # a second run over the state the first one left gives the same result
    c.run(copy_back=False, destroy=False)
    assert rel_err(c.logical_state(), want) <= 1e-10


def test_random_layers_full_vector(gpu):
    n = 18
    c = gpu.Circuit(n).add_random_layers(12, 3)
    c.compile()
    c.run(copy_back=False, destroy=False)
    want = oracle.run_plain(to_oracle(random_layers(n, 12, 3)), n)
    assert rel_err(c.logical_state(), want) <= 1e-10


def test_print_state_matches_oracle_format(gpu):
    n = 10
    spec = random_mixed(n, 80, seed=3)
    c = _circuit(gpu, n, spec)
    c.set_option("mode", 1)
    c.compile()
    c.run()
    want_psi = oracle.run_plain(to_oracle(spec), n)
    text = c.stateText()
    # same lines as the oracle's printState of the oracle state, within print precision
    ref = oracle.print_state(want_psi, n, list(range(n)))
    a, b = text.splitlines(), ref.splitlines()
    assert len(a) == len(b)
    for x, y in zip(a, b):
        ix, rx = x.split(" ", 1)
        iy, ry = y.split(" ", 1)
        assert ix == iy
        vx = [float(v.rstrip(":")) for v in rx.split()]
        vy = [float(v.rstrip(":")) for v in ry.split()]
        assert np.allclose(vx, vy, atol=2e-12)


def test_c64_fidelity(gpu):
    n = 16
    c = gpu.Circuit(n).add_random_layers(10, 11)
    c.set_option("precision", gpu.HQ_C64)
    c.compile()
    c.run(copy_back=False, destroy=False)
    got = c.logical_state(gpu.HQ_C64)
    want = oracle.run_plain(to_oracle(random_layers(n, 10, 11)), n)
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    assert 1 - fid <= 1e-5


@pytest.mark.parametrize("jit", [0, 1])
def test_c64_14_qubit_tiles(gpu, jit):
    # complex64 tiles of 14 qubits: 256 threads x 64 amplitudes (two-phase
    # exchange in the specialised kernel, one 128 KiB exchange in the interpreter)
    n = 18
    c = gpu.Circuit(n).add_random_layers(12, 13)
    c.set_option("precision", gpu.HQ_C64)
    c.set_option("mode", gpu.HQ_MODE_GROUP)
    c.set_option("share_qubits", 14)
    c.set_option("jit", jit)
    c.set_option("tune", jit)
    c.compile()
    assert all(g["backend"] != gpu.HQ_BACKEND_PERGATE or g["num_stages"] > 0 for g in c.schedule())
    c.run(copy_back=False, destroy=False)
    got = c.logical_state(gpu.HQ_C64)
    want = oracle.run_plain(to_oracle(random_layers(n, 12, 13)), n)
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    assert 1 - fid <= 1e-5
    assert np.max(np.abs(got - want)) <= 1e-5


@pytest.mark.parametrize("n,layers,prec", [(30, 30, "c128"), (33, 24, "c64"), (34, 24, "c64")])
def test_round_trip_at_full_size(gpu, n, layers, prec):
    """The headline circuits at their full sizes, followed by their inverse,
    must return |0...0>: a size-independent check of the whole path (beam
    partition, specialised kernels, scalar factoring, fused first group)."""
    spec = random_layers(n, layers, 0)
    c = _circuit(gpu, n, spec + inverse(spec))
    p = gpu.HQ_C128 if prec == "c128" else gpu.HQ_C64
    c.set_option("precision", p)
    c.run(copy_back=False, destroy=False)
    st = c.state_view(p)
    norm = st.norm2()
    pos = c.final_pos()
    assert sorted(pos) == list(range(n))
    a0 = st.get_amp(0)
    fid = abs(a0) ** 2
    if prec == "c128":
        assert abs(norm - 1) < 1e-12
        assert 1 - fid < 1e-12
    else:
        assert abs(norm - 1) < 2e-5
        assert 1 - fid / norm < 1e-6
    c.close()


@pytest.mark.parametrize("n,mode", [(16, "mix"), (14, "blas"), (24, "mix")])
def test_schedule_cache_round_trip(gpu, tmp_path, n, mode):
    """A saved schedule loaded into a fresh circuit runs to the same amplitudes
    (bit for bit) as the circuit that partitioned it."""
    spec = random_layers(n, 8, 21)
    m = dict(mix=4, blas=3)[mode]
    a = _circuit(gpu, n, spec)
    a.set_option("mode", m)
    a.compile()
    a.run(copy_back=False, destroy=False)
    want = a.logical_state()
    path = str(tmp_path / "s.bin")
    a.save_schedule(path)
    b = _circuit(gpu, n, spec)
    b.set_option("mode", m)
    b.load_schedule(path)
    b.run(copy_back=False, destroy=False)
    got = b.logical_state()
    assert np.array_equal(got, want)
    assert b.schedule_order() == a.schedule_order()
    ref = oracle.run_plain(to_oracle(spec), n)
    assert rel_err(got, ref) <= 1e-10


def test_known_answers_at_full_size(gpu):
    """Circuit families of the reference's benchmark inputs (bv, qft, ghz),
    read from qasm text in the reference dialect and run at full size with
    the default (hybrid, specialised-kernel) path; each has a closed-form
    answer independent of the partition."""
    import math
    from helpers import bv_spec, ghz_spec, qft_spec, to_qasm
    hq = gpu
    # Bernstein-Vazirani, 28 qubits: the data register holds the secret
    n, secret = 28, 0b101101110010110101001011011
    c = hq.Circuit.from_qasm(to_qasm(n, bv_spec(n, secret)))
    c.run(copy_back=False, destroy=False)
    st = c.state_view()
    pos = c.final_pos()
    hit = 0
    for anc in (0, 1):
        logical = secret | (anc << (n - 1))
        phys = sum(((logical >> q) & 1) << pos[q] for q in range(n))
        hit += abs(st.get_amp(phys)) ** 2
    assert abs(hit - 1) < 1e-12
    c.close()
    # GHZ, 30 qubits
    n = 30
    c = hq.Circuit.from_qasm(to_qasm(n, ghz_spec(n)))
    c.run(copy_back=False, destroy=False)
    st = c.state_view()
    pos = c.final_pos()
    ones = sum(1 << pos[q] for q in range(n))
    assert abs(st.get_amp(0) - 2 ** -0.5) < 1e-14 and abs(st.get_amp(ones) - 2 ** -0.5) < 1e-14
    assert abs(st.norm2() - 1) < 1e-13
    c.close()
    # QFT of a basis state, 26 qubits: amplitude k (bit-reversed) = exp(2 pi i x k / N) / sqrt(N)
    n, x = 26, 0b10110011100011110000101101
    N = 1 << n
    c = hq.Circuit.from_qasm(to_qasm(n, qft_spec(n, x)))
    c.run(copy_back=False, destroy=False)
    st = c.state_view()
    pos = c.final_pos()
    rng = np.random.default_rng(0)
    for k in rng.integers(0, N, 40).tolist() + [0, 1, N - 1]:
        rev = int(format(k, f"0{n}b")[::-1], 2)
        phys = sum(((rev >> q) & 1) << pos[q] for q in range(n))
        ph = 2 * math.pi * ((x * k) % N) / N   # exact reduction before the trigonometry
        want = complex(math.cos(ph), math.sin(ph)) / math.sqrt(N)
        assert abs(st.get_amp(phys) - want) < 1e-12 / math.sqrt(N) * 100
    c.close()


def test_circuits_share_a_resident_state(gpu):
    """hq_circuit_use_state: circuits run one after another on one
    caller-owned state (with a work buffer: mixed schedules), which outlives
    run(destroy=1) and close(); a state of the wrong size is refused."""
    n = 20
    st = gpu.State(n, gpu.HQ_C128, with_buffer=True)
    cost = None
    for k, spec in enumerate((random_layers(n, 6, 41), blocks_and_layers(n))):
        c = _circuit(gpu, n, spec)
        c.set_option("mode", gpu.HQ_MODE_MIX)
        if k == 1:
            import tempfile
            cost = tempfile.NamedTemporaryFile("w", suffix=".out", delete=False)
            cost.write(flat_cost_tables(0.03))
            cost.close()
            c.load_cost_model(cost.name)
        c.use_state(st)
        c.run(copy_back=False, destroy=True)
        assert c.device_state().value == st.handle.value
        want = oracle.run_plain(to_oracle(spec), n)
        assert rel_err(c.logical_state(), want) <= 1e-10
        c.close()
        assert abs(st.norm2() - 1) < 1e-12   # still allocated after the circuit is gone
    small = gpu.State(n - 1)
    c = _circuit(gpu, n, random_layers(n, 1, 1))
    with pytest.raises(gpu.HQError):
        c.use_state(small)
    c.close()
    small.close()
    st.close()


def test_in_place_transmm_needs_no_work_buffer(gpu, tmp_path):
    """A mixed schedule whose TransMM groups run in place runs on a borrowed
    state without a work buffer (against the oracle); the same circuit in the
    reference's transpose form asks for the buffer and refuses that state."""
    n = 20
    spec = blocks_and_layers(n)
    f = tmp_path / "cost.out"
    f.write_text(flat_cost_tables(0.03))
    want = oracle.run_plain(to_oracle(spec), n)
    st = gpu.State(n, gpu.HQ_C128)            # no work buffer
    c = _circuit(gpu, n, spec)
    c.set_option("mode", gpu.HQ_MODE_MIX)
    c.load_cost_model(str(f))
    c.use_state(st)
    c.compile()
    assert any(g["backend"] == gpu.HQ_BACKEND_BLAS for g in c.schedule())
    c.run(copy_back=False, destroy=False)
    assert rel_err(c.logical_state(), want) <= 1e-10
    c.close()
    c = _circuit(gpu, n, spec)
    c.set_option("mode", gpu.HQ_MODE_MIX)
    c.set_option("transmm_inplace", 0)
    c.load_cost_model(str(f))
    c.use_state(st)
    with pytest.raises(gpu.HQError, match="work buffer"):
        c.compile()
    c.close()
    st.close()


_JIT_CHILD = r"""
import sys
sys.path.insert(0, sys.argv[1])
import hyquas_amd as hq
c = hq.Circuit(22).add_random_layers(4, 7)
c.set_option("jit", 1)
c.set_option("tune", 0)
c.compile()
c.run(copy_back=False, destroy=False)
idx, prob = c.state_view().topk(4)
print("OK", c.state_view().norm2(), " ".join(str(int(i)) for i in idx), " ".join(repr(float(p)) for p in prob))
"""


@pytest.mark.parametrize("helper", ["none", "/nonexistent/hq_rtc"])
def test_specialised_kernels_without_compile_helper(gpu, tmp_path, helper):
    """Specialised kernels still build when the compile helper is switched off
    (HQ_RTC_HELPER=none: hiprtc in the process) or cannot start (a missing
    path: one warning, then in-process compiles), with a cold kernel cache;
    the results equal the helper-built kernels' bit for bit."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = {}
    for name, env_helper in (("helper", None), ("alt", helper)):
        # cold caches: ours and the compiler library's own
        env = dict(os.environ, HQ_JIT_CACHE=str(tmp_path / name), AMD_COMGR_CACHE="0")
        env.pop("HQ_RTC_HELPER", None)
        if env_helper is not None:
            env["HQ_RTC_HELPER"] = env_helper
        r = subprocess.run([sys.executable, "-c", _JIT_CHILD, root], env=env, capture_output=True, text=True,
                           timeout=240)
        print(name, r.returncode, r.stdout[-300:], r.stderr[-300:], flush=True)
        assert r.returncode == 0
        line = [l for l in r.stdout.splitlines() if l.startswith("OK")][-1]
        outs[name] = line
        if name == "alt" and helper != "none":
            assert "compile helper" in r.stderr and "in-process" in r.stderr
    assert outs["helper"] == outs["alt"]
    assert abs(float(outs["helper"].split()[1]) - 1) < 1e-12
