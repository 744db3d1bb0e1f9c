"""Host-side logic of the product against the oracle, CPU only: gate
factories, workload generators, the partitioner (reference parameters give
the reference grouping; gfx950 parameters keep the invariants), the qasm
reader and the error paths of the C ABI."""
import numpy as np
import pytest

import oracle
from helpers import (CCX, CNOT, H, RX, hea_layers, random_layers, random_mixed, rel_err, to_oracle)
from test_oracle import (MIX_CASES, MIX_FLAT_COST, MIX_SPEC, NPAR, QASM, _check_partition_invariants,
                         _execute_partition)

ALL_TYPES = list(range(22)) + [23, 24, 25, 26, 27]
CTRL = {0: 2, 1: 1, 2: 1, 3: 1, 4: 1, 5: 1, 6: 1, 7: 1}
PARAMS = dict(NPAR)
PARAMS.update({26: 2, 27: 2})


@pytest.mark.parametrize("t", [x for x in ALL_TYPES if x != 22])
def test_gate_factories_match_oracle(hq, t):
    rs = np.random.default_rng(t)
    p = tuple(rs.uniform(-3, 3, PARAMS.get(t, 0)))
    nc = CTRL.get(t, 0)
    c1, c2 = (1 if nc >= 1 else -1), (2 if nc == 2 else -1)
    g = hq.Gate.make(t, c2, c1, 0, p)
    o = oracle.gate(t, 0, c1, c2, p)
    assert list(g.rec.mat) == list(o["mat"])
    assert (g.rec.type, g.rec.target, g.rec.control, g.rec.control2) == (t, 0, c1, c2)


def test_gate_factory_errors(hq):
    with pytest.raises(hq.HQError):
        hq.Gate.make(RX, -1, -1, 0, ())
    with pytest.raises(hq.HQError):
        hq.Gate.make(CNOT, -1, -1, 0, ())
    with pytest.raises(hq.HQError):
        hq.Gate.make(22, -1, -1, 0, ())


def _spec_of(c):
    return [(r.type, r.target, r.control, r.control2, tuple(r.mat)) for r in c.gates()]


@pytest.mark.parametrize("n,layers,seed", [(5, 3, 0), (20, 20, 0), (30, 2, 7), (11, 4, 123)])
def test_random_layer_generator(hq, n, layers, seed):
    c = hq.Circuit(n).add_random_layers(layers, seed)
    want = to_oracle(random_layers(n, layers, seed))
    got = c.gate_array()
    assert len(got) == len(want) == layers * (n + n // 2)
    for a, b in zip(got, want):
        assert (a["type"], a["target"], a["control"]) == (b["type"], b["target"], b["control"])
        assert np.array_equal(a["mat"], b["mat"])


def test_hea_generator(hq):
    n, layers = 13, 5
    got = hq.Circuit(n).add_hea_layers(layers, 9).gate_array()
    want = to_oracle(hea_layers(n, layers, 9))
    assert len(got) == len(want) == (3 * n - 1) * layers
    for a, b in zip(got, want):
        assert (a["type"], a["target"], a["control"]) == (b["type"], b["target"], b["control"])
        assert np.array_equal(a["mat"], b["mat"])


def _product_partition(hq, n, spec, mode, reference=True, **opts):
    c = hq.Circuit(n)
    for s in spec:
        c.addGate(hq.Gate.make(s[0], s[4] if len(s) > 4 else -1, s[2], s[1], s[3] if len(s) > 3 else ()))
    if reference:
        c.set_option("profile", 1)
    c.set_option("mode", mode)
    for k, v in opts.items():
        c.set_option(k, v)
    c.partition()
    order = c.schedule_order()
    groups, k = [], 0
    for g in c.schedule():
        groups.append(dict(backend=g["backend"], gates=order[k:k + g["num_gates"]], related=g["related"],
                           mat_qubit=g["mat_qubit"]))
        k += g["num_gates"]
    return groups, c.final_pos()


@pytest.mark.parametrize("mode", [1, 3, 4])
@pytest.mark.parametrize("n,seed,kind", [(12, 0, "layers"), (16, 1, "mixed"), (20, 2, "layers"), (28, 3, "mixed"),
                                         (28, 4, "layers"), (9, 5, "mixed")])
def test_partition_matches_reference_restatement(hq, mode, n, seed, kind):
    spec = random_layers(n, 8, seed) if kind == "layers" else random_mixed(n, 300, seed, ccx=False)
    og, ofpos, _ = oracle.partition(to_oracle(spec), n, oracle.partition_params(n, mode))
    pg, pfpos = _product_partition(hq, n, spec, mode)
    assert [(g["backend"], g["gates"], g["related"], g["mat_qubit"]) for g in og] == \
           [(g["backend"], g["gates"], g["related"], g["mat_qubit"]) for g in pg]
    assert ofpos == pfpos


def test_partition_hand_traced_cases_in_product(hq):
    n = 12
    spec = [(H, q, -1, ()) for q in range(n)] + [(CNOT, 11, 0, ())]
    pg, _ = _product_partition(hq, n, spec, 1)
    assert [g["gates"] for g in pg] == [[0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 12], [9, 10]]
    n = 8
    spec = [(H, q, -1, ()) for q in range(n)] + [(CNOT, 1, 0, ())]
    pg, fpos = _product_partition(hq, n, spec, 3)
    assert [g["gates"] for g in pg] == [[0, 1, 2, 3, 4, 5, 8], [6, 7]]
    assert fpos == [6, 7, 2, 3, 4, 5, 0, 1]


@pytest.mark.parametrize("share", [10, 12, 13])
def test_gfx950_partition_invariants(hq, share):
    n = 24
    spec = random_mixed(n, 400, 11, ccx=True)
    pg, _ = _product_partition(hq, n, spec, 4, reference=False, share_qubits=share)
    _check_partition_invariants(spec, n, pg, share, 7)


@pytest.mark.parametrize("n,share,kind", [(14, 8, "layers"), (16, 9, "mixed"), (18, 10, "hea")])
def test_beam_partition_is_valid_and_not_larger(hq, n, share, kind):
    """Beam search over OShareMem-only partitions: every gate once, per-qubit
    order kept, tiles within the limit, the same state as the circuit on the
    oracle, and never more groups than the greedy rule."""
    spec = (random_layers(n, 10, 3) if kind == "layers" else
            random_mixed(n, 300, 4, ccx=True) if kind == "mixed" else hea_layers(n, 6, 5))
    greedy, _ = _product_partition(hq, n, spec, 1, reference=False, share_qubits=share, beam=1, starts=1)
    beam, _ = _product_partition(hq, n, spec, 1, reference=False, share_qubits=share, beam=8, starts=4)
    _check_partition_invariants(spec, n, beam, share, 7)
    assert len(beam) <= len(greedy)
    got, pos = _execute_partition(spec, n, [dict(g, perm=None) for g in beam])
    assert pos == list(range(n))
    want = oracle.run_plain(to_oracle(spec), n)
    assert rel_err(got, want) <= 1e-12


def test_beam_search_reduces_30q_groups(hq):
    c = hq.Circuit(30).add_random_layers(30, 0)
    c.partition()
    beam = len(c.schedule())
    c.set_option("beam", 1)
    c.set_option("starts", 1)
    c.partition()
    assert beam < len(c.schedule())


V100_SINGLES = [225, 469, 469, 352, 350, 350, 194, 209, 209, 216, 216, 370, 367, 369]   # U1 .. RZ
V100_CTRLS = [195, 193, 137, 224, 225, 225, 224]                                      # CNOT .. CRZ


def test_cost_model_reference_format(hq, tmp_path):
    """The reference's <n>qubits.out format (src/evaluator.cpp:60-103) holding
    its own built-in V100 numbers gives the same partition as the built-in
    tables."""
    rows = ["1"] + [str(v) for v in V100_SINGLES] + [""] + [str(v) for v in V100_CTRLS] + [""]
    rows += [f"{1 << i} {23.068396 if i == 6 else 0.0}" for i in range(10)] + ["", "11.367814"]
    f = tmp_path / "28qubits.out"
    f.write_text("\n".join(rows) + "\n")
    spec = random_mixed(28, 300, 3, ccx=False)
    want, wpos = _product_partition(hq, 28, spec, 4)
    c = hq.Circuit(28)
    for s in spec:
        c.addGate(hq.Gate.make(s[0], -1, s[2], s[1], s[3] if len(s) > 3 else ()))
    c.set_option("profile", 1)
    c.set_option("mode", 4)
    c.load_cost_model(str(f))
    c.partition()
    got = [(g["backend"], g["num_gates"], g["related"]) for g in c.schedule()]
    assert got == [(g["backend"], len(g["gates"]), g["related"]) for g in want]
    assert c.final_pos() == wpos
    # a BLAS-heavy file moves the choice to OShareMem everywhere
    rows2 = list(rows)
    rows2[-1] = "1000000"
    f.write_text("\n".join(rows2) + "\n")
    c.load_cost_model(str(f))
    c.partition()
    assert all(g["backend"] == hq.HQ_BACKEND_PERGATE for g in c.schedule())


def _reference_format(singles, ctrls, blas, cutt):
    rows = ["1"] + [str(v) for v in singles] + [""] + [str(v) for v in ctrls] + [""]
    rows += [f"{1 << i} {blas.get(i, 1.0)}" for i in range(10)] + ["", str(cutt)]
    return "\n".join(rows) + "\n"


@pytest.mark.parametrize("case", sorted(MIX_CASES))
def test_partition_hand_traced_mix_mode_in_product(hq, tmp_path, case):
    """The hand-traced MIX-mode partitions of the oracle's tests (backend
    chosen per group by cost per gate, src/compiler.cpp:248-279): the product
    with the reference's rules and the same tables, given as a reference-format
    parameter file (or the built-in V100 tables), gives the same groups,
    backends, matrix sizes and layout."""
    want = MIX_CASES[case]
    c = hq.Circuit(12)
    for s in MIX_SPEC:
        c.addGate(hq.Gate.make(s[0], -1, s[2], s[1], s[3]))
    c.set_option("profile", 1)
    c.set_option("mode", hq.HQ_MODE_MIX)
    if want["blas"] is not None:
        f = tmp_path / "12qubits.out"
        f.write_text(_reference_format([MIX_FLAT_COST] * 14, [MIX_FLAT_COST] * 7, want["blas"], 0.0))
        c.load_cost_model(str(f))
    c.partition()
    order = c.schedule_order()
    got, k = [], 0
    for g in c.schedule():
        got.append((g["backend"], order[k:k + g["num_gates"]], g["related"],
                    g["mat_qubit"] if g["backend"] == hq.HQ_BACKEND_BLAS else 0))
        k += g["num_gates"]
    assert got == want["groups"]
    assert c.final_pos() == want["perm"]


def test_cost_model_key_value_format(hq, tmp_path):
    f = tmp_path / "m.params"
    kv = ["# comment", "valu_rate 3.8e13", "valu_rate_f32 3.8e13"] + \
         [f"shm_bw_run_{L} {1e12 * (1 + L)}" for L in range(13)] + [f"shm_bw_run_f32_{L} 4e12" for L in range(13)]
    f.write_text("\n".join(kv) + "\n")
    c = hq.Circuit(24).add_random_layers(10, 1)
    c.load_cost_model(str(f))
    c.partition()
    assert len(c.schedule()) > 0
    f.write_text("nonsense_key 1\n")
    with pytest.raises(hq.HQError):
        c.load_cost_model(str(f))
    f.write_text("7\n1 2\n")
    with pytest.raises(hq.HQError):
        c.load_cost_model(str(f))


def test_ccx_second_control_is_a_dependency(hq):
    """H(2) before CCX(0, 2 -> 1) must stay before it (the reference's
    getGroupOpt ignores controlQubit2 and may reorder them)."""
    n = 14
    spec = [(RX, 2, -1, (0.3,))] * 1 + [(CCX, 1, 0, (), 2), (H, 2, -1, ())]
    spec = [(H, q, -1, ()) for q in range(4, n)] + spec
    pg, _ = _product_partition(hq, n, spec, 1)
    order = [i for g in pg for i in g["gates"]]
    assert order.index(len(spec) - 3) < order.index(len(spec) - 2) < order.index(len(spec) - 1)


def test_qasm_reader_matches_oracle(hq):
    c = hq.Circuit.from_qasm(QASM)
    n, want = oracle.parse_qasm(QASM)
    assert c.numQubits == n
    got = c.gate_array()
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert (a["type"], a["target"], a["control"], a["control2"]) == \
               (b["type"], b["target"], b["control"], b["control2"])
        assert np.allclose(a["mat"], b["mat"], atol=1e-15)


@pytest.mark.parametrize("bad", ["qreg q[2];\nmeasure q[0];\n", "qreg q[2];\nrx q[0];\n", "h q[0];\n",
                                 "qreg q[2];\ncx q[0];\n", "qreg q[2];\nu3(0.1,0.2 q[0];\n", "qreg q[2];\nh q[5];\n"])
def test_qasm_reader_rejects(hq, bad):
    with pytest.raises(hq.HQError) as e:
        hq.Circuit.from_qasm(bad)
    assert e.value.status == hq.HQ_ERR_PARSE


def test_options_validation(hq):
    c = hq.Circuit(10)
    for k, v in [("mode", 2), ("precision", 5), ("share_qubits", 20), ("nonsense", 1)]:
        with pytest.raises(hq.HQError):
            c.set_option(k, v)


def test_no_device_error_path(hq):
    if hq.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(hq.HQError) as e:
        hq.State(10)
    assert e.value.status == hq.HQ_ERR_NO_DEVICE
    c = hq.Circuit(10).add_random_layers(2, 0)
    with pytest.raises(hq.HQError):
        c.compile()


def _parse_schedule(b):
    """Independent reader of the schedule byte layout (reference
    Schedule/LocalGroup/GateGroup/State/Gate::serialize, packed, little-endian)."""
    import struct
    cur = [0]

    def take(fmt):
        v = struct.unpack_from("<" + fmt, b, cur[0])
        cur[0] += struct.calcsize("<" + fmt)
        return v

    def state():
        (n,) = take("Q")
        return list(take(f"{n}i")), list(take(f"{n}i"))

    def gate():
        gid, typ = take("ii")
        mat = take("8d")
        (nl,) = take("Q")
        name = b[cur[0]:cur[0] + nl].decode()
        assert b[cur[0] + nl] == 0
        cur[0] += nl + 1
        t, c1, c2 = take("iii")
        return dict(id=gid, type=typ, mat=mat, name=name, target=t, control=c1, control2=c2)

    def group():
        ng, rel, nperm, matq, backend = take("QqQii")
        st = state()
        gs = [gate() for _ in range(ng)]
        perm = list(take(f"{nperm}i"))
        return dict(related=rel, mat_qubit=matq, backend=backend, state=st, gates=gs, perm=perm)

    assert b[:8] == b"HQSCHED1"
    nq, zero = struct.unpack_from("<II", b, 8)
    cur[0] = 16
    (nlg,) = take("Q")
    final = state()
    lgs = []
    for _ in range(nlg):
        a2a, na, np_, no, nf, rel = take("iQQQQq")
        take(f"{na}i"), take(f"{np_}i")
        st = state()
        assert no == 0
        lgs.append(dict(related=rel, state=st, groups=[group() for _ in range(nf)]))
    assert cur[0] == len(b)
    return nq, final, lgs


def test_schedule_bytes_follow_the_reference_layout(hq, tmp_path):
    n = 10
    c = hq.Circuit(n).add_random_layers(3, 5)
    c.addGate(hq.Gate.CCX(0, 1, 2))
    c.set_option("mode", 3)          # TransMM groups: permutations and layouts in the bytes
    c.set_option("blas_mat", 4)
    c.partition()
    path = str(tmp_path / "s.bin")
    c.save_schedule(path)
    nq, final, lgs = _parse_schedule(open(path, "rb").read())
    assert nq == n and len(lgs) == 1
    sched = c.schedule()
    groups = lgs[0]["groups"]
    assert [g["backend"] for g in groups] == [s["backend"] for s in sched]
    assert [g["related"] for g in groups] == [s["related"] for s in sched]
    assert [g["mat_qubit"] for g in groups] == [s["mat_qubit"] for s in sched]
    assert [gt["id"] for g in groups for gt in g["gates"]] == c.schedule_order()
    assert final[0] == c.final_pos()
    arr = c.gate_array()
    for g in groups:
        assert sorted(g["perm"]) == (list(range(n)) if g["perm"] else [])
        for gt in g["gates"]:
            ref = arr[gt["id"]]
            assert (gt["type"], gt["target"], gt["control"], gt["control2"]) == \
                   (ref["type"], ref["target"], ref["control"], ref["control2"])
            assert list(gt["mat"]) == list(ref["mat"])
            assert gt["name"]


def test_schedule_load_validates(hq, tmp_path):
    n = 12
    c = hq.Circuit(n).add_random_layers(4, 1)
    c.partition()
    path = str(tmp_path / "s.bin")
    c.save_schedule(path)
    other = hq.Circuit(n).add_random_layers(4, 2)          # different angles and qubits
    with pytest.raises(hq.HQError) as e:
        other.load_schedule(path)
    assert e.value.status == hq.HQ_ERR_ARG
    wrong_n = hq.Circuit(n + 1).add_random_layers(4, 1)
    with pytest.raises(hq.HQError):
        wrong_n.load_schedule(path)
    data = open(path, "rb").read()
    bad = tmp_path / "bad.bin"
    bad.write_bytes(data[:len(data) // 2])
    same = hq.Circuit(n).add_random_layers(4, 1)
    with pytest.raises(hq.HQError) as e:
        same.load_schedule(str(bad))
    assert e.value.status == hq.HQ_ERR_PARSE
    if hq.device_count() == 0:
        with pytest.raises(hq.HQError) as e:
            same.load_schedule(path)      # valid; building programs needs the GPU
        assert e.value.status == hq.HQ_ERR_NO_DEVICE


def test_qasm_writer_round_trips_through_the_reader(hq):
    from helpers import qft_spec, to_qasm
    n = 12
    spec = random_layers(n, 4, 2) + qft_spec(n, 0b101101) + [(CCX, 3, 1, (), 2)]
    c = hq.Circuit.from_qasm(to_qasm(n, spec))
    got = c.gate_array()
    want = to_oracle(spec)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert (a["type"], a["target"], a["control"], a["control2"]) == \
               (b["type"], b["target"], b["control"], b["control2"])
        assert np.array_equal(a["mat"], b["mat"])


@pytest.mark.parametrize("n,prec", [(20, 0), (22, 1)])
def test_transmm_in_place_planning(hq, tmp_path, n, prec):
    """Option "transmm_inplace": with it (the default) the partitioner plans
    TransMM groups that keep the layout (the final layout is the identity),
    without it the reference's transposes move the group's qubits to the
    lowest positions; both partitions hold every gate once and interleave
    TransMM and OShareMem groups under tables that favour dense groups."""
    from helpers import blocks_and_layers, flat_cost_tables
    spec = blocks_and_layers(n)
    f = tmp_path / "cost.out"
    f.write_text(flat_cost_tables(0.03))
    out = {}
    for inplace in (1, 0):
        c = hq.Circuit(n)
        for s in spec:
            c.addGate(hq.Gate.make(s[0], s[4] if len(s) > 4 else -1, s[2], s[1], s[3] if len(s) > 3 else ()))
        c.set_option("mode", hq.HQ_MODE_MIX)
        c.set_option("precision", prec)
        c.set_option("transmm_inplace", inplace)
        c.load_cost_model(str(f))
        c.partition()
        sched = c.schedule()
        kinds = "".join("B" if g["backend"] == hq.HQ_BACKEND_BLAS else "P" for g in sched)
        assert kinds.count("B") >= 2 and "BP" in kinds
        assert sum(g["num_gates"] for g in sched) == len(spec)
        assert sorted(c.schedule_order()) == list(range(len(spec)))
        out[inplace] = c.final_pos()
    assert out[1] == list(range(n))
    assert out[0] != list(range(n)) and sorted(out[0]) == list(range(n))
