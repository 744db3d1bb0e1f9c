// executor.cpp -- lowering of schedule groups to device programs.
//
// PerGate (OShareMem) groups: the tile is the group's related qubits topped
// up to the tile size from the lowest physical positions, gates are mapped to
// physical positions through the current layout (src/executor.cpp:462-493,
// 585-632); the device layer then splits the tile into register/thread bits
// and plans the register stages.  BLAS (TransMM) groups: FP64 fusion on the
// host (src/schedule.cpp:578-702) and a transpose + dense program, or one
// in-place dense pass over the group's qubits where they are (the layout
// then stays; Schedule::runState records the layout a run leaves).

#include "executor.h"

#include <algorithm>
#include <cmath>

#include "hq_internal.h"

qindex Executor::fillRelated(qindex related, const State& st, int size)
{
    int cnt = __builtin_popcountll((unsigned long long)related);
    for (int i = 0; i < (int)st.layout.size() && cnt < size; i++) {
        qindex b = qindex(1) << st.layout[i];
        if (!(related & b)) { related |= b; cnt++; }
    }
    return related;
}

// A TransMM group runs in place when its plan kept the layout and the dense
// kernel takes its matrix on this state.
static bool runsInPlace(const GateGroup& gg, const State& st, int n, int precision)
{
    return gg.state.pos == st.pos && n >= hq::dense_at_min_qubits(precision, gg.matQubit);
}

bool Executor::needsBuffer(const Schedule& sch, int n, int precision)
{
    State st(n);
    for (auto& lg : sch.localGroups)
        for (auto& gg : lg.fullGroups) {
            if (gg.backend != Backend::BLAS) continue;
            if (!runsInPlace(gg, st, n, precision)) return true;
            st = gg.state;
        }
    return false;
}

// Tiles smaller than 7 qubits (states below 7 qubits) run the group's gates
// through the plain kernels: the register kernel needs 64 lanes x 2 registers.
static const int kMinTile = 7;

hq_status Executor::prepare(Schedule& sch, int n, hq_state s, const ExecParams& p)
{
    State st(n);
    int nq = n, precision = HQ_C128;
    hq_status irc = hq_state_info(s, &nq, &precision);
    if (irc != HQ_OK) return irc;
    std::vector<hq::SharememRequest> reqs;
    std::vector<GateGroup*> owners;
    bool first = true;
    sch.initFused = false;
    // a run starts from |0...0>: until a non-diagonal gate has targeted a
    // qubit (diagonal gates and controls leave a |0> qubit at |0>), every
    // amplitude with that qubit set is zero, so tiles whose block address
    // sets such a qubit are zero before the group and stay zero.  `support`
    // holds the physical qubits some earlier non-diagonal gate (or an
    // in-place TransMM group) touched; it is tracked until the first TransMM
    // group in the transpose form (its permutation moves qubits).
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    uint64_t support = 0;
    bool track = true;
    for (auto& lg : sch.localGroups) {
        for (auto& gg : lg.fullGroups) {
            const bool isFirst = first;
            first = false;
            if (gg.program) { hq_program_destroy(gg.program); gg.program = nullptr; }
            if (gg.backend == Backend::PerGate) {
                const int k = std::min(p.localSize, n);
                if (__builtin_popcountll((unsigned long long)gg.relatedQubits) > k) {
                    hq::set_error("executor: group has more related qubits than the tile");
                    return HQ_ERR_ARG;
                }
                gg.state = st;
                gg.numStages = 0;   // tiles below kMinTile: plain launches at run time
                if (k >= kMinTile) {
                    qindex rel = fillRelated(gg.relatedQubits, st, k);
                    hq::SharememRequest r;
                    for (int q = 0; q < n; q++)
                        if ((rel >> q) & 1) r.mask |= 1ull << st.pos[q];
                    r.gates.reserve(gg.gates.size());
                    for (const Gate& g : gg.gates) {
                        hq_gate h = g.toHQ();
                        h.target = st.pos[g.targetQubit];
                        h.control = g.controlQubit >= 0 ? st.pos[g.controlQubit] : -1;
                        h.control2 = g.controlQubit2 >= 0 ? st.pos[g.controlQubit2] : -1;
                        r.gates.push_back(h);
                    }
                    // the run starts from |0...0>: the first pass builds it in registers
                    r.init = isFirst;
                    sch.initFused = sch.initFused || isFirst;
                    // share of a full read + write of the state the pass moves:
                    // the first pass only writes (its non-origin tiles store zeros)
                    gg.liveFraction = isFirst ? 0.5 : 1.0;
                    if (track && !isFirst) {
                        r.zero_mask = all & ~support;
                        const int dead = __builtin_popcountll((unsigned long long)(r.zero_mask & ~r.mask));
                        gg.liveFraction = std::ldexp(1.0, -dead);
                    }
                    for (const Gate& g : gg.gates)
                        if (!g.isDiagonal()) support |= 1ull << st.pos[g.targetQubit];
                    reqs.push_back(std::move(r));
                    owners.push_back(&gg);
                } else {
                    track = false;
                }
            } else if (gg.backend == Backend::BLAS) {
                gg.liveFraction = 1.0;
                gg.initCPUMatrix();
                const size_t K = (size_t)1 << gg.matQubit;
                std::vector<double> m(2 * K * K);
                for (size_t i = 0; i < K * K; i++) { m[2 * i] = gg.matrix[i].real(); m[2 * i + 1] = gg.matrix[i].imag(); }
                const std::vector<int> ml = gg.matLogical();
                bool low = true;   // matrix bit j already at position j
                for (int j = 0; j < gg.matQubit; j++) low = low && st.pos[ml[j]] == j;
                if (runsInPlace(gg, st, n, precision)) {
                    // one pass where the qubits are; the layout stays
                    std::vector<int> positions(gg.matQubit);
                    for (int j = 0; j < gg.matQubit; j++) {
                        positions[j] = st.pos[ml[j]];
                        support |= 1ull << positions[j];   // U mixes amplitudes across its qubits
                    }
                    hq_status rc = hq_program_create_dense(s, positions.data(), m.data(), gg.matQubit, &gg.program);
                    if (rc != HQ_OK) return rc;
                } else if (gg.state.pos == st.pos && !low) {
                    hq::set_error("executor: a TransMM group planned in place is too large for the dense kernel on "
                                  "this state (schedule loaded for another precision?)");
                    return HQ_ERR_UNSUPPORTED;
                } else {
                    // the reference's transpose (identity when the qubits are
                    // already lowest) + GEMM on runs of K amplitudes
                    hq_status rc =
                        hq_program_create_blas(s, gg.cuttPerm.data(), m.data(), gg.matQubit, &gg.program);
                    if (rc != HQ_OK) return rc;
                    st = gg.state;
                    track = false;   // the permutation moves the qubits the tracking is about
                }
            }
        }
    }
    sch.runState = st;
    const bool jit = p.jit == 1 || (p.jit < 0 && n >= p.jitMinQubits);
    std::vector<hq_program> progs;
    const int tune = !jit ? 0 : p.tune >= 0 ? p.tune : (n >= p.tuneMinQubits ? 1 : 0);
    hq_status rc = hq::build_sharemem_programs(s, reqs, p.regQubits, jit, &progs, tune);
    if (rc != HQ_OK) return rc;
    for (size_t i = 0; i < owners.size(); i++) {
        owners[i]->program = progs[i];
        hq_program_info(progs[i], nullptr, &owners[i]->numStages, nullptr);
    }
    return HQ_OK;
}

hq_status Executor::run(Schedule& sch, hq_state s)
{
    for (auto& lg : sch.localGroups) {
        for (auto& gg : lg.fullGroups) {
            hq_status rc;
            if (gg.program) {
                rc = hq_program_launch(s, gg.program);
            } else {
                // tiny state: same group, plain kernels, physical positions
                std::vector<hq_gate> hg;
                for (const Gate& g : gg.gates) {
                    hq_gate h = g.toHQ();
                    h.target = gg.state.pos[g.targetQubit];
                    h.control = g.controlQubit >= 0 ? gg.state.pos[g.controlQubit] : -1;
                    h.control2 = g.controlQubit2 >= 0 ? gg.state.pos[g.controlQubit2] : -1;
                    hg.push_back(h);
                }
                rc = hq_apply_gates_plain(s, hg.data(), (int)hg.size());
            }
            if (rc != HQ_OK) return rc;
        }
    }
    return HQ_OK;
}
