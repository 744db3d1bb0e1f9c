// schedule.cpp -- State/GateGroup/Schedule (src/schedule.cpp, single GPU).

#include "schedule.h"

#include "hq_internal.h"

#include <algorithm>
#include <cstdio>
#include <utility>

std::string to_string(Backend b)
{
    switch (b) {
    case Backend::None: return "None";
    case Backend::PerGate: return "PerGate";
    case Backend::BLAS: return "BLAS";
    }
    return "?";
}

State::State(int n)
{
    for (int i = 0; i < n; i++) { pos.push_back(i); layout.push_back(i); }
}

GateGroup::GateGroup(GateGroup&& o) noexcept { *this = std::move(o); }

GateGroup& GateGroup::operator=(GateGroup&& o) noexcept
{
    if (this != &o) {
        if (program) hq_program_destroy(program);
        gates = std::move(o.gates);
        gateIndex = std::move(o.gateIndex);
        relatedQubits = o.relatedQubits;
        state = std::move(o.state);
        cuttPerm = std::move(o.cuttPerm);
        matQubit = o.matQubit;
        backend = o.backend;
        matrix = std::move(o.matrix);
        program = o.program;
        numStages = o.numStages;
        o.program = nullptr;
    }
    return *this;
}

GateGroup::~GateGroup()
{
    if (program) hq_program_destroy(program);
}

qindex GateGroup::newRelated(qindex rel, const Gate& g, qindex localQubits, bool enableGlobal)
{
    // src/schedule.cpp:37-51: with enableGlobal (OShareMem) only non-diagonal
    // targets need to be in the tile; without it (TransMM) every touched
    // local qubit is part of the matrix
    if (enableGlobal) {
        if (!g.isDiagonal()) rel |= qindex(1) << g.targetQubit;
    } else {
        if (!g.isDiagonal() || ((localQubits >> g.targetQubit) & 1)) rel |= qindex(1) << g.targetQubit;
        if (g.controlQubit != -1 && ((localQubits >> g.controlQubit) & 1)) rel |= qindex(1) << g.controlQubit;
        if (g.controlQubit2 != -1 && ((localQubits >> g.controlQubit2) & 1)) rel |= qindex(1) << g.controlQubit2;
    }
    return rel;
}

void GateGroup::addGate(const Gate& g, int index, qindex localQubits, bool enableGlobal)
{
    gates.push_back(g);
    gateIndex.push_back(index);
    relatedQubits = newRelated(relatedQubits, g, localQubits, enableGlobal);
}

State GateGroup::initPerGateState(const State& old)
{
    state = old;
    return state;
}

State GateGroup::initBlasState(const State& old, int numLocalQubits, int minMatSize, int inPlacePrecision)
{
    if (inPlacePrecision >= 0) {
        const int nm = __builtin_popcountll((unsigned long long)relatedQubits);
        const int mq = std::min(std::max(nm, minMatSize), numLocalQubits);
        if (numLocalQubits >= hq::dense_at_min_qubits(inPlacePrecision, mq)) {
            matQubit = mq;
            cuttPerm.resize(numLocalQubits);
            for (int i = 0; i < numLocalQubits; i++) cuttPerm[i] = i;
            state = old;
            return state;
        }
    }
    // src/schedule.cpp:339-382: swap the group's qubits that sit above the
    // first |group| positions with the non-group qubits below it, pairwise in
    // position order; cuttPerm[newPos] = oldPos
    std::vector<int> pos = old.pos, layout = old.layout;
    std::vector<int> toGlobal, toLocal;
    int nm = __builtin_popcountll((unsigned long long)relatedQubits);
    for (int i = 0; i < nm; i++) {
        int q = layout[i];
        if (!((relatedQubits >> q) & 1)) toGlobal.push_back(q);
    }
    for (int i = nm; i < numLocalQubits; i++) {
        int q = layout[i];
        if ((relatedQubits >> q) & 1) toLocal.push_back(q);
    }
    cuttPerm.resize(numLocalQubits);
    for (int i = 0; i < numLocalQubits; i++) cuttPerm[i] = i;
    for (size_t i = 0; i < toGlobal.size() && i < toLocal.size(); i++) {
        int x = toGlobal[i], y = toLocal[i];
        std::swap(cuttPerm[pos[x]], cuttPerm[pos[y]]);
        std::swap(pos[x], pos[y]);
        layout[pos[x]] = x;
        layout[pos[y]] = y;
    }
    matQubit = std::min(std::max(nm, minMatSize), numLocalQubits);
    state = State(pos, layout);
    return state;
}

State GateGroup::initState(const State& old, int numLocalQubits, int minMatSize, int inPlacePrecision)
{
    if (backend == Backend::BLAS) return initBlasState(old, numLocalQubits, minMatSize, inPlacePrecision);
    return initPerGateState(old);
}

std::vector<int> GateGroup::matLogical() const
{
    const int n = (int)state.pos.size();
    std::vector<std::pair<int, int>> sel;   // (position, logical qubit)
    for (int q = 0; q < n; q++)
        if ((relatedQubits >> q) & 1) sel.push_back({state.pos[q], q});
    for (int p = 0; p < n && (int)sel.size() < matQubit; p++) {
        const int q = state.layout[p];
        if (!((relatedQubits >> q) & 1)) sel.push_back({p, q});
    }
    std::sort(sel.begin(), sel.end());
    std::vector<int> out;
    for (auto& x : sel) out.push_back(x.second);
    return out;
}

void GateGroup::initCPUMatrix()
{
    // Column i of U is the group applied to basis vector e_i; stored at
    // matrix[i*K .. i*K+K-1] (column-major), as src/schedule.cpp:578-702.
    const int K = 1 << matQubit;
    matrix.assign((size_t)K * K, std::complex<double>(0, 0));
    for (int i = 0; i < K; i++) matrix[(size_t)i * K + i] = 1;
    struct Op {
        int t, c1, c2;
        std::complex<double> m00, m01, m10, m11;
    };
    std::vector<Op> ops;
    // matrix bit of each logical qubit (the position after the reference's
    // transpose; the rank among the group's qubits when it runs in place)
    const std::vector<int> ml = matLogical();
    std::vector<int> bit(state.pos.size(), -1);
    for (size_t j = 0; j < ml.size(); j++) bit[ml[j]] = (int)j;
    for (auto& g : gates) {
        Op o;
        o.t = bit[g.targetQubit];
        o.c1 = g.controlQubit >= 0 ? bit[g.controlQubit] : -1;
        o.c2 = g.controlQubit2 >= 0 ? bit[g.controlQubit2] : -1;
        o.m00 = {g.mat[0][0].x, g.mat[0][0].y};
        o.m01 = {g.mat[0][1].x, g.mat[0][1].y};
        o.m10 = {g.mat[1][0].x, g.mat[1][0].y};
        o.m11 = {g.mat[1][1].x, g.mat[1][1].y};
        ops.push_back(o);
    }
#pragma omp parallel for schedule(static)
    for (int col = 0; col < K; col++) {
        std::complex<double>* v = matrix.data() + (size_t)col * K;
        for (const Op& o : ops) {
            const int half = K >> 1;
            const int mask = (1 << o.t) - 1;
            for (int i = 0; i < half; i++) {
                int lo = ((i >> o.t) << (o.t + 1)) | (i & mask);
                if (o.c1 >= 0 && !((lo >> o.c1) & 1)) continue;
                if (o.c2 >= 0 && !((lo >> o.c2) & 1)) continue;
                int hi = lo | (1 << o.t);
                std::complex<double> a = v[lo], b = v[hi];
                v[lo] = o.m00 * a + o.m01 * b;
                v[hi] = o.m10 * a + o.m11 * b;
            }
        }
    }
}

void Schedule::clear()
{
    localGroups.clear();
    finalState = State();
    runState = State();
    initFused = false;
}

void Schedule::dump(int numQubits) const
{
    const int L = 3;
    for (const auto& lg : localGroups) {
        for (int i = 0; i < numQubits * L; i++) putchar('-');
        printf("\n");
        for (const auto& gg : lg.fullGroups) {
            printf("<%s>\n", to_string(gg.backend).c_str());
            for (const Gate& g : gg.gates) {
                for (int i = 0; i < numQubits; i++) {
                    if (i == g.controlQubit || i == g.controlQubit2) {
                        printf(".");
                        for (int j = 1; j < L; j++) printf(" ");
                    } else if (i == g.targetQubit) {
                        printf("%s", g.name.c_str());
                        for (int j = (int)g.name.length(); j < L; j++) printf(" ");
                    } else {
                        putchar(gg.contains(i) ? '+' : '|');
                        for (int j = 1; j < L; j++) printf(" ");
                    }
                }
                printf("\n");
            }
            printf("\n");
        }
        for (int i = 0; i < numQubits * L; i++) putchar('#');
        printf("\n\n");
    }
    fflush(stdout);
}
