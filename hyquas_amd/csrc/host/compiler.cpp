// compiler.cpp -- the hybrid partitioner.
//
// Restates OneLayerCompiler::getGroupOpt and AdvanceCompiler::run
// (src/compiler.cpp:220-452) for a single GPU, with every limit a runtime
// parameter (the reference fixes LOCAL_QUBIT_SIZE = 10 for its 16 KiB tile;
// this build's OShareMem tile is sized for gfx950's registers + 160 KiB LDS).
// Deviation: CCX's second control is tracked as a dependency (the reference
// ignores controlQubit2 in both sweeps, which can reorder gates on it).

#include "compiler.h"

#include <algorithm>
#include <cstring>

namespace {

int popc(qindex x) { return __builtin_popcountll((unsigned long long)x); }

struct Bits {
    std::vector<uint64_t> w;
    explicit Bits(int nbits = 0) : w((nbits + 63) / 64, 0) {}
    void set(int i) { w[i >> 6] |= 1ull << (i & 63); }
    bool test(int i) const { return (w[i >> 6] >> (i & 63)) & 1; }
    int count() const
    {
        int c = 0;
        for (auto x : w) c += __builtin_popcountll(x);
        return c;
    }
    bool any() const
    {
        for (auto x : w) if (x) return true;
        return false;
    }
    void orWith(const Bits& o) { for (size_t i = 0; i < w.size(); i++) w[i] |= o.w[i]; }
    void andNot(const Bits& o) { for (size_t i = 0; i < w.size(); i++) w[i] &= ~o.w[i]; }
};

}  // namespace

GroupFinder::GroupFinder(int numQubits, const std::vector<Gate>& gates, int window)
    : n_(numQubits), window_(window), gates_(gates)
{
    remain_.resize(gates.size());
    for (size_t i = 0; i < gates.size(); i++) remain_[i] = (int)i;
}

void GroupFinder::remove(const std::vector<int>& idx)
{
    // idx ascending, a subset of remain_
    std::vector<int> out;
    out.reserve(remain_.size());
    size_t s = 0;
    for (int r : remain_) {
        if (s < idx.size() && idx[s] == r) { s++; continue; }
        out.push_back(r);
    }
    remain_.swap(out);
}

std::vector<int> GroupFinder::find(std::vector<char>& full, std::vector<qindex>& related, bool enableGlobal,
                                   int localSize, qindex localQubits) const
{
    const int n = n_;
    const int gate_num = std::min<int>((int)remain_.size(), window_);
    std::vector<Bits> cur(n, Bits(gate_num));
    Bits selected(gate_num);

    // 1. per-qubit chains of gates that can join the group
    for (int id = 0; id < gate_num; id++) {
        if (id % 100 == 0) {
            bool live = false;
            for (int i = 0; i < n; i++) if (!full[i]) live = true;
            if (!live) break;
        }
        const Gate& g = gates_[remain_[id]];
        if (g.isControlGate()) {
            const int c = g.controlQubit, t = g.targetQubit, c2 = g.controlQubit2;
            if (!full[c] && !full[t] && (c2 < 0 || !full[c2])) {
                qindex nr = related[c] | related[t] | (c2 >= 0 ? related[c2] : 0);
                nr = GateGroup::newRelated(nr, g, localQubits, enableGlobal);
                if (popc(nr) <= localSize) {
                    Bits nc = cur[c];
                    nc.orWith(cur[t]);
                    if (c2 >= 0) nc.orWith(cur[c2]);
                    nc.set(id);
                    cur[c] = nc;
                    cur[t] = nc;
                    if (c2 >= 0) cur[c2] = nc;
                    related[c] = related[t] = nr;
                    if (c2 >= 0) related[c2] = nr;
                    continue;
                }
            }
            full[c] = full[t] = 1;
            if (c2 >= 0) full[c2] = 1;
        } else {
            if (!full[g.targetQubit]) {
                cur[g.targetQubit].set(id);
                related[g.targetQubit] = GateGroup::newRelated(related[g.targetQubit], g, localQubits, enableGlobal);
            }
        }
    }

    // 2. greedily merge the longest chains that fit
    std::vector<char> blocked(n, 0);
    qindex selRel = 0;
    while (true) {
        int mx = 0, best = -1;
        for (int i = 0; i < n; i++) {
            int cnt = cur[i].count();
            if (!blocked[i] && cnt > mx) {
                if (popc(selRel | related[i]) <= localSize) { mx = cnt; best = i; }
                else blocked[i] = 1;
            }
        }
        if (mx == 0) break;
        selected.orWith(cur[best]);
        selRel |= related[best];
        blocked[best] = 1;
        for (int i = 0; i < n; i++) {
            if (!blocked[i] && cur[i].any()) {
                if ((related[i] | selRel) == selRel) {
                    selected.orWith(cur[i]);
                    blocked[i] = 1;
                } else {
                    cur[i].andNot(cur[best]);
                }
            }
        }
    }

    // 3. diagonal gates that commute into the group (OShareMem only)
    if (enableGlobal) {
        std::fill(blocked.begin(), blocked.end(), 0);
        int cnt = 0;
        for (int id = 0; id < gate_num; id++) {
            cnt++;
            if (cnt % 100 == 0) {
                bool live = false;
                for (int i = 0; i < n; i++) if (!blocked[i]) live = true;
                if (!live) break;
            }
            if (selected.test(id)) continue;
            const Gate& g = gates_[remain_[id]];
            if (g.isDiagonal()) {
                if (g.isControlGate()) {
                    if (!blocked[g.controlQubit] && !blocked[g.targetQubit]) selected.set(id);
                    else blocked[g.controlQubit] = blocked[g.targetQubit] = 1;
                } else {
                    if (!blocked[g.targetQubit]) selected.set(id);
                }
            } else {
                if (g.isControlGate()) blocked[g.controlQubit] = 1;
                if (g.isC2Gate()) blocked[g.controlQubit2] = 1;
                blocked[g.targetQubit] = 1;
            }
        }
    }

    std::vector<int> ret;
    for (int id = 0; id < gate_num; id++)
        if (selected.test(id)) ret.push_back(remain_[id]);
    return ret;
}

Schedule compileSchedule(int n, const std::vector<Gate>& gates, const CompilerParams& p)
{
    Schedule sch;
    LocalGroup lg;
    State state(n);
    const qindex all = n >= 63 ? ~qindex(0) : ((qindex(1) << n) - 1);
    GroupFinder finder(n, gates, p.window);
    const bool usePerGate = p.mode == HQ_MODE_GROUP || p.mode == HQ_MODE_MIX;
    const bool useBLAS = p.mode == HQ_MODE_BLAS || p.mode == HQ_MODE_MIX;
    const int coalesce = std::min(p.coalesce, n);
    const int localSize = std::min(p.localSize, n);

    std::vector<char> full(n);
    std::vector<qindex> related(n);
    while (!finder.empty()) {
        std::vector<int> ggIdx;
        Backend backend = Backend::PerGate;
        qindex cacheRelated = 0;
        double bestEff = 1e10;
        if (usePerGate) {
            std::fill(full.begin(), full.end(), 0);
            qindex co = 0;
            for (int j = 0; j < coalesce; j++) co |= qindex(1) << state.layout[j];
            std::fill(related.begin(), related.end(), co);
            cacheRelated = co;
            ggIdx = finder.find(full, related, true, localSize, all);
            backend = Backend::PerGate;
            if (useBLAS && !ggIdx.empty()) {
                std::vector<const Gate*> gs;
                for (int x : ggIdx) gs.push_back(&gates[x]);
                bestEff = p.cost.perfPerGate(n, p.precision, gs) / ggIdx.size();
            }
        }
        if (usePerGate && useBLAS) {
            for (int ms = p.matLo; ms < p.matHi; ms++) {
                if (ms > n) break;
                std::fill(full.begin(), full.end(), 0);
                std::fill(related.begin(), related.end(), 0);
                std::vector<int> idx = finder.find(full, related, false, ms, all);
                if (idx.empty()) continue;
                double eff = p.cost.perfBLAS(n, p.precision, ms) / idx.size();
                if (eff < bestEff) {
                    ggIdx = idx;
                    backend = Backend::BLAS;
                    bestEff = eff;
                }
            }
        } else if (useBLAS) {
            std::fill(full.begin(), full.end(), 0);
            std::fill(related.begin(), related.end(), 0);
            ggIdx = finder.find(full, related, false, std::min(p.blasMat, n), all);
            backend = Backend::BLAS;
        }
        if (ggIdx.empty()) break;   // cannot happen for valid parameters
        GateGroup gg;
        gg.backend = backend;
        for (int x : ggIdx) gg.addGate(gates[x], x, all, backend == Backend::PerGate);
        if (backend == Backend::PerGate) gg.relatedQubits |= cacheRelated;
        state = gg.initState(state, n, p.minMat);
        finder.remove(ggIdx);
        lg.relatedQubits |= gg.relatedQubits;
        lg.fullGroups.push_back(std::move(gg));
    }
    lg.state = State(n);
    sch.localGroups.push_back(std::move(lg));
    sch.finalState = state;
    return sch;
}
