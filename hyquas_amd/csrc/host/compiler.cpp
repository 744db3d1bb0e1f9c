// compiler.cpp -- the hybrid partitioner.
//
// Restates OneLayerCompiler::getGroupOpt and AdvanceCompiler::run
// (src/compiler.cpp:220-452) for a single GPU, with every limit a runtime
// parameter (the reference fixes LOCAL_QUBIT_SIZE = 10 for its 16 KiB tile;
// this build's OShareMem tile is sized for gfx950's registers + 160 KiB LDS).
// Deviation: CCX's second control is tracked as a dependency (the reference
// ignores controlQubit2 in both sweeps, which can reorder gates on it).

#include "compiler.h"

#include "hq_internal.h"

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

std::vector<std::vector<int>> GroupFinder::findAll(std::vector<char>& full, std::vector<qindex>& related,
                                                   bool enableGlobal, int localSize, qindex localQubits,
                                                   int starts) const
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

    // 2. greedily merge the longest chains that fit.  With starts > 1 the
    // merge is repeated with each of the `starts` longest chains taken first,
    // and the largest group wins (a one-group lookahead; starts = 1 is the
    // reference's rule).
    auto merge = [&](int first, Bits& sel) {
        std::vector<Bits> cc = cur;
        std::vector<char> blocked(n, 0);
        qindex selRel = 0;
        bool forced = first >= 0;
        while (true) {
            int mx = 0, best = -1;
            if (forced) {
                best = first;
                mx = cc[first].count();
                forced = false;
            } else {
                for (int i = 0; i < n; i++) {
                    int cnt = cc[i].count();
                    if (!blocked[i] && cnt > mx) {
                        if (popc(selRel | related[i]) <= localSize) { mx = cnt; best = i; }
                        else blocked[i] = 1;
                    }
                }
            }
            if (mx == 0) break;
            sel.orWith(cc[best]);
            selRel |= related[best];
            blocked[best] = 1;
            for (int i = 0; i < n; i++) {
                if (!blocked[i] && cc[i].any()) {
                    if ((related[i] | selRel) == selRel) {
                        sel.orWith(cc[i]);
                        blocked[i] = 1;
                    } else {
                        cc[i].andNot(cc[best]);
                    }
                }
            }
        }
    };
    std::vector<Bits> cands;
    merge(-1, selected);
    cands.push_back(selected);
    if (starts > 1) {
        std::vector<int> order;
        for (int i = 0; i < n; i++)
            if (cur[i].any() && popc(related[i]) <= localSize) order.push_back(i);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cur[a].count() > cur[b].count(); });
        for (int k = 1; k < starts && k < (int)order.size(); k++) {
            Bits sel(gate_num);
            merge(order[k], sel);
            cands.push_back(sel);
        }
    }
    std::vector<char> blocked(n, 0);

    // 3. diagonal gates that commute into the group (OShareMem only)
    auto sweep = [&](Bits& sel) {
        std::fill(blocked.begin(), blocked.end(), 0);
        int cnt = 0;
        for (int id = 0; id < gate_num; id++) {
            cnt++;
            if (cnt % 100 == 0) {
                bool live = false;
                for (int i = 0; i < n; i++) if (!blocked[i]) live = true;
                if (!live) break;
            }
            if (sel.test(id)) continue;
            const Gate& g = gates_[remain_[id]];
            if (g.isDiagonal()) {
                if (g.isControlGate()) {
                    if (!blocked[g.controlQubit] && !blocked[g.targetQubit]) sel.set(id);
                    else blocked[g.controlQubit] = blocked[g.targetQubit] = 1;
                } else {
                    if (!blocked[g.targetQubit]) sel.set(id);
                }
            } else {
                if (g.isControlGate()) blocked[g.controlQubit] = 1;
                if (g.isC2Gate()) blocked[g.controlQubit2] = 1;
                blocked[g.targetQubit] = 1;
            }
        }
    };
    std::vector<std::vector<int>> out;
    for (auto& sel : cands) {
        if (enableGlobal) sweep(sel);
        std::vector<int> ret;
        for (int id = 0; id < gate_num; id++)
            if (sel.test(id)) ret.push_back(remain_[id]);
        bool dup = false;
        for (auto& o : out) if (o == ret) dup = true;
        if (!dup) out.push_back(std::move(ret));
    }
    return out;
}

std::vector<int> GroupFinder::find(std::vector<char>& full, std::vector<qindex>& related, bool enableGlobal,
                                   int localSize, qindex localQubits, int starts) const
{
    auto c = findAll(full, related, enableGlobal, localSize, localQubits, starts);
    size_t best = 0;
    for (size_t i = 1; i < c.size(); i++)
        if (c[i].size() > c[best].size()) best = i;
    return c.empty() ? std::vector<int>() : c[best];
}

// Physical tile of an OShareMem group: its related (logical) qubits at their
// physical positions, topped up with the lowest physical positions
// (fillRelatedQubits, src/executor.cpp:585-596).
// Whether a TransMM group of `ms` matrix qubits runs in place (no transpose pass).
static bool blasInPlace(const CompilerParams& p, int n, int ms)
{
    return p.blasInPlace && n >= hq::dense_at_min_qubits(p.precision, std::max(ms, p.minMat));
}

static qindex tileMask(qindex relatedLogical, const State& st, int size)
{
    qindex m = 0;
    int cnt = 0;
    for (int q = 0; q < (int)st.pos.size(); q++)
        if ((relatedLogical >> q) & 1) { m |= qindex(1) << st.pos[q]; cnt++; }
    for (int ph = 0; ph < (int)st.pos.size() && cnt < size; ph++)
        if (!((m >> ph) & 1)) { m |= qindex(1) << ph; cnt++; }
    return m;
}

static Schedule greedySchedule(int n, const std::vector<Gate>& gates, const CompilerParams& p, double* cost)
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
    *cost = 0;
    while (!finder.empty()) {
        std::vector<int> ggIdx;
        Backend backend = Backend::PerGate;
        qindex cacheRelated = 0;
        double bestEff = 1e10;
        int blasMs = 0;
        if (usePerGate) {
            std::fill(full.begin(), full.end(), 0);
            qindex co = 0;
            for (int j = 0; j < coalesce; j++) co |= qindex(1) << state.layout[j];
            std::fill(related.begin(), related.end(), co);
            cacheRelated = co;
            ggIdx = finder.find(full, related, true, localSize, all, std::max(1, p.starts));
            backend = Backend::PerGate;
            if (useBLAS && !ggIdx.empty()) {
                std::vector<const Gate*> gs;
                GateGroup tmp;
                for (int x : ggIdx) { gs.push_back(&gates[x]); tmp.addGate(gates[x], x, all, true); }
                bestEff = p.cost.perfPerGate(n, p.precision, gs,
                                             tileMask(tmp.relatedQubits | co, state, localSize)) / ggIdx.size();
            }
        }
        if (usePerGate && useBLAS) {
            for (int ms = p.matLo; ms < p.matHi; ms++) {
                if (ms > n) break;
                std::fill(full.begin(), full.end(), 0);
                std::fill(related.begin(), related.end(), 0);
                std::vector<int> idx = finder.find(full, related, false, ms, all);
                if (idx.empty()) continue;
                double eff = p.cost.perfBLAS(n, p.precision, ms, blasInPlace(p, n, ms)) / idx.size();
                if (eff < bestEff) {
                    ggIdx = idx;
                    backend = Backend::BLAS;
                    bestEff = eff;
                    blasMs = ms;
                }
            }
        } else if (useBLAS) {
            std::fill(full.begin(), full.end(), 0);
            std::fill(related.begin(), related.end(), 0);
            ggIdx = finder.find(full, related, false, std::min(p.blasMat, n), all);
            backend = Backend::BLAS;
            blasMs = std::min(p.blasMat, n);
        }
        if (ggIdx.empty()) break;   // cannot happen for valid parameters
        GateGroup gg;
        gg.backend = backend;
        for (int x : ggIdx) gg.addGate(gates[x], x, all, backend == Backend::PerGate);
        if (backend == Backend::PerGate) gg.relatedQubits |= cacheRelated;
        {
            std::vector<const Gate*> gs;
            for (int x : ggIdx) gs.push_back(&gates[x]);
            *cost += backend == Backend::PerGate
                         ? p.cost.perfPerGate(n, p.precision, gs, tileMask(gg.relatedQubits, state, localSize))
                         : p.cost.perfBLAS(n, p.precision, blasMs, blasInPlace(p, n, blasMs));
        }
        state = gg.initState(state, n, p.minMat, p.blasInPlace ? p.precision : -1);
        finder.remove(ggIdx);
        lg.relatedQubits |= gg.relatedQubits;
        lg.fullGroups.push_back(std::move(gg));
    }
    lg.state = State(n);
    sch.localGroups.push_back(std::move(lg));
    sch.finalState = state;
    return sch;
}


// Beam search over OShareMem-only partitions: every step expands each kept
// partial partition by the candidate groups of GroupFinder::findAll and keeps
// the `beam` partitions that have placed the most gates; the first to place
// every gate has the fewest groups found.  The layout never changes without
// TransMM groups, so the coalesced qubits are physical 0..coalesce-1.
static std::vector<std::vector<int>> beamPerGate(int n, const std::vector<Gate>& gates, const CompilerParams& p)
{
    struct Node {
        GroupFinder f;
        std::vector<std::vector<int>> groups;
        size_t done;
    };
    const qindex all = n >= 63 ? ~qindex(0) : ((qindex(1) << n) - 1);
    const int coalesce = std::min(p.coalesce, n);
    const int localSize = std::min(p.localSize, n);
    qindex co = 0;
    for (int j = 0; j < coalesce; j++) co |= qindex(1) << j;
    std::vector<Node> beam;
    beam.push_back(Node{GroupFinder(n, gates, p.window), {}, 0});
    std::vector<char> full(n);
    std::vector<qindex> related(n);
    const int width = std::max(1, p.beam);
    const int starts = std::max(2, p.starts);
    while (true) {
        for (auto& nd : beam)
            if (nd.f.empty()) return nd.groups;
        std::vector<Node> next;
        for (auto& nd : beam) {
            std::fill(full.begin(), full.end(), 0);
            std::fill(related.begin(), related.end(), co);
            auto cands = nd.f.findAll(full, related, true, localSize, all, starts);
            for (auto& c : cands) {
                if (c.empty()) continue;
                Node ch{nd.f, nd.groups, nd.done + c.size()};
                ch.f.remove(c);
                ch.groups.push_back(c);
                next.push_back(std::move(ch));
            }
        }
        if (next.empty()) return beam.front().groups;   // cannot happen for valid parameters
        std::vector<size_t> ord(next.size());
        for (size_t i = 0; i < ord.size(); i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return next[a].done > next[b].done; });
        std::vector<Node> kept;
        for (size_t i : ord) {
            if ((int)kept.size() >= width) break;
            bool dup = false;
            for (auto& k : kept)
                if (k.done == next[i].done && k.f.remain() == next[i].f.remain()) { dup = true; break; }
            if (!dup) kept.push_back(std::move(next[i]));
        }
        beam.swap(kept);
    }
}

static Schedule perGateSchedule(int n, const std::vector<Gate>& gates, const CompilerParams& p,
                                const std::vector<std::vector<int>>& groups, double* cost)
{
    Schedule sch;
    LocalGroup lg;
    State state(n);
    const qindex all = n >= 63 ? ~qindex(0) : ((qindex(1) << n) - 1);
    qindex co = 0;
    for (int j = 0; j < std::min(p.coalesce, n); j++) co |= qindex(1) << j;
    *cost = 0;
    const int localSize = std::min(p.localSize, n);
    for (auto& idx : groups) {
        std::vector<const Gate*> gs;
        for (int x : idx) gs.push_back(&gates[x]);
        GateGroup gg;
        gg.backend = Backend::PerGate;
        for (int x : idx) gg.addGate(gates[x], x, all, true);
        gg.relatedQubits |= co;
        *cost += p.cost.perfPerGate(n, p.precision, gs, tileMask(gg.relatedQubits, state, localSize));
        state = gg.initState(state, n, p.minMat);
        lg.relatedQubits |= gg.relatedQubits;
        lg.fullGroups.push_back(std::move(gg));
    }
    lg.state = State(n);
    sch.localGroups.push_back(std::move(lg));
    sch.finalState = state;
    return sch;
}

static Schedule compileOne(int n, const std::vector<Gate>& gates, const CompilerParams& p, double* cost)
{
    double greedyCost = 0;
    Schedule g = greedySchedule(n, gates, p, &greedyCost);
    *cost = greedyCost;
    if (p.beam <= 1 || (p.mode != HQ_MODE_GROUP && p.mode != HQ_MODE_MIX)) return g;
    double beamCost = 0;
    Schedule b = perGateSchedule(n, gates, p, beamPerGate(n, gates, p), &beamCost);
    if (beamCost < greedyCost) {
        *cost = beamCost;
        return b;
    }
    return g;
}

Schedule compileSchedule(int n, const std::vector<Gate>& gates, const CompilerParams& p)
{
    double cost = 0;
    const bool search = p.coalesceAuto && !p.cost.reference && (p.mode == HQ_MODE_GROUP || p.mode == HQ_MODE_MIX);
    if (!search) return compileOne(n, gates, p, &cost);
    // the coalesced (always-in-tile) low qubits trade run length against free
    // tile slots: the other coalesce replaces the default only when its
    // predicted time is lower by more than the model's error (3 %; passes of
    // one partition vary by 1-2 % between runs and boxes)
    Schedule best = compileOne(n, gates, p, &cost);
    double bestCost = cost;
    const double base = cost;
    // runs shorter than 16 amplitudes were never faster once the groups carry
    // their gates (12- and 13-qubit tiles, 30 qubits), so 4 and 5 are tried
    for (int c = 4; c <= 5; c++) {
        if (c == p.coalesce || c >= std::min(p.localSize, n)) continue;
        CompilerParams q = p;
        q.coalesce = c;
        Schedule s = compileOne(n, gates, q, &cost);
        if (cost < bestCost && cost < 0.97 * base) {
            best = std::move(s);
            bestCost = cost;
        }
    }
    return best;
}
