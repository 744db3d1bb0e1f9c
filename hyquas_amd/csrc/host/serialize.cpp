// serialize.cpp -- the schedule's byte format (src/schedule.cpp:156-325,
// src/gate.cpp:558-589): every field copied in native x86-64 layout with no
// padding, counts as size_t ahead of their vectors, in the reference's order:
//
//   Gate       int gateID, int type, 4 x {double re, im} (mat, row-major),
//              size_t name_len, name bytes + NUL, int target, control, control2
//   State      size_t n, int pos[n], int layout[n]
//   GateGroup  size_t num_gates, qindex relatedQubits, size_t num_perm,
//              int matQubit, int backend, State, Gate[num_gates], int cuttPerm[num_perm]
//   LocalGroup int a2aCommSize, size_t num_a2a, size_t num_perm, size_t num_og,
//              size_t num_fg, qindex relatedQubits, int a2aComm[], int cuttPerm[],
//              State, GateGroup overlap[num_og], GateGroup full[num_fg]
//   Schedule   size_t num_lg, State finalState, LocalGroup[num_lg]
//
// On one GPU the exchange fields are empty.  A gate's gateID carries its index
// in the circuit's addGate order.  The file form adds a 16-byte header
// ("HQSCHED1", uint32 qubit count, uint32 zero).

#include <cstring>
#include <stdexcept>

#include "schedule.h"

namespace {

struct Writer {
    std::vector<unsigned char> out;
    template <typename T>
    void put(const T& v)
    {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        out.insert(out.end(), p, p + sizeof(T));
    }
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        out.insert(out.end(), c, c + n);
    }
    void ints(const std::vector<int>& v) { bytes(v.data(), v.size() * sizeof(int)); }
};

struct Reader {
    const unsigned char* p;
    size_t len, cur = 0;
    void need(size_t n) const
    {
        if (cur + n > len) throw std::runtime_error("schedule bytes truncated");
    }
    template <typename T>
    T get()
    {
        need(sizeof(T));
        T v;
        memcpy(&v, p + cur, sizeof(T));
        cur += sizeof(T);
        return v;
    }
    std::vector<int> ints(size_t n)
    {
        if (n > (1u << 26)) throw std::runtime_error("schedule vector too long");
        need(n * sizeof(int));
        std::vector<int> v(n);
        if (n) memcpy(v.data(), p + cur, n * sizeof(int));
        cur += n * sizeof(int);
        return v;
    }
};

void putState(Writer& w, const State& s)
{
    w.put<size_t>(s.pos.size());
    w.ints(s.pos);
    w.ints(s.layout);
}

State getState(Reader& r)
{
    size_t n = r.get<size_t>();
    State s;
    s.pos = r.ints(n);
    s.layout = r.ints(n);
    return s;
}

void putGate(Writer& w, const Gate& g, int index)
{
    w.put<int>(index);
    w.put<int>((int)g.type);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) { w.put<double>(g.mat[i][j].x); w.put<double>(g.mat[i][j].y); }
    w.put<size_t>(g.name.size());
    w.bytes(g.name.c_str(), g.name.size() + 1);
    w.put<int>(g.targetQubit);
    w.put<int>(g.controlQubit);
    w.put<int>(g.controlQubit2);
}

Gate getGate(Reader& r, int* index)
{
    Gate g;
    *index = r.get<int>();
    g.gateID = *index;
    g.type = (GateType)r.get<int>();
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) { g.mat[i][j].x = r.get<double>(); g.mat[i][j].y = r.get<double>(); }
    size_t nl = r.get<size_t>();
    if (nl > 64) throw std::runtime_error("gate name too long");
    r.need(nl + 1);
    g.name.assign(reinterpret_cast<const char*>(r.p + r.cur), nl);
    r.cur += nl + 1;
    g.targetQubit = r.get<int>();
    g.controlQubit = r.get<int>();
    g.controlQubit2 = r.get<int>();
    return g;
}

void putGroup(Writer& w, const GateGroup& gg)
{
    w.put<size_t>(gg.gates.size());
    w.put<qindex>(gg.relatedQubits);
    w.put<size_t>(gg.cuttPerm.size());
    w.put<int>(gg.matQubit);
    w.put<int>((int)gg.backend);
    putState(w, gg.state);
    for (size_t i = 0; i < gg.gates.size(); i++)
        putGate(w, gg.gates[i], i < gg.gateIndex.size() ? gg.gateIndex[i] : gg.gates[i].gateID);
    w.ints(gg.cuttPerm);
}

GateGroup getGroup(Reader& r)
{
    GateGroup gg;
    size_t ng = r.get<size_t>();
    gg.relatedQubits = r.get<qindex>();
    size_t np = r.get<size_t>();
    gg.matQubit = r.get<int>();
    int be = r.get<int>();
    if (be < 0 || be > 2) throw std::runtime_error("bad backend");
    gg.backend = (Backend)be;
    gg.state = getState(r);
    if (ng > (1u << 24)) throw std::runtime_error("group too large");
    for (size_t i = 0; i < ng; i++) {
        int idx;
        gg.gates.push_back(getGate(r, &idx));
        gg.gateIndex.push_back(idx);
    }
    gg.cuttPerm = r.ints(np);
    return gg;
}

}  // namespace

std::vector<unsigned char> Schedule::serialize() const
{
    Writer w;
    w.put<size_t>(localGroups.size());
    putState(w, finalState);
    for (const LocalGroup& lg : localGroups) {
        w.put<int>(0);                          // a2aCommSize
        w.put<size_t>(0);                       // a2aComm
        w.put<size_t>(0);                       // cuttPerm
        w.put<size_t>(0);                       // overlap groups
        w.put<size_t>(lg.fullGroups.size());
        w.put<qindex>(lg.relatedQubits);
        putState(w, lg.state);
        for (const GateGroup& gg : lg.fullGroups) putGroup(w, gg);
    }
    return w.out;
}

Schedule Schedule::deserialize(const unsigned char* arr, size_t len, size_t& cur)
{
    Reader r{arr, len, cur};
    Schedule s;
    size_t nl = r.get<size_t>();
    if (nl > 64) throw std::runtime_error("too many local groups");
    s.finalState = getState(r);
    for (size_t i = 0; i < nl; i++) {
        LocalGroup lg;
        (void)r.get<int>();
        size_t na = r.get<size_t>(), np = r.get<size_t>(), no = r.get<size_t>(), nf = r.get<size_t>();
        lg.relatedQubits = r.get<qindex>();
        (void)r.ints(na);
        (void)r.ints(np);
        lg.state = getState(r);
        if (no != 0) throw std::runtime_error("overlap groups are a multi-GPU feature");
        if (nf > (1u << 24)) throw std::runtime_error("too many groups");
        for (size_t k = 0; k < nf; k++) lg.fullGroups.push_back(getGroup(r));
        s.localGroups.push_back(std::move(lg));
    }
    cur = r.cur;
    return s;
}
