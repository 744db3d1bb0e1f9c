// gate.cpp -- gate matrices and factories (src/gate.cpp:9-552).
//
// All factories go through one table-driven builder.  Deviation from the
// reference: Gate::TDG keeps type TDG (the reference types it as T,
// src/gate.cpp:239, which makes its OShareMem and plain paths apply T).

#include "gate.h"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>

qComplex operator*(const qComplex& a, const qComplex& b)
{
    return qComplex{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

qComplex operator+(const qComplex& a, const qComplex& b) { return qComplex{a.x + b.x, a.y + b.y}; }

namespace {

std::atomic<int> g_gate_id{0};

struct TypeInfo {
    const char* name;
    int controls;
    int params;
    bool diagonal;
};

const TypeInfo& info(GateType t)
{
    static const TypeInfo table[] = {
        {"CCX", 2, 0, false}, {"CN", 1, 0, false},  {"CY", 1, 0, false},  {"CZ", 1, 0, true},
        {"CRX", 1, 1, false}, {"CRY", 1, 1, false}, {"CU1", 1, 1, true},  {"CRZ", 1, 1, true},
        {"U1", 0, 1, true},   {"U2", 0, 2, false},  {"U3", 0, 3, false},  {"H", 0, 0, false},
        {"X", 0, 0, false},   {"Y", 0, 0, false},   {"Z", 0, 0, true},    {"S", 0, 0, true},
        {"SDG", 0, 0, true},  {"T", 0, 0, true},    {"TDG", 0, 0, true},  {"RX", 0, 1, false},
        {"RY", 0, 1, false},  {"RZ", 0, 1, true},   {"TOTAL", 0, 0, false},
        {"ID", 0, 0, false},  {"GII", 0, 0, false}, {"GZZ", 0, 0, false}, {"GOC", 0, 2, false},
        {"GCC", 0, 2, false},
    };
    return table[(int)t];
}

void set(qComplex m[2][2], qComplex a, qComplex b, qComplex c, qComplex d)
{
    m[0][0] = a; m[0][1] = b; m[1][0] = c; m[1][1] = d;
}

// Closed-form matrices of src/gate.cpp:9-348.
void build_matrix(GateType t, const double* p, qComplex m[2][2])
{
    const qComplex zero{0, 0}, one{1, 0};
    const double r2 = 1 / std::sqrt(2.0);
    switch (t) {
    case GateType::CCX: case GateType::CNOT: case GateType::X: set(m, zero, one, one, zero); break;
    case GateType::CY: case GateType::Y: set(m, zero, {0, -1}, {0, 1}, zero); break;
    case GateType::CZ: case GateType::Z: set(m, one, zero, zero, {-1, 0}); break;
    case GateType::CRX: case GateType::RX: {
        double c = std::cos(p[0] / 2.0), s = std::sin(p[0] / 2.0);
        set(m, {c, 0}, {0, -s}, {0, -s}, {c, 0});
        break;
    }
    case GateType::CRY: case GateType::RY: {
        double c = std::cos(p[0] / 2.0), s = std::sin(p[0] / 2.0);
        set(m, {c, 0}, {-s, 0}, {s, 0}, {c, 0});
        break;
    }
    case GateType::CU1: case GateType::U1: set(m, one, zero, zero, {std::cos(p[0]), std::sin(p[0])}); break;
    case GateType::CRZ: case GateType::RZ: {
        double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        set(m, {c, -s}, zero, zero, {c, s});
        break;
    }
    case GateType::U2: {
        double phi = p[0], lam = p[1], k = std::sqrt(2.0);
        set(m, {1.0 / k, 0}, {-std::cos(lam) / k, -std::sin(lam) / k}, {std::cos(phi) / k, std::sin(phi) / k},
            {std::cos(lam + phi) / k, std::sin(lam + phi) / k});
        break;
    }
    case GateType::U3: {
        double th = p[0], phi = p[1], lam = p[2];
        double c = std::cos(th / 2), s = std::sin(th / 2);
        set(m, {c, 0}, {-std::cos(lam) * s, -std::sin(lam) * s}, {std::cos(phi) * s, std::sin(phi) * s},
            {std::cos(phi + lam) * c, std::sin(phi + lam) * c});
        break;
    }
    case GateType::H: set(m, {r2, 0}, {r2, 0}, {r2, 0}, {-r2, 0}); break;
    case GateType::S: set(m, one, zero, zero, {0, 1}); break;
    case GateType::SDG: set(m, one, zero, zero, {0, -1}); break;
    case GateType::T: set(m, one, zero, zero, {r2, r2}); break;
    case GateType::TDG: set(m, one, zero, zero, {r2, -r2}); break;
    case GateType::ID: set(m, one, zero, zero, one); break;
    case GateType::GII: set(m, {0, 1}, zero, zero, {0, 1}); break;
    case GateType::GZZ: set(m, {-1, 0}, zero, zero, {-1, 0}); break;
    case GateType::GOC: set(m, one, zero, zero, {p[0], p[1]}); break;
    case GateType::GCC: set(m, {p[0], p[1]}, zero, zero, {p[0], p[1]}); break;
    default: set(m, one, zero, zero, one); break;
    }
}

Gate build(GateType t, int c2, int c1, int tgt, const double* p)
{
    Gate g;
    g.gateID = ++g_gate_id;
    g.type = t;
    build_matrix(t, p, g.mat);
    g.name = info(t).name;
    g.targetQubit = tgt;
    g.controlQubit = c1;
    g.controlQubit2 = c2;
    return g;
}

}  // namespace

bool Gate::isDiagonal() const { return info(type).diagonal; }

hq_status make_gate(int type, int control2, int control, int target, const double* params, int nparams, Gate* out)
{
    if (type < 0 || type > (int)GateType::GCC || type == (int)GateType::TOTAL) return HQ_ERR_ARG;
    const TypeInfo& ti = info((GateType)type);
    if (nparams != ti.params || (ti.params && !params)) return HQ_ERR_ARG;
    if ((ti.controls >= 1) != (control >= 0)) return HQ_ERR_ARG;
    if ((ti.controls == 2) != (control2 >= 0)) return HQ_ERR_ARG;
    if (target < 0) return HQ_ERR_ARG;
    for (int i = 0; i < nparams; i++)
        if (!std::isfinite(params[i])) return HQ_ERR_ARG;
    *out = build((GateType)type, control2, control, target, params);
    return HQ_OK;
}

Gate Gate::CCX(int c1, int c2, int t) { return build(GateType::CCX, c2, c1, t, nullptr); }
Gate Gate::CNOT(int c, int t) { return build(GateType::CNOT, -1, c, t, nullptr); }
Gate Gate::CY(int c, int t) { return build(GateType::CY, -1, c, t, nullptr); }
Gate Gate::CZ(int c, int t) { return build(GateType::CZ, -1, c, t, nullptr); }
Gate Gate::CRX(int c, int t, qreal a) { return build(GateType::CRX, -1, c, t, &a); }
Gate Gate::CRY(int c, int t, qreal a) { return build(GateType::CRY, -1, c, t, &a); }
Gate Gate::CU1(int c, int t, qreal l) { return build(GateType::CU1, -1, c, t, &l); }
Gate Gate::CRZ(int c, int t, qreal a) { return build(GateType::CRZ, -1, c, t, &a); }
Gate Gate::U1(int t, qreal l) { return build(GateType::U1, -1, -1, t, &l); }
Gate Gate::U2(int t, qreal phi, qreal lam)
{
    double p[2] = {phi, lam};
    return build(GateType::U2, -1, -1, t, p);
}
Gate Gate::U3(int t, qreal th, qreal phi, qreal lam)
{
    double p[3] = {th, phi, lam};
    return build(GateType::U3, -1, -1, t, p);
}
Gate Gate::H(int t) { return build(GateType::H, -1, -1, t, nullptr); }
Gate Gate::X(int t) { return build(GateType::X, -1, -1, t, nullptr); }
Gate Gate::Y(int t) { return build(GateType::Y, -1, -1, t, nullptr); }
Gate Gate::Z(int t) { return build(GateType::Z, -1, -1, t, nullptr); }
Gate Gate::S(int t) { return build(GateType::S, -1, -1, t, nullptr); }
Gate Gate::SDG(int t) { return build(GateType::SDG, -1, -1, t, nullptr); }
Gate Gate::T(int t) { return build(GateType::T, -1, -1, t, nullptr); }
Gate Gate::TDG(int t) { return build(GateType::TDG, -1, -1, t, nullptr); }
Gate Gate::RX(int t, qreal a) { return build(GateType::RX, -1, -1, t, &a); }
Gate Gate::RY(int t, qreal a) { return build(GateType::RY, -1, -1, t, &a); }
Gate Gate::RZ(int t, qreal a) { return build(GateType::RZ, -1, -1, t, &a); }
Gate Gate::ID(int t) { return build(GateType::ID, -1, -1, t, nullptr); }
Gate Gate::GII(int t) { return build(GateType::GII, -1, -1, t, nullptr); }
Gate Gate::GZZ(int t) { return build(GateType::GZZ, -1, -1, t, nullptr); }
Gate Gate::GOC(int t, qreal re, qreal im)
{
    double p[2] = {re, im};
    return build(GateType::GOC, -1, -1, t, p);
}
Gate Gate::GCC(int t, qreal re, qreal im)
{
    double p[2] = {re, im};
    return build(GateType::GCC, -1, -1, t, p);
}

static double rand_angle() { return std::rand() * 1.0 / RAND_MAX * std::acos(-1) * 2; }

Gate Gate::random(int lo, int hi)
{
    int type = std::rand() % int(GateType::TOTAL);
    return random(lo, hi, GateType(type));
}

Gate Gate::random(int lo, int hi, GateType type)
{
    int span = hi - lo;
    int c2 = -1, c1 = -1, t = lo;
    const TypeInfo& ti = info(type);
    do {
        if (ti.controls == 2) c2 = std::rand() % span + lo;
        if (ti.controls >= 1) c1 = std::rand() % span + lo;
        t = std::rand() % span + lo;
    } while ((ti.controls >= 1 && c1 == t) || (ti.controls == 2 && (c2 == c1 || c2 == t)));
    double p[3];
    for (int i = 0; i < ti.params; i++) p[i] = rand_angle();
    return build(type, c2, c1, t, p);
}

Gate Gate::control(int c, int t, GateType type)
{
    double p = rand_angle();
    return build(type, -1, c, t, info(type).params ? &p : nullptr);
}

GateType Gate::toCU(GateType type)
{
    // src/gate.cpp:516-522
    return type == GateType::CCX ? GateType::CNOT : type;
}

GateType Gate::toU(GateType type)
{
    // src/gate.cpp:524-546
    switch (type) {
    case GateType::CCX: case GateType::CNOT: return GateType::X;
    case GateType::CY: return GateType::Y;
    case GateType::CZ: return GateType::Z;
    case GateType::CRX: return GateType::RX;
    case GateType::CRY: return GateType::RY;
    case GateType::CU1: return GateType::U1;
    case GateType::CRZ: return GateType::RZ;
    default: return type;
    }
}

std::string Gate::get_name(GateType ty) { return info(ty).name; }

Gate Gate::fromHQ(const hq_gate& h)
{
    Gate g;
    g.gateID = ++g_gate_id;
    g.type = (GateType)h.type;
    for (int i = 0; i < 4; i++) g.mat[i / 2][i % 2] = qComplex{h.mat[2 * i], h.mat[2 * i + 1]};
    g.name = (h.type >= 0 && h.type <= (int)GateType::GCC) ? info(g.type).name : "?";
    g.targetQubit = h.target;
    g.controlQubit = h.control;
    g.controlQubit2 = h.control2;
    return g;
}

hq_gate Gate::toHQ() const
{
    hq_gate h;
    h.type = (int32_t)type;
    h.target = targetQubit;
    h.control = controlQubit;
    h.control2 = controlQubit2;
    for (int i = 0; i < 4; i++) {
        h.mat[2 * i] = mat[i / 2][i % 2].x;
        h.mat[2 * i + 1] = mat[i / 2][i % 2].y;
    }
    return h;
}

namespace hq {

static uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next()
{
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

uint64_t Rng::below(uint64_t m) { return ((next() >> 11) * m) >> 53; }

}  // namespace hq
