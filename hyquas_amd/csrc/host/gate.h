// gate.h -- Gate IR of the simulator (API of src/gate.h:11-66).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "hyquas_amd.h"

typedef double qreal;
typedef long long qindex;

// Amplitude / matrix entry with the field names of cuDoubleComplex (.x, .y),
// so code written against the reference's qComplex keeps compiling.
struct qComplex {
    qreal x = 0, y = 0;
};
inline qComplex make_qComplex(qreal x, qreal y = 0) { return qComplex{x, y}; }
qComplex operator*(const qComplex& a, const qComplex& b);
qComplex operator+(const qComplex& a, const qComplex& b);

// Same enumerators, same values as the reference (src/gate.h:7-9).
enum class GateType {
    CCX, CNOT, CY, CZ, CRX, CRY, CU1, CRZ, U1, U2, U3, H, X, Y, Z, S, SDG, T, TDG, RX, RY, RZ, TOTAL,
    ID, GII, GZZ, GOC, GCC
};

struct Gate {
    int gateID = 0;
    GateType type = GateType::ID;
    qComplex mat[2][2];
    std::string name;
    int targetQubit = 0;
    int controlQubit = -1;   // -1 if no control
    int controlQubit2 = -1;  // -1 if no control

    Gate() = default;
    Gate(const Gate&) = default;
    Gate& operator=(const Gate&) = default;

    bool isControlGate() const { return controlQubit != -1; }
    bool isC2Gate() const { return controlQubit2 != -1; }
    bool isDiagonal() const;

    static Gate CCX(int c1, int c2, int targetQubit);
    static Gate CNOT(int controlQubit, int targetQubit);
    static Gate CY(int controlQubit, int targetQubit);
    static Gate CZ(int controlQubit, int targetQubit);
    static Gate CRX(int controlQubit, int targetQubit, qreal angle);
    static Gate CRY(int controlQubit, int targetQubit, qreal angle);
    static Gate CU1(int controlQubit, int targetQubit, qreal lambda);
    static Gate CRZ(int controlQubit, int targetQubit, qreal angle);
    static Gate U1(int targetQubit, qreal lambda);
    static Gate U2(int targetQubit, qreal phi, qreal lambda);
    static Gate U3(int targetQubit, qreal theta, qreal phi, qreal lambda);
    static Gate H(int targetQubit);
    static Gate X(int targetQubit);
    static Gate Y(int targetQubit);
    static Gate Z(int targetQubit);
    static Gate S(int targetQubit);
    static Gate SDG(int targetQubit);
    static Gate T(int targetQubit);
    static Gate TDG(int targetQubit);
    static Gate RX(int targetQubit, qreal angle);
    static Gate RY(int targetQubit, qreal angle);
    static Gate RZ(int targetQubit, qreal angle);
    static Gate ID(int targetQubit);
    static Gate GII(int targetQubit);
    static Gate GZZ(int targetQubit);
    static Gate GOC(int targetQubit, qreal real, qreal imag);
    static Gate GCC(int targetQubit, qreal real, qreal imag);
    // Random gates from the C library rand(), as src/gate.cpp:351-492 (used by
    // micro-benchmarks; the reproducible workloads use hq::Rng instead).
    static Gate random(int lo, int hi);
    static Gate random(int lo, int hi, GateType type);
    static Gate control(int controlQubit, int targetQubit, GateType type);
    static GateType toCU(GateType type);
    static GateType toU(GateType type);
    static std::string get_name(GateType ty);

    // From/to the C ABI record.
    static Gate fromHQ(const hq_gate& g);
    hq_gate toHQ() const;
};

// Every gate factory through one entry point (hq_gate_make).
hq_status make_gate(int type, int control2, int control, int target, const double* params, int nparams, Gate* out);

namespace hq {
// Portable random generator of the synthetic workloads: xoshiro256**
// seeded by SplitMix64; uniform doubles as (x >> 11) * 2^-53; bounded
// integers as ((x >> 11) * m) >> 53.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();            // [0, 1)
    uint64_t below(uint64_t m);  // [0, m), m < 2^11
};
}  // namespace hq
