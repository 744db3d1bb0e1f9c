// schedule.h -- schedule IR between the partitioner and the executor
// (src/schedule.h:9-104, single GPU).
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "gate.h"

enum class Backend { None, PerGate, BLAS };
std::string to_string(Backend b);

// pos[logical] = physical bit, layout[physical] = logical qubit.
struct State {
    std::vector<int> pos;
    std::vector<int> layout;
    State() = default;
    explicit State(int numQubits);
    State(const std::vector<int>& p, const std::vector<int>& l) : pos(p), layout(l) {}
};

struct GateGroup {
    std::vector<Gate> gates;
    std::vector<int> gateIndex;   // index of each gate in the circuit's addGate order
    qindex relatedQubits = 0;     // logical qubits
    State state;                  // layout after this group
    std::vector<int> cuttPerm;    // output bit i <- input bit cuttPerm[i] (BLAS)
    int matQubit = 0;             // log2 K (BLAS)
    Backend backend = Backend::None;
    std::vector<std::complex<double>> matrix;  // K x K column-major (BLAS)
    hq_program program = nullptr; // device-resident program (owned)
    int numStages = 0;            // OShareMem register stages
    double liveFraction = 1.0;    // share of a full read + write of the state the pass moves (0.5: write only; <1: zero tiles skipped)

    GateGroup() = default;
    GateGroup(GateGroup&& o) noexcept;
    GateGroup& operator=(GateGroup&& o) noexcept;
    GateGroup(const GateGroup&) = delete;
    GateGroup& operator=(const GateGroup&) = delete;
    ~GateGroup();

    bool contains(int i) const { return (relatedQubits >> i) & 1; }
    static qindex newRelated(qindex old, const Gate& g, qindex localQubits, bool enableGlobal);
    void addGate(const Gate& g, int index, qindex localQubits, bool enableGlobal);

    State initState(const State& oldState, int numLocalQubits, int minMatSize);
    State initPerGateState(const State& oldState);
    State initBlasState(const State& oldState, int numLocalQubits, int minMatSize);
    // FP64 fusion of the group into `matrix` (src/schedule.cpp:578-702)
    void initCPUMatrix();
};

struct LocalGroup {
    std::vector<GateGroup> fullGroups;
    qindex relatedQubits = 0;
    State state;
};

struct Schedule {
    std::vector<LocalGroup> localGroups;   // one on a single GPU
    State finalState;
    bool initFused = false;                // the first group's program builds |0...0> itself
    void dump(int numQubits) const;        // SHOW_SCHEDULE print (src/schedule.cpp:66-154)
    void clear();
    // Byte form (src/schedule.cpp:156-325); deserialize throws on malformed input
    // and advances cur.
    std::vector<unsigned char> serialize() const;
    static Schedule deserialize(const unsigned char* arr, size_t len, size_t& cur);
};
