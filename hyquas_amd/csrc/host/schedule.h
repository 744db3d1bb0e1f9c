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

    // inPlacePrecision >= 0: a BLAS group the in-place dense kernel can run
    // (matrices of 4..7 qubits, a state of at least one tile in that
    // precision) keeps the layout (identity cuttPerm); otherwise, and with
    // -1, the reference's transpose brings its qubits to the lowest positions.
    State initState(const State& oldState, int numLocalQubits, int minMatSize, int inPlacePrecision = -1);
    State initPerGateState(const State& oldState);
    State initBlasState(const State& oldState, int numLocalQubits, int minMatSize, int inPlacePrecision = -1);
    // Logical qubit of each matrix index bit (BLAS): the group's qubits and,
    // up to matQubit, the lowest-placed other qubits, in order of their
    // position in `state`.  After the reference's transpose these are the
    // qubits at positions 0..matQubit-1 (bit j = state.layout[j]); a group
    // run in place finds them where they are.
    std::vector<int> matLogical() const;
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
    State finalState;                      // the partition's final layout (serialized)
    State runState;                        // the layout a run leaves (set when the executor lowers the groups)
    bool initFused = false;                // the first group's program builds |0...0> itself
    // the layout of the state after a run: runState once lowered, else the partition's
    const State& stateLayout() const { return runState.pos.empty() ? finalState : runState; }
    void dump(int numQubits) const;        // SHOW_SCHEDULE print (src/schedule.cpp:66-154)
    void clear();
    // Byte form (src/schedule.cpp:156-325); deserialize throws on malformed input
    // and advances cur.
    std::vector<unsigned char> serialize() const;
    static Schedule deserialize(const unsigned char* arr, size_t len, size_t& cur);
};
