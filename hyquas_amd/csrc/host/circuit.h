// circuit.h -- public Circuit API (src/circuit.h:9-46), single MI355X GPU.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "compiler.h"
#include "executor.h"
#include "gate.h"
#include "logger.h"
#include "schedule.h"
#include "utils.h"

qreal zero_wrapper(qreal x);   // src/utils.cpp:77-84

struct ResultItem {
    ResultItem() = default;
    ResultItem(const qindex& idx, const qComplex& amp) : idx(idx), amp(amp) {}
    qindex idx = 0;
    qComplex amp;
    std::string str() const;   // "%lld %.12f: %.12f %.12f\n" (src/circuit.h:14-16)
    void print() const;
    bool operator<(const ResultItem& b) const { return idx < b.idx; }
};

namespace hq {
// Raised by the C++ API where the reference calls exit(1).
struct Error : std::runtime_error {
    hq_status status;
    Error(hq_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
}  // namespace hq

class Circuit {
public:
    explicit Circuit(int numQubits);
    ~Circuit();
    Circuit(const Circuit&) = delete;
    Circuit& operator=(const Circuit&) = delete;

    void addGate(const Gate& gate) { gates.push_back(gate); compiled = false; }
    void compile();
    int run(bool copy_back = true, bool destroy = true);
    void dumpGates() const;
    void printState();
    ResultItem ampAt(qindex idx);
    qComplex ampAtGPU(qindex idx);
    const int numQubits;

    // ---- status-returning forms used by the C ABI ----------------------------
    hq_status partitionChecked();   // host-only: schedule without device programs
    hq_status compileChecked();
    hq_status runChecked(bool copy_back, bool destroy, long long* us);
    // Compile-once cache: the partitioned schedule as bytes (16-byte header +
    // the reference's layout, src/schedule.cpp:156-325); loading checks every
    // gate against this circuit and builds the device programs.
    hq_status saveSchedule(std::vector<unsigned char>* out) const;
    hq_status loadSchedule(const unsigned char* data, size_t len);
    std::string stateText(bool showSchedule) const;

    // compile-time options of the reference, runtime here
    CompilerParams params;
    ExecParams exec;
    bool referenceProfile = false;
    bool localSizeAuto = true;   // tile qubits from the kernel kind unless "share_qubits" is set
    bool beamAuto = true;        // beam width from the qubit count unless "beam"/"starts" is set
    int autoLocalSize() const;

    const std::vector<Gate>& getGates() const { return gates; }
    const Schedule& getSchedule() const { return schedule; }
    hq_state deviceState() const { return state; }
    // Run on a caller-owned state (borrowed: never destroyed here); nullptr
    // returns to states of the circuit's own.
    hq_status useState(hq_state s);
    // logical -> physical position of each qubit in the state a run leaves
    const std::vector<int>& finalPos() const { return schedule.stateLayout().pos; }
    double lastTimes[4] = {0, 0, 0, 0};   // ms by launch kind (profile on)
    std::vector<std::pair<int, double>> lastLaunches;
    bool profileEvents = false;
    double compileMs = 0;

    qindex toPhysicalID(qindex idx) const;   // src/circuit.cpp:95-103
    qindex toLogicID(qindex idx) const;      // src/circuit.cpp:105-113

private:
    bool needsBuffer() const;
    hq_status ensureState();
    void freeState();

    std::vector<Gate> gates;
    Schedule schedule;
    bool compiled = false;
    hq_state state = nullptr;
    bool stateHasBuffer = false;
    bool stateBorrowed = false;
    std::vector<qComplex> result;
};

// Opaque handle of the C ABI (hq_circuit): owns one Circuit.
struct hq_circuit_s {
    std::unique_ptr<Circuit> c;
};
