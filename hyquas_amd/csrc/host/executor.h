// executor.h -- turns a Schedule into device programs and runs them
// (src/executor.cpp:11-57, 462-551, 585-640, single GPU).
#pragma once

#include <vector>

#include "schedule.h"

struct ExecParams {
    int localSize = 12;   // OShareMem tile qubits (fillRelatedQubits target)
    int regQubits = 0;    // register qubits per thread (0: default for the tile)
    int jit = -1;         // specialised kernels: 0 off, 1 on, -1 from jitMinQubits
    int jitMinQubits = 22;
    // per-group choice between the plain specialised kernel and the whole-line
    // non-temporal variant, by timing both on the state: 0 off (plain), 1 on,
    // 2 whole-line variant everywhere (no timing), -1: 1 with specialised
    // kernels from tuneMinQubits, else 0
    int tune = -1;
    int tuneMinQubits = 26;
};

class Executor {
public:
    // Lower every group and upload its program (once per compile).
    static hq_status prepare(Schedule& sch, int numQubits, hq_state s, const ExecParams& p);
    // Launch every group in order on s (asynchronous).
    static hq_status run(Schedule& sch, hq_state s);
    // fillRelatedQubits (src/executor.cpp:585-596): top the group's qubits up
    // to `size` with the lowest physical positions.
    static qindex fillRelated(qindex related, const State& st, int size);
    // Whether a run of sch needs the state's work buffer: a TransMM group
    // whose plan moves qubits (the reference's transpose) or whose matrix the
    // in-place kernel cannot take runs out of place.
    static bool needsBuffer(const Schedule& sch, int numQubits, int precision);
};
