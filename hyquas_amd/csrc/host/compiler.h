// compiler.h -- the hybrid partitioner (src/compiler.h / compiler.cpp), single GPU.
#pragma once

#include <vector>

#include "evaluator.h"
#include "schedule.h"

struct CompilerParams {
    int mode = HQ_MODE_MIX;   // HQ_MODE_GROUP / HQ_MODE_BLAS / HQ_MODE_MIX
    int precision = HQ_C128;
    int localSize = 12;       // OShareMem tile qubits (LOCAL_QUBIT_SIZE)
    int coalesce = 4;         // COALESCE_GLOBAL: lowest physical qubits always in the tile (reference: 3)
    bool coalesceAuto = true; // search coalesce in 2..5 by predicted time (off once "coalesce" is set)
    int minMat = 4;           // MIN_MAT_SIZE
    int blasMat = 6;          // BLAS_MAT_LIMIT (HQ_MODE_BLAS)
    int matLo = 4, matHi = 8; // mix mode tries matLo <= matSize < matHi
    int window = 512;         // gates considered per group (OneLayerCompiler<512>)
    int starts = 4;           // OShareMem merge lookahead: 1 = reference rule
    int beam = 8;             // > 1: beam search over OShareMem-only partitions (kept if better)
    // TransMM groups the in-place dense kernel can run keep the layout (no
    // transpose planned, OShareMem groups after them coalesce on the actual
    // low qubits); false: the reference's transposes (src/schedule.cpp:339-382)
    bool blasInPlace = true;
    CostModel cost;
};

// Greedy grouping of the remaining gates (OneLayerCompiler::getGroupOpt,
// src/compiler.cpp:322-452).
class GroupFinder {
public:
    GroupFinder(int numQubits, const std::vector<Gate>& gates, int window);
    // Selected gate indices (ascending) for a group with at most localSize
    // related qubits.  full/related are updated like the reference's arrays.
    std::vector<int> find(std::vector<char>& full, std::vector<qindex>& related, bool enableGlobal, int localSize,
                          qindex localQubits, int starts = 1) const;
    // Candidate groups: the reference's merge plus one per forced first chain
    // among the `starts` longest (distinct groups only, reference's first).
    std::vector<std::vector<int>> findAll(std::vector<char>& full, std::vector<qindex>& related, bool enableGlobal,
                                          int localSize, qindex localQubits, int starts) const;
    void remove(const std::vector<int>& idx);
    bool empty() const { return remain_.empty(); }
    const std::vector<int>& remain() const { return remain_; }

private:
    int n_;
    int window_;
    const std::vector<Gate>& gates_;
    std::vector<int> remain_;   // ascending indices of ungrouped gates
};

// AdvanceCompiler::run (src/compiler.cpp:220-319) on one GPU: every qubit is
// local, one LocalGroup holds all groups.
Schedule compileSchedule(int numQubits, const std::vector<Gate>& gates, const CompilerParams& p);
