// evaluator.h -- cost model of the hybrid partitioner (src/evaluator.h/.cpp).
//
// Two profiles:
//  * Reference: the reference's built-in V100 tables at n = 28 (per-gate
//    OShareMem costs, pergate group overhead 2^(n-27) ms, BLAS_perf[28][6]
//    and cutt_cost[28]), used to reproduce the reference partition.
//  * MI355X: a roofline model of this build's kernels whose constants come
//    from microbenchmarks on the GPU (hq_calibrate) and can be loaded from a
//    parameter file.  OShareMem group time = max(HBM pass, VALU work of the
//    gates) + LDS exchanges; TransMM = transpose pass + max(HBM pass, MFMA).
#pragma once

#include <string>
#include <vector>

#include "gate.h"

struct CostModel {
    // --- reference profile (evaluator.h:18-117, evaluator.cpp:5-35) -------------
    bool reference = false;
    double single_cost[(int)GateType::TOTAL] = {};  // pergate_single_perf[n][type][1]
    double ctr_cost[(int)GateType::TOTAL] = {};     // pergate_ctr_perf[n][type][0][2]
    double blas_perf[16] = {};                      // BLAS_perf[n][matSize]
    double cutt_cost = 0;                           // cutt_cost[n]
    double group_overhead_per_amp = 1.0 / (1 << 27);

    // --- MI355X profile (seconds, bytes) ---------------------------------------------
    double hbm_bw = 5.6e12;         // achieved streaming bandwidth, B/s
    double valu_rate = 39.3e12 * 0.75;   // FP64 lane-ops/s achieved (complex128)
    double valu_rate_f32 = 78.6e12 * 0.75;
    double mfma_f64 = 78.6e12 * 0.6;     // achieved FP64 MFMA flop/s
    double mfma_f32 = 157.3e12 * 0.6;    // achieved FP32 MFMA flop/s
    double lds_per_amp_stage = 0.0;      // s per amplitude per LDS exchange (0: from bandwidth)
    double launch_overhead = 6e-6;       // s per kernel launch
    double lds_bw = 256 * 128 * 2.4e9;   // B/s whole chip (128 B/clk/CU)

    static CostModel referenceV100(int numQubits);
    static CostModel mi355x();
    bool load(const std::string& path, std::string* err);  // parameter file
    bool save(const std::string& path) const;

    // VALU instructions per amplitude for one gate in the register kernel
    static double valuPerAmp(const Gate& g);
    // ms, like the reference's Evaluator
    double perfPerGate(int numQubits, int precision, const std::vector<const Gate*>& gates) const;
    double perfBLAS(int numQubits, int precision, int matSize) const;
};
