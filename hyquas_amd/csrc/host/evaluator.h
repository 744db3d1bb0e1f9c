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
    double hbm_bw = 5.457e+12;       // achieved transpose bandwidth, B/s (TransMM passes)
    // Achieved bandwidth of one OShareMem pass (read + write of the state) by
    // the run length L of its tile (the number of lowest physical qubits in the
    // tile: 2^L consecutive amplitudes per run), complex128 and complex64.
    // Measured by the calibration tool (specialised kernels holding one
    // diagonal gate, random 13-qubit tiles) at 30 qubits.
    double shm_bw_run[13] = {1.323e+12, 1.491e+12, 4.122e+12, 5.198e+12, 5.054e+12, 4.958e+12, 5.3e+12,
                             5.369e+12, 5.401e+12, 5.23e+12, 5.398e+12, 5.227e+12, 5.063e+12};
    double shm_bw_run_f32[13] = {1.141e+12, 1.355e+12, 2.199e+12, 3.945e+12, 4.553e+12, 4.869e+12, 4.978e+12,
                                 5.428e+12, 4.831e+12, 5.215e+12, 5.285e+12, 5.377e+12, 5.198e+12};
    double valu_rate = 3.793e+13;    // FP64 lane-ops/s achieved (complex128)
    double valu_rate_f32 = 3.84e+13;
    double mfma_f64 = 7.730e+13;     // achieved FP64 MFMA flop/s (dense kernel, K = 64; algorithmic 8K per amplitude)
    double mfma_f32 = 1.303e+14;     // achieved FP32 MFMA flop/s (dense kernel, K = 128)
    double lds_per_amp_stage = 0.0;      // s per amplitude per LDS exchange (0: from bandwidth)
    double launch_overhead = 2.642e-06;   // s per kernel launch
    double lds_bw = 256 * 128 * 2.4e9;   // B/s whole chip (128 B/clk/CU)

    static CostModel referenceV100(int numQubits);
    static CostModel mi355x();
    // Parameter file: this build's "key value" lines, or the reference's
    // <n>qubits.out tables (evaluator-preprocess format), which select the
    // reference's evaluation rules with the file's numbers.
    bool load(const std::string& path, std::string* err);
    bool save(const std::string& path) const;

    // VALU instructions per amplitude for one gate in the register kernel
    static double valuPerAmp(const Gate& g);
    // ms, like the reference's Evaluator.  tileMask: physical qubits of the
    // group's tile (0: a tile whose run length is 4).
    double perfPerGate(int numQubits, int precision, const std::vector<const Gate*>& gates,
                       qindex tileMask = 0) const;
    // OShareMem pass bandwidth for a tile with run length L
    double shmBandwidth(int precision, int runLength) const;
    static int runLength(qindex tileMask);
    // inPlace: the group's matrix runs where its qubits are, one pass, no transpose
    double perfBLAS(int numQubits, int precision, int matSize, bool inPlace = false) const;
};
