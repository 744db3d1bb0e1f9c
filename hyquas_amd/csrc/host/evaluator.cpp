// evaluator.cpp -- cost model (src/evaluator.cpp:111-229 and an MI355X model).

#include "evaluator.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

CostModel CostModel::referenceV100(int n)
{
    // The reference's default tables (compiled in unless
    // USE_EVALUATOR_PREPROCESS): only n = 28 exists, other n exit(1) there.
    // perfPerGate reads entry [1] of each single-gate table and [0][2] of
    // each control table (evaluator.cpp:168-223).
    CostModel c;
    c.reference = true;
    auto S = [&](GateType t, double v) { c.single_cost[(int)t] = v; };
    auto C = [&](GateType t, double v) { c.ctr_cost[(int)t] = v; };
    S(GateType::U1, 225); S(GateType::U2, 469); S(GateType::U3, 469); S(GateType::H, 352);
    S(GateType::X, 350); S(GateType::Y, 350); S(GateType::Z, 194); S(GateType::S, 209);
    S(GateType::SDG, 209); S(GateType::T, 216); S(GateType::TDG, 216); S(GateType::RX, 370);
    S(GateType::RY, 367); S(GateType::RZ, 369);
    C(GateType::CCX, 195); C(GateType::CNOT, 195); C(GateType::CY, 193); C(GateType::CZ, 137);
    C(GateType::CRX, 224); C(GateType::CRY, 225); C(GateType::CU1, 225); C(GateType::CRZ, 224);
    // BLAS_perf[28][6] = 23.068396, cutt_cost[28] = 11.367814; the other
    // BLAS_perf entries are never assigned and read as zero.
    c.blas_perf[6] = 23.068396;
    c.cutt_cost = 11.367814;
    (void)n;
    return c;
}

CostModel CostModel::mi355x() { return CostModel(); }

double CostModel::valuPerAmp(const Gate& g)
{
    // Instruction counts of the register-kernel update forms (hq_ops.h):
    // complex 2x2: 8, real / RX forms: 4, diagonal: 4, phase: 2 (half the
    // amplitudes), swap: 2 register moves; a control halves the work.
    double base;
    switch (g.type) {
    case GateType::CCX: case GateType::CNOT: case GateType::X: base = 2; break;
    case GateType::CY: case GateType::Y: base = 2; break;
    case GateType::H: case GateType::RX: case GateType::RY: case GateType::CRX: case GateType::CRY: base = 4; break;
    case GateType::RZ: case GateType::CRZ: base = 4; break;
    case GateType::U1: case GateType::CU1: case GateType::Z: case GateType::CZ: case GateType::S:
    case GateType::SDG: case GateType::T: case GateType::TDG: base = 2; break;
    default: base = 8; break;
    }
    if (g.controlQubit >= 0) base *= 0.5;
    if (g.controlQubit2 >= 0) base *= 0.5;
    return base;
}

double CostModel::perfPerGate(int n, int precision, const std::vector<const Gate*>& gates) const
{
    if (reference) {
        double t = 0;
        for (const Gate* g : gates)
            t += g->isControlGate() ? ctr_cost[(int)g->type] : single_cost[(int)g->type];
        return t / 1000 / 512 + group_overhead_per_amp * std::ldexp(1.0, n);
    }
    const double amps = std::ldexp(1.0, n);
    const double bytes = 2 * amps * (precision == HQ_C128 ? 16 : 8);
    double instr = 0;
    for (const Gate* g : gates) instr += valuPerAmp(*g);
    const double rate = precision == HQ_C128 ? valu_rate : valu_rate_f32;
    double t = std::max(bytes / hbm_bw, instr * amps / rate) + launch_overhead;
    return t * 1e3;
}

double CostModel::perfBLAS(int n, int precision, int matSize) const
{
    if (reference) return blas_perf[matSize] + cutt_cost;
    const double amps = std::ldexp(1.0, n);
    const double bytes = 2 * amps * (precision == HQ_C128 ? 16 : 8);
    const double flops = 8.0 * std::ldexp(1.0, matSize) * amps;
    const double mf = precision == HQ_C128 ? mfma_f64 : mfma_f32;
    double t = bytes / hbm_bw + std::max(bytes / hbm_bw, flops / mf) + 2 * launch_overhead;
    return t * 1e3;
}

bool CostModel::load(const std::string& path, std::string* err)
{
    // "key value" lines; '#' starts a comment
    std::ifstream f(path);
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        auto h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        std::istringstream is(line);
        std::string key;
        double v;
        if (!(is >> key)) continue;
        if (!(is >> v)) {
            if (err) *err = "bad value for " + key;
            return false;
        }
        if (key == "hbm_bw") hbm_bw = v;
        else if (key == "valu_rate") valu_rate = v;
        else if (key == "valu_rate_f32") valu_rate_f32 = v;
        else if (key == "mfma_f64") mfma_f64 = v;
        else if (key == "mfma_f32") mfma_f32 = v;
        else if (key == "launch_overhead") launch_overhead = v;
        else if (key == "lds_bw") lds_bw = v;
        else {
            if (err) *err = "unknown key " + key;
            return false;
        }
    }
    reference = false;
    return true;
}

bool CostModel::save(const std::string& path) const
{
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return false;
    fprintf(f, "# MI355X cost model (seconds, bytes, flop)\n");
    fprintf(f, "hbm_bw %.6e\nvalu_rate %.6e\nvalu_rate_f32 %.6e\nmfma_f64 %.6e\nmfma_f32 %.6e\n", hbm_bw,
            valu_rate, valu_rate_f32, mfma_f64, mfma_f32);
    fprintf(f, "launch_overhead %.6e\nlds_bw %.6e\n", launch_overhead, lds_bw);
    fclose(f);
    return true;
}
