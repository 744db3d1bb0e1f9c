// evaluator.cpp -- cost model (src/evaluator.cpp:111-229 and an MI355X model).

#include "evaluator.h"

#include <algorithm>
#include <cctype>
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

// The MI355X defaults are the member initializers (zero-gate OShareMem passes
// at 30 qubits by run length, this build's specialised kernels); the
// shm_bw_run_* keys of a parameter file written by `python -m
// hyquas_amd.calibrate` replace them.
CostModel CostModel::mi355x() { return CostModel(); }

int CostModel::runLength(qindex tileMask)
{
    int L = 0;
    while (L < 63 && ((tileMask >> L) & 1)) L++;
    return L;
}

double CostModel::shmBandwidth(int precision, int L) const
{
    const double* t = precision == HQ_C128 ? shm_bw_run : shm_bw_run_f32;
    const double b = t[std::min(std::max(L, 0), 12)];
    return b > 0 ? b : hbm_bw;
}

double CostModel::valuPerAmp(const Gate& g)
{
    // FP64 (or FP32) instructions per amplitude of the register-kernel update
    // forms (hq_ops.h): an uncontrolled H, RX, RY or RZ loses a scalar factor
    // and needs 4 per pair (2 per amplitude), a phase 4 per pair, a general
    // complex 2x2 16 per pair; swaps are register renames; a control halves
    // the pairs touched.
    const bool ctl = g.controlQubit >= 0;
    double base;
    switch (g.type) {
    case GateType::CCX: case GateType::CNOT: case GateType::X: base = ctl ? 1 : 0; break;
    case GateType::CY: case GateType::Y: base = 1; break;
    case GateType::H: case GateType::RX: case GateType::RY: base = 2; break;
    case GateType::CRX: case GateType::CRY: base = 4; break;
    case GateType::RZ: base = 2; break;
    case GateType::CRZ: base = 4; break;
    case GateType::U1: case GateType::CU1: case GateType::Z: case GateType::CZ: case GateType::S:
    case GateType::SDG: case GateType::T: case GateType::TDG: base = 2; break;
    default: base = 8; break;
    }
    if (ctl) base *= 0.5;
    if (g.controlQubit2 >= 0) base *= 0.5;
    return base;
}

double CostModel::perfPerGate(int n, int precision, const std::vector<const Gate*>& gates, qindex tileMask) const
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
    const double bw = shmBandwidth(precision, tileMask ? runLength(tileMask) : 4);
    double t = std::max(bytes / bw, instr * amps / rate) + launch_overhead;
    return t * 1e3;
}

double CostModel::perfBLAS(int n, int precision, int matSize, bool inPlace) const
{
    if (reference) return blas_perf[matSize] + cutt_cost;
    const double amps = std::ldexp(1.0, n);
    const double bytes = 2 * amps * (precision == HQ_C128 ? 16 : 8);
    const double flops = 8.0 * std::ldexp(1.0, matSize) * amps;
    const double mf = precision == HQ_C128 ? mfma_f64 : mfma_f32;
    // transpose pass (unless in place) + the GEMM pass
    double t = (inPlace ? 0.0 : bytes / hbm_bw + launch_overhead) + std::max(bytes / hbm_bw, flops / mf) +
               launch_overhead;
    return t * 1e3;
}

bool CostModel::load(const std::string& path, std::string* err)
{
    std::ifstream f(path);
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::stringstream all;
    all << f.rdbuf();
    const std::string text = all.str();
    size_t p0 = text.find_first_not_of(" \t\r\n");
    if (p0 != std::string::npos && (isdigit((unsigned char)text[p0]) || text[p0] == '-')) {
        // the reference's <n>qubits.out (src/evaluator.cpp:60-103): param_type,
        // 14 single-gate rows, 7 controlled-gate rows, 10 "K ms" rows
        // (K = 1..512), the transpose cost
        std::istringstream is(text);
        int ptype;
        if (!(is >> ptype) || (ptype != 0 && ptype != 1)) {
            if (err) *err = "bad param_type";
            return false;
        }
        const int LQ = 10;   // LOCAL_QUBIT_SIZE of the reference
        static const GateType singles[14] = {GateType::U1, GateType::U2, GateType::U3, GateType::H, GateType::X,
                                             GateType::Y, GateType::Z, GateType::S, GateType::SDG, GateType::T,
                                             GateType::TDG, GateType::RX, GateType::RY, GateType::RZ};
        static const GateType ctrls[7] = {GateType::CNOT, GateType::CY, GateType::CZ, GateType::CRX,
                                          GateType::CRY, GateType::CU1, GateType::CRZ};
        CostModel c;
        c.reference = true;
        // CCX has no row in the file: it keeps the current model's entry
        // (the reference's built-in table value at n = 28)
        c.ctr_cost[(int)GateType::CCX] = ctr_cost[(int)GateType::CCX];
        for (GateType t : singles) {
            double v[LQ];
            const int cnt = ptype == 0 ? LQ : 1;
            for (int i = 0; i < cnt; i++)
                if (!(is >> v[i])) { if (err) *err = "truncated single-gate table"; return false; }
            c.single_cost[(int)t] = ptype == 0 ? v[1] : v[0];
        }
        for (GateType t : ctrls) {
            const int cnt = ptype == 0 ? LQ * LQ : 1;
            double keep = 0;
            for (int i = 0; i < cnt; i++) {
                double x;
                if (!(is >> x)) { if (err) *err = "truncated control-gate table"; return false; }
                if (ptype == 1 || i == 2) keep = x;   // entry [0][2]
            }
            c.ctr_cost[(int)t] = keep;
        }
        for (int i = 0; i < 10; i++) {
            double K, ms;
            if (!(is >> K >> ms)) { if (err) *err = "truncated BLAS table"; return false; }
            c.blas_perf[i] = ms;
        }
        if (!(is >> c.cutt_cost)) { if (err) *err = "missing transpose cost"; return false; }
        *this = c;
        return true;
    }
    // "key value" lines; '#' starts a comment
    std::istringstream lines(text);
    std::string line;
    while (std::getline(lines, line)) {
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
        int L = -1;
        if (key == "hbm_bw") hbm_bw = v;
        else if (key == "valu_rate") valu_rate = v;
        else if (key == "valu_rate_f32") valu_rate_f32 = v;
        else if (key == "mfma_f64") mfma_f64 = v;
        else if (key == "mfma_f32") mfma_f32 = v;
        else if (key == "launch_overhead") launch_overhead = v;
        else if (key == "lds_bw") lds_bw = v;
        else if (sscanf(key.c_str(), "shm_bw_run_f32_%d", &L) == 1 && L >= 0 && L <= 12) shm_bw_run_f32[L] = v;
        else if (sscanf(key.c_str(), "shm_bw_run_%d", &L) == 1 && L >= 0 && L <= 12) shm_bw_run[L] = v;
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
    for (int L = 0; L <= 12; L++) fprintf(f, "shm_bw_run_%d %.6e\n", L, shm_bw_run[L]);
    for (int L = 0; L <= 12; L++) fprintf(f, "shm_bw_run_f32_%d %.6e\n", L, shm_bw_run_f32[L]);
    fclose(f);
    return true;
}
