// circuit.cpp -- Circuit (src/circuit.cpp:16-309, single GPU).

#include "circuit.h"

#include <cstring>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "hq_internal.h"

qreal zero_wrapper(qreal x)
{
    const qreal eps = 1e-14;
    return (x > -eps && x < eps) ? 0 : x;
}

std::string ResultItem::str() const
{
    char buf[160];
    snprintf(buf, sizeof(buf), "%lld %.12f: %.12f %.12f\n", idx, amp.x * amp.x + amp.y * amp.y, zero_wrapper(amp.x),
             zero_wrapper(amp.y));
    return buf;
}

void ResultItem::print() const { fputs(str().c_str(), stdout); }

Circuit::Circuit(int n) : numQubits(n)
{
    params.localSize = 12;
    exec.localSize = 12;
}

Circuit::~Circuit() { freeState(); }

void Circuit::freeState()
{
    // a borrowed state belongs to the caller: only the reference is dropped
    if (state && !stateBorrowed) hq_state_destroy(state);
    state = nullptr;
    stateHasBuffer = false;
    stateBorrowed = false;
}

hq_status Circuit::useState(hq_state s)
{
    int hb = 0;
    if (s) {
        int n = 0, prec = 0;
        hq_status rc = hq_state_info(s, &n, &prec);
        if (rc == HQ_OK) rc = hq_state_has_buffer(s, &hb);
        if (rc != HQ_OK) return rc;
        if (n != numQubits) {
            hq::set_error("use_state: the state has " + std::to_string(n) + " qubits, the circuit " +
                          std::to_string(numQubits));
            return HQ_ERR_ARG;
        }
    }
    freeState();
    state = s;
    stateBorrowed = s != nullptr;
    stateHasBuffer = hb != 0;
    return HQ_OK;
}

bool Circuit::needsBuffer() const
{
    // only TransMM groups that run out of place write through the work buffer
    return Executor::needsBuffer(schedule, numQubits, params.precision);
}

hq_status Circuit::ensureState()
{
    bool buf = needsBuffer();
    if (stateBorrowed) {
        int n, prec;
        hq_status rc = hq_state_info(state, &n, &prec);
        if (rc != HQ_OK) return rc;
        if (prec != params.precision) {
            hq::set_error("the state given with use_state has another precision than the circuit");
            return HQ_ERR_ARG;
        }
        if (buf && !stateHasBuffer) {
            hq::set_error("the schedule has TransMM groups in transpose form: the state given with use_state needs a "
                          "work buffer");
            return HQ_ERR_ARG;
        }
        return HQ_OK;
    }
    if (state) {
        int n, prec;
        hq_state_info(state, &n, &prec);
        if (prec == params.precision && (stateHasBuffer || !buf)) return HQ_OK;
        freeState();
    }
    hq_status rc = hq_state_create(numQubits, params.precision, buf ? 1 : 0, &state);
    if (rc != HQ_OK) { state = nullptr; return rc; }
    stateHasBuffer = buf;
    return HQ_OK;
}

hq_status Circuit::partitionChecked()
{
    if (numQubits < 1 || numQubits > 36) {
        hq::set_error("Circuit: number of qubits must be in [1, 36]");
        return HQ_ERR_ARG;
    }
    for (size_t i = 0; i < gates.size(); i++) {
        const Gate& g = gates[i];
        bool bad = g.targetQubit < 0 || g.targetQubit >= numQubits || g.controlQubit >= numQubits ||
                   g.controlQubit2 >= numQubits || (g.controlQubit >= 0 && g.controlQubit == g.targetQubit) ||
                   (g.controlQubit2 >= 0 && (g.controlQubit2 == g.targetQubit || g.controlQubit2 == g.controlQubit));
        if (bad) {
            hq::set_error("Circuit: gate " + std::to_string(i) + " (" + g.name + ") has an invalid qubit");
            return HQ_ERR_ARG;
        }
    }
    schedule.clear();
    compiled = false;
    if (params.mode == HQ_MODE_SERIAL) {
        schedule.finalState = State(numQubits);
    } else {
        if (referenceProfile) params.cost = CostModel::referenceV100(numQubits);
        if (localSizeAuto) params.localSize = autoLocalSize();
        if (beamAuto) {
            // a wider search costs host time once and saves a pass per group it
            // removes on every run: widen it with the state size
            params.beam = numQubits < 26 ? 8 : numQubits < 31 ? 16 : 64;
            params.starts = numQubits < 26 ? 4 : numQubits < 31 ? 8 : 16;
        }
        exec.localSize = params.localSize;
        schedule = compileSchedule(numQubits, gates, params);
    }
    return HQ_OK;
}

hq_status Circuit::compileChecked()
{
    auto t0 = std::chrono::steady_clock::now();
    hq_status rc = partitionChecked();
    if (rc != HQ_OK) return rc;
    if (params.mode != HQ_MODE_SERIAL) {
        rc = ensureState();
        if (rc != HQ_OK) return rc;
        rc = Executor::prepare(schedule, numQubits, state, exec);
        if (rc != HQ_OK) return rc;
    }
    compiled = true;
    compileMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    int totalGroups = 0;
    for (auto& lg : schedule.localGroups) totalGroups += (int)lg.fullGroups.size();
    Logger::add("Total Gates %d", (int)gates.size());
    Logger::add("Total Groups: %d %d %d %d", (int)schedule.localGroups.size(), totalGroups, (int)gates.size(), 0);
    Logger::add("Compile Time: %d us", (int)(compileMs * 1000));
    return HQ_OK;
}

void Circuit::compile()
{
    hq_status rc = compileChecked();
    if (rc != HQ_OK) throw hq::Error(rc, hq_last_error());
}

int Circuit::autoLocalSize() const
{
    // specialised kernels exchange a 128 KiB tile (13 qubits of complex128,
    // 14 of complex64) in two LDS halves and keep two blocks per CU (fewer
    // passes, measured faster); the interpreter keeps 12-qubit tiles (64 KiB
    // LDS for complex128)
    const bool jit = exec.jit == 1 || (exec.jit < 0 && numQubits >= exec.jitMinQubits);
    if (!jit) return 12;
    return params.precision == HQ_C64 ? 14 : 13;
}

static const char kSchedMagic[8] = {'H', 'Q', 'S', 'C', 'H', 'E', 'D', '1'};

hq_status Circuit::saveSchedule(std::vector<unsigned char>* out) const
{
    if (params.mode == HQ_MODE_SERIAL || schedule.localGroups.empty()) {
        hq::set_error("saveSchedule: partition or compile the circuit first (not in serial mode)");
        return HQ_ERR_STATE;
    }
    std::vector<unsigned char> body = schedule.serialize();
    out->assign(kSchedMagic, kSchedMagic + 8);
    uint32_t hdr[2] = {(uint32_t)numQubits, 0};
    out->insert(out->end(), reinterpret_cast<unsigned char*>(hdr), reinterpret_cast<unsigned char*>(hdr) + 8);
    out->insert(out->end(), body.begin(), body.end());
    return HQ_OK;
}

hq_status Circuit::loadSchedule(const unsigned char* data, size_t len)
{
    if (params.mode == HQ_MODE_SERIAL) {
        hq::set_error("loadSchedule: serial mode runs without a schedule");
        return HQ_ERR_STATE;
    }
    if (len < 16 || memcmp(data, kSchedMagic, 8) != 0) {
        hq::set_error("loadSchedule: not a schedule file");
        return HQ_ERR_PARSE;
    }
    uint32_t hdr[2];
    memcpy(hdr, data + 8, 8);
    if ((int)hdr[0] != numQubits) {
        hq::set_error("loadSchedule: the schedule is for " + std::to_string(hdr[0]) + " qubits");
        return HQ_ERR_ARG;
    }
    Schedule s;
    size_t cur = 16;
    try {
        s = Schedule::deserialize(data, len, cur);
    } catch (const std::exception& e) {
        hq::set_error(std::string("loadSchedule: ") + e.what());
        return HQ_ERR_PARSE;
    }
    if (cur != len) {
        hq::set_error("loadSchedule: trailing bytes");
        return HQ_ERR_PARSE;
    }
    // every gate of the circuit exactly once, identical to the circuit's
    std::vector<char> seen(gates.size(), 0);
    for (auto& lg : s.localGroups) {
        for (auto& gg : lg.fullGroups) {
            if (gg.gateIndex.size() != gg.gates.size()) { hq::set_error("loadSchedule: bad group"); return HQ_ERR_PARSE; }
            for (size_t i = 0; i < gg.gates.size(); i++) {
                const int idx = gg.gateIndex[i];
                if (idx < 0 || idx >= (int)gates.size() || seen[idx]) {
                    hq::set_error("loadSchedule: gate index " + std::to_string(idx) + " missing or repeated");
                    return HQ_ERR_ARG;
                }
                seen[idx] = 1;
                const Gate& a = gates[idx];
                const Gate& b = gg.gates[i];
                if (a.type != b.type || a.targetQubit != b.targetQubit || a.controlQubit != b.controlQubit ||
                    a.controlQubit2 != b.controlQubit2 || memcmp(a.mat, b.mat, sizeof(a.mat)) != 0) {
                    hq::set_error("loadSchedule: gate " + std::to_string(idx) + " differs from the circuit's");
                    return HQ_ERR_ARG;
                }
            }
        }
    }
    for (size_t i = 0; i < seen.size(); i++)
        if (!seen[i]) { hq::set_error("loadSchedule: gate " + std::to_string(i) + " is not scheduled"); return HQ_ERR_ARG; }
    // the tile must hold every group's related qubits
    int need = 0;
    for (auto& lg : s.localGroups)
        for (auto& gg : lg.fullGroups)
            if (gg.backend == Backend::PerGate)
                need = std::max(need, __builtin_popcountll((unsigned long long)gg.relatedQubits));
    if (localSizeAuto) params.localSize = std::max(autoLocalSize(), need);
    schedule = std::move(s);
    compiled = false;
    auto t0 = std::chrono::steady_clock::now();
    exec.localSize = params.localSize;
    hq_status rc = ensureState();
    if (rc != HQ_OK) return rc;
    rc = Executor::prepare(schedule, numQubits, state, exec);
    if (rc != HQ_OK) return rc;
    compiled = true;
    compileMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return HQ_OK;
}

hq_status Circuit::runChecked(bool copy_back, bool destroy, long long* us)
{
    hq_status rc;
    if (!compiled) {
        rc = compileChecked();
        if (rc != HQ_OK) return rc;
    }
    rc = ensureState();
    if (rc != HQ_OK) return rc;
    // a fused first group writes |0...0> evolved by its gates over the whole state
    if (params.mode == HQ_MODE_SERIAL || !schedule.initFused) {
        rc = hq_state_reset(state);
        if (rc != HQ_OK) return rc;
    }
    hq_profile_enable(state, profileEvents ? 1 : 0);
    rc = hq_state_sync(state);
    if (rc != HQ_OK) return rc;
    auto t0 = std::chrono::steady_clock::now();
    if (params.mode == HQ_MODE_SERIAL) {
        std::vector<hq_gate> hg;
        hg.reserve(gates.size());
        for (const Gate& g : gates) hg.push_back(g.toHQ());
        rc = hq_apply_gates_plain(state, hg.data(), (int)hg.size());
    } else {
        rc = Executor::run(schedule, state);
    }
    if (rc != HQ_OK) return rc;
    rc = hq_state_sync(state);
    if (rc != HQ_OK) return rc;
    auto t1 = std::chrono::steady_clock::now();
    long long dur = std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count();
    if (us) *us = dur;
    Logger::add("Time Cost: %d us", (int)dur);
    lastLaunches.clear();
    for (double& t : lastTimes) t = 0;
    if (profileEvents) {
        std::vector<int> kinds(1 << 16);
        std::vector<double> ms(1 << 16);
        int cnt = 0;
        rc = hq_profile_read(state, kinds.data(), ms.data(), (int)kinds.size(), &cnt);
        if (rc != HQ_OK) return rc;
        for (int i = 0; i < std::min(cnt, (int)kinds.size()); i++) {
            lastLaunches.push_back({kinds[i], ms[i]});
            if (kinds[i] >= 0 && kinds[i] < 4) lastTimes[kinds[i]] += ms[i];
        }
    }
    if (copy_back) {
        const uint64_t N = 1ull << numQubits;
        result.assign(N, qComplex{});
        if (params.precision == HQ_C128) {
            rc = hq_state_download(state, result.data(), N);
        } else {
            std::vector<float> tmp(2 * N);
            rc = hq_state_download(state, tmp.data(), N);
            for (uint64_t i = 0; i < N; i++) result[i] = qComplex{tmp[2 * i], tmp[2 * i + 1]};
        }
        if (rc != HQ_OK) return rc;
    }
    if (destroy && !stateBorrowed) freeState();
    return HQ_OK;
}

int Circuit::run(bool copy_back, bool destroy)
{
    long long us = 0;
    hq_status rc = runChecked(copy_back, destroy, &us);
    if (rc != HQ_OK) throw hq::Error(rc, hq_last_error());
    return (int)us;
}

void Circuit::dumpGates() const
{
    // src/circuit.cpp:73-93
    printf("total Gates: %d\n", (int)gates.size());
    const int L = 3;
    for (const Gate& g : gates) {
        for (int i = 0; i < numQubits; i++) {
            if (i == g.controlQubit) {
                printf(".");
                for (int j = 1; j < L; j++) printf(" ");
            } else if (i == g.targetQubit) {
                printf("%s", g.name.c_str());
                for (int j = (int)g.name.length(); j < L; j++) printf(" ");
            } else {
                printf("|");
                for (int j = 1; j < L; j++) printf(" ");
            }
        }
        printf("\n");
    }
}

qindex Circuit::toPhysicalID(qindex idx) const
{
    qindex id = 0;
    const auto& pos = schedule.stateLayout().pos;
    for (int i = 0; i < numQubits; i++)
        if ((idx >> i) & 1) id |= qindex(1) << pos[i];
    return id;
}

qindex Circuit::toLogicID(qindex idx) const
{
    qindex id = 0;
    const auto& pos = schedule.stateLayout().pos;
    for (int i = 0; i < numQubits; i++)
        if ((idx >> pos[i]) & 1) id |= qindex(1) << i;
    return id;
}

ResultItem Circuit::ampAt(qindex idx)
{
    if (result.empty()) throw hq::Error(HQ_ERR_STATE, "ampAt: run with copy_back first");
    qindex id = toPhysicalID(idx);
    return ResultItem(idx, result[id]);
}

qComplex Circuit::ampAtGPU(qindex idx)
{
    if (!state) throw hq::Error(HQ_ERR_STATE, "ampAtGPU: run with destroy=false first");
    double re, im;
    hq_status rc = hq_state_get_amp(state, (uint64_t)toPhysicalID(idx), &re, &im);
    if (rc != HQ_OK) throw hq::Error(rc, hq_last_error());
    return qComplex{re, im};
}

std::string Circuit::stateText(bool showSchedule) const
{
    // src/circuit.cpp:284-308 (single-GPU branch)
    std::string out;
    if (result.empty()) return out;
    auto at = [&](qindex i) { return ResultItem(i, result[toPhysicalID(i)]); };
    for (int i = 0; i < 128; i++) out += at(i).str();
    if (showSchedule) {
        for (int i = 0; i < numQubits; i++) out += at(qindex(1) << i).str();
        out += at((qindex(1) << numQubits) - 1).str();
    }
    std::vector<ResultItem> big;
    const qindex N = qindex(1) << numQubits;
    for (qindex i = 0; i < N; i++) {
        const qComplex& a = result[i];
        if (a.x * a.x + a.y * a.y > 0.001) {
            qindex lid = toLogicID(i);
            if (lid >= 128) big.push_back(ResultItem(lid, a));
        }
    }
    std::sort(big.begin(), big.end());
    for (auto& r : big) out += r.str();
    return out;
}

void Circuit::printState()
{
    std::string s = stateText(false);
    fputs(s.c_str(), stdout);
}

namespace {
std::mutex g_log_mu;
bool g_log_on = false;
std::vector<std::string>& log_lines()
{
    static std::vector<std::string> v;
    return v;
}
}  // namespace

void Logger::enable(bool on)
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    g_log_on = on;
}

void Logger::add(const char* format, ...)
{
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        if (!g_log_on) return;
    }
    char buf[1024];
    va_list args;
    va_start(args, format);
    vsnprintf(buf, sizeof(buf), format, args);
    va_end(args);
    std::lock_guard<std::mutex> lk(g_log_mu);
    log_lines().push_back(buf);
}

void Logger::print()
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    for (auto& s : log_lines()) printf("Logger: %s\n", s.c_str());
    log_lines().clear();
}

namespace MyGlobalVars {
int numGPUs = 1;
int localGPUs = 1;
int bit = 0;

void init()
{
    int count = 0;
    hq_status st = hq_device_count(&count);
    if (st != HQ_OK) throw hq::Error(st, hq_last_error());
    numGPUs = localGPUs = 1;
    bit = 0;
}
}  // namespace MyGlobalVars
