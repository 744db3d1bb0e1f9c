// capi.cpp -- C ABI of the Circuit layer (include/hyquas_amd.h, Circuit part).

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "circuit.h"
#include "hq_internal.h"

namespace hq {
hq_status parse_qasm(const std::string& text, std::unique_ptr<Circuit>* out);
hq_status load_qasm(const std::string& path, std::unique_ptr<Circuit>* out);
void add_random_layers(Circuit& c, int layers, uint64_t seed);
void add_hea_layers(Circuit& c, int layers, uint64_t seed);
}  // namespace hq

#define HQ_CARG(cond, msg)                              \
    do {                                                \
        if (!(cond)) { hq::set_error(msg); return HQ_ERR_ARG; } \
    } while (0)

#define HQ_GUARD(expr)                                                     \
    try {                                                                  \
        expr;                                                              \
    } catch (const hq::Error& e) {                                         \
        hq::set_error(e.what());                                           \
        return e.status;                                                   \
    } catch (const std::bad_alloc&) {                                      \
        hq::set_error("host allocation failed");                           \
        return HQ_ERR_OOM;                                                 \
    } catch (const std::exception& e) {                                    \
        hq::set_error(e.what());                                           \
        return HQ_ERR_ARG;                                                 \
    }

extern "C" {

hq_status hq_gate_make(int type, int control2, int control, int target, const double* params, int nparams,
                       hq_gate* out)
{
    HQ_CARG(out, "hq_gate_make: null output");
    Gate g;
    hq_status rc = make_gate(type, control2, control, target, params, nparams, &g);
    if (rc != HQ_OK) {
        hq::set_error("hq_gate_make: invalid type, qubits or parameter count");
        return rc;
    }
    *out = g.toHQ();
    return HQ_OK;
}

hq_status hq_circuit_create(int num_qubits, hq_circuit* out)
{
    HQ_CARG(out, "hq_circuit_create: null output");
    HQ_CARG(num_qubits >= 1 && num_qubits <= 36, "hq_circuit_create: num_qubits must be in [1, 36]");
    HQ_GUARD(*out = new hq_circuit_s{std::unique_ptr<Circuit>(new Circuit(num_qubits))});
    return HQ_OK;
}

hq_status hq_circuit_destroy(hq_circuit c)
{
    delete c;
    return HQ_OK;
}

hq_status hq_circuit_add_gate(hq_circuit c, const hq_gate* g)
{
    HQ_CARG(c && g, "hq_circuit_add_gate: null argument");
    const int n = c->c->numQubits;
    HQ_CARG(g->type >= 0 && g->type <= HQ_GCC && g->type != HQ_TOTAL, "hq_circuit_add_gate: unknown gate type");
    HQ_CARG(g->target >= 0 && g->target < n && g->control < n && g->control2 < n && g->control >= -1 &&
                g->control2 >= -1,
            "hq_circuit_add_gate: qubit out of range");
    HQ_GUARD(c->c->addGate(Gate::fromHQ(*g)));
    return HQ_OK;
}

hq_status hq_circuit_num_gates(hq_circuit c, int* count)
{
    HQ_CARG(c && count, "hq_circuit_num_gates: null argument");
    *count = (int)c->c->getGates().size();
    return HQ_OK;
}

hq_status hq_circuit_num_qubits(hq_circuit c, int* num_qubits)
{
    HQ_CARG(c && num_qubits, "hq_circuit_num_qubits: null argument");
    *num_qubits = c->c->numQubits;
    return HQ_OK;
}

hq_status hq_circuit_load_qasm(const char* path, hq_circuit* out)
{
    HQ_CARG(path && out, "hq_circuit_load_qasm: null argument");
    std::unique_ptr<Circuit> c;
    hq_status rc;
    HQ_GUARD(rc = hq::load_qasm(path, &c));
    if (rc != HQ_OK) return rc;
    *out = new hq_circuit_s{std::move(c)};
    return HQ_OK;
}

hq_status hq_circuit_parse_qasm(const char* text, hq_circuit* out)
{
    HQ_CARG(text && out, "hq_circuit_parse_qasm: null argument");
    std::unique_ptr<Circuit> c;
    hq_status rc;
    HQ_GUARD(rc = hq::parse_qasm(text, &c));
    if (rc != HQ_OK) return rc;
    *out = new hq_circuit_s{std::move(c)};
    return HQ_OK;
}

hq_status hq_circuit_set_option(hq_circuit c, const char* key, int64_t v)
{
    HQ_CARG(c && key, "hq_circuit_set_option: null argument");
    Circuit& C = *c->c;
    std::string k = key;
    if (k == "mode") {
        HQ_CARG(v == HQ_MODE_SERIAL || v == HQ_MODE_GROUP || v == HQ_MODE_BLAS || v == HQ_MODE_MIX, "bad mode");
        C.params.mode = (int)v;
    } else if (k == "precision") {
        HQ_CARG(v == HQ_C128 || v == HQ_C64, "bad precision");
        C.params.precision = (int)v;
    } else if (k == "profile") {
        HQ_CARG(v == 0 || v == 1, "profile must be 0 (gfx950) or 1 (reference)");
        C.referenceProfile = v == 1;
        if (v == 1) {
            C.localSizeAuto = false;
            C.params.localSize = 10;
            C.params.coalesce = 3;
            C.params.starts = 1;
            C.params.beam = 1;
            C.beamAuto = false;
            C.params.coalesceAuto = false;
            C.params.minMat = 4;
            C.params.blasMat = 6;
            C.params.matLo = 4;
            C.params.matHi = 8;
            C.params.blasInPlace = false;   // the reference's transposes and layouts
        } else {
            C.params.cost = CostModel::mi355x();
        }
    } else if (k == "share_qubits") {
        HQ_CARG(v >= 7 && v <= 14, "share_qubits must be in [7, 14]");
        C.params.localSize = (int)v;
        C.localSizeAuto = false;
    } else if (k == "coalesce") {
        HQ_CARG(v >= 0 && v <= 8, "coalesce must be in [0, 8]");
        C.params.coalesce = (int)v;
        C.params.coalesceAuto = false;
    } else if (k == "mat_min") {
        HQ_CARG(v >= 1 && v <= 7, "mat_min must be in [1, 7]");
        C.params.matLo = (int)v;
    } else if (k == "mat_max") {
        HQ_CARG(v >= 1 && v <= 7, "mat_max must be in [1, 7]");
        C.params.matHi = (int)v + 1;
    } else if (k == "blas_mat") {
        HQ_CARG(v >= 1 && v <= 7, "blas_mat must be in [1, 7]");
        C.params.blasMat = (int)v;
    } else if (k == "min_mat") {
        HQ_CARG(v >= 1 && v <= 7, "min_mat must be in [1, 7]");
        C.params.minMat = (int)v;
    } else if (k == "reg_qubits") {
        HQ_CARG(v >= 0 && v <= 5, "reg_qubits must be in [0, 5]");
        C.exec.regQubits = (int)v;
    } else if (k == "window") {
        HQ_CARG(v >= 1 && v <= 4096, "window must be in [1, 4096]");
        C.params.window = (int)v;
    } else if (k == "starts") {
        HQ_CARG(v >= 1 && v <= 64, "starts must be in [1, 64]");
        C.params.starts = (int)v;
        C.beamAuto = false;
    } else if (k == "beam") {
        HQ_CARG(v >= 1 && v <= 1024, "beam must be in [1, 1024]");
        C.params.beam = (int)v;
        C.beamAuto = false;
    } else if (k == "jit") {
        HQ_CARG(v >= -1 && v <= 1, "jit must be -1 (auto), 0 or 1");
        C.exec.jit = (int)v;
    } else if (k == "tune") {
        HQ_CARG(v >= -1 && v <= 2, "tune must be -1 (auto), 0, 1 or 2");
        C.exec.tune = (int)v;
    } else if (k == "transmm_inplace") {
        HQ_CARG(v == 0 || v == 1, "transmm_inplace must be 0 or 1");
        C.params.blasInPlace = v == 1;
    } else if (k == "profile_events") {
        C.profileEvents = v != 0;
    } else {
        hq::set_error("hq_circuit_set_option: unknown option " + k);
        return HQ_ERR_ARG;
    }
    return HQ_OK;
}

hq_status hq_circuit_load_cost_model(hq_circuit c, const char* path)
{
    HQ_CARG(c && path, "hq_circuit_load_cost_model: null argument");
    std::string err;
    if (!c->c->params.cost.load(path, &err)) {
        hq::set_error("cost model: " + err);
        return HQ_ERR_ARG;
    }
    c->c->referenceProfile = false;
    return HQ_OK;
}

hq_status hq_circuit_compile(hq_circuit c)
{
    HQ_CARG(c, "hq_circuit_compile: null circuit");
    hq_status rc;
    HQ_GUARD(rc = c->c->compileChecked());
    return rc;
}

hq_status hq_circuit_partition(hq_circuit c)
{
    HQ_CARG(c, "hq_circuit_partition: null circuit");
    hq_status rc;
    HQ_GUARD(rc = c->c->partitionChecked());
    return rc;
}

hq_status hq_circuit_save_schedule(hq_circuit c, const char* path)
{
    HQ_CARG(c && path, "hq_circuit_save_schedule: null argument");
    std::vector<unsigned char> bytes;
    hq_status rc;
    HQ_GUARD(rc = c->c->saveSchedule(&bytes));
    if (rc != HQ_OK) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) { hq::set_error(std::string("hq_circuit_save_schedule: cannot open ") + path); return HQ_ERR_ARG; }
    const bool ok = fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    if (fclose(f) != 0 || !ok) { hq::set_error("hq_circuit_save_schedule: write failed"); return HQ_ERR_ARG; }
    return HQ_OK;
}

hq_status hq_circuit_load_schedule(hq_circuit c, const char* path)
{
    HQ_CARG(c && path, "hq_circuit_load_schedule: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) { hq::set_error(std::string("hq_circuit_load_schedule: cannot open ") + path); return HQ_ERR_ARG; }
    std::vector<unsigned char> bytes;
    unsigned char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) bytes.insert(bytes.end(), buf, buf + k);
    fclose(f);
    hq_status rc;
    HQ_GUARD(rc = c->c->loadSchedule(bytes.data(), bytes.size()));
    return rc;
}

hq_status hq_circuit_run(hq_circuit c, int copy_back, int destroy, int64_t* us)
{
    HQ_CARG(c, "hq_circuit_run: null circuit");
    long long t = 0;
    hq_status rc;
    HQ_GUARD(rc = c->c->runChecked(copy_back != 0, destroy != 0, &t));
    if (us) *us = t;
    return rc;
}

hq_status hq_circuit_amp_at(hq_circuit c, uint64_t index, double* re, double* im)
{
    HQ_CARG(c && re && im, "hq_circuit_amp_at: null argument");
    HQ_CARG(index < (1ull << c->c->numQubits), "hq_circuit_amp_at: index out of range");
    ResultItem r;
    HQ_GUARD(r = c->c->ampAt((qindex)index));
    *re = r.amp.x;
    *im = r.amp.y;
    return HQ_OK;
}

hq_status hq_circuit_amp_at_gpu(hq_circuit c, uint64_t index, double* re, double* im)
{
    HQ_CARG(c && re && im, "hq_circuit_amp_at_gpu: null argument");
    HQ_CARG(index < (1ull << c->c->numQubits), "hq_circuit_amp_at_gpu: index out of range");
    qComplex a;
    HQ_GUARD(a = c->c->ampAtGPU((qindex)index));
    *re = a.x;
    *im = a.y;
    return HQ_OK;
}

hq_status hq_circuit_print_state(hq_circuit c, int show_schedule, char* buf, size_t cap, size_t* len)
{
    HQ_CARG(c, "hq_circuit_print_state: null circuit");
    std::string s;
    HQ_GUARD(s = c->c->stateText(show_schedule != 0));
    if (s.empty()) {
        hq::set_error("hq_circuit_print_state: run with copy_back first");
        return HQ_ERR_STATE;
    }
    if (len) *len = s.size();
    if (buf && cap) {
        size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
        memcpy(buf, s.data(), k);
        buf[k] = 0;
    }
    return HQ_OK;
}

hq_status hq_circuit_state(hq_circuit c, hq_state* out)
{
    HQ_CARG(c && out, "hq_circuit_state: null argument");
    *out = c->c->deviceState();
    if (!*out) {
        hq::set_error("hq_circuit_state: no device state (run with destroy=0)");
        return HQ_ERR_STATE;
    }
    return HQ_OK;
}

hq_status hq_circuit_use_state(hq_circuit c, hq_state s)
{
    HQ_CARG(c, "hq_circuit_use_state: null circuit");
    return c->c->useState(s);
}

hq_status hq_circuit_final_pos(hq_circuit c, int* pos)
{
    HQ_CARG(c && pos, "hq_circuit_final_pos: null argument");
    const auto& p = c->c->finalPos();
    if ((int)p.size() != c->c->numQubits) {
        hq::set_error("hq_circuit_final_pos: compile first");
        return HQ_ERR_STATE;
    }
    for (size_t i = 0; i < p.size(); i++) pos[i] = p[i];
    return HQ_OK;
}

hq_status hq_circuit_schedule(hq_circuit c, int* num_groups, int cap, int* backend, int* num_gates,
                              uint64_t* related, int* mat_qubit, int* num_stages)
{
    HQ_CARG(c && num_groups, "hq_circuit_schedule: null argument");
    int k = 0;
    for (const auto& lg : c->c->getSchedule().localGroups) {
        for (const auto& gg : lg.fullGroups) {
            if (k < cap) {
                if (backend) backend[k] = gg.backend == Backend::PerGate ? HQ_BACKEND_PERGATE : HQ_BACKEND_BLAS;
                if (num_gates) num_gates[k] = (int)gg.gates.size();
                if (related) related[k] = (uint64_t)gg.relatedQubits;
                if (mat_qubit) mat_qubit[k] = gg.backend == Backend::BLAS ? gg.matQubit : 0;
                if (num_stages) num_stages[k] = gg.numStages;
            }
            k++;
        }
    }
    *num_groups = k;
    return HQ_OK;
}

hq_status hq_circuit_schedule_live(hq_circuit c, int cap, double* fraction, int* num_groups)
{
    HQ_CARG(c && num_groups, "hq_circuit_schedule_live: null argument");
    int k = 0;
    for (const auto& lg : c->c->getSchedule().localGroups)
        for (const auto& gg : lg.fullGroups) {
            if (fraction && k < cap) fraction[k] = gg.liveFraction;
            k++;
        }
    *num_groups = k;
    return HQ_OK;
}

hq_status hq_circuit_schedule_order(hq_circuit c, int* order, int cap)
{
    HQ_CARG(c && order, "hq_circuit_schedule_order: null argument");
    int k = 0;
    for (const auto& lg : c->c->getSchedule().localGroups)
        for (const auto& gg : lg.fullGroups)
            for (int idx : gg.gateIndex) {
                if (k < cap) order[k] = idx;
                k++;
            }
    if (k > cap) {
        hq::set_error("hq_circuit_schedule_order: capacity too small");
        return HQ_ERR_ARG;
    }
    return HQ_OK;
}

hq_status hq_circuit_last_times(hq_circuit c, double* plain_ms, double* sharemem_ms, double* transpose_ms,
                                double* dense_ms)
{
    HQ_CARG(c, "hq_circuit_last_times: null circuit");
    if (plain_ms) *plain_ms = c->c->lastTimes[0];
    if (sharemem_ms) *sharemem_ms = c->c->lastTimes[1];
    if (transpose_ms) *transpose_ms = c->c->lastTimes[2];
    if (dense_ms) *dense_ms = c->c->lastTimes[3];
    return HQ_OK;
}

hq_status hq_circuit_launch_times(hq_circuit c, int* kinds, double* ms, int cap, int* count)
{
    HQ_CARG(c && count, "hq_circuit_launch_times: null argument");
    const auto& v = c->c->lastLaunches;
    for (int i = 0; i < (int)v.size() && i < cap; i++) {
        if (kinds) kinds[i] = v[i].first;
        if (ms) ms[i] = v[i].second;
    }
    *count = (int)v.size();
    return HQ_OK;
}

hq_status hq_circuit_add_random_layers(hq_circuit c, int layers, uint64_t seed)
{
    HQ_CARG(c && layers >= 0, "hq_circuit_add_random_layers: bad argument");
    HQ_GUARD(hq::add_random_layers(*c->c, layers, seed));
    return HQ_OK;
}

hq_status hq_circuit_add_hea_layers(hq_circuit c, int layers, uint64_t seed)
{
    HQ_CARG(c && layers >= 0, "hq_circuit_add_hea_layers: bad argument");
    HQ_GUARD(hq::add_hea_layers(*c->c, layers, seed));
    return HQ_OK;
}

hq_status hq_circuit_get_gates(hq_circuit c, hq_gate* out, int cap, int* count)
{
    HQ_CARG(c && count, "hq_circuit_get_gates: null argument");
    const auto& g = c->c->getGates();
    for (int i = 0; i < (int)g.size() && i < cap; i++) out[i] = g[i].toHQ();
    *count = (int)g.size();
    return HQ_OK;
}

}  // extern "C"
