// qasm.cpp -- reader for the OpenQASM subset of the reference CLI (main.cpp:12-231).
//
// Dialect: whitespace-separated statements; the rest of a line after a
// statement is ignored.  "//", "OPENQASM", "include" lines are skipped;
// "qreg" skips three characters and reads the register size; gates are
// cx ccx cy cz h x y z s sdg t tdg (no parameters) and crx cry crz cu1 u1 u3
// rx ry rz with a parenthesised parameter list written without spaces.  A
// parameter "pi*x" is pi*x, "pi/x" is pi/x, anything else is read as a
// number prefix.  Qubit ids are all digit runs of the operand token.  Any
// other token is an error (the reference exits).

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "circuit.h"
#include "hq_internal.h"

namespace {

struct Reader {
    const std::string& s;
    size_t p = 0;
    explicit Reader(const std::string& text) : s(text) {}
    bool token(std::string& out)
    {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
        if (p >= s.size()) return false;
        size_t b = p;
        while (p < s.size() && !std::isspace((unsigned char)s[p])) p++;
        out = s.substr(b, p - b);
        return true;
    }
    void restOfLine()
    {
        size_t k = 0;
        while (p < s.size() && k < 999) {
            char c = s[p++];
            k++;
            if (c == '\n') break;
        }
    }
};

std::vector<int> qubit_ids(const std::string& tok)
{
    std::vector<int> ids;
    for (size_t i = 0; i < tok.size(); i++) {
        if (std::isdigit((unsigned char)tok[i])) {
            long x = 0;
            while (i < tok.size() && std::isdigit((unsigned char)tok[i])) x = x * 10 + (tok[i++] - '0');
            i--;
            ids.push_back((int)x);
        }
    }
    return ids;
}

bool gate_params(const std::string& tok, std::string& name, std::vector<double>& params)
{
    const double pi = std::acos(-1.0);
    size_t open = tok.find('(');
    name = tok.substr(0, open);
    params.clear();
    if (open == std::string::npos) return true;
    size_t i = open + 1;
    while (true) {
        size_t e = tok.find_first_of(",)", i);
        if (e == std::string::npos) return false;
        std::string st = tok.substr(i, e - i);
        double scale = 1;
        bool divide = false;
        if (st.compare(0, 3, "pi*") == 0) { scale = pi; st = st.substr(3); }
        else if (st.compare(0, 3, "pi/") == 0) { divide = true; st = st.substr(3); }
        char* end = nullptr;
        double v = std::strtod(st.c_str(), &end);
        if (end == st.c_str()) return false;
        params.push_back(divide ? pi / v : scale * v);
        if (tok[e] == ')') break;
        i = e + 1;
    }
    return true;
}

struct GateSpec {
    const char* name;
    GateType type;
    int qubits;
    int params;
};

const GateSpec kGates[] = {
    {"cx", GateType::CNOT, 2, 0}, {"ccx", GateType::CCX, 3, 0}, {"cy", GateType::CY, 2, 0},
    {"cz", GateType::CZ, 2, 0},   {"h", GateType::H, 1, 0},     {"x", GateType::X, 1, 0},
    {"y", GateType::Y, 1, 0},     {"z", GateType::Z, 1, 0},     {"s", GateType::S, 1, 0},
    {"sdg", GateType::SDG, 1, 0}, {"t", GateType::T, 1, 0},     {"tdg", GateType::TDG, 1, 0},
    {"crx", GateType::CRX, 2, 1}, {"cry", GateType::CRY, 2, 1}, {"crz", GateType::CRZ, 2, 1},
    {"cu1", GateType::CU1, 2, 1}, {"u1", GateType::U1, 1, 1},   {"u3", GateType::U3, 1, 3},
    {"rx", GateType::RX, 1, 1},   {"ry", GateType::RY, 1, 1},   {"rz", GateType::RZ, 1, 1},
};

}  // namespace

namespace hq {

hq_status parse_qasm(const std::string& text, std::unique_ptr<Circuit>* out)
{
    Reader r(text);
    std::string tok;
    std::unique_ptr<Circuit> c;
    while (r.token(tok)) {
        if (tok == "//" || tok == "OPENQASM" || tok == "include") {
            r.restOfLine();
            continue;
        }
        if (tok == "qreg") {
            if (r.p + 3 > r.s.size()) { set_error("qasm: truncated qreg"); return HQ_ERR_PARSE; }
            r.p += 3;
            const char* b = r.s.c_str() + r.p;
            char* end = nullptr;
            long n = std::strtol(b, &end, 10);
            if (end == b || n < 1 || n > 36) { set_error("qasm: bad qreg size"); return HQ_ERR_PARSE; }
            r.p += (size_t)(end - b);
            c.reset(new Circuit((int)n));
            r.restOfLine();
            continue;
        }
        std::string name;
        std::vector<double> params;
        if (!gate_params(tok, name, params)) { set_error("qasm: bad parameter list in " + tok); return HQ_ERR_PARSE; }
        const GateSpec* spec = nullptr;
        for (const auto& g : kGates)
            if (name == g.name && ((g.params == 0 && tok.find('(') == std::string::npos) || g.params > 0)) spec = &g;
        if (!spec) { set_error("qasm: unrecognized token " + tok); return HQ_ERR_PARSE; }
        if ((int)params.size() != spec->params) { set_error("qasm: wrong parameter count in " + tok); return HQ_ERR_PARSE; }
        if (!c) { set_error("qasm: gate before qreg"); return HQ_ERR_PARSE; }
        if (!r.token(tok)) { set_error("qasm: missing operands"); return HQ_ERR_PARSE; }
        std::vector<int> q = qubit_ids(tok);
        if ((int)q.size() != spec->qubits) { set_error("qasm: wrong operand count in " + tok); return HQ_ERR_PARSE; }
        for (int x : q)
            if (x >= c->numQubits) { set_error("qasm: qubit out of range in " + tok); return HQ_ERR_PARSE; }
        int c1 = -1, c2 = -1, t;
        if (spec->qubits == 3) { c1 = q[0]; c2 = q[1]; t = q[2]; }
        else if (spec->qubits == 2) { c1 = q[0]; t = q[1]; }
        else t = q[0];
        Gate g;
        hq_status rc = make_gate((int)spec->type, c2, c1, t, params.data(), (int)params.size(), &g);
        if (rc != HQ_OK) { set_error("qasm: invalid gate " + tok); return rc; }
        if (g.controlQubit == g.targetQubit || (c2 >= 0 && (c2 == c1 || c2 == t))) {
            set_error("qasm: repeated qubit in " + tok);
            return HQ_ERR_PARSE;
        }
        c->addGate(g);
        r.restOfLine();
    }
    if (!c) { set_error("qasm: fail to load circuit (no qreg)"); return HQ_ERR_PARSE; }
    *out = std::move(c);
    return HQ_OK;
}

hq_status load_qasm(const std::string& path, std::unique_ptr<Circuit>* out)
{
    std::ifstream f(path);
    if (!f) { set_error("qasm: fail to open " + path); return HQ_ERR_ARG; }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_qasm(ss.str(), out);
}

// ---- synthetic workloads ---------------------------------------------------------------

void add_random_layers(Circuit& c, int layers, uint64_t seed)
{
    const int n = c.numQubits;
    Rng rng(seed);
    const double two_pi = 2 * std::acos(-1.0);
    std::vector<int> perm(n);
    for (int l = 0; l < layers; l++) {
        for (int q = 0; q < n; q++) {
            switch (rng.below(5)) {
            case 0: c.addGate(Gate::H(q)); break;
            case 1: c.addGate(Gate::RX(q, two_pi * rng.uniform())); break;
            case 2: c.addGate(Gate::RY(q, two_pi * rng.uniform())); break;
            case 3: c.addGate(Gate::RZ(q, two_pi * rng.uniform())); break;
            default: c.addGate(Gate::T(q)); break;
            }
        }
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) {
            int j = (int)rng.below((uint64_t)i + 1);
            std::swap(perm[i], perm[j]);
        }
        for (int i = 0; i + 1 < n; i += 2) c.addGate(Gate::CNOT(perm[i], perm[i + 1]));
    }
}

void add_hea_layers(Circuit& c, int layers, uint64_t seed)
{
    const int n = c.numQubits;
    Rng rng(seed);
    const double two_pi = 2 * std::acos(-1.0);
    for (int l = 0; l < layers; l++) {
        for (int q = 0; q < n; q++) {
            c.addGate(Gate::RY(q, two_pi * rng.uniform()));
            c.addGate(Gate::RZ(q, two_pi * rng.uniform()));
        }
        for (int q = 0; q + 1 < n; q++) c.addGate(Gate::CZ(q, q + 1));
    }
}

}  // namespace hq
