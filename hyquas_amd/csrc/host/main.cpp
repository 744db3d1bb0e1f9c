// main.cpp -- command-line driver (main.cpp:233-251 of the reference):
//   hyquas_main [--mode serial|group|blas|mix] [--c64] [--reference] [--show-schedule] circuit.qasm
// parses the circuit, compiles, runs on the GPU, prints the state in the
// reference's format and the Logger lines.

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "circuit.h"

namespace hq {
hq_status load_qasm(const std::string& path, std::unique_ptr<Circuit>* out);
}

int main(int argc, char** argv)
{
    std::string path;
    int mode = HQ_MODE_MIX, precision = HQ_C128;
    bool reference = false, show = false;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--mode" && i + 1 < argc) {
            std::string m = argv[++i];
            mode = m == "serial" ? HQ_MODE_SERIAL : m == "group" ? HQ_MODE_GROUP : m == "blas" ? HQ_MODE_BLAS : HQ_MODE_MIX;
        } else if (a == "--c64") {
            precision = HQ_C64;
        } else if (a == "--reference") {
            reference = true;
        } else if (a == "--show-schedule") {
            show = true;
        } else {
            path = a;
        }
    }
    if (path.empty()) {
        printf("./hyquas_main [--mode serial|group|blas|mix] [--c64] [--reference] [--show-schedule] qasmfile\n");
        return 1;
    }
    Logger::enable(true);
    std::unique_ptr<Circuit> c;
    if (hq::load_qasm(path, &c) != HQ_OK) {
        printf("%s\n", hq_last_error());
        return 1;
    }
    c->params.mode = mode;
    c->params.precision = precision;
    if (reference) {
        c->referenceProfile = true;
        c->params.localSize = 10;
        c->params.blasInPlace = false;
    }
    try {
        c->compile();
        if (show) c->getSchedule().dump(c->numQubits);
        c->run();
        std::string s = c->stateText(show);
        fputs(s.c_str(), stdout);
    } catch (const hq::Error& e) {
        printf("error: %s\n", e.what());
        return 1;
    }
    Logger::print();
    return 0;
}
