// A driver written against the reference's public C++ API the way its own
// drivers are (main.cpp:233-251, micro-benchmark/*.cpp): MyGlobalVars::init,
// Circuit, Gate factories and Gate::control, compile, run, printState,
// ampAt, Logger.  Builds with this library's host headers unchanged.
//
//   ref_style_driver <n>   -- prints the state of a fixed n-qubit circuit
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "circuit.h"
#include "logger.h"

int main(int argc, char* argv[])
{
    MyGlobalVars::init();
    int n = argc > 1 ? atoi(argv[1]) : 12;
    Circuit c(n);
    for (int i = 0; i < n; i++) c.addGate(Gate::H(i));
    for (int i = 0; i + 1 < n; i++) c.addGate(Gate::control(i, i + 1, GateType::CNOT));
    for (int i = 0; i < n; i++) c.addGate(Gate::RY(i, 0.1 * (i + 1)));
    for (int i = 0; i + 2 < n; i += 2) c.addGate(Gate::CCX(i, i + 1, i + 2));
    c.addGate(Gate::CU1(0, n - 1, M_PI / 3));
    c.addGate(Gate::U3(1, 0.3, 0.2, 0.1));
    c.addGate(Gate::TDG(2));
    c.compile();
    int time = c.run();
    c.printState();
    ResultItem r = c.ampAt(3);
    fprintf(stderr, "%s %s amp[3] = %.17g %.17g, %d us\n", Gate::get_name(GateType::CNOT).c_str(),
            Gate::get_name(GateType::CCX).c_str(), r.amp.x, r.amp.y, time);
    Logger::print();
    return 0;
}
