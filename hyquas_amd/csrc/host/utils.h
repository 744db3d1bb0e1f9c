// utils.h -- the process-wide setup drivers of the reference call before
// building circuits (MyGlobalVars, src/utils.h:146-157).  One GPU: no global
// qubits (bit = 0), the state lives on the current HIP device, every
// stream/handle the reference keeps here is owned by the state instead.
#pragma once

namespace MyGlobalVars {
extern int numGPUs;     // 1
extern int localGPUs;   // 1
extern int bit;         // log2(numGPUs) = 0: no global qubits
// Checks that a HIP device is visible (throws hq::Error otherwise).
void init();
}  // namespace MyGlobalVars
