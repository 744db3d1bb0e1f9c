// hq_internal.h -- device-layer entry points used by the host layer that are
// not part of the C ABI (no HIP types, so host files build with any C++17 compiler).
#pragma once

#include "hyquas_amd.h"

#include <string>

namespace hq {
void set_error(const std::string& msg);
// hq_program_create_sharemem with an explicit register/thread split
// (reg_qubits = 0: default for the tile size).
hq_status build_sharemem_program(hq_state s, const hq_gate* gates, int num_gates, uint64_t shared_mask,
                                 int reg_qubits, hq_program* out);
}  // namespace hq
