// hq_internal.h -- device-layer entry points used by the host layer that are
// not part of the C ABI (no HIP types, so host files build with any C++17 compiler).
#pragma once

#include "hyquas_amd.h"

#include <string>
#include <vector>

namespace hq {
void set_error(const std::string& msg);
// hq_program_create_sharemem with an explicit register/thread split
// (reg_qubits = 0: default for the tile size); init: the group builds
// |0...0> in registers instead of loading the state; variant (specialised
// kernel) bit 0: whole-line HBM stages (added stages without gates where
// needed); bit 1: non-temporal stores; bit 2: non-temporal loads.
hq_status build_sharemem_program(hq_state s, const hq_gate* gates, int num_gates, uint64_t shared_mask,
                                 int reg_qubits, hq_program* out, bool init = false, int variant = 0,
                                 uint64_t zero_mask = 0);

// Many OShareMem programs at once; with jit, every group also gets a
// structure-specialised kernel, compiled in parallel.  tune_mode (jit only):
// 1 compiles two program variants of every group (plain stages; whole-line
// stages with non-temporal loads and stores), times them on the state (its
// contents are overwritten) and keeps the faster; 2 uses the whole-line
// variant for every group; 0 the plain one.
struct SharememRequest {
    std::vector<hq_gate> gates;
    uint64_t mask = 0;
    bool init = false;   // first group of a run: build |0...0> instead of loading the state
    uint64_t zero_mask = 0;  // physical qubits still at |0> before this group (their tiles are skipped)
};
// Smallest state for hq_program_create_dense with `mat_qubits` matrix qubits
// (one whole tile of the in-place GEMM); mat_qubits outside 4..7: never.
int dense_at_min_qubits(int precision, int mat_qubits);

hq_status build_sharemem_programs(hq_state s, const std::vector<SharememRequest>& reqs, int reg_qubits, bool jit,
                                  std::vector<hq_program>* out, int tune_mode = 0);
}  // namespace hq
