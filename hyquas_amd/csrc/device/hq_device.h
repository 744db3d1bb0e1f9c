// hq_device.h -- internal types of the HIP device layer (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "hyquas_amd.h"
#include "hq_internal.h"

namespace hq {

// ---- errors ------------------------------------------------------------------
void set_error(const std::string& msg);
const char* get_error();

#define HQ_HIP(call)                                                               \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::hq::set_error(std::string(#call) + ": " + hipGetErrorString(e_));    \
            return HQ_ERR_HIP;                                                     \
        }                                                                          \
    } while (0)

#define HQ_CHECK(cond, code, msg)                                                  \
    do {                                                                           \
        if (!(cond)) { ::hq::set_error(msg); return (code); }                      \
    } while (0)

// ---- complex amplitude -----------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

// ---- state -------------------------------------------------------------------------
struct ProfileEntry {
    int kind;
    hipEvent_t start, stop;
};

}  // namespace hq

struct hq_state_s {
    int n = 0;
    int precision = HQ_C128;
    int device = 0;
    uint64_t amps = 0;
    size_t amp_bytes = 16;
    void* data = nullptr;     // current state vector
    void* buf = nullptr;      // work buffer (TransMM), may be null
    hipStream_t stream = nullptr;
    double* d_red = nullptr;  // reduction scratch (device)
    double* h_red = nullptr;  // reduction scratch (pinned host)
    size_t red_cap = 0;       // doubles
    bool profiling = false;
    std::vector<hq::ProfileEntry> prof;
    std::vector<hipEvent_t> free_events;
};

namespace hq {

// Profiling helpers: bracket one launch with events when enabled.
hq_status prof_begin(hq_state s, int kind, int* slot);
hq_status prof_end(hq_state s, int slot);

// ---- OShareMem program (device-resident) ---------------------------------------------
// Op classes of the register kernel.
enum ShmClass : uint8_t {
    SHM_SWAP = 0,   // lo <-> hi                       (X, CNOT, CCX)
    SHM_Y = 1,      // lo' = -i hi, hi' = i lo         (Y, CY)
    SHM_REAL = 2,   // real 2x2                        (H, RY, CRY, ...)
    SHM_RX = 3,     // real diagonal, imaginary off-diagonal (RX, CRX)
    SHM_GEN = 4,    // complex 2x2                     (U2, U3, ...)
    SHM_DIAG = 5,   // lo *= d0, hi *= d1              (RZ, CRZ, ...)
    SHM_PHASE = 6,  // hi *= d1                        (T, S, Z, U1, CZ, CU1, ...)
    // uncontrolled gates with a scalar factored out (the group's product of
    // factors is applied once, by a final SHM_SCALE op); t, u = m[0]
    SHM_HADD = 7,   // lo' = lo + hi, hi' = lo - hi               (H / r)
    SHM_ROTA = 8,   // lo' = lo - t hi, hi' = t lo + hi           (RY / cos)
    SHM_ROTB = 9,   // lo' = u lo - hi, hi' = lo + u hi           (RY / sin)
    SHM_RXA = 10,   // lo' = lo + i t hi, hi' = i t lo + hi       (RX / cos)
    SHM_RXB = 11,   // lo' = u lo + i hi, hi' = i lo + u hi       (RX / sin)
    SHM_SCALE = 12, // every amplitude *= (m[0], m[1])
    // rotations in specialised kernels shared by same-structure groups (one
    // form for every angle; the scalar factor goes to SHM_SCALE): three
    // shears, a = m[0], b = m[2], |a|, |b| <= 1
    SHM_SHEAR = 13,  // lo += a hi, hi += b lo, lo += a hi        (RY)
    SHM_SHEARX = 14, // lo += i a hi, hi -= i b lo, lo += i a hi  (RX)
};

enum ShmWhere : uint8_t { SHM_IN_REG = 0, SHM_IN_THREAD = 1, SHM_IN_BLOCK = 2 };

struct ShmOp {
    uint8_t cls;
    uint8_t where;      // ShmWhere of the target
    uint8_t tpos;       // register-index bit / thread-index bit / block-index bit
    uint8_t pad0;
    uint32_t creg;      // control mask over register-index bits
    uint32_t cthr;      // control mask over thread-index bits
    uint32_t flags;     // SHM_FLAG_*
    uint64_t cblk;      // control mask over block-index bits
    double m[8];        // coefficients, re/im interleaved, row-major 2x2
};
static_assert(sizeof(ShmOp) == 88, "ShmOp layout");

// the op's coefficients are fixed by its gate type (H, T, CNOT, ...), so a
// specialised kernel may embed them as constants
constexpr uint32_t SHM_FLAG_CONST = 1;

constexpr int SHM_MAX_TILE = 14;
constexpr int SHM_MAX_TB = 9;
constexpr int SHM_MAX_RB = 6;   // 6 only for complex64 (64 amplitudes = 128 VGPRs per thread)

struct ShmStage {
    uint32_t reg_bits[SHM_MAX_RB];  // tile bit of register-index bit i (RB used)
    uint32_t thr_bits[SHM_MAX_TB];  // tile bit of thread-index bit i (TB used)
    uint32_t op_begin, op_end;
};

struct ShmHeader {
    uint64_t block_phys[64];   // physical bit (as a mask) of block-index bit j
    uint64_t tile_phys[SHM_MAX_TILE];  // physical bit (as a mask) of tile bit i
    uint32_t swz_col[SHM_MAX_TILE];    // LDS swizzle column of tile bit i
    uint32_t num_block_bits;
    uint32_t num_stages;
    uint32_t num_ops;
    uint32_t pad;
    // tiles whose block address has a bit in zero_mask hold only zeros before
    // the pass (the run started from |0...0> and no earlier group reached
    // those qubits) and keep them: the pass skips them
    uint64_t zero_mask;
    // followed by ShmStage[num_stages], ShmOp[num_ops] (8-byte aligned)
};

// Host-side result of lowering + register-stage planning of one group.
struct ShmPlan {
    int n = 0, precision = 0, k = 0, rb = 0, tb = 0;
    bool init = false;        // start from |0...0> instead of loading (first group of a run)
    bool nt_load = false;     // specialised kernel: non-temporal HBM loads
    bool nt_store = false;    // specialised kernel: non-temporal HBM stores
    bool whole_line = false;  // planned with the whole-line HBM stages (io = 3)
    uint64_t shared_mask = 0;
    ShmHeader hdr;
    std::vector<ShmStage> stages;
    std::vector<ShmOp> ops;
};

}  // namespace hq

struct hq_program_s {
    int kind = 0;            // 1 sharemem, 2 blas (transpose + GEMM), 3 dense in place
    int n = 0;
    int precision = HQ_C128;
    // sharemem
    void* d_prog = nullptr;
    size_t prog_bytes = 0;
    int tile = 0, rb = 0, tb = 0, num_stages = 0, num_ops = 0;
    uint64_t shared_mask = 0;
    // blas
    bool identity_perm = true;
    std::vector<int> perm;
    void* d_targs = nullptr; // transpose launch arguments (device)
    void* d_mat = nullptr;   // K x K, column-major, split planes re | im, in state precision
    int mat_qubits = 0;
    uint64_t rows_mask = 0;  // dense in place: the qubits U acts on (row bit j = j-th lowest)
    // specialised (JIT) OShareMem kernel; null when the interpreter runs the group
    hipModule_t jit_module = nullptr;
    hipFunction_t jit_fn = nullptr;
    void* d_coef = nullptr;  // runtime coefficients of the specialised kernel
    hq::ShmPlan* plan = nullptr;  // kept for specialisation (owned)
};

namespace hq {
// Launchers implemented in the .hip files.
hq_status launch_sharemem(hq_state s, hq_program p);
hq_status transpose_prepare(hq_state s, const int* perm, void** dev_args);
hq_status transpose_launch(hq_state s, const void* dev_args, void* in, void* out);
hq_status launch_dense(hq_state s, const void* d_mat, int mat_qubits, void* in, void* out);
// U (rows in the order of the set bits of `rows`) on those qubits, in place.
hq_status launch_dense_at(hq_state s, const void* d_mat, int mat_qubits, uint64_t rows);
int sharemem_default_rb(int precision, int tile);
// io: bit 0 adds a leading, bit 1 a trailing stage without gates where the
// first / last stage holds one of the three lowest tile bits in registers
// (the HBM access then covers whole 128-byte lines per wave instruction).
// shared_code: op classes that do not depend on gate angles (rotations keep
// their full 2x2 form), so same-structure groups share a specialised kernel.
hq_status plan_sharemem(int n, int precision, const hq_gate* gates, int num_gates, uint64_t shared_mask,
                        int rb, ShmPlan* out, bool init = false, int io = 0, bool shared_code = false);
// Specialised kernel for a planned group: generate, compile (cached), load.
hq_status jit_sharemem(const ShmPlan& plan, hq_program p);
hq_status launch_sharemem_jit(hq_state s, hq_program p);
hq_status jit_sharemem_many(const std::vector<const ShmPlan*>& plans, const std::vector<hq_program>& progs,
                            int threads);
// batch: the batched form hq_group_batch(const BatchItem* items, unsigned count)
// (blocks map to (circuit, tile) by the items' first_block, the item's prog
// field holds the group's coefficient array); readout: it also folds the
// tile's sum p, p Z, p ZZ into the item's partials (the circuit's last group).
std::string jit_source(const ShmPlan& plan, std::vector<double>* coef, bool batch = false, bool readout = false);
// Compile a generated source (cached like every specialised kernel) and load
// its kernel `name` as a module of the current device.
hq_status jit_compile_load(const std::string& src, const char* name, hipModule_t* mod, hipFunction_t* fn);
}  // namespace hq
