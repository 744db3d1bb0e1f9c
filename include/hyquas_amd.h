/*
 * hyquas_amd.h -- C ABI of the MI355X state-vector simulator.
 *
 * Two layers, both plain C (no C++/HIP/torch types cross this boundary):
 *
 *  1. Device layer (hq_state_* / hq_apply_* / hq_program_*): the functions
 *     the reference's executor reaches through src/kernel.h:11-32 plus its
 *     cuTT/cuBLAS calls.  Qubit numbers here are PHYSICAL bit positions of
 *     the amplitude index.
 *  2. Circuit layer (hq_circuit_*): the public Circuit/Gate API of
 *     src/circuit.h:20-46 and src/gate.h:30-59 for hosts that are not C++
 *     (ctypes, cgo, JNI).  Qubit numbers here are LOGICAL.
 *
 * Every entry point returns an hq_status; none exits the process (the
 * reference calls exit(1) on any error, src/utils.h:106-144).  The text of
 * the most recent error on the calling thread is in hq_last_error().
 *
 * Threading: a state or circuit handle must be used by one host thread at
 * a time; distinct handles are independent (each state owns its stream and
 * device buffers, so nothing is shared through a global gate table as in
 * src/kernelOpt.cu:15).
 */
#ifndef HYQUAS_AMD_H
#define HYQUAS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
typedef enum hq_status {
    HQ_OK = 0,
    HQ_ERR_ARG = 1,          /* invalid argument (qubit out of range, ...)      */
    HQ_ERR_NO_DEVICE = 2,    /* no HIP device visible                           */
    HQ_ERR_HIP = 3,          /* a HIP runtime call failed                       */
    HQ_ERR_OOM = 4,          /* device or host allocation failed                */
    HQ_ERR_PARSE = 5,        /* malformed circuit text                          */
    HQ_ERR_STATE = 6,        /* call out of order (e.g. run before compile)     */
    HQ_ERR_UNSUPPORTED = 7   /* valid request this build does not implement     */
} hq_status;

/* ---- gate types: the values of GateType, src/gate.h:7-9 ------------------ */
enum {
    HQ_CCX = 0, HQ_CNOT, HQ_CY, HQ_CZ, HQ_CRX, HQ_CRY, HQ_CU1, HQ_CRZ,
    HQ_U1, HQ_U2, HQ_U3, HQ_H, HQ_X, HQ_Y, HQ_Z, HQ_S, HQ_SDG, HQ_T, HQ_TDG,
    HQ_RX, HQ_RY, HQ_RZ, HQ_TOTAL, HQ_ID, HQ_GII, HQ_GZZ, HQ_GOC, HQ_GCC
};

/* ---- amplitude precision ------------------------------------------------- */
enum { HQ_C128 = 0,   /* complex128 state (qreal = double, USE_DOUBLE)       */
       HQ_C64 = 1 };  /* complex64 state (qreal = float), FP64 host fusion   */

/* ---- schedule backends (Backend, src/schedule.h:9-11) -------------------- */
enum { HQ_BACKEND_NONE = 0, HQ_BACKEND_PERGATE = 1, HQ_BACKEND_BLAS = 2 };

/* ---- compile modes (the CMake BACKEND option, CMakeLists.txt:16-42) ------- */
enum { HQ_MODE_SERIAL = 0,   /* plain per-gate kernels                       */
       HQ_MODE_GROUP = 1,    /* OShareMem groups only                        */
       HQ_MODE_BLAS = 3,     /* TransMM groups only                          */
       HQ_MODE_MIX = 4 };    /* hybrid: cost model picks per group           */

/* A gate.  Mirrors Gate (src/gate.h:11-19) with the matrix in the order of
 * KernelGate (src/gate.h:75): re00 im00 re01 im01 re10 im10 re11 im11. */
typedef struct hq_gate {
    int32_t type;       /* HQ_* gate type                                   */
    int32_t target;
    int32_t control;    /* -1 if none                                       */
    int32_t control2;   /* -1 if none (only CCX has two controls)           */
    double  mat[8];
} hq_gate;

/* Opaque handles. */
typedef struct hq_state_s   *hq_state;
typedef struct hq_program_s *hq_program;
typedef struct hq_circuit_s *hq_circuit;
typedef struct hq_engine_s  *hq_engine;
typedef struct hq_batch_s   *hq_batch;

/* ======================================================================== */
/* Library                                                                  */
/* ======================================================================== */

/* Text of the last error on this thread ("" if none). */
const char *hq_last_error(void);

/* Number of visible HIP devices (replaces the device query inside
 * MyGlobalVars::init, src/utils.cpp:17-24).  HQ_ERR_NO_DEVICE if none. */
hq_status hq_device_count(int *count);

/* Build a gate from its type and parameters: the Gate factories of
 * src/gate.cpp:9-342 (Gate::H(t), Gate::RX(t, angle), Gate::CCX(c1, c2, t)...).
 * params: RX/RY/RZ/U1/CRX/CRY/CRZ/CU1 one angle; U2 (phi, lambda);
 * U3 (theta, phi, lambda); GOC/GCC (re, im). */
hq_status hq_gate_make(int type, int control2, int control, int target,
                       const double *params, int nparams, hq_gate *out);

/* ======================================================================== */
/* Device layer                                                             */
/* ======================================================================== */

/* Allocate a 2^num_qubits state on the current device, with an equal-size
 * work buffer when with_buffer != 0 (needed by hq_apply_transpose /
 * hq_apply_dense).  Replaces the allocation half of kernelInit
 * (src/kernelSimple.cu:9-27), which doubles the allocation for blas/mix. */
hq_status hq_state_create(int num_qubits, int precision, int with_buffer, hq_state *out);

/* Free a state (kernelDestroy, src/kernelSimple.cu:528-529). */
hq_status hq_state_destroy(hq_state s);

/* Set |0...0> (the zero + amp[0]=1 half of kernelInit,
 * src/kernelSimple.cu:21-31).  Asynchronous on the state's stream. */
hq_status hq_state_reset(hq_state s);

/* Copy the whole state to/from host memory; count must be 2^n amplitudes of
 * the state's precision (interleaved re, im).  Synchronous.  Download
 * replaces kernelDeviceToHost (src/kernelSimple.cu:524-526). */
hq_status hq_state_download(hq_state s, void *host, uint64_t count);
hq_status hq_state_upload(hq_state s, const void *host, uint64_t count);

/* One amplitude, widened to double (kernelGetAmp, src/kernelSimple.cu:518-522). */
hq_status hq_state_get_amp(hq_state s, uint64_t index, double *re, double *im);

/* Wait for all work queued on the state (Executor::finalize's stream sync,
 * src/executor.cpp:634-640). */
hq_status hq_state_sync(hq_state s);

/* Number of qubits / precision of a state. */
hq_status hq_state_info(hq_state s, int *num_qubits, int *precision);

/* Whether the state was created with a work buffer (TransMM groups need one). */
hq_status hq_state_has_buffer(hq_state s, int *has_buffer);

/* Plain method: one HBM pass per gate, in order (kernelExecSimple,
 * src/kernelSimple.cu:335-446).  Asynchronous. */
hq_status hq_apply_gates_plain(hq_state s, const hq_gate *gates, int num_gates);

/* OShareMem group: compile the gates (physical qubits) into a device-resident
 * program for the tile spanned by shared_mask (physical positions held in
 * LDS/registers; every non-diagonal target must be in it).  Replaces the
 * lowering + copyGatesToSymbol of src/executor.cpp:462-493 and
 * src/kernelOpt.cu:499-501.  The program stays valid until destroyed and may
 * be launched on any state with the same qubit count and precision. */
hq_status hq_program_create_sharemem(hq_state s, const hq_gate *gates, int num_gates,
                                     uint64_t shared_mask, hq_program *out);

/* TransMM group as one program: bit-permutation transpose by perm (output
 * bit i takes input bit perm[i], as cuttPlan is called at
 * src/schedule.cpp:767-774) followed by the column-major K x K matrix
 * (K = 2^mat_qubits, FP64, re/im interleaved, mat[(col*K + row)*2]) applied
 * to every run of K consecutive amplitudes (cublasZgemm at
 * src/executor.cpp:544-549).  perm may be NULL for the identity.  Needs a
 * state created with a buffer. */
hq_status hq_program_create_blas(hq_state s, const int *perm, const double *mat_colmajor,
                                 int mat_qubits, hq_program *out);

/* A TransMM group's matrix applied where its qubits are, in one pass and in
 * place: no transpose, no work buffer, the layout unchanged.  Replaces the
 * cuttPlan + cublasZgemm pair above (src/executor.cpp:533-551) when the
 * caller does not need the reference's permuted layout.  Matrix index bit j
 * (row and column, same format as hq_program_create_blas) is the state's
 * qubit positions[j]; mat_qubits in 4..7 and a state of at least one tile:
 * mat_qubits + 5 qubits in complex128 (+ 4 at 7), + 6 in complex64 (+ 5 at
 * 7); 12 qubits always suffice. */
hq_status hq_program_create_dense(hq_state s, const int *positions, const double *mat_colmajor,
                                  int mat_qubits, hq_program *out);

/* Launch a program on a state (launchExecutor, src/kernelOpt.cu:503-506, or
 * applyBlasGroup, src/executor.cpp:533-551).  Asynchronous. */
hq_status hq_program_launch(hq_state s, hq_program p);
hq_status hq_program_destroy(hq_program p);

/* Replace the interpreter of an OShareMem program by a kernel specialised to
 * its structure (register/thread/block bit layout, pairs, controls), emitted
 * as HIP source and compiled for gfx950 with hiprtc; parametric angles stay
 * runtime data.  Compiled code is cached (HQ_JIT_CACHE, default
 * ~/.cache/hyquas_amd/jit).  No-op for TransMM programs.  This build's
 * replacement of the compile-time specialisation the reference gets from
 * its per-type kernel switch (src/kernelOpt.cu:214-386). */
hq_status hq_program_specialize(hq_program p);
hq_status hq_program_is_specialized(hq_program p, int *yes);

/* Host only (no device needed): plan an OShareMem group and emit the HIP
 * source of its specialised kernel (extern "C" hq_group(C *psi, const R *cf),
 * one block of 2^thread_qubits threads per tile, 2^(n - tile) blocks).
 * coef receives the runtime coefficient array (cf), shape[0..2] the tile,
 * register and thread qubit counts.  variant bit 0: 0 is the plain kernel,
 * 1 the whole-line non-temporal one (stages without gates added where the
 * first or last stage would access HBM in pieces shorter than 128 bytes);
 * the two kernels the "tune" option chooses between; bit 5: non-temporal
 * stores only (loads through the caches), on the plain or the whole-line
 * stages; bit 1: the kernel of a
 * run's first group, which builds |0...0> instead of loading the state;
 * bit 2: the batched form of the engine, hq_group_batch(items, count), whose
 * blocks map to (circuit, tile) through an item table (40-byte items:
 * state, coefficients, partials, first block, qubits, mode; mode bit 1 = no
 * store); bit 3 (with bit 2): the circuit's last group, which also writes
 * each tile's sum p, sum p Z and sum p ZZ to the item's partials; bit 4:
 * op classes independent of the gate angles (same-structure groups share
 * one source; RY and RX as three shears, their scalars in the final scale
 * op).  For inspection, caching and CPU tests. */
hq_status hq_sharemem_source(int num_qubits, int precision, const hq_gate *gates, int num_gates,
                             uint64_t shared_mask, int reg_qubits, int variant, char *buf, size_t cap,
                             size_t *len, double *coef, size_t coef_cap, size_t *num_coef, int *shape);

/* Describe a compiled OShareMem program: tile qubits, register stages and
 * lowered op count (for tests and the cost model). */
hq_status hq_program_info(hq_program p, int *tile_qubits, int *num_stages, int *num_ops);

/* One-shot forms of the above (create + launch + destroy). */
hq_status hq_apply_sharemem(hq_state s, const hq_gate *gates, int num_gates, uint64_t shared_mask);
hq_status hq_apply_transpose(hq_state s, const int *perm);                  /* cuttExecute  */
hq_status hq_apply_dense(hq_state s, const double *mat_colmajor, int mat_qubits); /* cublasZgemm */

/* Probability that qubit `target` (physical) reads 0 (kernelMeasure,
 * src/kernelSimple.cu:497-516).  FP64 accumulation.  Synchronous. */
hq_status hq_measure_zero(hq_state s, int target, double *p0);

/* Squared norm sum |a_i|^2, FP64 accumulation (Kahan-compensated per thread
 * for complex64 states).  Synchronous. */
hq_status hq_state_norm(hq_state s, double *norm2);

/* <Z_q> for every physical qubit q (z[n]) and <Z_q Z_{q+1}> (zz[n-1]); either
 * pointer may be NULL.  One HBM pass, FP64 accumulation.  Synchronous. */
hq_status hq_expect_z(hq_state s, double *z, double *zz);

/* hq_expect_z without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The per-block partials and their
 * column sums live in a fixed device array of the library (257 rows of 73
 * doubles, 147 KiB, one copy per device, held until the process ends) instead
 * of the state's lazily allocated reduction buffers; the caller passes the
 * state and the output arrays as for hq_expect_z.  Same values as
 * hq_expect_z bit for bit (same partials, same summation order).  Every call
 * on one device shares that device's array: calls that overlap on one device
 * (several host threads, or states on different streams) race.  States of at
 * most 36 qubits.  Synchronous. */
hq_status hq_expect_z_noalloc(hq_state s, double *z, double *zz);

/* Overlap <a|b> of two states with the same qubit count (any precisions),
 * FP64 accumulation.  Synchronous. */
hq_status hq_state_inner(hq_state a, hq_state b, double *re, double *im);

/* Device-side timing of the most recent launches (HIP events on the state's
 * stream).  enable != 0 starts recording one event pair per program launch
 * or plain gate batch; hq_profile_read returns up to cap entries of
 * {kind, ms} (kind: 0 plain, 1 sharemem, 2 transpose, 3 dense) and clears
 * the log.  Synchronizes the stream. */
hq_status hq_profile_enable(hq_state s, int enable);
hq_status hq_profile_read(hq_state s, int *kinds, double *ms, int cap, int *count);

/* ======================================================================== */
/* Circuit layer (src/circuit.h:20-46)                                      */
/* ======================================================================== */

hq_status hq_circuit_create(int num_qubits, hq_circuit *out);          /* Circuit(n)   */
hq_status hq_circuit_destroy(hq_circuit c);
hq_status hq_circuit_add_gate(hq_circuit c, const hq_gate *g);         /* addGate      */
hq_status hq_circuit_num_gates(hq_circuit c, int *count);
hq_status hq_circuit_num_qubits(hq_circuit c, int *num_qubits);

/* Parse the OpenQASM subset of main.cpp:68-231 from a file or a string. */
hq_status hq_circuit_load_qasm(const char *path, hq_circuit *out);
hq_status hq_circuit_parse_qasm(const char *text, hq_circuit *out);

/* Options (the compile-time CMake options of the reference, runtime here):
 *   "mode"        HQ_MODE_* (default HQ_MODE_MIX)
 *   "precision"   HQ_C128 / HQ_C64 (default HQ_C128)
 *   "profile"     0 = gfx950 parameters (default), 1 = reference parameters
 *                 (LOCAL_QUBIT_SIZE 10, COALESCE 3, V100 cost tables)
 *   "share_qubits" OShareMem tile qubits (default with specialised kernels:
 *                 13 for complex128, 14 for complex64; interpreter: 12; 7..14)
 *   "coalesce"    COALESCE_GLOBAL (default 4; the reference uses 3)
 *   "mat_min", "mat_max"  TransMM group qubit range in mix mode (4, 7)
 *   "blas_mat"    BLAS_MAT_LIMIT for HQ_MODE_BLAS (default 6)
 *   "min_mat"     MIN_MAT_SIZE (default 4)
 *   "reg_qubits"  register qubits per thread in OShareMem (0 = auto)
 *   "window"      gates considered per group (default 512)
 *   "jit"         specialised OShareMem kernels: 0 off, 1 on, -1 auto
 *                 (default: on from 22 qubits)
 *   "tune"        per-group choice, timed on the state at compile, between
 *                 the plain specialised kernel and a whole-line non-temporal
 *                 one: 0 off, 1 on, 2 the whole-line kernel everywhere
 *                 (untimed), -1 auto (default: on with specialised kernels
 *                 from 26 qubits); timing overwrites the state
 *   "transmm_inplace" TransMM groups: 1 (default) planned and run as one
 *                 in-place dense pass over the group's qubits where they are
 *                 (no transpose, no work buffer, the layout unchanged, so later
 *                 OShareMem tiles coalesce on the qubits actually lowest); 0
 *                 the reference's transpose + GEMM and permuted layouts
 *                 (src/schedule.cpp:339-382, src/executor.cpp:533-551;
 *                 "profile" = 1 sets 0); groups with matrices beyond 7 qubits
 *                 or states below one dense tile use the transpose form
 *   "profile_events" 1 = time every launch with HIP events
 * Must be set before compile. */
hq_status hq_circuit_set_option(hq_circuit c, const char *key, int64_t value);

/* Load an MI355X cost-model parameter file (the format of
 * evaluator-preprocess/parameter-files, extended; see hq_calibrate). */
hq_status hq_circuit_load_cost_model(hq_circuit c, const char *path);

/* Partition + lower + fuse (Circuit::compile, src/circuit.cpp:177-210).
 * Uploads every group program to the device. */
hq_status hq_circuit_compile(hq_circuit c);

/* The partition half of compile only (Compiler::run, src/compiler.cpp:70-165):
 * host work, no device needed; hq_circuit_schedule* then describe it. */
hq_status hq_circuit_partition(hq_circuit c);

/* Compile-once cache of the partition: the schedule in the reference's byte
 * layout (Schedule::serialize, src/schedule.cpp:156-325, src/gate.cpp:558-589)
 * after a 16-byte header ("HQSCHED1", qubit count).  Save needs a partitioned
 * or compiled circuit; load checks every gate against c (type, qubits,
 * matrix, each index once) and builds the device programs, as compile does. */
hq_status hq_circuit_save_schedule(hq_circuit c, const char *path);
hq_status hq_circuit_load_schedule(hq_circuit c, const char *path);

/* Circuit::run (src/circuit.cpp:16-71): allocate (first call) and reset the
 * state, run every group, wait.  *us receives the wall time of the group
 * loop in microseconds (the reference's "Time Cost").  copy_back keeps a
 * host copy for hq_circuit_amp_at / hq_circuit_print_state; destroy frees
 * the device state afterwards (keep it for hq_circuit_amp_at_gpu and the
 * device readouts). */
hq_status hq_circuit_run(hq_circuit c, int copy_back, int destroy, int64_t *us);

/* ampAt (host copy, needs copy_back) / ampAtGPU (device, needs destroy=0);
 * logical index (src/circuit.cpp:115-151). */
hq_status hq_circuit_amp_at(hq_circuit c, uint64_t index, double *re, double *im);
hq_status hq_circuit_amp_at_gpu(hq_circuit c, uint64_t index, double *re, double *im);

/* printState text (src/circuit.cpp:250-309) into buf (NUL-terminated when
 * it fits); *len receives the full length.  show_schedule adds the 1<<i and
 * 2^n-1 lines of the SHOW_SCHEDULE build. */
hq_status hq_circuit_print_state(hq_circuit c, int show_schedule, char *buf,
                                 size_t cap, size_t *len);

/* Borrow the device state after run(destroy=0) for the device readouts
 * (physical indices: combine with hq_circuit_final_pos). */
hq_status hq_circuit_state(hq_circuit c, hq_state *out);

/* Run c on a state the caller owns (a resident state reused by circuit after
 * circuit) instead of one that run() allocates: s must have c's qubit count
 * and, at compile/run time, c's precision and a work buffer if the schedule
 * has TransMM groups that run out of place (the transpose form).  c borrows s: run(destroy=1) and hq_circuit_destroy
 * leave it allocated, and s must outlive its use by c.  s = NULL returns c to
 * states of its own.  No counterpart in the reference, whose Circuit
 * allocates its state in run() (src/circuit.cpp:16-71). */
hq_status hq_circuit_use_state(hq_circuit c, hq_state s);

/* Logical->physical map of the state a run leaves, pos[n]: after compile the
 * layout the lowered groups produce (TransMM groups in place keep the layout;
 * in the transpose form, option "transmm_inplace" = 0, they leave the
 * partition's permuted layout, schedule.finalState.pos of the reference);
 * after partition only, the partition's final layout. */
hq_status hq_circuit_final_pos(hq_circuit c, int *pos);

/* Schedule summary after compile: number of groups, and per group its
 * backend, gate count, related (logical) qubit mask, TransMM matrix qubits,
 * OShareMem register stages; any array may be NULL. */
hq_status hq_circuit_schedule(hq_circuit c, int *num_groups, int cap, int *backend,
                              int *num_gates, uint64_t *related, int *mat_qubit,
                              int *num_stages);

/* Share of one full read + write of the state that each group's pass moves,
 * in schedule order (compiled circuits).  A run starts from |0...0>: the first
 * OShareMem pass builds it in registers and only writes (0.5); until some
 * tile has covered a qubit, tiles whose block address sets that qubit hold
 * zeros and the next passes skip them (the live tiles' share); 1 for every
 * other pass. */
hq_status hq_circuit_schedule_live(hq_circuit c, int cap, double *fraction, int *num_groups);

/* Gate order of the schedule: order[i] = index (in addGate order) of the
 * i-th gate executed; cap >= number of gates. */
hq_status hq_circuit_schedule_order(hq_circuit c, int *order, int cap);

/* Timing of the last run with option "profile_events" = 1: total device ms
 * per launch kind (kinds as in hq_profile_read). */
hq_status hq_circuit_last_times(hq_circuit c, double *plain_ms, double *sharemem_ms,
                                double *transpose_ms, double *dense_ms);

/* Per-launch device times of the last run with option "profile_events" = 1:
 * kinds[i] as in hq_profile_read, ms[i] its duration; *count = launches. */
hq_status hq_circuit_launch_times(hq_circuit c, int *kinds, double *ms, int cap, int *count);

/* The circuit's gates in addGate order (logical qubits). */
hq_status hq_circuit_get_gates(hq_circuit c, hq_gate *out, int cap, int *count);

/* Generators of the synthetic workloads (portable: xoshiro256** seeded by
 * SplitMix64, doubles as (x >> 11) * 2^-53).
 *  random layers: each layer puts one of {H, RX, RY, RZ, T} (uniform) on
 *    every qubit, theta ~ U[0, 2pi), then CNOTs on a uniformly random perfect
 *    matching (Fisher-Yates; pairs (p[2i], p[2i+1]) as (control, target)).
 *  hardware-efficient ansatz: per layer RY(theta), RZ(phi) on every qubit,
 *    then CZ(i, i+1) for i < n-1.
 * Append to c. */
hq_status hq_circuit_add_random_layers(hq_circuit c, int layers, uint64_t seed);
hq_status hq_circuit_add_hea_layers(hq_circuit c, int layers, uint64_t seed);

/* ---- many circuits: batching and serving -------------------------------------
 *
 * No counterpart in the reference, which runs one circuit per process
 * (src/main.cpp:233-251).  An engine compiles circuits on host threads
 * (OShareMem groups only; circuits of at most "whole_max" qubits are one group
 * whose tile is the whole state), admits them in submission order while their
 * device memory fits `hbm_budget` bytes, and advances every admitted circuit by
 * one group per round (hq_engine_step), all circuits of one tile shape in one
 * launch.  Circuits submitted while others run join at the next round.
 * Readouts per circuit: {norm, sum_q <Z_q>, sum_{q<n-1} <Z_q Z_{q+1}>} over
 * logical qubits (always) and, with HQ_OUT_TOPK, the 16 most probable basis
 * states (probability descending, ties by smaller index).  Reductions run in a
 * fixed order, so a result does not depend on the other circuits.  One host
 * thread drives an engine (its calls are not synchronised with each other); a
 * circuit that cannot fit the budget even alone fails instead of waiting; a HIP
 * error during a step is sticky (later steps return it). */
enum { HQ_OUT_EXPECT = 1, HQ_OUT_TOPK = 2 };
enum { HQ_JOB_QUEUED = 0, HQ_JOB_COMPILED = 1, HQ_JOB_ADMITTED = 2, HQ_JOB_DONE = 3, HQ_JOB_FAILED = 4 };

hq_status hq_engine_create(int precision, uint64_t hbm_budget, hq_engine *out);
hq_status hq_engine_destroy(hq_engine e);
/* "whole_max" (13), "share_qubits" (12; 13 with "jit" in complex128), "coalesce" (4), "compile_threads" (4,
 * before the first submit), "max_inflight" (2 rounds queued ahead of the host),
 * "jit" (0: interpreter; 1: specialised batched kernels), "streams" (4: the
 * streams a round's group launches spread over, before the first round or
 * batch; 1: one stream). */
hq_status hq_engine_set_option(hq_engine e, const char *key, int64_t value);
/* Specialised batched kernels the engine has compiled so far (option "jit";
 * one per distinct group structure).  Safe to call while another thread
 * compiles circuits into the engine (hq_batch_create, hq_engine_submit). */
hq_status hq_engine_kernels(hq_engine e, int64_t *compiled);

/* Tile qubits for circuits of num_qubits qubits (0: the whole_max/share_qubits
 * rule); hq_engine_get_tile reports the tile in effect. */
hq_status hq_engine_set_tile(hq_engine e, int num_qubits, int tile);
hq_status hq_engine_get_tile(hq_engine e, int num_qubits, int *tile);
/* Queue a copy of c's gates (7..34 qubits); outputs = HQ_OUT_* bits. */
hq_status hq_engine_submit(hq_engine e, hq_circuit c, int outputs, int64_t *job);
/* Retire finished rounds, admit what fits, launch one round.  Returns without
 * waiting for the round (at most max_inflight rounds are queued). *pending =
 * circuits submitted and not yet done. */
hq_status hq_engine_step(hq_engine e, int *pending);
/* Step until every submitted circuit is done. */
hq_status hq_engine_drain(hq_engine e);
/* Circuits finished (or failed) since the last poll. */
hq_status hq_engine_poll(hq_engine e, int64_t *jobs, int cap, int *count);
hq_status hq_engine_job_status(hq_engine e, int64_t job, int *status);
/* expect3 = {norm, sum Z, sum ZZ}; topk_* [16] (needs HQ_OUT_TOPK); times3 =
 * {submitted, admitted, done} in seconds on the engine clock.  A failed
 * circuit returns its error. */
hq_status hq_engine_result(hq_engine e, int64_t job, double *expect3, uint64_t *topk_idx, double *topk_prob,
                           double *times3);
hq_status hq_engine_release(hq_engine e, int64_t job);
hq_status hq_engine_stats(hq_engine e, uint64_t *bytes_used, uint64_t *bytes_peak, uint64_t *bytes_reserved,
                          int64_t *rounds);
hq_status hq_engine_now(hq_engine e, double *seconds);

/* A fixed set of circuits on an idle engine: compiled, resident and captured
 * as one hipGraph at creation; hq_batch_run replays it (device_ms = its
 * duration by HIP events) and results are read with hq_batch_results
 * (expect3[3 * count], topk_*[16 * count]). */
hq_status hq_batch_create(hq_engine e, const hq_circuit *circuits, int count, int outputs, hq_batch *out);
hq_status hq_batch_run(hq_batch b, double *device_ms);
hq_status hq_batch_results(hq_batch b, double *expect3, uint64_t *topk_idx, double *topk_prob);
hq_status hq_batch_info(hq_batch b, int *count, int *rounds, int *launches, uint64_t *device_bytes);
/* Algorithmic HBM bytes of one replay (no counterpart in the reference): every
 * pass reads and writes its circuit's state once, except that the first pass
 * only writes and later passes skip the tiles still known to be zero; divided
 * by device_ms this is the batched pass's roofline figure. */
hq_status hq_batch_pass_bytes(hq_batch b, uint64_t *bytes);
hq_status hq_batch_destroy(hq_batch b);

/* The k (<= 16) most probable physical basis states of s, probability
 * descending, ties by smaller index; FP64.  Synchronous. */
hq_status hq_state_topk(hq_state s, int k, uint64_t *idx, double *prob);

/* `shots` measurement samples of s (physical basis indices, in draw order):
 * shot k is the smallest i with sum_{j<=i} |a_j|^2 > u_k * sum_j |a_j|^2, u_k
 * the k-th uniform of the portable generator seeded with `seed`.  No
 * counterpart in the reference (it reads P(q = 0), src/kernelSimple.cu:497-516).
 * Synchronous. */
hq_status hq_state_sample(hq_state s, int shots, uint64_t seed, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* HYQUAS_AMD_H */
