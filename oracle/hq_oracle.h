/*
 * hq_oracle.h -- CPU FP64 restatement of the reference simulator's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (hyquas_amd/) links,
 * loads or calls this code.  It is used by tests/ (as the checker),
 * __graft_entry__.smoke() (as the checker) and bench.py (as the timed
 * CPU baseline "port").  Every function cites the reference file:line
 * whose behaviour it restates (paths relative to the HyQuas source tree).
 *
 * State vectors are interleaved complex128: psi[2*i] = Re, psi[2*i+1] = Im,
 * indexed by the PHYSICAL amplitude index.
 *
 * Parity status of each function (how it is pinned) is listed in the
 * header comment of hq_oracle.c.
 */
#ifndef HQ_ORACLE_H
#define HQ_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate type codes: same order as GateType in src/gate.h:7-9. */
enum {
    ORA_CCX = 0, ORA_CNOT, ORA_CY, ORA_CZ, ORA_CRX, ORA_CRY, ORA_CU1, ORA_CRZ,
    ORA_U1, ORA_U2, ORA_U3, ORA_H, ORA_X, ORA_Y, ORA_Z, ORA_S, ORA_SDG, ORA_T,
    ORA_TDG, ORA_RX, ORA_RY, ORA_RZ, ORA_TOTAL, ORA_ID, ORA_GII, ORA_GZZ,
    ORA_GOC, ORA_GCC
};

/* One gate; same memory layout as hq_gate in include/hyquas_amd.h. */
typedef struct ora_gate {
    int32_t type;
    int32_t target;
    int32_t control;   /* -1: none */
    int32_t control2;  /* -1: none */
    double mat[8];     /* row-major 2x2: re00 im00 re01 im01 re10 im10 re11 im11 */
} ora_gate;

/* ---- gates (src/gate.cpp:9-342, src/gate.h:27-29) ---------------------- */
int  ora_gate_make(int type, int control2, int control, int target,
                   const double *params, int nparams, ora_gate *out);
int  ora_is_diagonal(int type);

/* ---- plain per-gate path (src/kernelSimple.cu:9-446) -------------------- */
void ora_state_zero(double *psi, int n);
int  ora_apply_gate(double *psi, int n, const ora_gate *g);
int  ora_run_plain(double *psi, int n, const ora_gate *g, int m, int nthreads);
int  ora_max_threads(void);

/* ---- brute-force dense unitary, n <= 10 (independent pin) --------------- */
int  ora_dense_unitary(int n, const ora_gate *g, int m, double *U);

/* ---- OShareMem group semantics (src/executor.cpp:190-403,462-493,585-632,
 *      src/kernelOpt.cu:214-433), single GPU --------------------------- */
int  ora_sharemem_group(double *psi, int n, const ora_gate *g, int m,
                        uint64_t shared_mask);

/* ---- TransMM pieces (src/schedule.cpp:578-702, src/executor.cpp:533-551) */
int  ora_fuse_group(const ora_gate *g, int m, const int *pos, int mat_qubit,
                    double *mat_colmajor);
int  ora_transpose(const double *in, double *out, int n, const int *perm);
int  ora_apply_dense_runs(const double *in, double *out, int n,
                          const double *mat_colmajor, int mat_qubit);

/* ---- partitioner (src/compiler.cpp:220-452, src/schedule.cpp:37-51,
 *      339-382, src/evaluator.cpp:111-229) ----------------------------- */
enum { ORA_BACKEND_PERGATE = 1, ORA_BACKEND_BLAS = 2 };
enum { ORA_MODE_GROUP = 1, ORA_MODE_BLAS = 3, ORA_MODE_MIX = 4 };

typedef struct ora_partition_params {
    int mode;            /* ORA_MODE_*                                       */
    int local_size;      /* LOCAL_QUBIT_SIZE (10)                            */
    int coalesce;        /* COALESCE_GLOBAL (3)                              */
    int min_mat;         /* MIN_MAT_SIZE (4)                                 */
    int blas_mat_limit;  /* BLAS_MAT_LIMIT (MAT, 6 or 7)                     */
    int mat_lo, mat_hi;  /* mix-mode matSize loop: mat_lo <= m < mat_hi (4,8) */
    int window;          /* gate window of getGroupOpt (512)                 */
    /* cost model (Evaluator), only used by ORA_MODE_MIX                     */
    double single_cost[ORA_TOTAL];   /* pergate_single_perf[n][type][1]      */
    double ctr_cost[ORA_TOTAL];      /* pergate_ctr_perf[n][type][0][2]      */
    double group_overhead;           /* pergate_group_overhead * 2^n         */
    double blas_cost[64];            /* BLAS_perf[n][matSize]                */
    double cutt_cost;                /* cutt_cost[n]                         */
} ora_partition_params;

/* Defaults of the reference build (utils.h:56-61, CMakeLists.txt) and its
 * built-in V100 cost tables at n = 28 (evaluator.h:18-117, evaluator.cpp:5-35). */
void ora_partition_defaults(ora_partition_params *p, int n, int mode);

/* Output arrays are caller-allocated with capacity max_groups (and
 * max_groups*n for perms).  order[] receives the gate indices group by
 * group; group_start[g]..group_start[g+1] delimits group g.
 * Returns the number of groups, or a negative value on error.           */
int  ora_partition(int n, const ora_gate *g, int m, const ora_partition_params *p,
                   int max_groups, int *order, int *group_start, int *backend,
                   uint64_t *related, int *mat_qubit, int *perms,
                   int *final_pos, int *final_layout);

/* ---- readout (src/circuit.cpp:95-118,250-309, src/circuit.h:14-16,
 *      src/utils.cpp:77-84) ------------------------------------------- */
size_t ora_print_state(const double *psi, int n, const int *pos,
                       int show_schedule, char *buf, size_t cap);
double ora_zero_wrapper(double x);

/* ---- OpenQASM-subset parser (main.cpp:12-231) --------------------------- */
/* Returns number of gates (>=0) and sets *num_qubits, or -1 on a parse error. */
int  ora_parse_qasm(const char *text, int *num_qubits, ora_gate *out, int cap);

#ifdef __cplusplus
}
#endif
#endif
