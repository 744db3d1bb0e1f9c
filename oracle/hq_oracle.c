/*
 * hq_oracle.c -- CPU FP64 restatement of the reference simulator's hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the CPU
 * baseline leg of bench.py.  The product (hyquas_amd/) never links it.
 *
 * How each part is pinned (the reference ships its fixtures as Git-LFS
 * pointers with no content and cannot be compiled here, so no golden vector
 * or reference build exists; every function is pinned by an independent
 * check instead -- see tests/test_oracle.py):
 *   ora_gate_make        closed forms (textbook matrices), unitarity 1e-15.
 *   ora_apply_gate /     brute-force dense unitary (ora_dense_unitary is a
 *   ora_run_plain        Kronecker-definition product, a different
 *                        algorithm), known-answer circuits (H^n, GHZ, QFT of
 *                        a basis state, Bernstein-Vazirani), norm.
 *   ora_dense_unitary    closed forms on 1-2 qubits, unitarity.
 *   ora_sharemem_group   equality with ora_run_plain for every block/shared
 *                        split (block controls, block diagonal targets).
 *   ora_fuse_group       product of embedded gate matrices, unitarity.
 *   ora_transpose        bijectivity, identity perm, composition with its
 *                        inverse, single-swap closed form.
 *   ora_apply_dense_runs naive complex matmul on tiny inputs; equality with
 *                        per-gate application.
 *   ora_partition        invariants (every gate once, per-qubit order kept,
 *                        group qubit limits) + hand-traced small cases +
 *                        end-to-end amplitude equality.  Exact grouping on
 *                        large circuits: parity unpinned beyond those.
 *   ora_print_state      hand-written expected text for small states.
 *   ora_parse_qasm       hand-written circuits in the main.cpp dialect.
 *
 * Deliberate deviations from the reference, applied identically in the
 * product:
 *   - TDG and SDG apply their own matrices.  The reference types TDG as T
 *     (src/gate.cpp:239) and its plain path dispatches SDG to the S kernel
 *     (src/kernelSimple.cu:413-416), so its backends disagree with each
 *     other on circuits that use them.
 *   - CCX's second control counts as a dependency in the partitioner
 *     (src/compiler.cpp:344-363 and 425-441 ignore controlQubit2, which
 *     lets later gates on that qubit be scheduled before the CCX).
 *   - Malformed input returns an error instead of exit()/UB.
 */
#include "hq_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXQ 64

static int popcount64(uint64_t x) { return __builtin_popcountll(x); }

/* ------------------------------------------------------------------------ */
/* Gates: src/gate.cpp:9-342 (factories), src/gate.h:27-29 (isDiagonal).   */
/* ------------------------------------------------------------------------ */

int ora_is_diagonal(int type)
{
    /* src/gate.h:27-29 */
    switch (type) {
    case ORA_CZ: case ORA_CU1: case ORA_CRZ: case ORA_U1: case ORA_Z:
    case ORA_S: case ORA_SDG: case ORA_T: case ORA_TDG: case ORA_RZ:
        return 1;
    default:
        return 0;
    }
}

static int num_controls(int type)
{
    switch (type) {
    case ORA_CCX: return 2;
    case ORA_CNOT: case ORA_CY: case ORA_CZ: case ORA_CRX: case ORA_CRY:
    case ORA_CU1: case ORA_CRZ: return 1;
    default: return 0;
    }
}

static int num_params(int type)
{
    switch (type) {
    case ORA_CRX: case ORA_CRY: case ORA_CU1: case ORA_CRZ: case ORA_U1:
    case ORA_RX: case ORA_RY: case ORA_RZ: return 1;
    case ORA_U2: case ORA_GOC: case ORA_GCC: return 2;
    case ORA_U3: return 3;
    default: return 0;
    }
}

static void setm(double *m, double r00, double i00, double r01, double i01,
                 double r10, double i10, double r11, double i11)
{
    m[0] = r00; m[1] = i00; m[2] = r01; m[3] = i01;
    m[4] = r10; m[5] = i10; m[6] = r11; m[7] = i11;
}

int ora_gate_make(int type, int control2, int control, int target,
                  const double *p, int np, ora_gate *g)
{
    if (type < 0 || type > ORA_GCC || type == ORA_TOTAL) return -1;
    if (np != num_params(type)) return -1;
    int nc = num_controls(type);
    if ((nc >= 1) != (control >= 0)) return -1;
    if ((nc == 2) != (control2 >= 0)) return -1;
    g->type = type; g->target = target; g->control = control; g->control2 = control2;
    double *m = g->mat;
    const double r2 = 1 / sqrt(2);           /* host value, gate.cpp:156,229 */
    switch (type) {
    case ORA_CCX: case ORA_CNOT: case ORA_X:       /* gate.cpp:9-32,170-181 */
        setm(m, 0, 0, 1, 0, 1, 0, 0, 0); break;
    case ORA_CY: case ORA_Y:                       /* gate.cpp:34-44,183-193 */
        setm(m, 0, 0, 0, -1, 0, 1, 0, 0); break;
    case ORA_CZ: case ORA_Z:                       /* gate.cpp:46-56,195-205 */
        setm(m, 1, 0, 0, 0, 0, 0, -1, 0); break;
    case ORA_CRX: case ORA_RX: {                   /* gate.cpp:58-68,254-264 */
        double c = cos(p[0] / 2.0), s = sin(p[0] / 2.0);
        setm(m, c, 0, 0, -s, 0, -s, c, 0); break;
    }
    case ORA_CRY: case ORA_RY: {                   /* gate.cpp:70-80,266-276 */
        double c = cos(p[0] / 2.0), s = sin(p[0] / 2.0);
        setm(m, c, 0, -s, 0, s, 0, c, 0); break;
    }
    case ORA_CU1: case ORA_U1:                     /* gate.cpp:82-95,109-122 */
        setm(m, 1, 0, 0, 0, 0, 0, cos(p[0]), sin(p[0])); break;
    case ORA_CRZ: case ORA_RZ: {                   /* gate.cpp:97-107,278-288 */
        double c = cos(p[0] / 2), s = sin(p[0] / 2);
        setm(m, c, -s, 0, 0, 0, 0, c, s); break;
    }
    case ORA_U2: {                                 /* gate.cpp:124-136 */
        double phi = p[0], lam = p[1];
        setm(m, 1.0 / sqrt(2), 0, -cos(lam) / sqrt(2), -sin(lam) / sqrt(2),
             cos(phi) / sqrt(2), sin(phi) / sqrt(2),
             cos(lam + phi) / sqrt(2), sin(lam + phi) / sqrt(2));
        break;
    }
    case ORA_U3: {                                 /* gate.cpp:138-150 */
        double th = p[0], phi = p[1], lam = p[2];
        setm(m, cos(th / 2), 0, -cos(lam) * sin(th / 2), -sin(lam) * sin(th / 2),
             cos(phi) * sin(th / 2), sin(phi) * sin(th / 2),
             cos(phi + lam) * cos(th / 2), sin(phi + lam) * cos(th / 2));
        break;
    }
    case ORA_H:                                    /* gate.cpp:152-162 */
        setm(m, r2, 0, r2, 0, r2, 0, -r2, 0); break;
    case ORA_S:                                    /* gate.cpp:207-217 */
        setm(m, 1, 0, 0, 0, 0, 0, 0, 1); break;
    case ORA_SDG:                                  /* gate.cpp:219-229 */
        setm(m, 1, 0, 0, 0, 0, 0, 0, -1); break;
    case ORA_T:                                    /* gate.cpp:231-241 */
        setm(m, 1, 0, 0, 0, 0, 0, r2, r2); break;
    case ORA_TDG:                                  /* gate.cpp:243-252 */
        setm(m, 1, 0, 0, 0, 0, 0, r2, -r2); break;
    case ORA_ID:                                   /* gate.cpp:290-300 */
        setm(m, 1, 0, 0, 0, 0, 0, 1, 0); break;
    case ORA_GII:                                  /* gate.cpp:302-312 */
        setm(m, 0, 1, 0, 0, 0, 0, 0, 1); break;
    case ORA_GZZ:                                  /* gate.cpp:314-324 */
        setm(m, -1, 0, 0, 0, 0, 0, -1, 0); break;
    case ORA_GOC:                                  /* gate.cpp:326-336 */
        setm(m, 1, 0, 0, 0, 0, 0, p[0], p[1]); break;
    case ORA_GCC:                                  /* gate.cpp:338-348 */
        setm(m, p[0], p[1], 0, 0, 0, 0, p[0], p[1]); break;
    default:
        return -1;
    }
    return 0;
}

static int gate_check(const ora_gate *g, int n)
{
    if (g->type < 0 || g->type > ORA_GCC || g->type == ORA_TOTAL) return -1;
    int nc = num_controls(g->type);
    if (g->target < 0 || g->target >= n) return -1;
    if (nc >= 1) {
        if (g->control < 0 || g->control >= n || g->control == g->target) return -1;
    } else if (g->control != -1) return -1;
    if (nc == 2) {
        if (g->control2 < 0 || g->control2 >= n || g->control2 == g->target ||
            g->control2 == g->control) return -1;
    } else if (g->control2 != -1) return -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Plain per-gate path: src/kernelSimple.cu:9-37 (init), 40-332 (kernels),  */
/* 335-446 (dispatch).  Each gate is one pass over the 2^(n-1) pairs.       */
/* ------------------------------------------------------------------------ */

void ora_state_zero(double *psi, int n)
{
    /* kernelInit: memset + amp[0] = 1 (kernelSimple.cu:9-37) */
    size_t N = (size_t)1 << n;
    memset(psi, 0, 2 * N * sizeof(double));
    psi[0] = 1.0;
}

int ora_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Update of one (lo, hi) pair; formulas of the kernels at
 * kernelSimple.cu:74-332 (alpha/beta arguments as passed at 335-446). */
static inline void pair_update(const ora_gate *g, double *a, size_t lo, size_t hi)
{
    const double *m = g->mat;
    const double r2 = 1 / sqrt(2);
    double lr = a[2 * lo], li = a[2 * lo + 1], hr = a[2 * hi], hi_ = a[2 * hi + 1];
    switch (g->type) {
    case ORA_CCX: case ORA_CNOT: case ORA_X:       /* CCXKernel/CNOTKernel/XKernel */
        a[2 * lo] = hr; a[2 * lo + 1] = hi_; a[2 * hi] = lr; a[2 * hi + 1] = li;
        break;
    case ORA_CY: case ORA_Y:                       /* CYKernel/YKernel */
        a[2 * lo] = hi_; a[2 * lo + 1] = -hr; a[2 * hi] = -li; a[2 * hi + 1] = lr;
        break;
    case ORA_CZ: case ORA_Z:                       /* CZKernel/ZKernel */
        a[2 * hi] = -hr; a[2 * hi + 1] = -hi_;
        break;
    case ORA_CRX: case ORA_RX: {                   /* alpha=r00, beta=-i01 */
        double al = m[0], be = -m[3];
        a[2 * lo] = al * lr + be * hi_; a[2 * lo + 1] = al * li - be * hr;
        a[2 * hi] = al * hr + be * li;  a[2 * hi + 1] = al * hi_ - be * lr;
        break;
    }
    case ORA_CRY: case ORA_RY: {                   /* alpha=r00, beta=r10 */
        double al = m[0], be = m[4];
        a[2 * lo] = al * lr - be * hr; a[2 * lo + 1] = al * li - be * hi_;
        a[2 * hi] = be * lr + al * hr; a[2 * hi + 1] = be * li + al * hi_;
        break;
    }
    case ORA_CRZ: case ORA_RZ: {                   /* alpha=r00, beta=-i00 */
        double al = m[0], be = -m[1];
        a[2 * lo] = al * lr + be * li; a[2 * lo + 1] = al * li - be * lr;
        a[2 * hi] = al * hr - be * hi_; a[2 * hi + 1] = al * hi_ + be * hr;
        break;
    }
    case ORA_CU1: case ORA_U1: {                   /* alpha=r11, beta=i11 */
        double al = m[6], be = m[7];
        a[2 * hi] = al * hr - be * hi_; a[2 * hi + 1] = al * hi_ + be * hr;
        break;
    }
    case ORA_H:                                    /* HKernel, recRoot2 = 1/sqrt(2) */
        a[2 * lo] = r2 * (lr + hr); a[2 * lo + 1] = r2 * (li + hi_);
        a[2 * hi] = r2 * (lr - hr); a[2 * hi + 1] = r2 * (li - hi_);
        break;
    case ORA_S:                                    /* SKernel */
        a[2 * hi] = -hi_; a[2 * hi + 1] = hr;
        break;
    case ORA_SDG:                                  /* SDGKernel */
        a[2 * hi] = hi_; a[2 * hi + 1] = -hr;
        break;
    case ORA_T:                                    /* TKernel */
        a[2 * hi] = r2 * (hr - hi_); a[2 * hi + 1] = r2 * (hr + hi_);
        break;
    case ORA_TDG:                                  /* TDGKernel */
        a[2 * hi] = r2 * (hr + hi_); a[2 * hi + 1] = r2 * (hi_ - hr);
        break;
    case ORA_ID:
        break;
    default: {                                     /* UKernel (U2, U3, GII, GZZ, GOC, GCC) */
        a[2 * lo]     = (lr * m[0] - li * m[1]) + (hr * m[2] - hi_ * m[3]);
        a[2 * lo + 1] = (lr * m[1] + li * m[0]) + (hr * m[3] + hi_ * m[2]);
        a[2 * hi]     = (lr * m[4] - li * m[5]) + (hr * m[6] - hi_ * m[7]);
        a[2 * hi + 1] = (lr * m[5] + li * m[4]) + (hr * m[7] + hi_ * m[6]);
        break;
    }
    }
}

static int apply_gate_threads(double *psi, int n, const ora_gate *g, int nthreads)
{
    if (gate_check(g, n)) return -1;
    const int t = g->target;
    const long long half = (long long)1 << (n - 1);
    const size_t mask = ((size_t)1 << t) - 1;
    const int c1 = g->control, c2 = g->control2;
    (void)nthreads;
    #pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long long i = 0; i < half; i++) {
        size_t lo = (((size_t)i >> t) << (t + 1)) | ((size_t)i & mask);
        if (c1 >= 0 && !((lo >> c1) & 1)) continue;   /* CONTROL_GATE_BEGIN */
        if (c2 >= 0 && !((lo >> c2) & 1)) continue;   /* CC_GATE_BEGIN */
        size_t hi = lo | ((size_t)1 << t);
        pair_update(g, psi, lo, hi);
    }
    return 0;
}

int ora_apply_gate(double *psi, int n, const ora_gate *g)
{
    return apply_gate_threads(psi, n, g, ora_max_threads());
}

int ora_run_plain(double *psi, int n, const ora_gate *g, int m, int nthreads)
{
    /* kernelExecSimple: one pass per gate, in order (kernelSimple.cu:335-446) */
    if (n < 1 || n > 40) return -1;
    if (nthreads <= 0) nthreads = ora_max_threads();
    for (int k = 0; k < m; k++)
        if (apply_gate_threads(psi, n, &g[k], nthreads)) return -1 - k;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Brute force: U = G_m ... G_1, each G_k built from its definition as a    */
/* 2^n x 2^n matrix (no pair loop).  Column-major: U[col*N + row].          */
/* ------------------------------------------------------------------------ */

int ora_dense_unitary(int n, const ora_gate *g, int m, double *U)
{
    if (n < 1 || n > 10) return -1;
    size_t N = (size_t)1 << n;
    double *G = (double *)malloc(2 * N * N * sizeof(double));
    double *T = (double *)malloc(2 * N * N * sizeof(double));
    if (!G || !T) { free(G); free(T); return -1; }
    memset(U, 0, 2 * N * N * sizeof(double));
    for (size_t i = 0; i < N; i++) U[2 * (i * N + i)] = 1.0;
    for (int k = 0; k < m; k++) {
        if (gate_check(&g[k], n)) { free(G); free(T); return -1; }
        const ora_gate *q = &g[k];
        int t = q->target;
        memset(G, 0, 2 * N * N * sizeof(double));
        for (size_t col = 0; col < N; col++) {
            int on = 1;
            if (q->control >= 0 && !((col >> q->control) & 1)) on = 0;
            if (q->control2 >= 0 && !((col >> q->control2) & 1)) on = 0;
            for (int b = 0; b < 2; b++) {
                size_t row = (col & ~((size_t)1 << t)) | ((size_t)b << t);
                int bc = (int)((col >> t) & 1);
                double re, im;
                if (on) { re = q->mat[2 * (2 * b + bc)]; im = q->mat[2 * (2 * b + bc) + 1]; }
                else { re = (b == bc) ? 1.0 : 0.0; im = 0.0; }
                G[2 * (col * N + row)] = re;
                G[2 * (col * N + row) + 1] = im;
            }
        }
        /* T = G * U */
        for (size_t c = 0; c < N; c++)
            for (size_t r = 0; r < N; r++) {
                double sr = 0, si = 0;
                for (size_t j = 0; j < N; j++) {
                    double ar = G[2 * (j * N + r)], ai = G[2 * (j * N + r) + 1];
                    double br = U[2 * (c * N + j)], bi = U[2 * (c * N + j) + 1];
                    sr += ar * br - ai * bi;
                    si += ar * bi + ai * br;
                }
                T[2 * (c * N + r)] = sr;
                T[2 * (c * N + r) + 1] = si;
            }
        memcpy(U, T, 2 * N * N * sizeof(double));
    }
    free(G); free(T);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* OShareMem group semantics (single GPU).                                  */
/*  - Shared set: the physical positions in shared_mask; shared bit i is the */
/*    i-th lowest shared position, block bit j the j-th lowest non-shared    */
/*    position (executor.cpp:598-632).                                       */
/*  - Each block gathers its 2^k amplitudes (kernelOpt.cu:388-411), applies  */
/*    every gate with block bits resolved per block (kernelOpt.cu:214-386):  */
/*    a control on a block bit skips the gate when that bit is 0 and is      */
/*    dropped when it is 1; a diagonal target on a block bit becomes a       */
/*    scalar factor mat[b][b] on the tile (restricted by shared controls),   */
/*    the same rewrite executor.cpp:287-398 applies to GPU-global qubits.    */
/*  - Non-diagonal targets must be shared (compiler.cpp / newRelated).       */
/* ------------------------------------------------------------------------ */

int ora_sharemem_group(double *psi, int n, const ora_gate *g, int m, uint64_t shared_mask)
{
    if (n < 1 || n > 40) return -1;
    if (n < 64 && (shared_mask >> n)) return -1;
    int k = popcount64(shared_mask);
    if (k < 1) return -1;
    int share_id[MAXQ], block_id[MAXQ];
    int sc = 0, bc = 0;
    for (int q = 0; q < n; q++) {
        if ((shared_mask >> q) & 1) { share_id[q] = sc++; block_id[q] = -1; }
        else { block_id[q] = bc++; share_id[q] = -1; }
    }
    for (int j = 0; j < m; j++) {
        if (gate_check(&g[j], n)) return -1;
        if (share_id[g[j].target] < 0 && !ora_is_diagonal(g[j].type)) return -2;
    }
    size_t T = (size_t)1 << k;
    long long nblocks = (long long)1 << (n - k);
    int err = 0;
    #pragma omp parallel
    {
        double *tile = (double *)malloc(2 * T * sizeof(double));
        if (!tile) {
            #pragma omp atomic write
            err = 1;
        }
        #pragma omp for schedule(static)
        for (long long b = 0; b < nblocks; b++) {
            if (!tile) continue;
            /* block bias = pdep(b, blockHot) (kernelOpt.cu:389-399) */
            size_t bias = 0;
            for (int q = 0, j = 0; q < n; q++)
                if (block_id[q] >= 0) { if ((b >> j) & 1) bias |= (size_t)1 << q; j++; }
            for (size_t e = 0; e < T; e++) {
                size_t addr = bias;
                for (int q = 0, i = 0; q < n; q++)
                    if (share_id[q] >= 0) { if ((e >> i) & 1) addr |= (size_t)1 << q; i++; }
                tile[2 * e] = psi[2 * addr]; tile[2 * e + 1] = psi[2 * addr + 1];
            }
            for (int j = 0; j < m; j++) {
                ora_gate x = g[j];
                int skip = 0;
                int ctl[2] = { x.control, x.control2 };
                int keep[2] = { -1, -1 }, nk = 0;
                for (int c = 0; c < 2; c++) {
                    if (ctl[c] < 0) continue;
                    if (share_id[ctl[c]] >= 0) keep[nk++] = share_id[ctl[c]];
                    else if (!((bias >> ctl[c]) & 1)) skip = 1;
                }
                if (skip) continue;
                if (share_id[x.target] < 0) {
                    /* diagonal on a block bit: factor mat[b][b] on the tile */
                    int hb = (int)((bias >> x.target) & 1);
                    double fr = x.mat[2 * (3 * hb)], fi = x.mat[2 * (3 * hb) + 1];
                    for (size_t e = 0; e < T; e++) {
                        if (keep[0] >= 0 && !((e >> keep[0]) & 1)) continue;
                        if (keep[1] >= 0 && !((e >> keep[1]) & 1)) continue;
                        double r = tile[2 * e], i = tile[2 * e + 1];
                        tile[2 * e] = r * fr - i * fi;
                        tile[2 * e + 1] = r * fi + i * fr;
                    }
                    continue;
                }
                /* tile-local gate: same type, renumbered, with the kept controls */
                ora_gate y = x;
                y.target = share_id[x.target];
                y.control = keep[0];
                y.control2 = keep[1];
                if (nk == 0) {
                    /* a controlled type whose controls were all on block bits acts
                       as its uncontrolled form (Gate::toU, gate.cpp:524-546) */
                    switch (x.type) {
                    case ORA_CCX: case ORA_CNOT: y.type = ORA_X; break;
                    case ORA_CY: y.type = ORA_Y; break;
                    case ORA_CZ: y.type = ORA_Z; break;
                    case ORA_CRX: y.type = ORA_RX; break;
                    case ORA_CRY: y.type = ORA_RY; break;
                    case ORA_CU1: y.type = ORA_U1; break;
                    case ORA_CRZ: y.type = ORA_RZ; break;
                    default: break;
                    }
                } else if (nk == 1 && x.type == ORA_CCX) {
                    y.type = ORA_CNOT;                 /* Gate::toCU (gate.cpp:516-522) */
                }
                /* apply on the tile */
                size_t half = T >> 1, mask = ((size_t)1 << y.target) - 1;
                for (size_t i = 0; i < half; i++) {
                    size_t lo = ((i >> y.target) << (y.target + 1)) | (i & mask);
                    if (y.control >= 0 && !((lo >> y.control) & 1)) continue;
                    if (y.control2 >= 0 && !((lo >> y.control2) & 1)) continue;
                    pair_update(&y, tile, lo, lo | ((size_t)1 << y.target));
                }
            }
            for (size_t e = 0; e < T; e++) {
                size_t addr = bias;
                for (int q = 0, i = 0; q < n; q++)
                    if (share_id[q] >= 0) { if ((e >> i) & 1) addr |= (size_t)1 << q; i++; }
                psi[2 * addr] = tile[2 * e]; psi[2 * addr + 1] = tile[2 * e + 1];
            }
        }
        free(tile);
    }
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* TransMM: fusion (schedule.cpp:510-702), transpose (cuTT as called at     */
/* schedule.cpp:767-774: dims all 2, output dimension i = input dimension   */
/* perm[i]) and GEMM (executor.cpp:533-551: column-major K x (2^n/K)).      */
/* ------------------------------------------------------------------------ */

int ora_fuse_group(const ora_gate *g, int m, const int *pos, int mat_qubit, double *mat)
{
    /* initCPUMatrix, single GPU: start from I and apply every gate to each
     * basis vector i stored at mat[i*K .. i*K+K-1]; the result is U in
     * column-major order (schedule.cpp:578-702). */
    if (mat_qubit < 1 || mat_qubit > 10) return -1;
    size_t K = (size_t)1 << mat_qubit;
    memset(mat, 0, 2 * K * K * sizeof(double));
    for (size_t i = 0; i < K; i++) mat[2 * (i * K + i)] = 1.0;
    for (int j = 0; j < m; j++) {
        ora_gate y = g[j];
        y.target = pos[g[j].target];
        y.control = g[j].control >= 0 ? pos[g[j].control] : -1;
        y.control2 = g[j].control2 >= 0 ? pos[g[j].control2] : -1;
        if (y.target >= mat_qubit || y.control >= mat_qubit || y.control2 >= mat_qubit) return -1;
        /* matrix entries, not the specialised formulas (APPLY_*_GATE macros) */
        for (size_t col = 0; col < K; col++) {
            double *v = mat + 2 * col * K;
            size_t half = K >> 1, mask = ((size_t)1 << y.target) - 1;
            for (size_t i = 0; i < half; i++) {
                size_t lo = ((i >> y.target) << (y.target + 1)) | (i & mask);
                if (y.control >= 0 && !((lo >> y.control) & 1)) continue;
                if (y.control2 >= 0 && !((lo >> y.control2) & 1)) continue;
                size_t hi = lo | ((size_t)1 << y.target);
                const double *u = y.mat;
                double lr = v[2 * lo], li = v[2 * lo + 1], hr = v[2 * hi], hh = v[2 * hi + 1];
                v[2 * lo]     = (lr * u[0] - li * u[1]) + (hr * u[2] - hh * u[3]);
                v[2 * lo + 1] = (lr * u[1] + li * u[0]) + (hr * u[3] + hh * u[2]);
                v[2 * hi]     = (lr * u[4] - li * u[5]) + (hr * u[6] - hh * u[7]);
                v[2 * hi + 1] = (lr * u[5] + li * u[4]) + (hr * u[7] + hh * u[6]);
            }
        }
    }
    return 0;
}

int ora_transpose(const double *in, double *out, int n, const int *perm)
{
    if (n < 1 || n > 40) return -1;
    uint64_t seen = 0;
    for (int i = 0; i < n; i++) {
        if (perm[i] < 0 || perm[i] >= n || ((seen >> perm[i]) & 1)) return -1;
        seen |= (uint64_t)1 << perm[i];
    }
    long long N = (long long)1 << n;
    #pragma omp parallel for schedule(static)
    for (long long j = 0; j < N; j++) {
        size_t src = 0;
        for (int b = 0; b < n; b++)
            if ((j >> b) & 1) src |= (size_t)1 << perm[b];
        out[2 * j] = in[2 * src];
        out[2 * j + 1] = in[2 * src + 1];
    }
    return 0;
}

int ora_apply_dense_runs(const double *in, double *out, int n, const double *U, int mq)
{
    if (mq < 1 || mq > n) return -1;
    long long K = (long long)1 << mq;
    long long cols = ((long long)1 << n) / K;
    #pragma omp parallel for schedule(static)
    for (long long c = 0; c < cols; c++) {
        const double *b = in + 2 * c * K;
        double *o = out + 2 * c * K;
        for (long long r = 0; r < K; r++) {
            double sr = 0, si = 0;
            for (long long j = 0; j < K; j++) {
                double ur = U[2 * (j * K + r)], ui = U[2 * (j * K + r) + 1];
                sr += ur * b[2 * j] - ui * b[2 * j + 1];
                si += ur * b[2 * j + 1] + ui * b[2 * j];
            }
            o[2 * r] = sr; o[2 * r + 1] = si;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Partitioner: OneLayerCompiler::getGroupOpt (compiler.cpp:322-452),        */
/* AdvanceCompiler::run (compiler.cpp:220-319), GateGroup::newRelated        */
/* (schedule.cpp:37-51), initBlasState (schedule.cpp:339-382), Evaluator     */
/* (evaluator.cpp:111-229).  Single GPU: all qubits local.                   */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t *w; int nw; } bitset_t;

static int bs_count(const uint64_t *w, int nw)
{
    int c = 0;
    for (int i = 0; i < nw; i++) c += popcount64(w[i]);
    return c;
}

static int bs_any(const uint64_t *w, int nw)
{
    for (int i = 0; i < nw; i++) if (w[i]) return 1;
    return 0;
}

static uint64_t new_related(uint64_t rel, const ora_gate *g, uint64_t local, int enable_global)
{
    /* schedule.cpp:37-51 */
    if (enable_global) {
        if (!ora_is_diagonal(g->type)) rel |= (uint64_t)1 << g->target;
    } else {
        if (!ora_is_diagonal(g->type) || ((local >> g->target) & 1))
            rel |= (uint64_t)1 << g->target;
        if (g->control != -1 && ((local >> g->control) & 1))
            rel |= (uint64_t)1 << g->control;
        if (g->control2 != -1 && ((local >> g->control2) & 1))
            rel |= (uint64_t)1 << g->control2;
    }
    return rel;
}

/* remain: sorted indices of the gates not yet grouped (count nrem).
 * Returns the selected gate indices (in order) into out, and their count. */
static int get_group_opt(int n, const ora_gate *gates, const int *remain, int nrem,
                         int window, int full[], uint64_t related[], int enable_global,
                         int local_size, uint64_t local_qubits, int *out)
{
    int gate_num = nrem < window ? nrem : window;
    int nw = (window + 63) / 64;
    uint64_t *cur = (uint64_t *)calloc((size_t)n * nw, sizeof(uint64_t));
    uint64_t *nc = (uint64_t *)calloc(nw, sizeof(uint64_t));
    uint64_t *sel = (uint64_t *)calloc(nw, sizeof(uint64_t));
    /* phase 1: per-qubit dependency chains (compiler.cpp:335-369) */
    for (int id = 0; id < gate_num; id++) {
        if (id % 100 == 0) {
            int live = 0;
            for (int i = 0; i < n; i++) if (!full[i]) live = 1;
            if (!live) break;
        }
        const ora_gate *g = &gates[remain[id]];
        if (g->control != -1) {
            int c = g->control, t = g->target, c2 = g->control2;
            int ok = !full[c] && !full[t] && (c2 < 0 || !full[c2]);
            if (ok) {
                uint64_t nr = related[c] | related[t] | (c2 >= 0 ? related[c2] : 0);
                nr = new_related(nr, g, local_qubits, enable_global);
                if (popcount64(nr) <= local_size) {
                    for (int w = 0; w < nw; w++)
                        nc[w] = cur[c * nw + w] | cur[t * nw + w] | (c2 >= 0 ? cur[c2 * nw + w] : 0);
                    nc[id / 64] |= (uint64_t)1 << (id % 64);
                    memcpy(cur + c * nw, nc, nw * sizeof(uint64_t));
                    memcpy(cur + t * nw, nc, nw * sizeof(uint64_t));
                    if (c2 >= 0) memcpy(cur + c2 * nw, nc, nw * sizeof(uint64_t));
                    related[c] = related[t] = nr;
                    if (c2 >= 0) related[c2] = nr;
                    continue;
                }
            }
            full[c] = full[t] = 1;
            if (c2 >= 0) full[c2] = 1;
        } else {
            if (!full[g->target]) {
                cur[g->target * nw + id / 64] |= (uint64_t)1 << (id % 64);
                related[g->target] = new_related(related[g->target], g, local_qubits, enable_global);
            }
        }
    }
    /* phase 2: greedy union of chains under the qubit limit (371-401) */
    int blocked[MAXQ];
    memset(blocked, 0, sizeof(blocked));
    uint64_t sel_rel = 0;
    for (;;) {
        int mx = 0, best = -1;
        for (int i = 0; i < n; i++) {
            int cnt = bs_count(cur + i * nw, nw);
            if (!blocked[i] && cnt > mx) {
                if (popcount64(sel_rel | related[i]) <= local_size) { mx = cnt; best = i; }
                else blocked[i] = 1;
            }
        }
        if (mx == 0) break;
        for (int w = 0; w < nw; w++) sel[w] |= cur[best * nw + w];
        sel_rel |= related[best];
        blocked[best] = 1;
        for (int i = 0; i < n; i++) {
            if (!blocked[i] && bs_any(cur + i * nw, nw)) {
                if ((related[i] | sel_rel) == sel_rel) {
                    for (int w = 0; w < nw; w++) sel[w] |= cur[i * nw + w];
                    blocked[i] = 1;
                } else {
                    for (int w = 0; w < nw; w++) cur[i * nw + w] &= ~cur[best * nw + w];
                }
            }
        }
    }
    /* phase 3: diagonal sweep (412-445), only with enableGlobal */
    if (enable_global) {
        memset(blocked, 0, sizeof(blocked));
        int cnt = 0;
        for (int id = 0; id < gate_num; id++) {
            cnt++;
            if (cnt % 100 == 0) {
                int live = 0;
                for (int i = 0; i < n; i++) if (!blocked[i]) live = 1;
                if (!live) break;
            }
            if ((sel[id / 64] >> (id % 64)) & 1) continue;
            const ora_gate *g = &gates[remain[id]];
            if (ora_is_diagonal(g->type)) {
                if (g->control != -1) {
                    if (!blocked[g->control] && !blocked[g->target])
                        sel[id / 64] |= (uint64_t)1 << (id % 64);
                    else
                        blocked[g->control] = blocked[g->target] = 1;
                } else {
                    if (!blocked[g->target]) sel[id / 64] |= (uint64_t)1 << (id % 64);
                }
            } else {
                if (g->control != -1) blocked[g->control] = 1;
                if (g->control2 != -1) blocked[g->control2] = 1;
                blocked[g->target] = 1;
            }
        }
    }
    int cntout = 0;
    for (int id = 0; id < gate_num; id++)
        if ((sel[id / 64] >> (id % 64)) & 1) out[cntout++] = remain[id];
    free(cur); free(nc); free(sel);
    return cntout;
}

static double perf_pergate(const ora_partition_params *p, const ora_gate *gates,
                           const int *idx, int cnt)
{
    /* Evaluator::perfPerGate (evaluator.cpp:168-223) */
    double t = 0;
    for (int i = 0; i < cnt; i++) {
        int ty = gates[idx[i]].type;
        t += num_controls(ty) ? p->ctr_cost[ty] : p->single_cost[ty];
    }
    return t / 1000 / 512 + p->group_overhead;
}

static void init_blas_state(int n, uint64_t rel, int *pos, int *layout, int *perm)
{
    /* GateGroup::initBlasState (schedule.cpp:339-382) */
    int to_global[MAXQ], to_local[MAXQ], ng = 0, nl = 0;
    int nm = popcount64(rel);
    for (int i = 0; i < nm; i++) {
        int q = layout[i];
        if (!((rel >> q) & 1)) to_global[ng++] = q;
    }
    for (int i = nm; i < n; i++) {
        int q = layout[i];
        if ((rel >> q) & 1) to_local[nl++] = q;
    }
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < ng && i < nl; i++) {
        int x = to_global[i], y = to_local[i];
        int tmp = perm[pos[x]]; perm[pos[x]] = perm[pos[y]]; perm[pos[y]] = tmp;
        tmp = pos[x]; pos[x] = pos[y]; pos[y] = tmp;
        layout[pos[x]] = x; layout[pos[y]] = y;
    }
}

void ora_partition_defaults(ora_partition_params *p, int n, int mode)
{
    memset(p, 0, sizeof(*p));
    p->mode = mode;
    p->local_size = 10;      /* LOCAL_QUBIT_SIZE, utils.h:56 */
    p->coalesce = 3;         /* COALESCE_GLOBAL default, CMakeLists.txt */
    p->min_mat = 4;          /* MIN_MAT_SIZE default */
    p->blas_mat_limit = 6;   /* BLAS_MAT_LIMIT (MAT) default */
    p->mat_lo = 4; p->mat_hi = 8;   /* compiler.cpp:265 */
    p->window = 512;         /* AdvanceCompiler: OneLayerCompiler<512> */
    /* V100 n=28 tables: entry [1] of each single-gate row and [0][2] of each
       control-gate table (evaluator.h:18-117), as read by perfPerGate. */
    p->single_cost[ORA_U1] = 225; p->single_cost[ORA_U2] = 469; p->single_cost[ORA_U3] = 469;
    p->single_cost[ORA_H] = 352;  p->single_cost[ORA_X] = 350;  p->single_cost[ORA_Y] = 350;
    p->single_cost[ORA_Z] = 194;  p->single_cost[ORA_S] = 209;  p->single_cost[ORA_SDG] = 209;
    p->single_cost[ORA_T] = 216;  p->single_cost[ORA_TDG] = 216; p->single_cost[ORA_RX] = 370;
    p->single_cost[ORA_RY] = 367; p->single_cost[ORA_RZ] = 369;
    p->ctr_cost[ORA_CCX] = 195; p->ctr_cost[ORA_CNOT] = 195; p->ctr_cost[ORA_CY] = 193;
    p->ctr_cost[ORA_CZ] = 137;  p->ctr_cost[ORA_CRX] = 224;  p->ctr_cost[ORA_CRY] = 225;
    p->ctr_cost[ORA_CU1] = 225; p->ctr_cost[ORA_CRZ] = 224;
    /* pergate_group_overhead = 1/2^27 per amplitude (evaluator.h:126) */
    p->group_overhead = ldexp(1.0, n - 27);
    /* BLAS_perf[28][6] and cutt_cost[28] (evaluator.cpp:32-33); the other
       BLAS_perf entries are never written by the constructor and read as 0
       in a fresh heap allocation. */
    p->blas_cost[6] = 23.068396;
    p->cutt_cost = 11.367814;
}

int ora_partition(int n, const ora_gate *gates, int m, const ora_partition_params *p,
                  int max_groups, int *order, int *group_start, int *backend,
                  uint64_t *related_out, int *mat_qubit, int *perms,
                  int *final_pos, int *final_layout)
{
    if (n < 1 || n > MAXQ - 1 || !p || p->window < 1) return -1;
    for (int i = 0; i < m; i++) if (gate_check(&gates[i], n)) return -1;
    int pos[MAXQ], layout[MAXQ];
    for (int i = 0; i < n; i++) pos[i] = layout[i] = i;
    int *remain = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
    int *idx = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
    int *best = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
    int nrem = m;
    for (int i = 0; i < m; i++) remain[i] = i;
    const uint64_t all = (n == 64) ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1);
    int ng = 0, nout = 0, rc = 0;
    group_start[0] = 0;
    while (nrem > 0) {
        if (ng >= max_groups) { rc = -2; break; }
        int full[MAXQ];
        uint64_t related[MAXQ];
        int nbest = 0, bk = 0;
        uint64_t cache_rel = 0;
        int use_pg = (p->mode == ORA_MODE_GROUP || p->mode == ORA_MODE_MIX);
        int use_blas = (p->mode == ORA_MODE_BLAS || p->mode == ORA_MODE_MIX);
        double best_eff = 1e10;
        if (use_pg) {
            /* fillRelated: the COALESCE lowest physical qubits (compiler.cpp:231-237) */
            memset(full, 0, sizeof(full));
            for (int i = 0; i < n; i++) {
                related[i] = 0;
                for (int j = 0; j < p->coalesce && j < n; j++) related[i] |= (uint64_t)1 << layout[j];
            }
            cache_rel = related[0];
            nbest = get_group_opt(n, gates, remain, nrem, p->window, full, related, 1,
                                  p->local_size, all, best);
            bk = ORA_BACKEND_PERGATE;
            if (use_blas && nbest > 0) best_eff = perf_pergate(p, gates, best, nbest) / nbest;
        }
        if (use_blas && use_pg) {
            for (int ms = p->mat_lo; ms < p->mat_hi; ms++) {
                memset(full, 0, sizeof(full));
                memset(related, 0, sizeof(related));
                int cnt = get_group_opt(n, gates, remain, nrem, p->window, full, related, 0,
                                        ms, all, idx);
                if (cnt == 0) continue;
                double eff = (p->blas_cost[ms] + p->cutt_cost) / cnt;
                if (eff < best_eff) {
                    memcpy(best, idx, sizeof(int) * cnt);
                    nbest = cnt; bk = ORA_BACKEND_BLAS; best_eff = eff;
                }
            }
        } else if (use_blas) {
            memset(full, 0, sizeof(full));
            memset(related, 0, sizeof(related));
            nbest = get_group_opt(n, gates, remain, nrem, p->window, full, related, 0,
                                  p->blas_mat_limit, all, best);
            bk = ORA_BACKEND_BLAS;
        }
        if (nbest == 0) { rc = -3; break; }
        uint64_t rel = 0;
        for (int i = 0; i < nbest; i++)
            rel = new_related(rel, &gates[best[i]], bk == ORA_BACKEND_PERGATE ? all : all,
                              bk == ORA_BACKEND_PERGATE);
        int *perm = perms ? perms + (size_t)ng * n : NULL;
        int tmp_perm[MAXQ];
        if (!perm) perm = tmp_perm;
        if (bk == ORA_BACKEND_PERGATE) {
            rel |= cache_rel;
            for (int i = 0; i < n; i++) perm[i] = i;
            mat_qubit[ng] = 0;
        } else {
            init_blas_state(n, rel, pos, layout, perm);
            int nm = popcount64(rel);
            mat_qubit[ng] = nm > p->min_mat ? nm : p->min_mat;
        }
        backend[ng] = bk;
        related_out[ng] = rel;
        for (int i = 0; i < nbest; i++) order[nout++] = best[i];
        group_start[ng + 1] = nout;
        ng++;
        /* removeGatesOpt: drop the selected indices from remain (both sorted) */
        int w = 0, s = 0;
        for (int r = 0; r < nrem; r++) {
            /* best[] is in remain order */
            if (s < nbest && remain[r] == best[s]) { s++; continue; }
            remain[w++] = remain[r];
        }
        nrem = w;
    }
    if (final_pos) memcpy(final_pos, pos, sizeof(int) * n);
    if (final_layout) memcpy(final_layout, layout, sizeof(int) * n);
    free(remain); free(idx); free(best);
    return rc ? rc : ng;
}

/* ------------------------------------------------------------------------ */
/* Readout: Circuit::printState (circuit.cpp:250-309, single GPU branch),   */
/* toPhysicalID/toLogicID (95-113), ResultItem::print (circuit.h:14-16),    */
/* zero_wrapper (utils.cpp:77-84).                                           */
/* ------------------------------------------------------------------------ */

double ora_zero_wrapper(double x)
{
    const double eps = 1e-14;
    return (x > -eps && x < eps) ? 0 : x;
}

typedef struct { long long idx; double re, im; } ora_item;

static int item_cmp(const void *a, const void *b)
{
    long long x = ((const ora_item *)a)->idx, y = ((const ora_item *)b)->idx;
    return x < y ? -1 : (x > y);
}

static size_t put_item(char *buf, size_t cap, size_t len, long long idx, double re, double im)
{
    char line[160];
    int k = snprintf(line, sizeof(line), "%lld %.12f: %.12f %.12f\n", idx,
                     re * re + im * im, ora_zero_wrapper(re), ora_zero_wrapper(im));
    if (buf && len + (size_t)k < cap) memcpy(buf + len, line, (size_t)k + 1);
    return len + (size_t)k;
}

static long long to_physical(long long idx, int n, const int *pos)
{
    /* Circuit::toPhysicalID (circuit.cpp:95-103) */
    long long id = 0;
    for (int q = 0; q < n; q++)
        if ((idx >> q) & 1) id |= 1ll << pos[q];
    return id;
}

size_t ora_print_state(const double *psi, int n, const int *pos, int show_schedule,
                       char *buf, size_t cap)
{
    size_t len = 0;
    if (buf && cap) buf[0] = 0;
    for (long long i = 0; i < 128; i++) {
        long long id = to_physical(i, n, pos);
        len = put_item(buf, cap, len, i, psi[2 * id], psi[2 * id + 1]);
    }
    if (show_schedule) {
        for (int i = 0; i < n; i++) {
            long long id = to_physical(1ll << i, n, pos);
            len = put_item(buf, cap, len, 1ll << i, psi[2 * id], psi[2 * id + 1]);
        }
        long long id = to_physical((1ll << n) - 1, n, pos);
        len = put_item(buf, cap, len, (1ll << n) - 1, psi[2 * id], psi[2 * id + 1]);
    }
    size_t N = (size_t)1 << n, cnt = 0, capi = 1024;
    ora_item *items = (ora_item *)malloc(capi * sizeof(ora_item));
    for (size_t i = 0; i < N; i++) {
        double re = psi[2 * i], im = psi[2 * i + 1];
        if (re * re + im * im > 0.001) {
            long long lid = 0;
            for (int q = 0; q < n; q++) if ((i >> pos[q]) & 1) lid |= 1ll << q;
            if (lid >= 128) {
                if (cnt == capi) { capi *= 2; items = (ora_item *)realloc(items, capi * sizeof(ora_item)); }
                items[cnt].idx = lid; items[cnt].re = re; items[cnt].im = im; cnt++;
            }
        }
    }
    qsort(items, cnt, sizeof(ora_item), item_cmp);
    for (size_t i = 0; i < cnt; i++) len = put_item(buf, cap, len, items[i].idx, items[i].re, items[i].im);
    free(items);
    return len;
}

/* ------------------------------------------------------------------------ */
/* Parser: main.cpp:12-231.  Whitespace-separated tokens (fscanf "%s"); the */
/* rest of the line after each statement is discarded (fgets, line 222).   */
/* ------------------------------------------------------------------------ */

typedef struct { const char *s; size_t p; } rd_t;

static int rd_token(rd_t *r, char *tok, size_t cap)
{
    while (r->s[r->p] && isspace((unsigned char)r->s[r->p])) r->p++;
    if (!r->s[r->p]) return 0;
    size_t k = 0;
    while (r->s[r->p] && !isspace((unsigned char)r->s[r->p])) {
        if (k + 1 < cap) tok[k++] = r->s[r->p];
        r->p++;
    }
    tok[k] = 0;
    return 1;
}

static void rd_rest_of_line(rd_t *r)
{
    /* fgets(buffer, 1000, f): up to 999 chars or through the newline */
    size_t k = 0;
    while (r->s[r->p] && k < 999) {
        char c = r->s[r->p++];
        k++;
        if (c == '\n') break;
    }
}

static int parse_qid(const char *buf, int *out, int cap)
{
    /* main.cpp:12-27: every run of digits is one qubit id */
    int cnt = 0;
    for (size_t i = 0; buf[i]; i++) {
        if (buf[i] >= '0' && buf[i] <= '9') {
            long x = 0;
            while (buf[i] >= '0' && buf[i] <= '9') { x = x * 10 + (buf[i] - '0'); i++; }
            i--;
            if (cnt < cap) out[cnt] = (int)x;
            cnt++;
        }
    }
    return cnt;
}

static int parse_gate(const char *buf, char *name, size_t ncap, double *params, int pcap)
{
    /* main.cpp:29-66 */
    const double pi = acos(-1);
    size_t l = strlen(buf), i = 0, k = 0;
    while (i < l && buf[i] != '(') { if (k + 1 < ncap) name[k++] = buf[i]; i++; }
    name[k] = 0;
    int np = 0;
    while (i < l) {
        i++;
        char st[256]; size_t sk = 0;
        while (i < l && buf[i] != ',' && buf[i] != ')') { if (sk + 1 < sizeof(st)) st[sk++] = buf[i]; i++; }
        st[sk] = 0;
        if (i >= l) return -1;                /* unterminated list: UB in the reference */
        double param = 1;
        const char *num = st;
        if (st[0] == 'p' && st[1] == 'i' && st[2] == '*') { param = pi; num = st + 3; }
        else if (st[0] == 'p' && st[1] == 'i' && st[2] == '/') { param = -pi; num = st + 3; }
        char *end;
        double v = strtod(num, &end);
        if (end == num) return -1;            /* std::stod throws */
        if (param > 0) param *= v; else param = pi / v;
        if (np < pcap) params[np] = param;
        np++;
        if (buf[i] == ')') break;
    }
    return np;
}

int ora_parse_qasm(const char *text, int *num_qubits, ora_gate *out, int cap)
{
    rd_t r = { text, 0 };
    char tok[1000];
    int n = -1, cnt = 0;
    int have = 0;
    while (rd_token(&r, tok, sizeof(tok))) {
        int q[8], nq;
        ora_gate g;
        int type = -1, need = 0, np = 0;
        double params[8];
        if (!strcmp(tok, "//") || !strcmp(tok, "OPENQASM") || !strcmp(tok, "include")) {
            rd_rest_of_line(&r);
            continue;
        } else if (!strcmp(tok, "qreg")) {
            /* fscanf("%*c%*c%*c%d"): skip exactly three characters, then an int */
            for (int s = 0; s < 3; s++) { if (!r.s[r.p]) return -1; r.p++; }
            char *end;
            long v = strtol(r.s + r.p, &end, 10);
            if (end == r.s + r.p) return -1;
            r.p = (size_t)(end - r.s);
            n = (int)v; have = 1;
            rd_rest_of_line(&r);
            continue;
        }
        static const struct { const char *s; int t; int nq; } simple[] = {
            {"cx", ORA_CNOT, 2}, {"ccx", ORA_CCX, 3}, {"cy", ORA_CY, 2}, {"cz", ORA_CZ, 2},
            {"h", ORA_H, 1}, {"x", ORA_X, 1}, {"y", ORA_Y, 1}, {"z", ORA_Z, 1},
            {"s", ORA_S, 1}, {"sdg", ORA_SDG, 1}, {"t", ORA_T, 1}, {"tdg", ORA_TDG, 1},
        };
        for (size_t s = 0; s < sizeof(simple) / sizeof(simple[0]); s++)
            if (!strcmp(tok, simple[s].s)) { type = simple[s].t; need = simple[s].nq; }
        if (type < 0) {
            char name[1000];
            np = parse_gate(tok, name, sizeof(name), params, 8);
            static const struct { const char *s; int t; int nq; int np; } par[] = {
                {"crx", ORA_CRX, 2, 1}, {"cry", ORA_CRY, 2, 1}, {"crz", ORA_CRZ, 2, 1},
                {"cu1", ORA_CU1, 2, 1}, {"u1", ORA_U1, 1, 1}, {"u3", ORA_U3, 1, 3},
                {"rx", ORA_RX, 1, 1}, {"ry", ORA_RY, 1, 1}, {"rz", ORA_RZ, 1, 1},
            };
            for (size_t s = 0; s < sizeof(par) / sizeof(par[0]); s++)
                if (!strcmp(name, par[s].s)) {
                    if (np != par[s].np) return -1;
                    type = par[s].t; need = par[s].nq;
                }
            if (type < 0) return -1;           /* "unrecognized token" -> exit(1) */
        } else {
            np = 0;
        }
        if (!have) return -1;                  /* gate before qreg: null Circuit */
        if (!rd_token(&r, tok, sizeof(tok))) return -1;
        nq = parse_qid(tok, q, 8);
        if (nq != need) return -1;
        int c2 = -1, c = -1, t;
        if (need == 3) { c = q[0]; c2 = q[1]; t = q[2]; }
        else if (need == 2) { c = q[0]; t = q[1]; }
        else t = q[0];
        if (ora_gate_make(type, c2, c, t, params, np, &g)) return -1;
        if (gate_check(&g, n)) return -1;
        if (cnt < cap) out[cnt] = g;
        cnt++;
        rd_rest_of_line(&r);
    }
    if (!have) return -1;                      /* "fail to load circuit" */
    *num_qubits = n;
    return cnt;
}
