// pergate.hip -- the plain method: one HBM pass per gate.
//
// Replaces kernelExecSimple and its 19 gate kernels (src/kernelSimple.cu:40-446).
// Differences by design: one templated kernel per update form instead of one
// per gate type; threads enumerate only the pairs whose controls are set
// (the reference launches 2^(n-1) threads and idles the others); 64-bit
// indexing so n >= 33 works.

#include "hq_ops.h"

#include <algorithm>
#include <cmath>

namespace hq {

bool gate_is_diagonal_type(int type)
{
    // Gate::isDiagonal, src/gate.h:27-29
    switch (type) {
    case HQ_CZ: case HQ_CU1: case HQ_CRZ: case HQ_U1: case HQ_Z:
    case HQ_S: case HQ_SDG: case HQ_T: case HQ_TDG: case HQ_RZ:
        return true;
    default:
        return false;
    }
}

int gate_num_controls(int type)
{
    switch (type) {
    case HQ_CCX: return 2;
    case HQ_CNOT: case HQ_CY: case HQ_CZ: case HQ_CRX: case HQ_CRY: case HQ_CU1: case HQ_CRZ: return 1;
    default: return 0;
    }
}

hq_status validate_gate(const hq_gate& g, int n)
{
    HQ_CHECK(g.type >= 0 && g.type <= HQ_GCC && g.type != HQ_TOTAL, HQ_ERR_ARG, "gate: unknown type");
    HQ_CHECK(g.target >= 0 && g.target < n, HQ_ERR_ARG, "gate: target out of range");
    int nc = gate_num_controls(g.type);
    if (nc >= 1) {
        HQ_CHECK(g.control >= 0 && g.control < n && g.control != g.target, HQ_ERR_ARG,
                 "gate: control missing, out of range or equal to the target");
    } else {
        HQ_CHECK(g.control == -1, HQ_ERR_ARG, "gate: this gate type takes no control");
    }
    if (nc == 2) {
        HQ_CHECK(g.control2 >= 0 && g.control2 < n && g.control2 != g.target && g.control2 != g.control,
                 HQ_ERR_ARG, "gate: second control missing, out of range or repeated");
    } else {
        HQ_CHECK(g.control2 == -1, HQ_ERR_ARG, "gate: this gate type takes no second control");
    }
    for (int i = 0; i < 8; i++)
        HQ_CHECK(std::isfinite(g.mat[i]), HQ_ERR_ARG, "gate: matrix entry is not finite");
    return HQ_OK;
}

int classify_gate(const hq_gate& g)
{
    const double* m = g.mat;
    bool offdiag_zero = m[2] == 0 && m[3] == 0 && m[4] == 0 && m[5] == 0;
    if (gate_is_diagonal_type(g.type) || offdiag_zero) {
        if (m[0] == 1 && m[1] == 0) return SHM_PHASE;
        return SHM_DIAG;
    }
    bool diag_zero = m[0] == 0 && m[1] == 0 && m[6] == 0 && m[7] == 0;
    if (diag_zero && m[2] == 1 && m[3] == 0 && m[4] == 1 && m[5] == 0) return SHM_SWAP;
    if (diag_zero && m[2] == 0 && m[3] == -1 && m[4] == 0 && m[5] == 1) return SHM_Y;
    if (m[1] == 0 && m[3] == 0 && m[5] == 0 && m[7] == 0) return SHM_REAL;
    if (m[1] == 0 && m[7] == 0 && m[2] == 0 && m[4] == 0) return SHM_RX;
    return SHM_GEN;
}

// Insert bits into a pair index: `nins` positions (ascending) get the values
// in `vals` (0 for the target, 1 for each control).
struct Inserts {
    int pos[3];
    int val[3];
    int n;
};

template <int CLS, typename R>
__global__ __launch_bounds__(256) void k_plain(cplx<R>* __restrict__ a, uint64_t work, Inserts ins, int t,
                                               Coef<R> c)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < work; i += stride) {
        uint64_t lo = i;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (k < ins.n) {
                int p = ins.pos[k];
                lo = ((lo >> p) << (p + 1)) | (lo & (((uint64_t)1 << p) - 1)) | ((uint64_t)ins.val[k] << p);
            }
        }
        uint64_t hi = lo | ((uint64_t)1 << t);
        cplx<R> x = a[lo], y = a[hi];
        pair_update<CLS, R>(x, y, c.m);
        if constexpr (CLS != SHM_PHASE) a[lo] = x;
        a[hi] = y;
    }
}

template <typename R>
static hq_status launch_plain_gate(hq_state s, const hq_gate& g)
{
    int cls = classify_gate(g);
    Inserts ins{};
    ins.n = 0;
    int ps[3], vs[3], np = 0;
    ps[np] = g.target; vs[np++] = 0;
    if (g.control >= 0) { ps[np] = g.control; vs[np++] = 1; }
    if (g.control2 >= 0) { ps[np] = g.control2; vs[np++] = 1; }
    // ascending insertion order
    for (int i = 0; i < np; i++)
        for (int j = i + 1; j < np; j++)
            if (ps[j] < ps[i]) { std::swap(ps[i], ps[j]); std::swap(vs[i], vs[j]); }
    for (int i = 0; i < np; i++) { ins.pos[i] = ps[i]; ins.val[i] = vs[i]; }
    ins.n = np;
    uint64_t work = s->amps >> np;
    Coef<R> c;
    for (int i = 0; i < 8; i++) c.m[i] = (R)g.mat[i];
    unsigned threads = 256;
    uint64_t blocks = (work + threads - 1) / threads;
    if (blocks > 1u << 20) blocks = 1u << 20;
    cplx<R>* a = (cplx<R>*)s->data;
    switch (cls) {
#define HQ_PLAIN_CASE(C) \
    case C: hipLaunchKernelGGL((k_plain<C, R>), dim3((unsigned)blocks), dim3(threads), 0, s->stream, a, work, ins, g.target, c); break;
        HQ_PLAIN_CASE(SHM_SWAP)
        HQ_PLAIN_CASE(SHM_Y)
        HQ_PLAIN_CASE(SHM_REAL)
        HQ_PLAIN_CASE(SHM_RX)
        HQ_PLAIN_CASE(SHM_GEN)
        HQ_PLAIN_CASE(SHM_DIAG)
        HQ_PLAIN_CASE(SHM_PHASE)
#undef HQ_PLAIN_CASE
    default: break;
    }
    HQ_HIP(hipGetLastError());
    return HQ_OK;
}

}  // namespace hq

using namespace hq;

extern "C" hq_status hq_apply_gates_plain(hq_state s, const hq_gate* gates, int num_gates)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_apply_gates_plain: null state");
    HQ_CHECK(num_gates >= 0 && (num_gates == 0 || gates), HQ_ERR_ARG, "hq_apply_gates_plain: bad gate array");
    for (int i = 0; i < num_gates; i++) {
        hq_status st = validate_gate(gates[i], s->n);
        if (st != HQ_OK) return st;
    }
    int slot;
    hq_status st = prof_begin(s, 0, &slot);
    if (st != HQ_OK) return st;
    for (int i = 0; i < num_gates; i++) {
        if (gates[i].type == HQ_ID) continue;
        st = s->precision == HQ_C128 ? launch_plain_gate<double>(s, gates[i]) : launch_plain_gate<float>(s, gates[i]);
        if (st != HQ_OK) return st;
    }
    return prof_end(s, slot);
}
