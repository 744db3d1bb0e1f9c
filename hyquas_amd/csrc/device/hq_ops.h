// hq_ops.h -- per-pair gate updates shared by the plain and OShareMem kernels.
//
// A gate is reduced on the host to one of the ShmClass forms below; each form
// is the 2x2 update of one (lo, hi) amplitude pair with the matrix entries
// m = {re00, im00, re01, im01, re10, im10, re11, im11}.  The reference writes
// one kernel per gate type with the same arithmetic (src/kernelSimple.cu:74-332,
// src/kernelOpt.cu:21-149); here the forms are chosen by matrix structure so a
// fused or user-supplied matrix gets the cheapest applicable update.
#pragma once

#include "hq_device.h"

namespace hq {

template <typename R>
struct Coef {
    R m[8];
};

template <int CLS, typename R>
__device__ __forceinline__ void pair_update(cplx<R>& lo, cplx<R>& hi, const R* __restrict__ m)
{
    const R lr = lo.x, li = lo.y, hr = hi.x, hm = hi.y;
    if constexpr (CLS == SHM_SWAP) {
        lo = hi;
        hi.x = lr; hi.y = li;
    } else if constexpr (CLS == SHM_Y) {
        lo.x = hm;  lo.y = -hr;
        hi.x = -li; hi.y = lr;
    } else if constexpr (CLS == SHM_REAL) {
        lo.x = m[0] * lr + m[2] * hr; lo.y = m[0] * li + m[2] * hm;
        hi.x = m[4] * lr + m[6] * hr; hi.y = m[4] * li + m[6] * hm;
    } else if constexpr (CLS == SHM_RX) {
        // lo' = a lo + i b hi, hi' = i c lo + d hi   (a=re00, b=im01, c=im10, d=re11)
        lo.x = m[0] * lr - m[3] * hm; lo.y = m[0] * li + m[3] * hr;
        hi.x = m[6] * hr - m[5] * li; hi.y = m[6] * hm + m[5] * lr;
    } else if constexpr (CLS == SHM_GEN) {
        lo.x = (m[0] * lr - m[1] * li) + (m[2] * hr - m[3] * hm);
        lo.y = (m[0] * li + m[1] * lr) + (m[2] * hm + m[3] * hr);
        hi.x = (m[4] * lr - m[5] * li) + (m[6] * hr - m[7] * hm);
        hi.y = (m[4] * li + m[5] * lr) + (m[6] * hm + m[7] * hr);
    } else if constexpr (CLS == SHM_DIAG) {
        lo.x = m[0] * lr - m[1] * li; lo.y = m[0] * li + m[1] * lr;
        hi.x = m[6] * hr - m[7] * hm; hi.y = m[6] * hm + m[7] * hr;
    } else if constexpr (CLS == SHM_HADD) {
        lo.x = lr + hr; lo.y = li + hm;
        hi.x = lr - hr; hi.y = li - hm;
    } else if constexpr (CLS == SHM_ROTA) {
        const R t = m[0];
        lo.x = lr - t * hr; lo.y = li - t * hm;
        hi.x = hr + t * lr; hi.y = hm + t * li;
    } else if constexpr (CLS == SHM_ROTB) {
        const R u = m[0];
        lo.x = u * lr - hr; lo.y = u * li - hm;
        hi.x = lr + u * hr; hi.y = li + u * hm;
    } else if constexpr (CLS == SHM_RXA) {
        const R t = m[0];
        lo.x = lr - t * hm; lo.y = li + t * hr;
        hi.x = hr - t * li; hi.y = hm + t * lr;
    } else if constexpr (CLS == SHM_RXB) {
        const R u = m[0];
        lo.x = u * lr - hm; lo.y = u * li + hr;
        hi.x = u * hr - li; hi.y = u * hm + lr;
    } else {  // SHM_PHASE
        hi.x = m[6] * hr - m[7] * hm; hi.y = m[6] * hm + m[7] * hr;
    }
}

// Multiply one amplitude by the complex factor (re, im).
template <typename R>
__device__ __forceinline__ void scale(cplx<R>& v, R re, R im)
{
    const R x = v.x, y = v.y;
    v.x = re * x - im * y;
    v.y = re * y + im * x;
}

}  // namespace hq

// Host-side classification (defined in pergate.hip).
namespace hq {
int classify_gate(const hq_gate& g);
bool gate_is_diagonal_type(int type);
int gate_num_controls(int type);
hq_status validate_gate(const hq_gate& g, int n);
}
