// sharemem_dev.h -- device body of the register-tiled OShareMem pass, shared by
// the single-state kernel (sharemem.hip) and the batched kernels (engine.hip).
//
// Reference behaviour (src/kernelOpt.cu:388-433): a block owns the 2^k
// amplitudes that differ only in the k tile ("shared") qubits, applies every
// gate of the group to them and writes them back.  Here the tile lives in
// registers (2^RB per thread, 2^TB threads); stages of gates on register bits
// are separated by one LDS exchange each.
#pragma once

#include "hq_ops.h"

namespace hq {

// Global-memory pointers.  The batched kernel reads its program and state
// pointers from an item table, which makes them generic (flat) pointers; the
// compiler takes every flat load for a per-lane value, so the op dispatch
// below would compile to divergent branches.  Casting to the global address
// space keeps the program's fields uniform (scalar loads, scalar branches).
// (The host pass of a HIP file sees the same templates: no address space there.)
#if defined(__HIP_DEVICE_COMPILE__)
#define HQ_GLOBAL __attribute__((address_space(1)))
#else
#define HQ_GLOBAL
#endif
template <typename T> __device__ __forceinline__ HQ_GLOBAL T* to_global(T* p) { return (HQ_GLOBAL T*)p; }

// tile_pass modes (bit set)
constexpr uint32_t TILE_INIT = 1;     // start from |0...0> instead of loading the tile
constexpr uint32_t TILE_NOSTORE = 2;  // do not write the tile back
constexpr uint32_t TILE_READ = 4;     // write {sum p, sum p*Z, sum p*ZZ} of the tile to part[3*blk]

template <typename R, int RB, int CLS, int P>
__device__ __forceinline__ void reg_pairs(cplx<R> (&a)[1 << RB], uint32_t creg, const R* m)
{
    // without register controls every pair is updated: no per-pair branch
    // (a branch per pair makes the compiler copy both amplitudes for the
    // path that skips the update)
    if (creg == 0) {
#pragma unroll
        for (int j = 0; j < (1 << RB); j++) {
            if (j & (1 << P)) continue;
            pair_update<CLS, R>(a[j], a[j | (1 << P)], m);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) {
        if (j & (1 << P)) continue;
        if ((j & creg) != creg) continue;
        pair_update<CLS, R>(a[j], a[j | (1 << P)], m);
    }
}

template <typename R, int RB, int P>
__device__ __forceinline__ void reg_op(cplx<R> (&a)[1 << RB], int cls, uint32_t creg, const R* m)
{
    switch (cls) {
    case SHM_SWAP: reg_pairs<R, RB, SHM_SWAP, P>(a, creg, m); break;
    case SHM_Y: reg_pairs<R, RB, SHM_Y, P>(a, creg, m); break;
    case SHM_REAL: reg_pairs<R, RB, SHM_REAL, P>(a, creg, m); break;
    case SHM_RX: reg_pairs<R, RB, SHM_RX, P>(a, creg, m); break;
    case SHM_GEN: reg_pairs<R, RB, SHM_GEN, P>(a, creg, m); break;
    case SHM_DIAG: reg_pairs<R, RB, SHM_DIAG, P>(a, creg, m); break;
    case SHM_HADD: reg_pairs<R, RB, SHM_HADD, P>(a, creg, m); break;
    case SHM_ROTA: reg_pairs<R, RB, SHM_ROTA, P>(a, creg, m); break;
    case SHM_ROTB: reg_pairs<R, RB, SHM_ROTB, P>(a, creg, m); break;
    case SHM_RXA: reg_pairs<R, RB, SHM_RXA, P>(a, creg, m); break;
    case SHM_RXB: reg_pairs<R, RB, SHM_RXB, P>(a, creg, m); break;
    default: reg_pairs<R, RB, SHM_PHASE, P>(a, creg, m); break;
    }
}

template <typename R, int RB>
__device__ __forceinline__ void apply_op(cplx<R> (&a)[1 << RB], const HQ_GLOBAL ShmOp* __restrict__ op, uint32_t tid,
                                         uint64_t blk)
{
    // Every branch below is uniform over the block (the op fields and the
    // block index are): a lane whose thread-bit controls are not all set gets
    // the identity's coefficients instead of leaving the op, so no branch
    // depends on the lane (a per-lane exit made every later dispatch a
    // divergent branch).  Thread-controlled ops only use classes that can
    // express the identity (the planner turns SWAP into REAL and Y into RX).
    const uint64_t cblk = op->cblk;
    if ((blk & cblk) != cblk) return;                 // block control not set (uniform)
    const uint32_t cthr = op->cthr;
    const bool on = (tid & cthr) == cthr;
    R m[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const R id = (i == 0 || i == 6) ? R(1) : R(0);
        const R v = (R)op->m[i];
        m[i] = on ? v : id;
    }
    const uint32_t creg = op->creg;
    const int cls = op->cls;
    const int tpos = op->tpos;
    if (cls == SHM_SCALE) {
#pragma unroll
        for (int j = 0; j < (1 << RB); j++) scale(a[j], m[0], m[1]);
        return;
    }
    if (op->where == SHM_IN_REG) {
        switch (tpos) {
        case 0: reg_op<R, RB, 0>(a, cls, creg, m); break;
        case 1: if constexpr (RB > 1) reg_op<R, RB, 1>(a, cls, creg, m); break;
        case 2: if constexpr (RB > 2) reg_op<R, RB, 2>(a, cls, creg, m); break;
        case 3: if constexpr (RB > 3) reg_op<R, RB, 3>(a, cls, creg, m); break;
        case 4: if constexpr (RB > 4) reg_op<R, RB, 4>(a, cls, creg, m); break;
        default: if constexpr (RB > 5) reg_op<R, RB, 5>(a, cls, creg, m); break;
        }
        static_assert(RB <= 6, "register positions above 5 are not dispatched");
        return;
    }
    // diagonal gate whose target is a thread bit or a block bit: one factor
    // per lane for every register amplitude (a phase leaves the |0> half alone)
    const uint32_t hb = op->where == SHM_IN_THREAD ? (tid >> tpos) & 1u : (uint32_t)((blk >> tpos) & 1u);
    const R fr = hb ? m[6] : (cls == SHM_PHASE ? R(1) : m[0]);
    const R fi = hb ? m[7] : (cls == SHM_PHASE ? R(0) : m[1]);
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) {
        if ((j & creg) != creg) continue;
        scale(a[j], fr, fi);
    }
}

template <int K>
__device__ __forceinline__ uint32_t swz(uint32_t x, const uint32_t* __restrict__ col)
{
    uint32_t f = 0;
#pragma unroll
    for (int i = 3; i < K; i++) f ^= ((x >> i) & 1u) ? col[i] : 0u;
    return x ^ f;
}

// Z-sum and nearest-neighbour ZZ-sum eigenvalues of basis state idx over n qubits:
// sum_q (1 - 2 b_q) and sum_{q<n-1} (1 - 2 (b_q xor b_{q+1})).
__device__ __forceinline__ void zsums(uint64_t idx, int n, double* z, double* zz)
{
    // __popcll returns unsigned: keep the arithmetic signed
    *z = (double)(n - 2 * (int)__popcll(idx));
    const uint64_t m = n >= 2 ? ((1ull << (n - 1)) - 1) : 0;
    *zz = (double)((n - 1) - 2 * (int)__popcll((idx ^ (idx >> 1)) & m));
}

// One OShareMem pass of tile `blk` of the group program `prog` over `psi`.
// `mode` is uniform over the grid; the single-state kernel passes 0.
template <typename R, int RB, int TB>
__device__ __forceinline__ void tile_pass(cplx<R>* __restrict__ psi_any, const char* __restrict__ prog_any,
                                          uint64_t blk, cplx<R>* __restrict__ lds, uint32_t mode, int n,
                                          double* __restrict__ part_any)
{
    HQ_GLOBAL cplx<R>* __restrict__ psi = to_global(psi_any);
    const HQ_GLOBAL char* __restrict__ prog = to_global(prog_any);
    HQ_GLOBAL double* __restrict__ part = to_global(part_any);
    constexpr int NR = 1 << RB;
    constexpr int K = RB + TB;

    const HQ_GLOBAL ShmHeader* __restrict__ H = (const HQ_GLOBAL ShmHeader*)prog;
    const HQ_GLOBAL ShmStage* __restrict__ ST = (const HQ_GLOBAL ShmStage*)(prog + sizeof(ShmHeader));
    const uint32_t nst = H->num_stages;
    const HQ_GLOBAL ShmOp* __restrict__ OP =
        (const HQ_GLOBAL ShmOp*)(prog + sizeof(ShmHeader) + nst * sizeof(ShmStage));
    const uint32_t tid = threadIdx.x;

    // block base address: pdep(blk, block bits)
    uint64_t base = 0;
    const uint32_t nbb = H->num_block_bits;
    for (uint32_t j = 0; j < nbb; j++)
        if ((blk >> j) & 1) base |= H->block_phys[j];
    // a tile known to be zero stays zero (uniform over the block)
    if (!(mode & TILE_READ) && (base & H->zero_mask)) return;

    uint32_t col[K];
#pragma unroll
    for (int i = 0; i < K; i++) col[i] = H->swz_col[i];

    cplx<R> a[NR];

    // mapping of stage 0
    uint32_t rbit[RB];
    uint64_t rphys[RB];
#pragma unroll
    for (int i = 0; i < RB; i++) {
        rbit[i] = 1u << ST[0].reg_bits[i];
        rphys[i] = H->tile_phys[ST[0].reg_bits[i]];
    }
    uint32_t et = 0;
    uint64_t pt = 0;
#pragma unroll
    for (int i = 0; i < TB; i++) {
        uint32_t b = ST[0].thr_bits[i];
        if ((tid >> i) & 1) { et |= 1u << b; pt |= H->tile_phys[b]; }
    }
    if (mode == TILE_INIT && base != 0) {
        // first pass of a run: a tile other than the origin's starts at zero
        // and stays zero under the group's linear gates
        HQ_GLOBAL cplx<R>* __restrict__ dst = psi + (base | pt);
        cplx<R> z;
        z.x = R(0);
        z.y = R(0);
#pragma unroll
        for (int j = 0; j < NR; j++) {
            uint64_t off = 0;
#pragma unroll
            for (int i = 0; i < RB; i++) if ((j >> i) & 1) off |= rphys[i];
            dst[off] = z;
        }
        return;
    }
    if (mode & TILE_INIT) {
        const bool origin = (base | pt) == 0;
#pragma unroll
        for (int j = 0; j < NR; j++) {
            a[j].x = (origin && j == 0) ? R(1) : R(0);
            a[j].y = R(0);
        }
    } else {
        const HQ_GLOBAL cplx<R>* __restrict__ src = psi + (base | pt);
#pragma unroll
        for (int j = 0; j < NR; j++) {
            uint64_t off = 0;
#pragma unroll
            for (int i = 0; i < RB; i++) if ((j >> i) & 1) off |= rphys[i];
            a[j] = src[off];
        }
    }

    for (uint32_t s = 0; s < nst; s++) {
        if (s > 0) {
            // exchange: write under the old mapping, read under the new one.
            // The swizzle is linear over xor, so a register's LDS index is the
            // lane's swizzled index xor the swizzled bits of the register index
            const uint32_t st = swz<K>(et, col);
            uint32_t srb[RB];
#pragma unroll
            for (int i = 0; i < RB; i++) srb[i] = swz<K>(rbit[i], col);
#pragma unroll
            for (int j = 0; j < NR; j++) {
                uint32_t so = 0;
#pragma unroll
                for (int i = 0; i < RB; i++) if ((j >> i) & 1) so ^= srb[i];
                lds[st ^ so] = a[j];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < RB; i++) rbit[i] = 1u << ST[s].reg_bits[i];
            et = 0;
#pragma unroll
            for (int i = 0; i < TB; i++)
                if ((tid >> i) & 1) et |= 1u << ST[s].thr_bits[i];
            const uint32_t st2 = swz<K>(et, col);
#pragma unroll
            for (int i = 0; i < RB; i++) srb[i] = swz<K>(rbit[i], col);
#pragma unroll
            for (int j = 0; j < NR; j++) {
                uint32_t so = 0;
#pragma unroll
                for (int i = 0; i < RB; i++) if ((j >> i) & 1) so ^= srb[i];
                a[j] = lds[st2 ^ so];
            }
            // the next exchange overwrites lds: every lane must have read first
            if (s + 1 < nst) __syncthreads();
        }
        const uint32_t ob = ST[s].op_begin, oe = ST[s].op_end;
        for (uint32_t o = ob; o < oe; o++) apply_op<R, RB>(a, OP + o, tid, blk);
    }

    // mapping of the last stage
    pt = 0;
#pragma unroll
    for (int i = 0; i < TB; i++)
        if ((et >> ST[nst - 1].thr_bits[i]) & 1) pt |= H->tile_phys[ST[nst - 1].thr_bits[i]];
#pragma unroll
    for (int i = 0; i < RB; i++) rphys[i] = H->tile_phys[ST[nst - 1].reg_bits[i]];

    if (mode & TILE_READ) {
        double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int j = 0; j < NR; j++) {
            uint64_t off = 0;
#pragma unroll
            for (int i = 0; i < RB; i++) if ((j >> i) & 1) off |= rphys[i];
            const double re = (double)a[j].x, im = (double)a[j].y;
            const double p = re * re + im * im;
            double z, zz;
            zsums(base | pt | off, n, &z, &zz);
            s0 += p;
            s1 += p * z;
            s2 += p * zz;
        }
        // fixed-order block reduction: xor butterfly in the wave, then waves in order
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
            s2 += __shfl_xor(s2, o);
        }
        __shared__ double red[3][(1 << SHM_MAX_TB) / 64];
        const uint32_t w = tid >> 6;
        if ((tid & 63) == 0) { red[0][w] = s0; red[1][w] = s1; red[2][w] = s2; }
        __syncthreads();
        if (tid == 0) {
            double t0 = 0, t1 = 0, t2 = 0;
            for (uint32_t i = 0; i < ((1u << TB) >> 6); i++) { t0 += red[0][i]; t1 += red[1][i]; t2 += red[2][i]; }
            part[3 * blk + 0] = t0;
            part[3 * blk + 1] = t1;
            part[3 * blk + 2] = t2;
        }
    }

    if (!(mode & TILE_NOSTORE)) {
        HQ_GLOBAL cplx<R>* __restrict__ dst = psi + (base | pt);
#pragma unroll
        for (int j = 0; j < NR; j++) {
            uint64_t off = 0;
#pragma unroll
            for (int i = 0; i < RB; i++) if ((j >> i) & 1) off |= rphys[i];
            dst[off] = a[j];
        }
    }
}

}  // namespace hq
