// transmm.hip -- TransMM: bit-permutation transpose + complex GEMM on MFMA.
//
// Reference behaviour (src/executor.cpp:533-551, src/schedule.cpp:754-783):
// cuttExecute moves the group's qubits to the low bit positions (state ->
// buffer), then cublasZgemm multiplies every run of K = 2^matQubit
// consecutive amplitudes by the fused K x K matrix (buffer -> state).
//
// MI355X design: the transpose is a tiled bit permutation through LDS (tile
// spans the low bits of both the input and the output so reads and writes
// are 256-byte runs).  The GEMM is written for MFMA: one wave per 16-row
// block of U, U's fragments held in registers for the whole (persistent)
// kernel, B tiles staged through LDS as re/im planes, the complex product as
// four real MFMAs (v_mfma_f64_16x16x4f64 for complex128,
// v_mfma_f32_16x16x4f32 with FP32 accumulation for complex64), and the C
// tile staged back through LDS for coalesced stores.  Both passes ping-pong
// between the state and the work buffer; the handle's data pointer follows.

#include "hq_ops.h"

#include <algorithm>
#include <cstring>

namespace hq {

// ---- transpose ---------------------------------------------------------------------

constexpr int TR_TILE_BITS = 10;    // 1024 amplitudes per block
constexpr int TR_THREADS = 256;

struct TransposeArgs {
    uint64_t out_block_phys[64];   // output bit (mask) of block-index bit j
    uint64_t in_block_phys[64];    // input bit (mask) of block-index bit j
    uint64_t in_tile_phys[TR_TILE_BITS];   // input bit of input-tile bit i
    uint64_t out_tile_phys[TR_TILE_BITS];  // output bit of output-tile bit i
    uint32_t in2out[TR_TILE_BITS];         // output-tile bit of input-tile bit i
    uint32_t num_block_bits;
};

__device__ __forceinline__ uint32_t tr_swz(uint32_t x) { return x ^ ((x >> 5) & 31); }

template <typename R>
__global__ __launch_bounds__(TR_THREADS) void k_transpose(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                          const TransposeArgs* __restrict__ A)
{
    __shared__ cplx<R> lds[1 << TR_TILE_BITS];
    const uint64_t blk = blockIdx.x;
    uint64_t bin = 0, bout = 0;
    const uint32_t nbb = A->num_block_bits;
    for (uint32_t j = 0; j < nbb; j++)
        if ((blk >> j) & 1) { bin |= A->in_block_phys[j]; bout |= A->out_block_phys[j]; }
    const uint32_t tid = threadIdx.x;
    constexpr int TBITS = 8;   // log2(TR_THREADS)
    constexpr int PER = (1 << TR_TILE_BITS) / TR_THREADS;
    // per-lane parts of the three index maps
    uint64_t lane_in = 0, lane_out = 0;
    uint32_t lane_lds = 0;
#pragma unroll
    for (int i = 0; i < TBITS; i++) {
        if ((tid >> i) & 1) {
            lane_in |= A->in_tile_phys[i];
            lane_lds |= 1u << A->in2out[i];
            lane_out |= A->out_tile_phys[i];
        }
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        uint64_t ui = 0;
        uint32_t ul = 0;
#pragma unroll
        for (int i = 0; i < TR_TILE_BITS - TBITS; i++)
            if ((r >> i) & 1) { ui |= A->in_tile_phys[TBITS + i]; ul |= 1u << A->in2out[TBITS + i]; }
        lds[tr_swz(lane_lds | ul)] = in[bin | lane_in | ui];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; r++) {
        uint64_t uo = 0;
#pragma unroll
        for (int i = 0; i < TR_TILE_BITS - TBITS; i++)
            if ((r >> i) & 1) uo |= A->out_tile_phys[TBITS + i];
        uint32_t e = tid | ((uint32_t)r << TBITS);
        out[bout | lane_out | uo] = lds[tr_swz(e)];
    }
}

// Small states (n < 10): one thread per amplitude, direct gather.
template <typename R>
__global__ void k_transpose_small(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int n,
                                  const int* __restrict__ perm_dev)
{
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ((uint64_t)1 << n)) return;
    uint64_t src = 0;
    for (int b = 0; b < n; b++)
        if ((j >> b) & 1) src |= (uint64_t)1 << perm_dev[b];
    out[j] = in[src];
}

struct TransposePlan {
    TransposeArgs args;
    int n;
};

static void plan_transpose(int n, const int* perm, int run_bits, TransposePlan* P)
{
    // output bit b takes input bit perm[b]
    memset(P, 0, sizeof(*P));
    P->n = n;
    int inv[64];
    for (int b = 0; b < n; b++) inv[perm[b]] = b;
    // tile bits over OUTPUT positions: output low bits, the output positions of
    // the input low bits, then more output low bits up to TR_TILE_BITS
    uint64_t T = 0;
    int a = std::min(run_bits, n);
    for (int b = 0; b < a; b++) T |= 1ull << b;
    for (int p = 0; p < a; p++) T |= 1ull << inv[p];
    for (int b = 0; b < n && __builtin_popcountll(T) < TR_TILE_BITS; b++) T |= 1ull << b;
    // output-tile bits in ascending output position
    int out_bits[TR_TILE_BITS], no = 0;
    for (int b = 0; b < n; b++) if ((T >> b) & 1) out_bits[no++] = b;
    // input-tile bits: the same qubits in ascending INPUT position
    int in_pos[TR_TILE_BITS], ni = 0;
    for (int p = 0; p < n; p++) if ((T >> inv[p]) & 1) in_pos[ni++] = p;
    for (int i = 0; i < TR_TILE_BITS; i++) {
        P->args.in_tile_phys[i] = 1ull << in_pos[i];
        P->args.out_tile_phys[i] = 1ull << out_bits[i];
        int ob = inv[in_pos[i]];
        int rank = 0;
        for (int k = 0; k < TR_TILE_BITS; k++) if (out_bits[k] == ob) rank = k;
        P->args.in2out[i] = (uint32_t)rank;
    }
    int j = 0;
    for (int b = 0; b < n; b++) {
        if ((T >> b) & 1) continue;
        P->args.out_block_phys[j] = 1ull << b;
        P->args.in_block_phys[j] = 1ull << perm[b];
        j++;
    }
    P->args.num_block_bits = (uint32_t)j;
}

hq_status transpose_prepare(hq_state s, const int* perm, void** dev_args)
{
    // device-resident launch arguments, built once per program
    *dev_args = nullptr;
    const int n = s->n;
    if (n < TR_TILE_BITS) {
        HQ_HIP(hipMalloc(dev_args, sizeof(int) * n));
        HQ_HIP(hipMemcpy(*dev_args, perm, sizeof(int) * n, hipMemcpyHostToDevice));
        return HQ_OK;
    }
    TransposePlan P;
    plan_transpose(n, perm, s->precision == HQ_C128 ? 4 : 5, &P);
    HQ_HIP(hipMalloc(dev_args, sizeof(TransposeArgs)));
    HQ_HIP(hipMemcpy(*dev_args, &P.args, sizeof(TransposeArgs), hipMemcpyHostToDevice));
    return HQ_OK;
}

hq_status transpose_launch(hq_state s, const void* dev_args, void* in, void* out)
{
    const int n = s->n;
    if (n < TR_TILE_BITS) {
        unsigned g = (unsigned)((s->amps + 255) / 256);
        if (s->precision == HQ_C128)
            hipLaunchKernelGGL(k_transpose_small<double>, dim3(g), dim3(256), 0, s->stream,
                               (const cplx<double>*)in, (cplx<double>*)out, n, (const int*)dev_args);
        else
            hipLaunchKernelGGL(k_transpose_small<float>, dim3(g), dim3(256), 0, s->stream,
                               (const cplx<float>*)in, (cplx<float>*)out, n, (const int*)dev_args);
        HQ_HIP(hipGetLastError());
        return HQ_OK;
    }
    uint64_t blocks = s->amps >> TR_TILE_BITS;
    const TransposeArgs* A = (const TransposeArgs*)dev_args;
    if (s->precision == HQ_C128)
        hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)blocks), dim3(TR_THREADS), 0, s->stream,
                           (const cplx<double>*)in, (cplx<double>*)out, A);
    else
        hipLaunchKernelGGL(k_transpose<float>, dim3((unsigned)blocks), dim3(TR_THREADS), 0, s->stream,
                           (const cplx<float>*)in, (cplx<float>*)out, A);
    HQ_HIP(hipGetLastError());
    return HQ_OK;
}

// ---- dense K x K on runs of K amplitudes -------------------------------------------------

// Generic path (any K <= 128): one thread per output amplitude, U from global.
template <typename R>
__global__ __launch_bounds__(256) void k_dense_simple(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                      const R* __restrict__ ure, const R* __restrict__ uim,
                                                      int mq, uint64_t n_amps)
{
    const uint64_t K = (uint64_t)1 << mq;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_amps; i += stride) {
        uint64_t col = i >> mq, row = i & (K - 1);
        const cplx<R>* b = in + col * K;
        double sr = 0, si = 0;
        for (uint64_t j = 0; j < K; j++) {
            double ar = ure[j * K + row], ai = uim[j * K + row];
            cplx<R> v = b[j];
            sr += ar * (double)v.x - ai * (double)v.y;
            si += ar * (double)v.y + ai * (double)v.x;
        }
        cplx<R> o;
        o.x = (R)sr;
        o.y = (R)si;
        out[i] = o;
    }
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R> struct MfmaAcc;
template <> struct MfmaAcc<double> { using T = v4d; };
template <> struct MfmaAcc<float> { using T = v4f; };

__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16x16x4(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Output row of accumulator register v in lane l (column l%16).  Measured on
// gfx950: v_mfma_f32_16x16x4f32 holds rows 4*(l/16)+v, v_mfma_f64_16x16x4f64
// holds rows l/16 + 4*v.  Inputs: A[l%16][l/16], B[l/16][l%16] for both.
template <typename R> __device__ __forceinline__ int mfma_row(int lane, int v);
template <> __device__ __forceinline__ int mfma_row<float>(int lane, int v) { return 4 * (lane >> 4) + v; }
template <> __device__ __forceinline__ int mfma_row<double>(int lane, int v) { return (lane >> 4) + 4 * v; }

// Scatter the low bits of x to the set bits of m, in order (x's bit j goes
// to m's j-th lowest set bit), and the inverse gather.
__device__ __forceinline__ uint64_t deposit(uint64_t x, uint64_t m)
{
    uint64_t r = 0;
    for (uint64_t b = 1; m; b <<= 1, m &= m - 1)
        if (x & b) r |= m & (~m + 1);
    return r;
}
__device__ __forceinline__ uint64_t extract(uint64_t x, uint64_t m)
{
    uint64_t r = 0;
    for (uint64_t b = 1; m; b <<= 1, m &= m - 1)
        if (x & m & (~m + 1)) r |= b;
    return r;
}

// Which amplitudes a GEMM tile holds: the K matrix rows are the amplitude
// bits `rows` (row index bit j = the j-th lowest bit of `rows`), the NT tile
// columns the bits `cols`, the tile index the remaining bits `blocks`.  With
// rows = the MQ lowest bits and cols the next ones this is the contiguous
// layout a transpose leaves; any other placement applies U to those qubits
// where they are, in one pass.
struct DenseMasks {
    uint64_t rows, cols, blocks;
};

// One wave per 16-row block of U (K/16 waves), NT columns per tile.
//
// GAUSS: the complex product with three real MFMAs instead of four (Gauss's
// method): with T1 = Ar Br, T2 = Ai Bi and T3 = (Ar + Ai)(Br + Bi) summed
// over k, Re C = T1 - T2 and Im C = T3 - T1 - T2.  A quarter fewer matrix-core
// operations for the same product; the sums Ar + Ai are formed once per
// kernel, Br + Bi once per B element read.  The extra rounding is one
// addition per operand and one per output (FP64: ~1e-16 relative).
//
// Pipeline: two LDS buffers.  A block loads its next tile's B from HBM into
// registers before the matrix-core loop of the current tile and stages it
// into the other buffer afterwards, so the load's latency hides behind the
// MFMAs; the C tile goes back through the current buffer for coalesced
// stores.  Two barriers per tile.
//
// Addressing: element t of a tile (t over the bits rows|cols in increasing
// order, so consecutive t are consecutive amplitudes as far as those bits
// are contiguous) sits at deposit(tile, blocks) + deposit(t, rows|cols);
// thread x handles t = x + i * NTH, whose offset splits into a per-thread
// part (deposit of x) and a uniform part (deposit of i * NTH), and likewise
// its LDS slot.  `in` may equal `out`: a block reads a whole tile before it
// writes it, and tiles are disjoint.
template <typename R, int MQ, int NT, bool GAUSS>
__global__ __launch_bounds__(64 * ((1 << MQ) / 16)) void k_dense_mfma(const cplx<R>* in, cplx<R>* out,
                                                                     const R* __restrict__ ure,
                                                                     const R* __restrict__ uim, DenseMasks dm,
                                                                     uint64_t ntiles)
{
    constexpr int K = 1 << MQ;
    constexpr int KS = K / 4;          // k-steps
    constexpr int NCB = NT / 16;       // column blocks per tile
    constexpr int NTP = NT + 1;        // padded plane row
    constexpr int NTH = 64 * (K / 16); // threads per block
    constexpr int PER = NT * K / NTH;  // B elements per thread per tile
    static_assert(NT * K % NTH == 0, "tile must split evenly over the block");
    using Acc = typename MfmaAcc<R>::T;
    __shared__ R bre[2][K * NTP];
    __shared__ R bim[2][K * NTP];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int arow = wave * 16 + (lane & 15);
    const int kq = lane >> 4;          // 0..3

    // U fragments of this wave's row block: A[row][4q + kq]; at K = 128 the
    // sums Ar + Ai are formed per use (registers are full)
    constexpr bool KEEP_AS = GAUSS;
    R ar[KS], ai[KS], as[KEEP_AS ? KS : 1];
#pragma unroll
    for (int q = 0; q < KS; q++) {
        int kk = 4 * q + kq;
        ar[q] = ure[kk * K + arow];     // column-major: U(row, col) at col*K + row
        ai[q] = uim[kk * K + arow];
        if constexpr (KEEP_AS) as[q] = ar[q] + ai[q];
    }

    uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;         // uniform: the grid never exceeds the tiles
    const uint64_t tm = dm.rows | dm.cols;
    const uint64_t aoff = deposit(threadIdx.x, tm);
    const int alds = (int)extract(aoff, dm.rows) * NTP + (int)extract(aoff, dm.cols);
    // i * NTH sets only the top log2(PER) tile-index bits: element i's
    // uniform offset is a sum over i's bits of one amplitude offset and one
    // LDS offset per bit (kept in scalar registers)
    constexpr int PB = __builtin_ctz((unsigned)PER);
    uint64_t hoff[PB > 0 ? PB : 1];
    int hlds[PB > 0 ? PB : 1];
#pragma unroll
    for (int j = 0; j < PB; j++) {
        hoff[j] = deposit((uint64_t)NTH << j, tm);
        hlds[j] = (int)extract(hoff[j], dm.rows) * NTP + (int)extract(hoff[j], dm.cols);
    }
    auto boff = [&](int i) {
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < PB; j++)
            if ((i >> j) & 1) o += hoff[j];
        return o;
    };
    auto blds = [&](int i) {
        int o = 0;
#pragma unroll
        for (int j = 0; j < PB; j++)
            if ((i >> j) & 1) o += hlds[j];
        return o;
    };
    // tile bases: this block's tiles are blockIdx.x + k * gridDim.x (gridDim a
    // power of two), so the base is deposit(blockIdx.x, blocks) plus the k-th
    // subset of the block bits above the lowest log2(gridDim) ones, stepped
    // with ((cur | ~hi) + 1) & hi
    uint64_t hi = dm.blocks;
    for (unsigned g = gridDim.x; g > 1; g >>= 1) hi &= hi - 1;
    const uint64_t base0 = deposit(tile, dm.blocks);
    auto next_sub = [hi](uint64_t cur) { return ((cur | ~hi) + 1) & hi; };
    cplx<R> pf[PER];
    auto load = [&](uint64_t base) {
        const cplx<R>* src = in + base + aoff;
#pragma unroll
        for (int i = 0; i < PER; i++) pf[i] = src[boff(i)];
    };
    // B (K x NT) as planes [k][col]
    auto stage = [&](int buf) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            bre[buf][alds + blds(i)] = pf[i].x;
            bim[buf][alds + blds(i)] = pf[i].y;
        }
    };
    uint64_t sub = 0;                   // the current tile's subset of hi
    load(base0);
    stage(0);
    __syncthreads();
    for (int buf = 0; tile < ntiles; tile += gridDim.x, buf ^= 1, sub = next_sub(sub)) {
        const bool more = tile + gridDim.x < ntiles;    // uniform over the block
        if (more) load(base0 + next_sub(sub));
        Acc cr[NCB], ci[NCB], cs[GAUSS ? NCB : 1];
#pragma unroll
        for (int c = 0; c < NCB; c++) {
            cr[c] = Acc{0, 0, 0, 0};
            ci[c] = Acc{0, 0, 0, 0};
            if constexpr (GAUSS) cs[c] = Acc{0, 0, 0, 0};
        }
        const R* __restrict__ pre = bre[buf];
        const R* __restrict__ pim = bim[buf];
#pragma unroll
        for (int q = 0; q < KS; q++) {
            const int kk = 4 * q + kq;
#pragma unroll
            for (int c = 0; c < NCB; c++) {
                R br = pre[kk * NTP + c * 16 + (lane & 15)];
                R bi = pim[kk * NTP + c * 16 + (lane & 15)];
                if constexpr (GAUSS) {
                    R asq;
                    if constexpr (KEEP_AS) asq = as[q];
                    else asq = ar[q] + ai[q];
                    cr[c] = mfma16x16x4(ar[q], br, cr[c]);        // T1
                    ci[c] = mfma16x16x4(ai[q], bi, ci[c]);        // T2
                    cs[c] = mfma16x16x4(asq, br + bi, cs[c]);     // T3
                } else {
                    cr[c] = mfma16x16x4(ar[q], br, cr[c]);
                    cr[c] = mfma16x16x4(ai[q], -bi, cr[c]);
                    ci[c] = mfma16x16x4(ar[q], bi, ci[c]);
                    ci[c] = mfma16x16x4(ai[q], br, ci[c]);
                }
            }
        }
        if constexpr (GAUSS) {
#pragma unroll
            for (int c = 0; c < NCB; c++) {
                const Acc t1 = cr[c], t2 = ci[c];
                cr[c] = t1 - t2;
                ci[c] = cs[c] - t1 - t2;
            }
        }
        __syncthreads();                // every wave is done reading B from `buf`
        // C tile into `buf` (as [row][col]); the next B into the other buffer
        // (its last readers finished before the barrier above)
#pragma unroll
        for (int c = 0; c < NCB; c++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int row = wave * 16 + mfma_row<R>(lane, v);
                int col = c * 16 + (lane & 15);
                bre[buf][row * NTP + col] = cr[c][v];
                bim[buf][row * NTP + col] = ci[c][v];
            }
        }
        if (more) stage(buf ^ 1);
        __syncthreads();
        cplx<R>* dst = out + base0 + sub + aoff;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            cplx<R> v;
            v.x = bre[buf][alds + blds(i)];
            v.y = bim[buf][alds + blds(i)];
            dst[boff(i)] = v;
        }
    }
}

// complex128 uses Gauss's three-product form (FP64 matrix cores are the
// limit at K >= 64); complex64 keeps four products (FP32 matrix cores run at
// twice the FP64 rate and Gauss's form adds rounding at single precision)
template <typename R, int MQ, int NT>
static hq_status launch_mfma(hq_state s, const R* ure, const R* uim, const void* in, void* out, DenseMasks dm)
{
    constexpr bool gauss = sizeof(R) == 8;
    constexpr int K = 1 << MQ;
    const uint64_t ntiles = s->amps / ((uint64_t)K * NT);
    const uint64_t grid = std::min<uint64_t>(ntiles, 2048);
    hipLaunchKernelGGL((k_dense_mfma<R, MQ, NT, gauss>), dim3((unsigned)grid), dim3(64 * (K / 16)), 0, s->stream,
                       (const cplx<R>*)in, (cplx<R>*)out, ure, uim, dm, ntiles);
    HQ_HIP(hipGetLastError());
    return HQ_OK;
}

// Column tile width: complex128 (three products, three accumulators per
// column block, the next tile's B held in registers) takes 32-column tiles to
// stay within 256 registers (two waves per SIMD)
template <typename R> constexpr int dense_nt(int mq) { return sizeof(R) == 8 || mq == 7 ? 32 : 64; }

template <typename R>
static hq_status launch_mfma_masks(hq_state s, const R* ure, const R* uim, int mq, const void* in, void* out,
                                   DenseMasks dm)
{
    switch (mq) {
    case 4: return launch_mfma<R, 4, dense_nt<R>(4)>(s, ure, uim, in, out, dm);
    case 5: return launch_mfma<R, 5, dense_nt<R>(5)>(s, ure, uim, in, out, dm);
    case 6: return launch_mfma<R, 6, dense_nt<R>(6)>(s, ure, uim, in, out, dm);
    case 7: return launch_mfma<R, 7, dense_nt<R>(7)>(s, ure, uim, in, out, dm);
    default: set_error("dense kernel: matrix qubits must be 4..7"); return HQ_ERR_UNSUPPORTED;
    }
}

static uint64_t all_bits(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

template <typename R>
static hq_status launch_dense_t(hq_state s, const void* d_mat, int mq, void* in, void* out)
{
    const uint64_t K = (uint64_t)1 << mq;
    const R* ure = (const R*)d_mat;
    const R* uim = ure + K * K;
    // MFMA path: the matrix qubits are the lowest, the tile columns the next
    const int nt = mq >= 4 && mq <= 7 ? dense_nt<R>(mq) : 0;
    if (nt && s->amps >= K * nt) {
        DenseMasks dm;
        dm.rows = K - 1;
        dm.cols = (uint64_t)(nt - 1) << mq;
        dm.blocks = all_bits(s->n) & ~(dm.rows | dm.cols);
        return launch_mfma_masks<R>(s, ure, uim, mq, in, out, dm);
    }
    uint64_t blocks = std::min<uint64_t>((s->amps + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(k_dense_simple<R>, dim3((unsigned)blocks), dim3(256), 0, s->stream, (const cplx<R>*)in,
                       (cplx<R>*)out, ure, uim, mq, s->amps);
    HQ_HIP(hipGetLastError());
    return HQ_OK;
}

hq_status launch_dense(hq_state s, const void* d_mat, int mq, void* in, void* out)
{
    return s->precision == HQ_C128 ? launch_dense_t<double>(s, d_mat, mq, in, out)
                                   : launch_dense_t<float>(s, d_mat, mq, in, out);
}

int dense_at_min_qubits(int precision, int mq)
{
    if (mq < 4 || mq > 7) return 1 << 20;
    const int nt = precision == HQ_C128 ? dense_nt<double>(mq) : dense_nt<float>(mq);
    return mq + __builtin_ctz((unsigned)nt);
}

// U on the qubits at the set bits of `rows` (row index bit j = the j-th
// lowest of them), in place: tile columns along the lowest other bits.
template <typename R>
static hq_status launch_dense_at_t(hq_state s, const void* d_mat, int mq, uint64_t rows)
{
    const uint64_t K = (uint64_t)1 << mq;
    const R* ure = (const R*)d_mat;
    const R* uim = ure + K * K;
    const int nt = dense_nt<R>(mq);
    DenseMasks dm;
    dm.rows = rows;
    dm.cols = 0;
    for (int b = 0, got = 0; b < s->n && got < __builtin_ctz((unsigned)nt); b++)
        if (!((rows >> b) & 1)) { dm.cols |= 1ull << b; got++; }
    dm.blocks = all_bits(s->n) & ~(dm.rows | dm.cols);
    return launch_mfma_masks<R>(s, ure, uim, mq, s->data, s->data, dm);
}

hq_status launch_dense_at(hq_state s, const void* d_mat, int mq, uint64_t rows)
{
    HQ_CHECK(__builtin_popcountll((unsigned long long)rows) == mq && (rows & ~all_bits(s->n)) == 0 &&
                 s->n >= dense_at_min_qubits(s->precision, mq),
             HQ_ERR_ARG, "dense kernel: rows must name mat_qubits qubits of a large enough state");
    return s->precision == HQ_C128 ? launch_dense_at_t<double>(s, d_mat, mq, rows)
                                   : launch_dense_at_t<float>(s, d_mat, mq, rows);
}

}  // namespace hq
