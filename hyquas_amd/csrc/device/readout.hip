// readout.hip -- measurement sampling of a device state.
//
// The reference reads a state through kernelMeasure (P(qubit = 0),
// src/kernelSimple.cu:497-516) and full copies; drawing measurement shots
// is the step after it.  A shot is the smallest basis index i whose
// cumulative probability exceeds u * total, u uniform in [0, 1) from the
// portable generator (xoshiro256** seeded by SplitMix64, u = (x >> 11) 2^-53).
//
// Device work: one pass computes the probability sum of every 4096-amplitude
// chunk (fixed order); the host forms the chunk prefix (FP64, in order),
// picks each shot's chunk and residual; one thread per shot then walks its
// chunk in index order.

#include "hq_device.h"

#include <algorithm>
#include <vector>

#include "gate.h"

namespace hq {

constexpr int SAMPLE_CHUNK = 4096;

template <typename R>
__global__ __launch_bounds__(256) void k_chunk_sums(const cplx<R>* __restrict__ a, uint64_t amps,
                                                     double* __restrict__ sums)
{
    const uint64_t c = blockIdx.x;
    const uint32_t t = threadIdx.x;
    double s = 0;
    for (uint32_t j = 0; j < SAMPLE_CHUNK / 256; j++) {
        const uint64_t i = c * SAMPLE_CHUNK + (uint64_t)j * 256 + t;
        if (i < amps) {
            const double re = (double)a[i].x, im = (double)a[i].y;
            s += re * re + im * im;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double red[4];
    if ((t & 63) == 0) red[t >> 6] = s;
    __syncthreads();
    if (t == 0) sums[c] = (red[0] + red[1]) + (red[2] + red[3]);
}

struct ShotQuery {
    uint64_t chunk;
    double residual;
};

template <typename R>
__global__ __launch_bounds__(256) void k_sample(const cplx<R>* __restrict__ a, uint64_t amps,
                                                 const ShotQuery* __restrict__ q, uint32_t shots,
                                                 uint64_t* __restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= shots) return;
    const uint64_t base = q[k].chunk * SAMPLE_CHUNK;
    const uint64_t end = base + SAMPLE_CHUNK < amps ? base + SAMPLE_CHUNK : amps;
    const double r = q[k].residual;
    double acc = 0;
    uint64_t last = base;   // last index with nonzero probability (rounding guard)
    uint64_t pick = ~0ull;
    for (uint64_t i = base; i < end; i++) {
        const double re = (double)a[i].x, im = (double)a[i].y;
        const double p = re * re + im * im;
        if (p > 0) last = i;
        acc += p;
        if (acc > r) { pick = i; break; }
    }
    out[k] = pick == ~0ull ? last : pick;
}

}  // namespace hq

using namespace hq;

extern "C" hq_status hq_state_sample(hq_state s, int shots, uint64_t seed, uint64_t* out)
{
    HQ_CHECK(s && (out || shots == 0), HQ_ERR_ARG, "hq_state_sample: null argument");
    HQ_CHECK(shots >= 0 && shots <= (1 << 24), HQ_ERR_ARG, "hq_state_sample: shots must be in [0, 2^24]");
    if (shots == 0) return HQ_OK;
    const uint64_t chunks = (s->amps + SAMPLE_CHUNK - 1) / SAMPLE_CHUNK;
    double* d_sums = nullptr;
    ShotQuery* d_q = nullptr;
    uint64_t* d_out = nullptr;
    auto cleanup = [&]() {
        if (d_sums) (void)hipFree(d_sums);
        if (d_q) (void)hipFree(d_q);
        if (d_out) (void)hipFree(d_out);
    };
    auto fail = [&](hipError_t e) {
        cleanup();
        set_error(std::string("hq_state_sample: ") + hipGetErrorString(e));
        return HQ_ERR_HIP;
    };
    hipError_t e = hipMalloc((void**)&d_sums, chunks * sizeof(double));
    if (e != hipSuccess) return fail(e);
    if (s->precision == HQ_C128)
        hipLaunchKernelGGL(k_chunk_sums<double>, dim3((unsigned)chunks), dim3(256), 0, s->stream,
                           (const cplx<double>*)s->data, s->amps, d_sums);
    else
        hipLaunchKernelGGL(k_chunk_sums<float>, dim3((unsigned)chunks), dim3(256), 0, s->stream,
                           (const cplx<float>*)s->data, s->amps, d_sums);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    std::vector<double> prefix(chunks);
    if ((e = hipMemcpyAsync(prefix.data(), d_sums, chunks * sizeof(double), hipMemcpyDeviceToHost, s->stream)) !=
        hipSuccess)
        return fail(e);
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return fail(e);
    for (uint64_t c = 1; c < chunks; c++) prefix[c] += prefix[c - 1];
    const double total = prefix.back();
    if (!(total > 0)) {
        cleanup();
        set_error("hq_state_sample: the state has zero norm");
        return HQ_ERR_STATE;
    }
    std::vector<ShotQuery> q(shots);
    Rng rng(seed);
    for (int k = 0; k < shots; k++) {
        const double t = rng.uniform() * total;
        uint64_t c = (uint64_t)(std::upper_bound(prefix.begin(), prefix.end(), t) - prefix.begin());
        if (c >= chunks) c = chunks - 1;
        q[k].chunk = c;
        q[k].residual = t - (c ? prefix[c - 1] : 0.0);
    }
    if ((e = hipMalloc((void**)&d_q, shots * sizeof(ShotQuery))) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&d_out, shots * sizeof(uint64_t))) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(d_q, q.data(), shots * sizeof(ShotQuery), hipMemcpyHostToDevice, s->stream)) !=
        hipSuccess)
        return fail(e);
    const unsigned blocks = (unsigned)((shots + 255) / 256);
    if (s->precision == HQ_C128)
        hipLaunchKernelGGL(k_sample<double>, dim3(blocks), dim3(256), 0, s->stream, (const cplx<double>*)s->data,
                           s->amps, d_q, (uint32_t)shots, d_out);
    else
        hipLaunchKernelGGL(k_sample<float>, dim3(blocks), dim3(256), 0, s->stream, (const cplx<float>*)s->data,
                           s->amps, d_q, (uint32_t)shots, d_out);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(out, d_out, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream)) != hipSuccess)
        return fail(e);
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return fail(e);
    cleanup();
    return HQ_OK;
}
