// state.hip -- state vectors, readout reductions and profiling of the device layer.
//
// Replaces kernelInit / kernelDestroy / kernelGetAmp / kernelDeviceToHost /
// kernelMeasure (src/kernelSimple.cu:9-37, 448-529).  Allocation is 64-bit
// throughout (the reference's int nVec overflows at n >= 33).

#include "hq_device.h"

#include <cmath>
#include <cstring>
#include <mutex>

namespace hq {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
const char* get_error() { return g_error.c_str(); }

// ---- kernels -------------------------------------------------------------------------

// |0...0>: a fill kernel rather than hipMemsetAsync so the reset can be
// captured in a HIP graph.
template <typename R>
__global__ void k_reset(cplx<R>* __restrict__ a, uint64_t n_amps)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n_amps; i += stride) {
        cplx<R> v;
        v.x = (i == 0) ? R(1) : R(0);
        v.y = R(0);
        a[i] = v;
    }
}

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
constexpr int RED_MAXQ = 36;

__device__ inline double block_sum(double v, double* sh)
{
    // wave64 shuffle reduction, then across the 4 waves through LDS
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
    return r;
}

// Kahan-compensated FP64 accumulation of |a_i|^2 over a subset (bit `skip`
// clear when skip >= 0), one partial per block (measure<>/reduce<>,
// kernelSimple.cu:448-495).
template <typename R>
__global__ void k_norm(const cplx<R>* __restrict__ a, uint64_t n_amps, int target, double* __restrict__ out)
{
    __shared__ double sh[16];
    double s = 0, c = 0;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_amps; i += stride) {
        if (target >= 0 && ((i >> target) & 1)) continue;
        cplx<R> v = a[i];
        double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        double y = p - c;
        double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    double r = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// Sum of p_i, p_i * bit_q(i) and p_i * (bit_q ^ bit_{q+1}) per block.
template <typename R>
__global__ void k_expect_z(const cplx<R>* __restrict__ a, int n, double* __restrict__ out)
{
    __shared__ double sh[16];
    double tot = 0, s1[RED_MAXQ], s2[RED_MAXQ];
#pragma unroll
    for (int q = 0; q < RED_MAXQ; q++) { s1[q] = 0; s2[q] = 0; }
    uint64_t n_amps = (uint64_t)1 << n;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_amps; i += stride) {
        cplx<R> v = a[i];
        double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        tot += p;
#pragma unroll
        for (int q = 0; q < RED_MAXQ; q++) {
            if (q < n) {
                uint64_t b = (i >> q) & 1, b2 = (i >> (q + 1)) & 1;
                s1[q] += b ? p : 0.0;
                s2[q] += (b ^ b2) ? p : 0.0;
            }
        }
    }
    int stride_out = 1 + 2 * RED_MAXQ;
    double r = block_sum(tot, sh);
    if (threadIdx.x == 0) out[blockIdx.x * stride_out] = r;
#pragma unroll
    for (int q = 0; q < RED_MAXQ; q++) {
        if (q < n) {
            double r1 = block_sum(s1[q], sh);
            if (threadIdx.x == 0) out[blockIdx.x * stride_out + 1 + q] = r1;
            double r2 = block_sum(s2[q], sh);
            if (threadIdx.x == 0) out[blockIdx.x * stride_out + 1 + RED_MAXQ + q] = r2;
        } else if (threadIdx.x == 0) {
            // columns of absent qubits hold zeros: every column is defined
            out[blockIdx.x * stride_out + 1 + q] = 0.0;
            out[blockIdx.x * stride_out + 1 + RED_MAXQ + q] = 0.0;
        }
    }
}

template <typename RA, typename RB>
__global__ void k_inner(const cplx<RA>* __restrict__ a, const cplx<RB>* __restrict__ b, uint64_t n_amps,
                        double* __restrict__ out)
{
    __shared__ double sh[16];
    double sr = 0, si = 0;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_amps; i += stride) {
        cplx<RA> x = a[i];
        cplx<RB> y = b[i];
        double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        sr += xr * yr + xi * yi;   // conj(x) * y
        si += xr * yi - xi * yr;
    }
    double r1 = block_sum(sr, sh);
    if (threadIdx.x == 0) out[2 * blockIdx.x] = r1;
    double r2 = block_sum(si, sh);
    if (threadIdx.x == 0) out[2 * blockIdx.x + 1] = r2;
}

// Fixed-order Kahan sum of column c of `count` partial rows (row stride
// `stride`) into fin[c], one thread per column: the same operations in the
// same order as the host-side sum of the partials (sum_partials), so both
// give the same FP64 result bit for bit.
__global__ void k_sum_columns(const double* __restrict__ part, int count, int stride, int cols,
                              double* __restrict__ fin)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double sum = 0, comp = 0;
    for (int b = 0; b < count; b++) {
        double y = part[(size_t)b * stride + c] - comp;
        double t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    fin[c] = sum;
}

// ---- allocation-free readout -----------------------------------------------------------
// Scratch of hq_expect_z_noalloc: the per-block partials of k_expect_z (at
// most RED_Z_BLOCKS rows of 1 + 2 * RED_MAXQ doubles) followed by one row of
// column sums.  One copy per device, held until the process ends; host code
// addresses it only through the pointer hipGetSymbolAddress returns for the
// device (noalloc_scratch below).
constexpr int RED_Z_BLOCKS = 256;
constexpr int RED_Z_STRIDE = 1 + 2 * RED_MAXQ;
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)(RED_Z_BLOCKS + 1) * RED_Z_STRIDE;   // 147 KiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device`, looked up once per device with that
// device current (each device holds its own copy at its own address).
static hq_status noalloc_scratch(int device, double** out)
{
    HQ_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, HQ_ERR_UNSUPPORTED,
             "hq_expect_z_noalloc: device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        HQ_HIP(hipGetDevice(&cur));
        if (cur != device) HQ_HIP(hipSetDevice(device));
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr) {
            set_error(std::string("hq_expect_z_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return HQ_ERR_HIP;
        }
        g_scratch_ptr[device] = (double*)ptr;
    }
    *out = g_scratch_ptr[device];
    return HQ_OK;
}

static hq_status ensure_red(hq_state s, size_t doubles)
{
    if (s->red_cap >= doubles) return HQ_OK;
    if (s->d_red) { (void)hipFree(s->d_red); s->d_red = nullptr; }
    if (s->h_red) { (void)hipHostFree(s->h_red); s->h_red = nullptr; }
    s->red_cap = 0;
    HQ_HIP(hipMalloc(&s->d_red, doubles * sizeof(double)));
    HQ_HIP(hipHostMalloc(&s->h_red, doubles * sizeof(double), hipHostMallocDefault));
    s->red_cap = doubles;
    return HQ_OK;
}

static unsigned grid_for(uint64_t work, unsigned threads, unsigned cap)
{
    uint64_t g = (work + threads - 1) / threads;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

hq_status prof_begin(hq_state s, int kind, int* slot)
{
    *slot = -1;
    if (!s->profiling) return HQ_OK;
    ProfileEntry e;
    e.kind = kind;
    for (hipEvent_t* ev : {&e.start, &e.stop}) {
        if (!s->free_events.empty()) { *ev = s->free_events.back(); s->free_events.pop_back(); }
        else HQ_HIP(hipEventCreate(ev));
    }
    HQ_HIP(hipEventRecord(e.start, s->stream));
    s->prof.push_back(e);
    *slot = (int)s->prof.size() - 1;
    return HQ_OK;
}

hq_status prof_end(hq_state s, int slot)
{
    if (slot < 0) return HQ_OK;
    HQ_HIP(hipEventRecord(s->prof[slot].stop, s->stream));
    return HQ_OK;
}

}  // namespace hq

using namespace hq;

extern "C" {

const char* hq_last_error(void) { return hq::get_error(); }

hq_status hq_device_count(int* count)
{
    HQ_CHECK(count, HQ_ERR_ARG, "hq_device_count: null pointer");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess || c == 0) {
        *count = 0;
        set_error(std::string("no HIP device visible: ") + hipGetErrorString(e));
        return HQ_ERR_NO_DEVICE;
    }
    *count = c;
    return HQ_OK;
}

hq_status hq_state_create(int num_qubits, int precision, int with_buffer, hq_state* out)
{
    HQ_CHECK(out, HQ_ERR_ARG, "hq_state_create: null output");
    *out = nullptr;
    HQ_CHECK(num_qubits >= 1 && num_qubits <= 36, HQ_ERR_ARG, "hq_state_create: num_qubits must be in [1, 36]");
    HQ_CHECK(precision == HQ_C128 || precision == HQ_C64, HQ_ERR_ARG, "hq_state_create: bad precision");
    int ndev = 0;
    hq_status st = hq_device_count(&ndev);
    if (st != HQ_OK) return st;
    hq_state s = new hq_state_s();
    s->n = num_qubits;
    s->precision = precision;
    s->amps = (uint64_t)1 << num_qubits;
    s->amp_bytes = precision == HQ_C128 ? 16 : 8;
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    size_t bytes = s->amps * s->amp_bytes;
    hipError_t e = hipMalloc(&s->data, bytes);
    if (e != hipSuccess) {
        set_error("hq_state_create: hipMalloc of " + std::to_string(bytes) + " B failed: " + hipGetErrorString(e));
        delete s;
        return HQ_ERR_OOM;
    }
    if (with_buffer) {
        e = hipMalloc(&s->buf, bytes);
        if (e != hipSuccess) {
            set_error("hq_state_create: hipMalloc of the work buffer failed: " + std::string(hipGetErrorString(e)));
            (void)hipFree(s->data);
            delete s;
            return HQ_ERR_OOM;
        }
    }
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        (void)hipFree(s->data);
        if (s->buf) (void)hipFree(s->buf);
        delete s;
        return HQ_ERR_HIP;
    }
    *out = s;
    return hq_state_reset(s);
}

hq_status hq_state_destroy(hq_state s)
{
    if (!s) return HQ_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->data) (void)hipFree(s->data);
    if (s->buf) (void)hipFree(s->buf);
    if (s->d_red) (void)hipFree(s->d_red);
    if (s->h_red) (void)hipHostFree(s->h_red);
    for (auto& e : s->prof) { (void)hipEventDestroy(e.start); (void)hipEventDestroy(e.stop); }
    for (auto& e : s->free_events) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return HQ_OK;
}

hq_status hq_state_reset(hq_state s)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_state_reset: null state");
    unsigned g = grid_for(s->amps, 256, 65536);
    if (s->precision == HQ_C128)
        hipLaunchKernelGGL(k_reset<double>, dim3(g), dim3(256), 0, s->stream, (cplx<double>*)s->data, s->amps);
    else
        hipLaunchKernelGGL(k_reset<float>, dim3(g), dim3(256), 0, s->stream, (cplx<float>*)s->data, s->amps);
    HQ_HIP(hipGetLastError());
    return HQ_OK;
}

hq_status hq_state_info(hq_state s, int* num_qubits, int* precision)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_state_info: null state");
    if (num_qubits) *num_qubits = s->n;
    if (precision) *precision = s->precision;
    return HQ_OK;
}

hq_status hq_state_has_buffer(hq_state s, int* has_buffer)
{
    HQ_CHECK(s && has_buffer, HQ_ERR_ARG, "hq_state_has_buffer: null argument");
    *has_buffer = s->buf ? 1 : 0;
    return HQ_OK;
}

hq_status hq_state_sync(hq_state s)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_state_sync: null state");
    HQ_HIP(hipStreamSynchronize(s->stream));
    return HQ_OK;
}

hq_status hq_state_download(hq_state s, void* host, uint64_t count)
{
    HQ_CHECK(s && host, HQ_ERR_ARG, "hq_state_download: null argument");
    HQ_CHECK(count == s->amps, HQ_ERR_ARG, "hq_state_download: count must be 2^n");
    HQ_HIP(hipMemcpyAsync(host, s->data, s->amps * s->amp_bytes, hipMemcpyDeviceToHost, s->stream));
    HQ_HIP(hipStreamSynchronize(s->stream));
    return HQ_OK;
}

hq_status hq_state_upload(hq_state s, const void* host, uint64_t count)
{
    HQ_CHECK(s && host, HQ_ERR_ARG, "hq_state_upload: null argument");
    HQ_CHECK(count == s->amps, HQ_ERR_ARG, "hq_state_upload: count must be 2^n");
    HQ_HIP(hipMemcpyAsync(s->data, host, s->amps * s->amp_bytes, hipMemcpyHostToDevice, s->stream));
    HQ_HIP(hipStreamSynchronize(s->stream));
    return HQ_OK;
}

hq_status hq_state_get_amp(hq_state s, uint64_t index, double* re, double* im)
{
    HQ_CHECK(s && re && im, HQ_ERR_ARG, "hq_state_get_amp: null argument");
    HQ_CHECK(index < s->amps, HQ_ERR_ARG, "hq_state_get_amp: index out of range");
    hq_status st = ensure_red(s, 2);
    if (st != HQ_OK) return st;
    HQ_HIP(hipMemcpyAsync(s->h_red, (char*)s->data + index * s->amp_bytes, s->amp_bytes,
                          hipMemcpyDeviceToHost, s->stream));
    HQ_HIP(hipStreamSynchronize(s->stream));
    if (s->precision == HQ_C128) { *re = s->h_red[0]; *im = s->h_red[1]; }
    else { float* f = (float*)s->h_red; *re = f[0]; *im = f[1]; }
    return HQ_OK;
}

static hq_status sum_partials(hq_state s, int count, int stride, int offset, double* out)
{
    // Kahan sum of the per-block partials on the host, in FP64.
    double sum = 0, c = 0;
    for (int b = 0; b < count; b++) {
        double y = s->h_red[b * stride + offset] - c;
        double t = sum + y;
        c = (t - sum) - y;
        sum = t;
    }
    *out = sum;
    return HQ_OK;
}

static hq_status run_norm(hq_state s, int target, double* out)
{
    hq_status st = ensure_red(s, RED_BLOCKS);
    if (st != HQ_OK) return st;
    unsigned g = grid_for(s->amps, RED_THREADS, RED_BLOCKS);
    if (s->precision == HQ_C128)
        hipLaunchKernelGGL(k_norm<double>, dim3(g), dim3(RED_THREADS), 0, s->stream,
                           (const cplx<double>*)s->data, s->amps, target, s->d_red);
    else
        hipLaunchKernelGGL(k_norm<float>, dim3(g), dim3(RED_THREADS), 0, s->stream,
                           (const cplx<float>*)s->data, s->amps, target, s->d_red);
    HQ_HIP(hipGetLastError());
    HQ_HIP(hipMemcpyAsync(s->h_red, s->d_red, g * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HQ_HIP(hipStreamSynchronize(s->stream));
    return sum_partials(s, (int)g, 1, 0, out);
}

hq_status hq_state_norm(hq_state s, double* norm2)
{
    HQ_CHECK(s && norm2, HQ_ERR_ARG, "hq_state_norm: null argument");
    return run_norm(s, -1, norm2);
}

hq_status hq_measure_zero(hq_state s, int target, double* p0)
{
    HQ_CHECK(s && p0, HQ_ERR_ARG, "hq_measure_zero: null argument");
    HQ_CHECK(target >= 0 && target < s->n, HQ_ERR_ARG, "hq_measure_zero: target out of range");
    return run_norm(s, target, p0);
}

static void launch_expect_z(hq_state s, unsigned g, double* partials)
{
    if (s->precision == HQ_C128)
        hipLaunchKernelGGL(k_expect_z<double>, dim3(g), dim3(RED_THREADS), 0, s->stream,
                           (const cplx<double>*)s->data, s->n, partials);
    else
        hipLaunchKernelGGL(k_expect_z<float>, dim3(g), dim3(RED_THREADS), 0, s->stream,
                           (const cplx<float>*)s->data, s->n, partials);
}

// z[q] = <Z_q>, zz[q] = <Z_q Z_{q+1}> from the column sums of k_expect_z
// (tot, sum p bit_q, sum p (bit_q xor bit_{q+1})).
static void finish_expect_z(int n, const double* col, double* z, double* zz)
{
    const double tot = col[0];
    for (int q = 0; q < n; q++) {
        if (z) z[q] = tot - 2 * col[1 + q];
        if (zz && q + 1 < n) zz[q] = tot - 2 * col[1 + RED_MAXQ + q];
    }
}

hq_status hq_expect_z(hq_state s, double* z, double* zz)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_expect_z: null state");
    HQ_CHECK(s->n <= RED_MAXQ, HQ_ERR_UNSUPPORTED, "hq_expect_z: at most 36 qubits");
    const int stride = RED_Z_STRIDE;
    unsigned g = grid_for(s->amps, RED_THREADS, RED_Z_BLOCKS);
    hq_status st = ensure_red(s, (size_t)g * stride);
    if (st != HQ_OK) return st;
    launch_expect_z(s, g, s->d_red);
    HQ_HIP(hipGetLastError());
    HQ_HIP(hipMemcpyAsync(s->h_red, s->d_red, (size_t)g * stride * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
    HQ_HIP(hipStreamSynchronize(s->stream));
    double col[RED_Z_STRIDE];
    for (int c = 0; c < stride; c++) sum_partials(s, (int)g, stride, c, &col[c]);
    finish_expect_z(s->n, col, z, zz);
    return HQ_OK;
}

hq_status hq_expect_z_noalloc(hq_state s, double* z, double* zz)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_expect_z_noalloc: null state");
    HQ_CHECK(s->n <= RED_MAXQ, HQ_ERR_UNSUPPORTED, "hq_expect_z_noalloc: at most 36 qubits");
    const int stride = RED_Z_STRIDE;
    unsigned g = grid_for(s->amps, RED_THREADS, RED_Z_BLOCKS);
    HQ_CHECK((size_t)(g + 1) * stride <= NOALLOC_SCRATCH_DOUBLES, HQ_ERR_ARG,
             "hq_expect_z_noalloc: partials exceed the " + std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                 "-double scratch");
    double* ptr = nullptr;
    hq_status st = noalloc_scratch(s->device, &ptr);
    if (st != HQ_OK) return st;
    // every partial row and the column-sum row are written before they are
    // read: nothing in the scratch has to start at zero
    launch_expect_z(s, g, ptr);
    HQ_HIP(hipGetLastError());
    double* fin = ptr + (size_t)g * stride;
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(128), 0, s->stream, (const double*)ptr, (int)g, stride,
                       stride, fin);
    HQ_HIP(hipGetLastError());
    double col[RED_Z_STRIDE];
    HQ_HIP(hipMemcpyAsync(col, fin, stride * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HQ_HIP(hipStreamSynchronize(s->stream));
    finish_expect_z(s->n, col, z, zz);
    return HQ_OK;
}

hq_status hq_state_inner(hq_state a, hq_state b, double* re, double* im)
{
    HQ_CHECK(a && b && re && im, HQ_ERR_ARG, "hq_state_inner: null argument");
    HQ_CHECK(a->n == b->n, HQ_ERR_ARG, "hq_state_inner: qubit counts differ");
    HQ_HIP(hipStreamSynchronize(b->stream));
    unsigned g = grid_for(a->amps, RED_THREADS, RED_BLOCKS);
    hq_status st = ensure_red(a, 2 * (size_t)g);
    if (st != HQ_OK) return st;
    if (a->precision == HQ_C128 && b->precision == HQ_C128)
        hipLaunchKernelGGL((k_inner<double, double>), dim3(g), dim3(RED_THREADS), 0, a->stream,
                           (const cplx<double>*)a->data, (const cplx<double>*)b->data, a->amps, a->d_red);
    else if (a->precision == HQ_C128)
        hipLaunchKernelGGL((k_inner<double, float>), dim3(g), dim3(RED_THREADS), 0, a->stream,
                           (const cplx<double>*)a->data, (const cplx<float>*)b->data, a->amps, a->d_red);
    else if (b->precision == HQ_C128)
        hipLaunchKernelGGL((k_inner<float, double>), dim3(g), dim3(RED_THREADS), 0, a->stream,
                           (const cplx<float>*)a->data, (const cplx<double>*)b->data, a->amps, a->d_red);
    else
        hipLaunchKernelGGL((k_inner<float, float>), dim3(g), dim3(RED_THREADS), 0, a->stream,
                           (const cplx<float>*)a->data, (const cplx<float>*)b->data, a->amps, a->d_red);
    HQ_HIP(hipGetLastError());
    HQ_HIP(hipMemcpyAsync(a->h_red, a->d_red, 2 * (size_t)g * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HQ_HIP(hipStreamSynchronize(a->stream));
    sum_partials(a, (int)g, 2, 0, re);
    sum_partials(a, (int)g, 2, 1, im);
    return HQ_OK;
}

hq_status hq_profile_enable(hq_state s, int enable)
{
    HQ_CHECK(s, HQ_ERR_ARG, "hq_profile_enable: null state");
    s->profiling = enable != 0;
    return HQ_OK;
}

hq_status hq_profile_read(hq_state s, int* kinds, double* ms, int cap, int* count)
{
    HQ_CHECK(s && count, HQ_ERR_ARG, "hq_profile_read: null argument");
    HQ_HIP(hipStreamSynchronize(s->stream));
    int k = 0;
    for (auto& e : s->prof) {
        float t = 0;
        HQ_HIP(hipEventElapsedTime(&t, e.start, e.stop));
        if (k < cap) {
            if (kinds) kinds[k] = e.kind;
            if (ms) ms[k] = t;
        }
        k++;
        s->free_events.push_back(e.start);
        s->free_events.push_back(e.stop);
    }
    s->prof.clear();
    *count = k;
    return HQ_OK;
}

}  // extern "C"
