// hbm_stream.hip -- streaming ceilings of the HBM for the OShareMem access
// pattern: fills, reads, out-of-place copies, and in-place read + write of a
// 2^n x 16 B state through 2^12-element tiles whose lowest L bits are
// contiguous and whose other 12 - L bits are the highest address bits, with
// and without non-temporal accesses, one tile per block or persistent blocks.
//
//     hyquas_amd/hbm_stream [n]      (n = 30: a 16 GiB state)
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e_), __LINE__); exit(1); } } while (0)

typedef double E __attribute__((ext_vector_type(2)));
// tile element e (12 bits) and block b -> address: low L bits of e at [0,L),
// block bits at [L, N-(12-L)), remaining 12-L bits of e at the top.
__device__ __forceinline__ uint64_t addr(uint32_t e, uint64_t b, int L, int N) {
    uint64_t lo = e & ((1u << L) - 1);
    uint64_t hi = e >> L;
    return lo | (b << L) | (hi << (N - (12 - L)));
}

template <int PER, bool NTM>
__global__ __launch_bounds__(256) void k_tile(E* __restrict__ a, double s, int L, int N) {
    extern __shared__ char dyn[];
    if (s == 12345.0) dyn[threadIdx.x] = 0;  // keep the LDS allocation alive
    uint64_t b = blockIdx.x;
    E v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        uint64_t i = addr(j * 256 + threadIdx.x, b, L, N);
        if (NTM) v[j] = __builtin_nontemporal_load(&a[i]); else v[j] = a[i];
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) { v[j].x *= s; v[j].y *= s; }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        uint64_t i = addr(j * 256 + threadIdx.x, b, L, N);
        if (NTM) __builtin_nontemporal_store(v[j], &a[i]); else a[i] = v[j];
    }
}

// persistent: each block walks tiles blockIdx.x, +gridDim.x, ...; loads of the
// next tile are issued before the stores of the current one.
template <int PER, bool NTM>
__global__ __launch_bounds__(256) void k_tile_pers(E* __restrict__ a, double s, int L, int N, uint64_t ntiles) {
    uint64_t b = blockIdx.x;
    if (b >= ntiles) return;
    E v[PER], w[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        uint64_t i = addr(j * 256 + threadIdx.x, b, L, N);
        v[j] = NTM ? __builtin_nontemporal_load(&a[i]) : a[i];
    }
    for (;;) {
        uint64_t nb = b + gridDim.x;
        bool more = nb < ntiles;
        if (more) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                uint64_t i = addr(j * 256 + threadIdx.x, nb, L, N);
                w[j] = NTM ? __builtin_nontemporal_load(&a[i]) : a[i];
            }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            v[j].x *= s; v[j].y *= s;
            uint64_t i = addr(j * 256 + threadIdx.x, b, L, N);
            if (NTM) __builtin_nontemporal_store(v[j], &a[i]); else a[i] = v[j];
        }
        if (!more) break;
#pragma unroll
        for (int j = 0; j < PER; ++j) v[j] = w[j];
        b = nb;
    }
}

__global__ void k_fill(E* a, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) a[i] = E{i == 0 ? 1.0 : 0.0, 0.0};
}
__global__ void k_fill_nt(E* a, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) __builtin_nontemporal_store(E{i == 0 ? 1.0 : 0.0, 0.0}, &a[i]);
}
// read-only: sum of x into one double per block (to defeat elision)
__global__ __launch_bounds__(256) void k_read(const E* __restrict__ a, uint64_t n, double* out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    double acc = 0;
    for (; i < n; i += st) { E v = a[i]; acc += v.x + v.y; }
    __shared__ double red[256];
    red[threadIdx.x] = acc; __syncthreads();
    for (int o = 128; o; o >>= 1) { if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o]; __syncthreads(); }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
__global__ void k_copy(const E* __restrict__ a, E* __restrict__ b, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) b[i] = a[i];
}

int main(int argc, char** argv) {
    const int N = argc > 1 ? atoi(argv[1]) : 30;
    const uint64_t n = 1ull << N;
    E* a; CK(hipMalloc(&a, n * sizeof(E)));
    double* out; CK(hipMalloc(&out, 65536 * sizeof(double)));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    const double bytes2 = 2.0 * n * sizeof(E), bytes1 = 1.0 * n * sizeof(E);
    auto timed = [&](const char* name, double bytes, auto launch) {
        launch(); CK(hipDeviceSynchronize());
        float best = 1e30f;
        for (int r = 0; r < 5; ++r) {
            CK(hipEventRecord(e0)); launch(); CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
            float ms; CK(hipEventElapsedTime(&ms, e0, e1)); if (ms < best) best = ms;
        }
        printf("%-44s %8.3f ms  %7.0f GB/s\n", name, best, bytes / (best * 1e-3) / 1e9);
        fflush(stdout);
    };
    const uint64_t ntiles = n >> 12;
    char nm[128];
    for (int g : {1024, 4096, 16384, 65536})
    { snprintf(nm, sizeof nm, "fill grid=%d", g); timed(nm, bytes1, [&] { k_fill<<<g, 256>>>(a, n); }); }
    for (int g : {4096, 16384, 65536})
    { snprintf(nm, sizeof nm, "fill_nt grid=%d", g); timed(nm, bytes1, [&] { k_fill_nt<<<g, 256>>>(a, n); }); }
    for (int g : {4096, 16384, 65536})
    { snprintf(nm, sizeof nm, "read grid=%d", g); timed(nm, bytes1, [&] { k_read<<<g, 256>>>(a, n, out); }); }
    {
        E* b; CK(hipMalloc(&b, (n / 2) * sizeof(E)));
        for (int g : {4096, 16384, 65536})
        { snprintf(nm, sizeof nm, "copy half->other grid=%d", g); timed(nm, bytes1, [&] { k_copy<<<g, 256>>>(a, b, n / 2); }); }
        CK(hipFree(b));
    }
    for (int L : {12, 8, 6, 4, 3}) {
        for (int lds : {0, 65536}) {
            snprintf(nm, sizeof nm, "tile16 L=%d lds=%d", L, lds);
            timed(nm, bytes2, [&] { k_tile<16, false><<<ntiles, 256, lds>>>(a, 1.0, L, N); });
            snprintf(nm, sizeof nm, "tile16 nt L=%d lds=%d", L, lds);
            timed(nm, bytes2, [&] { k_tile<16, true><<<ntiles, 256, lds>>>(a, 1.0, L, N); });
        }
        for (int g : {512, 1024, 2048}) {
            snprintf(nm, sizeof nm, "pers16 L=%d grid=%d", L, g);
            timed(nm, bytes2, [&] { k_tile_pers<16, false><<<g, 256>>>(a, 1.0, L, N, ntiles); });
            snprintf(nm, sizeof nm, "pers16 nt L=%d grid=%d", L, g);
            timed(nm, bytes2, [&] { k_tile_pers<16, true><<<g, 256>>>(a, 1.0, L, N, ntiles); });
        }
    }
    CK(hipFree(a));
    return 0;
}
