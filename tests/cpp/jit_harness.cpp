// CPU execution of a generated OShareMem group kernel (test infrastructure):
// one std::thread per GPU thread of a block, a std::barrier for
// __syncthreads, blocks run one after another, the block's LDS is a
// function-static array.  Build with -DGROUP_SRC=\"/abs/path/kernel.hip\".
// usage: harness n tile threads psi_in coef_in psi_out
#include <barrier>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
struct Dim { unsigned x; };
thread_local Dim threadIdx, blockIdx;
std::barrier<>* g_bar;
#define __syncthreads() g_bar->arrive_and_wait()
#define __global__
#define __launch_bounds__(n)
#define __shared__ static
#define __restrict__
#define __device__
#define __forceinline__ inline
#define amdgpu_waves_per_eu(n)
#include GROUP_SRC
int main(int argc, char** argv) {
  int n = atoi(argv[1]), k = atoi(argv[2]), nt = atoi(argv[3]);
  size_t N = 1ull << n;
  std::vector<C> psi(N); FILE* f = fopen(argv[4], "rb"); fread(psi.data(), sizeof(C), N, f); fclose(f);
  f = fopen(argv[5], "rb"); std::vector<double> cd; double v; while (fread(&v, 8, 1, f) == 1) cd.push_back(v); fclose(f);
  std::vector<R> cf(cd.begin(), cd.end()); if (cf.empty()) cf.push_back(0);
  for (unsigned long long b = 0; b < (1ull << (n - k)); b++) {
    std::barrier<> bar(nt); g_bar = &bar;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back([&, t, b]() { threadIdx.x = t; blockIdx.x = (unsigned)b; hq_group(psi.data(), cf.data()); });
    for (auto& x : th) x.join();
  }
  f = fopen(argv[6], "wb"); fwrite(psi.data(), sizeof(C), N, f); fclose(f);
  return 0;
}
