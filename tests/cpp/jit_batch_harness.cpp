// CPU execution of the batched form of a generated OShareMem group kernel
// (test infrastructure): several circuits of one structure share the kernel,
// each with its own state, coefficient array and partials; the blocks of all
// circuits run one after another, each with one std::thread per GPU thread
// and a std::barrier for __syncthreads.  Build with
// -DGROUP_SRC=\"/abs/path/kernel.hip\".
// usage: harness n tile threads circuits mode states_in coefs_in states_out parts_out
#include <barrier>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
struct Dim { unsigned x; };
thread_local Dim threadIdx, blockIdx;
std::barrier<>* g_bar;
#define HQ_G
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
  int n = atoi(argv[1]), k = atoi(argv[2]), nt = atoi(argv[3]), nc = atoi(argv[4]);
  unsigned mode = (unsigned)atoi(argv[5]);
  size_t N = 1ull << n, tiles = 1ull << (n - k);
  std::vector<C> psi(N * nc);
  FILE* f = fopen(argv[6], "rb"); if (fread(psi.data(), sizeof(C), N * nc, f) != N * nc) return 2; fclose(f);
  f = fopen(argv[7], "rb"); std::vector<double> cd; double v; while (fread(&v, 8, 1, f) == 1) cd.push_back(v); fclose(f);
  std::vector<R> cf(cd.begin(), cd.end()); if (cf.empty()) cf.push_back(0);
  const size_t per = cf.size() / nc;
  std::vector<double> part(3 * tiles * nc, -1.0);
  std::vector<BItem> items(nc);
  for (int c = 0; c < nc; c++) {
    items[c].psi = (unsigned long long)(psi.data() + c * N);
    items[c].cf = (unsigned long long)(cf.data() + c * per);
    items[c].part = (unsigned long long)(part.data() + 3 * tiles * c);
    items[c].first_block = c * tiles;
    items[c].n = n;
    items[c].mode = mode;
  }
  for (unsigned long long b = 0; b < tiles * nc; b++) {
    std::barrier<> bar(nt); g_bar = &bar;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back([&, t, b]() { threadIdx.x = t; blockIdx.x = (unsigned)b; hq_group_batch(items.data(), (unsigned)nc); });
    for (auto& x : th) x.join();
  }
  f = fopen(argv[8], "wb"); fwrite(psi.data(), sizeof(C), N * nc, f); fclose(f);
  f = fopen(argv[9], "wb"); fwrite(part.data(), 8, part.size(), f); fclose(f);
  return 0;
}
