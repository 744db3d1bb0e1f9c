// engine.hip -- many circuits on one GPU: the batching and serving forms.
//
// Circuits are compiled on host threads (hybrid partitioner in OShareMem
// mode, src/compiler.cpp:220-452; a circuit of at most `whole_max` qubits is
// one group whose tile is the whole state), admitted in submission order
// while their state vectors fit an HBM budget (a buddy arena over large
// chunks), and advanced in ROUNDS: one round launches the next group of every
// admitted circuit, all circuits of one register/thread shape in ONE kernel
// whose blocks map to (circuit, tile).  A circuit's first group builds
// |0...0> in registers (no reset pass), its last group folds the Z / ZZ
// readout into the pass and skips the store unless the state is needed for
// top-k.  New circuits join at any round (continuous batching); a fixed
// batch can instead be captured once as a hipGraph and replayed.
//
// Readouts (src/kernelSimple.cu:448-516 computes only P(q = 0); the rest is
// the step after the reference's path): sum_q <Z_q>, sum_q <Z_q Z_{q+1}>,
// the norm, and the 16 most probable basis states.  Every reduction has a
// fixed order, so a circuit's results do not depend on what else runs.

#include "hq_device.h"
#include "sharemem_dev.h"

#include "circuit.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>

namespace hq {

// ---- device side ---------------------------------------------------------------------------

struct BatchItem {
    uint64_t psi;          // state of the circuit (cplx<R>*)
    uint64_t prog;         // group program (ShmHeader blob)
    uint64_t part;         // per-tile partial readouts, 3 doubles per tile (TILE_READ)
    uint64_t first_block;  // first grid block of this item
    uint32_t n;
    uint32_t mode;         // TILE_* bits
};
static_assert(sizeof(BatchItem) == 40, "BatchItem layout");

struct FinishItem {
    uint64_t part, tiles, out;  // out: 3 doubles {norm, sum Z, sum ZZ}
};

struct Cand {
    double p;
    uint64_t idx;
};
static_assert(sizeof(Cand) == 16, "Cand layout");

struct TopkItem {
    uint64_t src;          // amplitudes (kind 0) or candidates (kind 1)
    uint64_t count;        // elements in src
    uint64_t dst;          // 16 candidates per block
    uint64_t first_block;
    uint32_t kind, pad;
};

template <typename T>
__device__ __forceinline__ uint32_t find_item(const T* __restrict__ items, uint32_t count, uint64_t b)
{
    uint32_t lo = 0, hi = count;  // items[lo].first_block <= b
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (items[mid].first_block <= b) lo = mid; else hi = mid;
    }
    return lo;
}

template <typename R, int RB, int TB>
__global__ __launch_bounds__(1 << TB) void k_batch_sharemem(const BatchItem* __restrict__ items, uint32_t count)
{
    __shared__ cplx<R> lds[1 << (RB + TB)];
    const uint64_t b = blockIdx.x;
    const BatchItem* it = items + find_item(items, count, b);
    tile_pass<R, RB, TB>((cplx<R>*)it->psi, (const char*)it->prog, b - it->first_block, lds, it->mode, (int)it->n,
                         (double*)it->part);
}

// per circuit: fixed-order sum of the per-tile partials
__global__ __launch_bounds__(256) void k_batch_finish(const FinishItem* __restrict__ items)
{
    const FinishItem it = items[blockIdx.x];
    const double* __restrict__ p = (const double*)it.part;
    const uint32_t t = threadIdx.x;
    double s0 = 0, s1 = 0, s2 = 0;
    for (uint64_t i = t; i < it.tiles; i += 256) { s0 += p[3 * i]; s1 += p[3 * i + 1]; s2 += p[3 * i + 2]; }
    __shared__ double red[3][256];
    red[0][t] = s0; red[1][t] = s1; red[2][t] = s2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < (uint32_t)o) { red[0][t] += red[0][t + o]; red[1][t] += red[1][t + o]; red[2][t] += red[2][t + o]; }
        __syncthreads();
    }
    if (t == 0) {
        double* out = (double*)it.out;
        out[0] = red[0][0]; out[1] = red[1][0]; out[2] = red[2][0];
    }
}

// (p, idx) order of the top-k: larger probability first, then smaller index
__device__ __forceinline__ bool better(double pa, uint64_t ia, double pb, uint64_t ib)
{
    return pa > pb || (pa == pb && ia < ib);
}

__device__ __forceinline__ void wave_best(double& p, uint64_t& i)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        double op = __shfl_xor(p, o);
        uint64_t oi = __shfl_xor(i, o);
        if (better(op, oi, p, i)) { p = op; i = oi; }
    }
}

constexpr int TOPK = 16;
constexpr int TOPK_CHUNK = 4096;  // elements per block: 256 threads x 16

// 16 best (p, idx) of each 4096-element chunk, in order.  Empty slots hold
// p = -1.  Applied level by level until one chunk remains.
template <typename R>
__global__ __launch_bounds__(256) void k_topk(const TopkItem* __restrict__ items, uint32_t count)
{
    const uint64_t b = blockIdx.x;
    const TopkItem* it = items + find_item(items, count, b);
    const uint64_t chunk = b - it->first_block;
    const uint64_t e0 = chunk * TOPK_CHUNK;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    double p[16];
    uint64_t id[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const uint64_t e = e0 + (uint64_t)j * 256 + t;
        p[j] = -1.0;
        id[j] = ~0ull;
        if (e < it->count) {
            if (it->kind == 0) {
                const cplx<R> a = ((const cplx<R>*)it->src)[e];
                const double re = (double)a.x, im = (double)a.y;
                p[j] = re * re + im * im;
                id[j] = e;
            } else {
                const Cand c = ((const Cand*)it->src)[e];
                p[j] = c.p;
                id[j] = c.idx;
            }
        }
    }
    double kp = -1.0;
    uint64_t ki = ~0ull;
    for (int r = 0; r < TOPK; r++) {
        double bp = p[0];
        uint64_t bi = id[0];
        int bj = 0;
#pragma unroll
        for (int j = 1; j < 16; j++)
            if (better(p[j], id[j], bp, bi)) { bp = p[j]; bi = id[j]; bj = j; }
        double wp = bp;
        uint64_t wi = bi;
        wave_best(wp, wi);
        if (bp == wp && bi == wi) {
#pragma unroll
            for (int j = 0; j < 16; j++) if (j == bj) p[j] = -2.0;
        }
        if (lane == (uint32_t)r) { kp = wp; ki = wi; }
    }
    __shared__ double sp[64];
    __shared__ uint64_t si[64];
    if (lane < TOPK) { sp[w * TOPK + lane] = kp; si[w * TOPK + lane] = ki; }
    __syncthreads();
    if (w == 0) {
        double q = sp[lane];
        uint64_t qi = si[lane];
        double op = -1.0;
        uint64_t oi = ~0ull;
        for (int r = 0; r < TOPK; r++) {
            double wp = q;
            uint64_t wi = qi;
            wave_best(wp, wi);
            if (q == wp && qi == wi) q = -2.0;
            if (lane == (uint32_t)r) { op = wp; oi = wi; }
        }
        if (lane < TOPK) {
            Cand c;
            c.p = op < 0 ? -1.0 : op;
            c.idx = oi;
            ((Cand*)it->dst)[chunk * TOPK + lane] = c;
        }
    }
}

// ---- launch dispatch -----------------------------------------------------------------------

#define HQ_BATCH_SHAPES(X) \
    X(1, 6) X(2, 6) X(3, 6) X(4, 6) X(5, 6) X(2, 7) X(3, 7) X(4, 7) X(5, 7) X(3, 8) X(4, 8) X(5, 8) X(3, 9) \
    X(4, 9)

template <typename R>
static hq_status launch_batch_shape(int rb, int tb, const BatchItem* items, uint32_t count, uint64_t blocks,
                                    hipStream_t st)
{
#define HQ_BCASE(RBV, TBV)                                                                                  \
    if (rb == RBV && tb == TBV) {                                                                            \
        hipLaunchKernelGGL((k_batch_sharemem<R, RBV, TBV>), dim3((unsigned)blocks), dim3(1 << TBV), 0, st, items, \
                           count);                                                                           \
        HQ_HIP(hipGetLastError());                                                                           \
        return HQ_OK;                                                                                        \
    }
    HQ_BATCH_SHAPES(HQ_BCASE)
#undef HQ_BCASE
    set_error("engine: no batched kernel for this register/thread split");
    return HQ_ERR_UNSUPPORTED;
}

// ---- buddy arena over large device chunks --------------------------------------------------

struct Block {
    int chunk = -1;
    int order = 0;
    uint64_t off = 0;
    uint64_t addr = 0;
};

class Arena {
public:
    static constexpr int kMinOrder = 12;   // 4 KiB

    void configure(uint64_t budget, uint64_t chunk_bytes)
    {
        budget_ = budget;
        top_ = kMinOrder;
        while (top_ < 40 && (1ull << (top_ + 1)) <= chunk_bytes) top_++;
    }
    uint64_t chunk_bytes() const { return 1ull << top_; }
    uint64_t budget() const { return budget_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    uint64_t reserved() const { return reserved_; }

    static int order_of(uint64_t bytes)
    {
        int o = kMinOrder;
        while ((1ull << o) < bytes) o++;
        return o;
    }
    // blocks of these power-of-two sizes fit an empty arena: each within a
    // chunk and their sum within the whole chunks the budget allows (buddy
    // blocks of power-of-two sizes pack without loss)
    bool can_ever_fit(uint64_t largest_block, uint64_t total_pow2) const
    {
        const uint64_t chunk = 1ull << top_;
        return order_of(largest_block) <= top_ && total_pow2 <= (budget_ / chunk) * chunk;
    }
    // false: no room within the budget now (or a real device OOM)
    bool alloc(uint64_t bytes, Block* out)
    {
        const int want = order_of(bytes);
        if (want > top_) return false;
        for (int pass = 0; pass < 2; pass++) {
            for (size_t c = 0; c < chunks_.size(); c++) {
                auto& fr = chunks_[c].free;
                for (int o = want; o <= top_; o++) {
                    if (fr[o].empty()) continue;
                    uint64_t off = *fr[o].begin();
                    fr[o].erase(fr[o].begin());
                    while (o > want) {
                        o--;
                        fr[o].insert(off + (1ull << o));
                    }
                    out->chunk = (int)c;
                    out->order = want;
                    out->off = off;
                    out->addr = (uint64_t)(uintptr_t)chunks_[c].base + off;
                    used_ += 1ull << want;
                    peak_ = std::max(peak_, used_);
                    return true;
                }
            }
            if (pass == 0 && !grow()) return false;
        }
        return false;
    }
    void free(const Block& b)
    {
        if (b.chunk < 0) return;
        auto& fr = chunks_[b.chunk].free;
        uint64_t off = b.off;
        int o = b.order;
        used_ -= 1ull << b.order;
        while (o < top_) {
            uint64_t buddy = off ^ (1ull << o);
            auto itb = fr[o].find(buddy);
            if (itb == fr[o].end()) break;
            fr[o].erase(itb);
            off = std::min(off, buddy);
            o++;
        }
        fr[o].insert(off);
    }
    void release_all()
    {
        for (auto& c : chunks_) (void)hipFree(c.base);
        chunks_.clear();
        reserved_ = used_ = 0;
    }
    ~Arena() { release_all(); }

private:
    bool grow()
    {
        const uint64_t cb = 1ull << top_;
        if (reserved_ + cb > budget_) return false;
        void* p = nullptr;
        if (hipMalloc(&p, cb) != hipSuccess) { (void)hipGetLastError(); return false; }
        Chunk c;
        c.base = (char*)p;
        c.free.resize(top_ + 1);
        c.free[top_].insert(0);
        chunks_.push_back(std::move(c));
        reserved_ += cb;
        return true;
    }
    struct Chunk {
        char* base = nullptr;
        std::vector<std::set<uint64_t>> free;  // free offsets per order
    };
    std::vector<Chunk> chunks_;
    uint64_t budget_ = 0, used_ = 0, peak_ = 0, reserved_ = 0;
    int top_ = kMinOrder;
};

// ---- compiled circuits ---------------------------------------------------------------------

struct EngineOptions {
    int whole_max = 13;       // circuits up to this many qubits run as one whole-state group
    int share_qubits = 0;     // tile of partitioned circuits (0: 13 for specialised complex128 kernels, else 12)
    int precision = HQ_C128;  // the engine's (set at create)
    int coalesce = 4;
    int compile_threads = 4;
    int max_inflight = 2;     // rounds queued on the device ahead of the host
    int jit = 0;              // 1: specialised batched kernels, shared by same-structure groups
    int streams = 4;          // streams a round's group launches spread over (they touch different circuits)
    int tile_by_n[65] = {0};  // autotuned tile per qubit count (0: the rule above)
};

// Streams a round's group launches fork onto from the engine stream and join
// back from (events), so launches of different group structures, each often
// a few blocks over small states, can run side by side.
struct SideStreams {
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hq_status create(int n)
    {
        if (!streams.empty() || n <= 1) return HQ_OK;
        if (!fork) HQ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        for (int i = 0; i < n; i++) {
            // a stream and its join event enter together (or neither does)
            hipEvent_t ev = nullptr;
            HQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            hipStream_t st = nullptr;
            const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);   // default priority
            if (e != hipSuccess) {
                (void)hipEventDestroy(ev);
                set_error(std::string("engine: side stream: ") + hipGetErrorString(e));
                return HQ_ERR_HIP;
            }
            streams.push_back(st);
            join.push_back(ev);
        }
        return HQ_OK;
    }
    ~SideStreams()
    {
        for (auto st : streams) (void)hipStreamSynchronize(st);
        for (auto st : streams) (void)hipStreamDestroy(st);
        for (auto ev : join) (void)hipEventDestroy(ev);
        if (fork) (void)hipEventDestroy(fork);
    }
};

struct GroupPlan {
    int rb = 0, tb = 0, k = 0;
    uint32_t nops = 0;
    uint64_t blob_off = 0;        // the interpreter's program, or the specialised kernel's coefficients
    hipFunction_t fn = nullptr;   // specialised batched kernel (null: the interpreter runs the group)
    double live = 1.0;            // share of one read + write of the state the pass moves
};

// Specialised batched kernels of an engine, one per distinct source: groups
// of the same structure (gate types, qubits, layout) emit the same source and
// share the kernel; their angles stay in each job's coefficient array.
struct KernelCache {
    struct Entry {
        hipModule_t mod = nullptr;
        hipFunction_t fn = nullptr;
        bool ready = false;
        hq_status st = HQ_OK;
        std::string err;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::map<std::string, std::unique_ptr<Entry>> map;
    int device = 0;
    size_t compiled = 0;

    ~KernelCache()
    {
        for (auto& kv : map)
            if (kv.second->mod) (void)hipModuleUnload(kv.second->mod);
    }
    // the kernel of `src`: compiled and loaded by the first caller, waited
    // for by concurrent ones
    hq_status get(const std::string& src, hipFunction_t* fn)
    {
        Entry* e;
        {
            std::unique_lock<std::mutex> lk(mu);
            auto it = map.find(src);
            if (it != map.end()) {
                e = it->second.get();
                cv.wait(lk, [e] { return e->ready; });
                if (e->st != HQ_OK) { set_error(e->err); return e->st; }
                *fn = e->fn;
                return HQ_OK;
            }
            e = (map[src] = std::make_unique<Entry>()).get();
        }
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
        hipModule_t mod = nullptr;
        hipFunction_t f = nullptr;
        hq_status st = jit_compile_load(src, "hq_group_batch", &mod, &f);
        std::lock_guard<std::mutex> lk(mu);
        e->mod = mod;
        e->fn = f;
        e->st = st;
        if (st != HQ_OK) e->err = get_error();
        else compiled++;
        e->ready = true;
        cv.notify_all();
        *fn = f;
        return st;
    }
};

struct JobPlan {
    int n = 0;
    int k_last = 0;               // tile of the last group (tiles = 2^(n - k_last))
    std::vector<GroupPlan> groups;
    std::vector<char> blobs;      // all group programs, 256-byte aligned
};

static int tile_for(const EngineOptions& o, int n)
{
    int k = (n >= 0 && n <= 64) ? o.tile_by_n[n] : 0;
    // measured on the ansatz batch: specialised complex128 kernels run
    // fastest on 13-qubit tiles (fewer passes), everything else on 12
    const int share = o.share_qubits > 0 ? o.share_qubits : (o.jit && o.precision == HQ_C128 ? 13 : 12);
    if (k <= 0) k = n <= o.whole_max ? n : share;
    return std::min(k, n);
}

static hq_status append_blob(JobPlan* jp, const ShmPlan& plan)
{
    for (const ShmOp& op : plan.ops)
        HQ_CHECK(op.cls <= SHM_SCALE, HQ_ERR_UNSUPPORTED, "engine: op class the interpreter does not implement");
    GroupPlan g;
    g.rb = plan.rb;
    g.tb = plan.tb;
    g.k = plan.k;
    g.nops = (uint32_t)plan.ops.size();
    size_t off = (jp->blobs.size() + 255) & ~(size_t)255;
    const size_t bytes = sizeof(ShmHeader) + plan.stages.size() * sizeof(ShmStage) + plan.ops.size() * sizeof(ShmOp);
    jp->blobs.resize(off + bytes);
    char* d = jp->blobs.data() + off;
    memcpy(d, &plan.hdr, sizeof(ShmHeader));
    memcpy(d + sizeof(ShmHeader), plan.stages.data(), plan.stages.size() * sizeof(ShmStage));
    if (!plan.ops.empty())
        memcpy(d + sizeof(ShmHeader) + plan.stages.size() * sizeof(ShmStage), plan.ops.data(),
               plan.ops.size() * sizeof(ShmOp));
    g.blob_off = off;
    jp->groups.push_back(g);
    return HQ_OK;
}

// A group run by a specialised batched kernel: its coefficients (state precision)
// take the program's place in the job's blob area.
static void append_jit(JobPlan* jp, const ShmPlan& plan, const std::vector<double>& coef, hipFunction_t fn)
{
    GroupPlan g;
    g.rb = plan.rb;
    g.tb = plan.tb;
    g.k = plan.k;
    g.nops = (uint32_t)plan.ops.size();
    g.fn = fn;
    const size_t rbytes = plan.precision == HQ_C128 ? 8 : 4;
    size_t off = (jp->blobs.size() + 255) & ~(size_t)255;
    jp->blobs.resize(off + std::max<size_t>(coef.size(), 1) * rbytes, 0);
    char* d = jp->blobs.data() + off;
    for (size_t i = 0; i < coef.size(); i++) {
        if (rbytes == 8) memcpy(d + 8 * i, &coef[i], 8);
        else { const float f = (float)coef[i]; memcpy(d + 4 * i, &f, 4); }
    }
    g.blob_off = off;
    jp->groups.push_back(g);
}

// One group's program: the specialised batched kernel when `kc` is given and
// it compiles, else the interpreter's op table.
static hq_status add_group(JobPlan* jp, int n, int precision, const std::vector<hq_gate>& hg, uint64_t mask,
                           bool first, bool last, uint64_t zero_mask, KernelCache* kc)
{
    ShmPlan plan;
    hq_status st = plan_sharemem(n, precision, hg.data(), (int)hg.size(), mask, 0, &plan, kc && first, 0, kc != nullptr);
    if (st != HQ_OK) return st;
    plan.hdr.zero_mask = zero_mask;
    // the first pass only writes; later ones skip the tiles whose block
    // address sets a qubit still known to be |0>
    const double live = first ? 0.5 : std::ldexp(1.0, -__builtin_popcountll(zero_mask & ~mask));
    // groups of more than kJitMaxOps ops stay with the interpreter (compile time)
    constexpr size_t kJitMaxOps = 512;
    if (kc && plan.ops.size() <= kJitMaxOps) {
        plan.init = first;
        std::vector<double> coef;
        const std::string src = jit_source(plan, &coef, true, last);
        hipFunction_t fn = nullptr;
        if (!src.empty() && kc->get(src, &fn) == HQ_OK && fn) {
            append_jit(jp, plan, coef, fn);
            jp->groups.back().live = live;
            return HQ_OK;
        }
        // the interpreter takes the group when its kernel cannot be built
    }
    if (kc) {
        // in the interpreter's op forms (the shared-code forms are the
        // specialised kernels' own)
        plan = ShmPlan();
        st = plan_sharemem(n, precision, hg.data(), (int)hg.size(), mask, 0, &plan, false, 0, false);
        if (st != HQ_OK) return st;
        plan.hdr.zero_mask = zero_mask;
    }
    st = append_blob(jp, plan);
    if (st != HQ_OK) return st;
    jp->groups.back().live = live;
    return HQ_OK;
}

static hq_status compile_job(const std::vector<Gate>& gates, int n, int precision, const EngineOptions& o,
                             JobPlan* jp, KernelCache* kc = nullptr)
{
    HQ_CHECK(n >= 7 && n <= 34, HQ_ERR_ARG, "engine: circuits must have 7..34 qubits");
    jp->n = n;
    const int k = tile_for(o, n);
    HQ_CHECK(k >= 7 && k <= 13, HQ_ERR_ARG, "engine: tile must have 7..13 qubits");
    // specialised kernels for partitioned circuits only: a whole-state group
    // holds every gate of its circuit (straight-line code of thousands of
    // gates compiles slowly and is rarely shared); such circuits are small
    // and stay with the interpreter
    if (!o.jit || k == n) kc = nullptr;
    if (k == n) {
        std::vector<hq_gate> hg;
        hg.reserve(gates.size());
        for (const Gate& g : gates) hg.push_back(g.toHQ());
        hq_status st = add_group(jp, n, precision, hg, (n >= 64 ? ~0ull : (1ull << n) - 1), true, true, 0, kc);
        if (st != HQ_OK) return st;
        jp->k_last = k;
        return HQ_OK;
    }
    CompilerParams p;
    p.mode = HQ_MODE_GROUP;
    p.precision = precision;
    p.localSize = k;
    p.coalesce = std::min(o.coalesce, k);
    Schedule sch = compileSchedule(n, gates, p);
    State ident(n);
    // the first pass builds |0...0>; until a non-diagonal gate has targeted
    // a qubit, the tiles whose block address sets it are zero and later
    // passes skip them
    const uint64_t all = (1ull << n) - 1;
    uint64_t support = 0;
    size_t total = 0, idx = 0;
    for (auto& lg : sch.localGroups) total += lg.fullGroups.size();
    for (auto& lg : sch.localGroups) {
        for (auto& gg : lg.fullGroups) {
            HQ_CHECK(gg.backend == Backend::PerGate, HQ_ERR_ARG, "engine: unexpected TransMM group");
            qindex rel = Executor::fillRelated(gg.relatedQubits, ident, k);
            std::vector<hq_gate> hg;
            hg.reserve(gg.gates.size());
            for (const Gate& g : gg.gates) hg.push_back(g.toHQ());
            const bool first = idx == 0, last = idx + 1 == total;
            hq_status st = add_group(jp, n, precision, hg, (uint64_t)rel, first, last, first ? 0 : all & ~support, kc);
            if (st != HQ_OK) return st;
            for (const Gate& g : gg.gates)
                if (!g.isDiagonal()) support |= 1ull << g.targetQubit;
            idx++;
        }
    }
    if (jp->groups.empty()) {
        // no gates: one empty pass still produces the readout
        hq_status st = add_group(jp, n, precision, std::vector<hq_gate>(), (1ull << k) - 1, true, true, 0, kc);
        if (st != HQ_OK) return st;
    }
    jp->k_last = jp->groups.back().k;
    return HQ_OK;
}

// ---- jobs and rounds -----------------------------------------------------------------------

constexpr size_t RES_BYTES = 320;  // {norm, sum Z, sum ZZ} (64 B) + 16 candidates

struct Job {
    int64_t id = 0;
    int n = 0;
    int outputs = 0;
    std::vector<Gate> gates;
    JobPlan plan;
    enum Status { QUEUED, COMPILED, ADMITTED, DONE, FAILED } status = QUEUED;
    std::string error;
    std::vector<Block> blocks;
    uint64_t d_blob = 0, d_state = 0, d_part = 0, d_cand[2] = {0, 0};
    int next_group = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double res[3] = {0, 0, 0};
    Cand topk[TOPK];

    bool wants_topk() const { return (outputs & HQ_OUT_TOPK) != 0; }
    uint64_t tiles_last() const { return 1ull << (plan.n - plan.k_last); }
    // the state has to live in HBM unless the circuit is one group that needs no state afterwards
    bool needs_state() const { return plan.groups.size() > 1 || wants_topk(); }
};

// top-k levels: element counts per level (the last level has one chunk)
static std::vector<uint64_t> topk_levels(uint64_t elems)
{
    std::vector<uint64_t> lv;
    uint64_t c = elems;
    for (;;) {
        lv.push_back(c);
        uint64_t blocks = (c + TOPK_CHUNK - 1) / TOPK_CHUNK;
        if (blocks <= 1) break;
        c = blocks * TOPK;
    }
    return lv;
}

struct LaunchDesc {
    int kind;         // 0 group pass (interpreter), 1 finish, 2 top-k, 3 group pass (specialised kernel)
    int rb, tb;
    uint64_t table_off;
    uint32_t count;
    uint64_t blocks;
    hipFunction_t fn = nullptr;
};

struct RoundBuild {
    std::vector<char> table;          // BatchItem / FinishItem / TopkItem arrays
    std::vector<LaunchDesc> launches;
    std::vector<Job*> finishing;      // results land at res_base + i * RES_BYTES
};

static size_t table_append(std::vector<char>& t, const void* p, size_t bytes)
{
    size_t off = (t.size() + 63) & ~(size_t)63;
    t.resize(off + bytes);
    memcpy(t.data() + off, p, bytes);
    return off;
}

// One round over `jobs` (each advances by one group).  res_dev: device
// address of the round's result area.
static void build_round(const std::vector<Job*>& jobs, uint64_t res_dev, RoundBuild* rb)
{
    struct Entry {
        BatchItem it;
        int rb, tb;
        uint64_t tiles;
        uint32_t cost;
        hipFunction_t fn;
    };
    std::vector<Entry> ents;
    for (Job* j : jobs) {
        const int g = j->next_group;
        const GroupPlan& gp = j->plan.groups[g];
        const bool last = g + 1 == (int)j->plan.groups.size();
        Entry e;
        e.it.psi = j->d_state;
        e.it.prog = j->d_blob + gp.blob_off;
        e.it.part = last ? j->d_part : 0;
        e.it.first_block = 0;
        e.it.n = (uint32_t)j->n;
        e.it.mode = (g == 0 ? TILE_INIT : 0) | (last ? TILE_READ : 0) | (last && !j->wants_topk() ? TILE_NOSTORE : 0);
        e.rb = gp.rb;
        e.tb = gp.tb;
        e.tiles = 1ull << (j->n - gp.k);
        e.cost = gp.nops;
        e.fn = gp.fn;
        ents.push_back(e);
        if (last) rb->finishing.push_back(j);
    }
    // one launch per interpreter shape or specialised kernel; long blocks first
    std::stable_sort(ents.begin(), ents.end(), [](const Entry& a, const Entry& b) {
        if (a.fn != b.fn) return std::less<const void*>()(a.fn, b.fn);
        if (a.rb != b.rb) return a.rb < b.rb;
        if (a.tb != b.tb) return a.tb < b.tb;
        return a.cost > b.cost;
    });
    for (size_t i = 0; i < ents.size();) {
        size_t e = i;
        uint64_t blocks = 0;
        std::vector<BatchItem> items;
        while (e < ents.size() && ents[e].fn == ents[i].fn && ents[e].rb == ents[i].rb && ents[e].tb == ents[i].tb) {
            BatchItem it = ents[e].it;
            it.first_block = blocks;
            blocks += ents[e].tiles;
            items.push_back(it);
            e++;
        }
        LaunchDesc d;
        d.kind = ents[i].fn ? 3 : 0;
        d.fn = ents[i].fn;
        d.rb = ents[i].rb;
        d.tb = ents[i].tb;
        d.table_off = table_append(rb->table, items.data(), items.size() * sizeof(BatchItem));
        d.count = (uint32_t)items.size();
        d.blocks = blocks;
        rb->launches.push_back(d);
        i = e;
    }
    if (rb->finishing.empty()) return;
    std::vector<FinishItem> fin;
    for (size_t i = 0; i < rb->finishing.size(); i++) {
        Job* j = rb->finishing[i];
        fin.push_back(FinishItem{j->d_part, j->tiles_last(), res_dev + i * RES_BYTES});
    }
    LaunchDesc d{1, 0, 0, table_append(rb->table, fin.data(), fin.size() * sizeof(FinishItem)),
                 (uint32_t)fin.size(), (uint64_t)fin.size()};
    rb->launches.push_back(d);
    // top-k, level by level across the finishing circuits
    size_t max_levels = 0;
    std::vector<std::vector<uint64_t>> lv(rb->finishing.size());
    for (size_t i = 0; i < rb->finishing.size(); i++) {
        if (!rb->finishing[i]->wants_topk()) continue;
        lv[i] = topk_levels(1ull << rb->finishing[i]->n);
        max_levels = std::max(max_levels, lv[i].size());
    }
    for (size_t L = 0; L < max_levels; L++) {
        std::vector<TopkItem> items;
        uint64_t blocks = 0;
        for (size_t i = 0; i < rb->finishing.size(); i++) {
            if (L >= lv[i].size()) continue;
            Job* j = rb->finishing[i];
            TopkItem t;
            t.kind = L == 0 ? 0 : 1;
            t.pad = 0;
            t.src = L == 0 ? j->d_state : j->d_cand[(L - 1) & 1];
            t.count = lv[i][L];
            t.dst = L + 1 == lv[i].size() ? res_dev + i * RES_BYTES + 64 : j->d_cand[L & 1];
            t.first_block = blocks;
            blocks += (t.count + TOPK_CHUNK - 1) / TOPK_CHUNK;
            items.push_back(t);
        }
        LaunchDesc t{2, 0, 0, table_append(rb->table, items.data(), items.size() * sizeof(TopkItem)),
                     (uint32_t)items.size(), blocks};
        rb->launches.push_back(t);
    }
}

static hq_status issue_round(const RoundBuild& rb, uint64_t table_dev, int precision, hipStream_t main_st,
                             SideStreams* side = nullptr)
{
    // the group launches of a round touch different circuits: with side
    // streams they fork from the engine stream, spread round-robin, and join
    // back before the round's finish and top-k launches (which read them)
    size_t ngroup = 0;
    for (const LaunchDesc& d : rb.launches) ngroup += d.kind == 0 || d.kind == 3;
    const bool fork = side && !side->streams.empty() && ngroup > 1;
    const size_t nside = fork ? std::min(side->streams.size(), ngroup) : 0;
    bool joined = !fork;
    auto join = [&]() -> hq_status {
        for (size_t k = 0; k < nside; k++) {
            HQ_HIP(hipEventRecord(side->join[k], side->streams[k]));
            HQ_HIP(hipStreamWaitEvent(main_st, side->join[k], 0));
        }
        joined = true;
        return HQ_OK;
    };
    if (fork) {
        HQ_HIP(hipEventRecord(side->fork, main_st));
        for (size_t k = 0; k < nside; k++) HQ_HIP(hipStreamWaitEvent(side->streams[k], side->fork, 0));
    }
    size_t next_side = 0;
    auto launch = [&](const LaunchDesc& d) -> hq_status {
        const void* tab = (const void*)(uintptr_t)(table_dev + d.table_off);
        HQ_CHECK(d.blocks < (1ull << 31), HQ_ERR_UNSUPPORTED, "engine: launch too large");
        hq_status rc = HQ_OK;
        const bool group = d.kind == 0 || d.kind == 3;
        if (!group && !joined) {
            rc = join();
            if (rc != HQ_OK) return rc;
        }
        // (build_round lists the group launches first; one after the join
        // would run on the engine stream, after everything before it)
        hipStream_t st = main_st;
        if (group && !joined) st = side->streams[next_side++ % nside];
        if (d.kind == 0) {
            rc = precision == HQ_C128
                     ? launch_batch_shape<double>(d.rb, d.tb, (const BatchItem*)tab, d.count, d.blocks, st)
                     : launch_batch_shape<float>(d.rb, d.tb, (const BatchItem*)tab, d.count, d.blocks, st);
        } else if (d.kind == 3) {
            const void* items = tab;
            uint32_t count = d.count;
            void* args[] = {&items, &count};
            HQ_HIP(hipModuleLaunchKernel(d.fn, (unsigned)d.blocks, 1, 1, 1u << d.tb, 1, 1, 0, st, args, nullptr));
        } else if (d.kind == 1) {
            hipLaunchKernelGGL(k_batch_finish, dim3((unsigned)d.blocks), dim3(256), 0, st, (const FinishItem*)tab);
            HQ_HIP(hipGetLastError());
        } else {
            if (precision == HQ_C128)
                hipLaunchKernelGGL(k_topk<double>, dim3((unsigned)d.blocks), dim3(256), 0, st, (const TopkItem*)tab,
                                   d.count);
            else
                hipLaunchKernelGGL(k_topk<float>, dim3((unsigned)d.blocks), dim3(256), 0, st, (const TopkItem*)tab,
                                   d.count);
            HQ_HIP(hipGetLastError());
        }
        return rc;
    };
    for (const LaunchDesc& d : rb.launches) {
        const hq_status rc = launch(d);
        if (rc != HQ_OK) {
            // the engine stream still waits for what the side streams hold
            if (!joined) (void)join();
            return rc;
        }
    }
    if (!joined) return join();
    return HQ_OK;
}

// device bytes a job needs: programs, state, partials, top-k candidates
static void job_sizes(const Job& j, int precision, uint64_t* blob, uint64_t* state, uint64_t* part,
                      uint64_t* cand0, uint64_t* cand1)
{
    const uint64_t amp = precision == HQ_C128 ? 16 : 8;
    *blob = j.plan.blobs.size();
    *state = j.needs_state() ? (1ull << j.n) * amp : 0;
    *part = j.tiles_last() * 3 * sizeof(double);
    *cand0 = *cand1 = 0;
    if (j.wants_topk()) {
        auto lv = topk_levels(1ull << j.n);
        if (lv.size() > 1) *cand0 = lv[1] * sizeof(Cand);
        if (lv.size() > 2) *cand1 = lv[2] * sizeof(Cand);
    }
}

static bool job_alloc(Job* j, Arena& a, int precision)
{
    uint64_t sz[5];
    job_sizes(*j, precision, &sz[0], &sz[1], &sz[2], &sz[3], &sz[4]);
    uint64_t* dst[5] = {&j->d_blob, &j->d_state, &j->d_part, &j->d_cand[0], &j->d_cand[1]};
    for (int i = 0; i < 5; i++) {
        *dst[i] = 0;
        if (!sz[i]) continue;
        Block b;
        if (!a.alloc(sz[i], &b)) {
            for (auto& x : j->blocks) a.free(x);
            j->blocks.clear();
            return false;
        }
        j->blocks.push_back(b);
        *dst[i] = b.addr;
    }
    return true;
}

static uint64_t job_bytes(const Job& j, int precision, uint64_t* largest)
{
    uint64_t sz[5], t = 0;
    job_sizes(j, precision, &sz[0], &sz[1], &sz[2], &sz[3], &sz[4]);
    *largest = 0;
    for (uint64_t s : sz)
        if (s) { t += 1ull << Arena::order_of(s); *largest = std::max(*largest, s); }
    return t;
}

// pinned host + device staging of one round
struct Slot {
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool busy = false;
};

static hq_status slot_reserve(Slot& s, size_t bytes)
{
    if (bytes <= s.cap) return HQ_OK;
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
    s.host = s.dev = nullptr;
    s.cap = 0;
    size_t cap = std::max<size_t>(bytes, 1 << 20);
    HQ_HIP(hipHostMalloc((void**)&s.host, cap, hipHostMallocDefault));
    HQ_HIP(hipMalloc((void**)&s.dev, cap));
    s.cap = cap;
    return HQ_OK;
}

static void slot_free(Slot& s)
{
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
    if (s.done) (void)hipEventDestroy(s.done);
    s = Slot();
}

static double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace hq

using namespace hq;

// ---- engine ---------------------------------------------------------------------------------

struct hq_engine_s {
    int precision = HQ_C128;
    EngineOptions opt;
    KernelCache kernels;               // specialised batched kernels (option "jit")
    Arena arena;
    hipStream_t stream = nullptr;
    SideStreams side;                  // created at the first round (option "streams")
    double epoch = 0;

    std::mutex mu;                     // jobs map, compile queue
    std::condition_variable cv;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job*> compile_q;
    std::vector<std::thread> workers;
    bool stopping = false;
    int64_t next_id = 1;

    std::deque<Job*> admit_q;          // submission order (compiled or not)
    std::vector<Job*> running;
    struct Round {
        int slot;
        std::vector<Job*> finishing;
        size_t res_off;
    };
    std::deque<Round> inflight;
    std::vector<Slot> slots;
    std::vector<int64_t> completed;    // not yet polled
    int64_t rounds = 0;
    std::set<hq_batch_s*> batches;     // live batches (they hold arena blocks)
    uint64_t batch_bytes = 0;          // arena bytes held by live batches
    std::string broken;                // sticky error: a round failed half-way

    void start_workers()
    {
        if (!workers.empty()) return;
        const int nt = std::max(1, std::min(16, opt.compile_threads));
        for (int i = 0; i < nt; i++) workers.emplace_back([this] { work(); });
    }
    void work()
    {
        for (;;) {
            Job* j;
            EngineOptions o;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [this] { return stopping || !compile_q.empty(); });
                if (stopping) return;
                j = compile_q.front();
                compile_q.pop_front();
                o = opt;
            }
            JobPlan jp;
            hq_status st;
            std::string err;
            try {
                st = compile_job(j->gates, j->n, precision, o, &jp, &kernels);
                if (st != HQ_OK) err = get_error();
            } catch (const std::exception& e) {
                st = HQ_ERR_ARG;
                err = e.what();
            }
            std::lock_guard<std::mutex> lk(mu);
            std::vector<Gate>().swap(j->gates);
            if (st == HQ_OK) {
                j->plan = std::move(jp);
                j->status = Job::COMPILED;
            } else {
                j->status = Job::FAILED;
                j->error = err;
            }
        }
    }
    ~hq_engine_s()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        cv.notify_all();
        for (auto& t : workers) t.join();
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto st : side.streams) (void)hipStreamSynchronize(st);
        for (auto& s : slots) slot_free(s);
        arena.release_all();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// 1: the oldest round in flight was retired, 0: it is still running, <0: HIP error
int retire_front(hq_engine e, bool block)
{
    auto& r = e->inflight.front();
    Slot& s = e->slots[r.slot];
    hipError_t q = block ? hipEventSynchronize(s.done) : hipEventQuery(s.done);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) {
        set_error(std::string("engine: ") + hipGetErrorString(q));
        return -1;
    }
    const double t = now_s() - e->epoch;
    for (size_t i = 0; i < r.finishing.size(); i++) {
        Job* j = r.finishing[i];
        const char* res = s.host + r.res_off + i * RES_BYTES;
        memcpy(j->res, res, sizeof(j->res));
        memcpy(j->topk, res + 64, sizeof(j->topk));
        for (auto& b : j->blocks) e->arena.free(b);
        j->blocks.clear();
        j->t_done = t;
        {
            std::lock_guard<std::mutex> lk(e->mu);
            j->status = Job::DONE;
        }
        e->completed.push_back(j->id);
    }
    s.busy = false;
    e->inflight.pop_front();
    return 1;
}

// retire every finished round; block while too many rounds are queued
hq_status retire(hq_engine e)
{
    while (!e->inflight.empty()) {
        int r = retire_front(e, false);
        if (r < 0) return HQ_ERR_HIP;
        if (r == 0) break;
    }
    while ((int)e->inflight.size() >= std::max(1, e->opt.max_inflight)) {
        if (retire_front(e, true) < 0) return HQ_ERR_HIP;
    }
    return HQ_OK;
}

hq_status admit(hq_engine e, std::vector<Job*>* admitted)
{
    while (!e->admit_q.empty()) {
        Job* j = e->admit_q.front();
        int status;
        {
            std::lock_guard<std::mutex> lk(e->mu);
            status = j->status;
        }
        if (status == Job::FAILED) { e->admit_q.pop_front(); e->completed.push_back(j->id); continue; }
        if (status != Job::COMPILED) break;  // submission order
        uint64_t largest = 0;
        const uint64_t need = job_bytes(*j, e->precision, &largest);
        bool fits = e->arena.can_ever_fit(largest, need);
        if (fits && !job_alloc(j, e->arena, e->precision)) {
            // nothing of this engine holds memory: the circuit will never fit
            // (budget not reachable on the device); otherwise wait for memory
            if (!e->running.empty() || !e->inflight.empty() || e->arena.used() > e->batch_bytes) break;
            fits = false;
        }
        if (!fits) {
            std::lock_guard<std::mutex> lk(e->mu);
            j->status = Job::FAILED;
            j->error = "engine: circuit needs more device memory than the budget allows (" + std::to_string(need) +
                       " B)";
            e->admit_q.pop_front();
            e->completed.push_back(j->id);
            continue;
        }
        j->t_admit = now_s() - e->epoch;
        {
            std::lock_guard<std::mutex> lk(e->mu);
            j->status = Job::ADMITTED;
        }
        e->admit_q.pop_front();
        admitted->push_back(j);
        e->running.push_back(j);
    }
    return HQ_OK;
}

// launch one round for every running job; uploads the programs of newly admitted ones
hq_status launch_round(hq_engine e, const std::vector<Job*>& admitted)
{
    if (e->running.empty()) return HQ_OK;
    int si = -1;
    for (size_t i = 0; i < e->slots.size(); i++)
        if (!e->slots[i].busy) { si = (int)i; break; }
    if (si < 0) {
        e->slots.emplace_back();
        si = (int)e->slots.size() - 1;
        HQ_HIP(hipEventCreateWithFlags(&e->slots[si].done, hipEventDisableTiming));
    }
    Slot& s = e->slots[si];
    // layout: [program blobs of new jobs][round table][results]
    size_t blob_bytes = 0;
    for (Job* j : admitted) blob_bytes += (j->plan.blobs.size() + 255) & ~(size_t)255;
    RoundBuild probe;
    build_round(e->running, 0, &probe);
    const size_t table_off = blob_bytes;
    const size_t res_off = (table_off + probe.table.size() + 255) & ~(size_t)255;
    const size_t total = res_off + probe.finishing.size() * RES_BYTES;
    hq_status rc = slot_reserve(s, total);
    if (rc != HQ_OK) return rc;
    RoundBuild rb;
    build_round(e->running, (uint64_t)(uintptr_t)s.dev + res_off, &rb);
    size_t off = 0;
    for (Job* j : admitted) {
        memcpy(s.host + off, j->plan.blobs.data(), j->plan.blobs.size());
        HQ_HIP(hipMemcpyAsync((void*)(uintptr_t)j->d_blob, s.host + off, j->plan.blobs.size(), hipMemcpyHostToDevice,
                              e->stream));
        off += (j->plan.blobs.size() + 255) & ~(size_t)255;
    }
    memcpy(s.host + table_off, rb.table.data(), rb.table.size());
    HQ_HIP(hipMemcpyAsync(s.dev + table_off, s.host + table_off, rb.table.size(), hipMemcpyHostToDevice, e->stream));
    rc = e->side.create(e->opt.streams);
    if (rc != HQ_OK) return rc;
    rc = issue_round(rb, (uint64_t)(uintptr_t)s.dev + table_off, e->precision, e->stream, &e->side);
    if (rc != HQ_OK) return rc;
    if (!rb.finishing.empty())
        HQ_HIP(hipMemcpyAsync(s.host + res_off, s.dev + res_off, rb.finishing.size() * RES_BYTES,
                              hipMemcpyDeviceToHost, e->stream));
    HQ_HIP(hipEventRecord(s.done, e->stream));
    s.busy = true;
    e->inflight.push_back({si, rb.finishing, res_off});
    e->rounds++;
    // advance
    std::vector<Job*> still;
    for (Job* j : e->running) {
        j->next_group++;
        if (j->next_group < (int)j->plan.groups.size()) still.push_back(j);
    }
    e->running.swap(still);
    return HQ_OK;
}

Job* find_job(hq_engine e, int64_t id)
{
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->jobs.find(id);
    return it == e->jobs.end() ? nullptr : it->second.get();
}

}  // namespace

// ---- fixed batch (hipGraph) ------------------------------------------------------------------

struct hq_batch_s {
    hq_engine e = nullptr;
    std::vector<std::unique_ptr<Job>> jobs;
    char* d_table = nullptr;
    char* d_res = nullptr;
    double* h_res = nullptr;  // pinned, RES_BYTES per circuit
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int rounds = 0;
    int launches = 0;
    ~hq_batch_s()
    {
        // only a batch that hq_batch_create registered counted its blocks
        // into batch_bytes (a failed create frees its blocks without it)
        if (e && e->batches.erase(this))
            for (auto& j : jobs) for (auto& b : j->blocks) e->batch_bytes -= 1ull << b.order;
        if (e && e->stream) (void)hipStreamSynchronize(e->stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_table) (void)hipFree(d_table);
        if (d_res) (void)hipFree(d_res);
        if (h_res) (void)hipHostFree(h_res);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (e) for (auto& j : jobs) for (auto& b : j->blocks) e->arena.free(b);
    }
};

extern "C" {

hq_status hq_engine_create(int precision, uint64_t hbm_budget, hq_engine* out)
{
    HQ_CHECK(out, HQ_ERR_ARG, "hq_engine_create: null output");
    *out = nullptr;
    HQ_CHECK(precision == HQ_C128 || precision == HQ_C64, HQ_ERR_ARG, "hq_engine_create: bad precision");
    HQ_CHECK(hbm_budget >= (1ull << 24), HQ_ERR_ARG, "hq_engine_create: budget below 16 MiB");
    int ndev = 0;
    hq_status st = hq_device_count(&ndev);
    if (st != HQ_OK) return st;
    std::unique_ptr<hq_engine_s> e(new hq_engine_s());
    if (hipGetDevice(&e->kernels.device) != hipSuccess) e->kernels.device = 0;
    e->precision = precision;
    e->opt.precision = precision;
    uint64_t chunk = 1ull << 33;  // 8 GiB
    while (chunk > hbm_budget) chunk >>= 1;
    e->arena.configure(hbm_budget, chunk);
    HQ_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->epoch = now_s();
    *out = e.release();
    return HQ_OK;
}

hq_status hq_engine_destroy(hq_engine e)
{
    if (!e) return HQ_OK;
    HQ_CHECK(e->batches.empty(), HQ_ERR_ARG, "hq_engine_destroy: destroy the engine's batches first");
    delete e;
    return HQ_OK;
}

hq_status hq_engine_set_option(hq_engine e, const char* key, int64_t value)
{
    HQ_CHECK(e && key, HQ_ERR_ARG, "hq_engine_set_option: null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    const std::string k(key);
    if (k == "whole_max") {
        HQ_CHECK(value >= 0 && value <= 13, HQ_ERR_ARG, "whole_max must be in [0, 13]");
        e->opt.whole_max = (int)value;
    } else if (k == "share_qubits") {
        HQ_CHECK(value >= 7 && value <= 13, HQ_ERR_ARG, "share_qubits must be in [7, 13]");
        e->opt.share_qubits = (int)value;
    } else if (k == "coalesce") {
        HQ_CHECK(value >= 0 && value <= 13, HQ_ERR_ARG, "coalesce must be in [0, 13]");
        e->opt.coalesce = (int)value;
    } else if (k == "compile_threads") {
        HQ_CHECK(value >= 1 && value <= 16, HQ_ERR_ARG, "compile_threads must be in [1, 16]");
        HQ_CHECK(e->workers.empty(), HQ_ERR_ARG, "compile_threads must be set before the first submit");
        e->opt.compile_threads = (int)value;
    } else if (k == "jit") {
        HQ_CHECK(value == 0 || value == 1, HQ_ERR_ARG, "jit must be 0 or 1");
        e->opt.jit = (int)value;
    } else if (k == "max_inflight") {
        HQ_CHECK(value >= 1 && value <= 8, HQ_ERR_ARG, "max_inflight must be in [1, 8]");
        e->opt.max_inflight = (int)value;
    } else if (k == "streams") {
        HQ_CHECK(value >= 1 && value <= 8, HQ_ERR_ARG, "streams must be in [1, 8]");
        HQ_CHECK(e->side.streams.empty(), HQ_ERR_ARG, "streams must be set before the first round or batch");
        e->opt.streams = (int)value;
    } else {
        set_error("hq_engine_set_option: unknown option " + k);
        return HQ_ERR_ARG;
    }
    return HQ_OK;
}

hq_status hq_engine_kernels(hq_engine e, int64_t* compiled)
{
    HQ_CHECK(e && compiled, HQ_ERR_ARG, "hq_engine_kernels: null argument");
    std::lock_guard<std::mutex> lk(e->kernels.mu);
    *compiled = (int64_t)e->kernels.compiled;
    return HQ_OK;
}

hq_status hq_engine_set_tile(hq_engine e, int num_qubits, int tile)
{
    HQ_CHECK(e, HQ_ERR_ARG, "hq_engine_set_tile: null engine");
    HQ_CHECK(num_qubits >= 7 && num_qubits <= 34, HQ_ERR_ARG, "hq_engine_set_tile: num_qubits must be in [7, 34]");
    HQ_CHECK(tile == 0 || (tile >= 7 && tile <= 13 && tile <= num_qubits), HQ_ERR_ARG,
             "hq_engine_set_tile: tile must be 0 or in [7, min(13, num_qubits)]");
    std::lock_guard<std::mutex> lk(e->mu);
    e->opt.tile_by_n[num_qubits] = tile;
    return HQ_OK;
}

hq_status hq_engine_get_tile(hq_engine e, int num_qubits, int* tile)
{
    HQ_CHECK(e && tile, HQ_ERR_ARG, "hq_engine_get_tile: null argument");
    HQ_CHECK(num_qubits >= 1 && num_qubits <= 64, HQ_ERR_ARG, "hq_engine_get_tile: bad qubit count");
    std::lock_guard<std::mutex> lk(e->mu);
    *tile = tile_for(e->opt, num_qubits);
    return HQ_OK;
}

hq_status hq_engine_submit(hq_engine e, hq_circuit c, int outputs, int64_t* job)
{
    HQ_CHECK(e && c && c->c && job, HQ_ERR_ARG, "hq_engine_submit: null argument");
    HQ_CHECK((outputs & ~(HQ_OUT_EXPECT | HQ_OUT_TOPK)) == 0, HQ_ERR_ARG, "hq_engine_submit: unknown output flags");
    const int n = c->c->numQubits;
    HQ_CHECK(n >= 7 && n <= 34, HQ_ERR_ARG, "hq_engine_submit: circuits must have 7..34 qubits");
    std::unique_ptr<Job> j(new Job());
    j->n = n;
    j->outputs = outputs;
    j->gates = c->c->getGates();
    j->t_submit = now_s() - e->epoch;
    Job* raw = j.get();
    {
        std::lock_guard<std::mutex> lk(e->mu);
        j->id = e->next_id++;
        *job = j->id;
        e->jobs[j->id] = std::move(j);
        e->compile_q.push_back(raw);
    }
    e->admit_q.push_back(raw);
    e->start_workers();
    e->cv.notify_one();
    return HQ_OK;
}

hq_status hq_engine_step(hq_engine e, int* pending)
{
    HQ_CHECK(e, HQ_ERR_ARG, "hq_engine_step: null engine");
    HQ_CHECK(e->broken.empty(), HQ_ERR_HIP, "hq_engine_step: the engine failed earlier: " + e->broken);
    hq_status rc = retire(e);
    if (rc != HQ_OK) { e->broken = get_error(); return rc; }
    std::vector<Job*> admitted;
    rc = admit(e, &admitted);
    if (rc != HQ_OK) { e->broken = get_error(); return rc; }
    rc = launch_round(e, admitted);
    if (rc != HQ_OK) { e->broken = get_error(); return rc; }
    if (pending) {
        // not yet done: queued, compiling, admitted or with results in flight
        int cnt = (int)e->admit_q.size() + (int)e->running.size();
        for (auto& r : e->inflight) cnt += (int)r.finishing.size();
        *pending = cnt;
    }
    return HQ_OK;
}

hq_status hq_engine_drain(hq_engine e)
{
    HQ_CHECK(e, HQ_ERR_ARG, "hq_engine_drain: null engine");
    for (;;) {
        int pending = 0;
        hq_status rc = hq_engine_step(e, &pending);
        if (rc != HQ_OK) return rc;
        if (pending == 0) return HQ_OK;
        if (e->running.empty() && e->inflight.empty()) std::this_thread::sleep_for(std::chrono::microseconds(50));
        else if (e->running.empty() && retire_front(e, true) < 0) return HQ_ERR_HIP;
    }
}

hq_status hq_engine_poll(hq_engine e, int64_t* jobs, int cap, int* count)
{
    HQ_CHECK(e && count && (jobs || cap == 0), HQ_ERR_ARG, "hq_engine_poll: null argument");
    hq_status rc = retire(e);
    if (rc != HQ_OK) return rc;
    const int k = std::min<int>(cap, (int)e->completed.size());
    for (int i = 0; i < k; i++) jobs[i] = e->completed[i];
    e->completed.erase(e->completed.begin(), e->completed.begin() + k);
    *count = k;
    return HQ_OK;
}

hq_status hq_engine_job_status(hq_engine e, int64_t job, int* status)
{
    HQ_CHECK(e && status, HQ_ERR_ARG, "hq_engine_job_status: null argument");
    Job* j = find_job(e, job);
    HQ_CHECK(j, HQ_ERR_ARG, "hq_engine_job_status: unknown job");
    std::lock_guard<std::mutex> lk(e->mu);
    *status = j->status;
    return HQ_OK;
}

hq_status hq_engine_result(hq_engine e, int64_t job, double* expect3, uint64_t* topk_idx, double* topk_prob,
                           double* times3)
{
    HQ_CHECK(e, HQ_ERR_ARG, "hq_engine_result: null engine");
    Job* j = find_job(e, job);
    HQ_CHECK(j, HQ_ERR_ARG, "hq_engine_result: unknown job");
    int status;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        status = j->status;
        if (status == Job::FAILED) { set_error(j->error); return HQ_ERR_ARG; }
    }
    HQ_CHECK(status == Job::DONE, HQ_ERR_ARG, "hq_engine_result: job not finished");
    if (expect3) memcpy(expect3, j->res, sizeof(j->res));
    if (topk_idx || topk_prob) {
        HQ_CHECK(j->wants_topk(), HQ_ERR_ARG, "hq_engine_result: job was submitted without HQ_OUT_TOPK");
        for (int i = 0; i < TOPK; i++) {
            if (topk_idx) topk_idx[i] = j->topk[i].idx;
            if (topk_prob) topk_prob[i] = j->topk[i].p;
        }
    }
    if (times3) { times3[0] = j->t_submit; times3[1] = j->t_admit; times3[2] = j->t_done; }
    return HQ_OK;
}

hq_status hq_engine_release(hq_engine e, int64_t job)
{
    HQ_CHECK(e, HQ_ERR_ARG, "hq_engine_release: null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->jobs.find(job);
    HQ_CHECK(it != e->jobs.end(), HQ_ERR_ARG, "hq_engine_release: unknown job");
    HQ_CHECK(it->second->status == Job::DONE || it->second->status == Job::FAILED, HQ_ERR_ARG,
             "hq_engine_release: job not finished");
    e->jobs.erase(it);
    return HQ_OK;
}

hq_status hq_engine_stats(hq_engine e, uint64_t* used, uint64_t* peak, uint64_t* reserved, int64_t* rounds)
{
    HQ_CHECK(e, HQ_ERR_ARG, "hq_engine_stats: null engine");
    if (used) *used = e->arena.used();
    if (peak) *peak = e->arena.peak();
    if (reserved) *reserved = e->arena.reserved();
    if (rounds) *rounds = e->rounds;
    return HQ_OK;
}

hq_status hq_engine_now(hq_engine e, double* seconds)
{
    HQ_CHECK(e && seconds, HQ_ERR_ARG, "hq_engine_now: null argument");
    *seconds = now_s() - e->epoch;
    return HQ_OK;
}

// ---- fixed batch ----

hq_status hq_batch_create(hq_engine e, const hq_circuit* circuits, int count, int outputs, hq_batch* out)
{
    HQ_CHECK(e && out && circuits && count > 0, HQ_ERR_ARG, "hq_batch_create: bad argument");
    HQ_CHECK((outputs & ~(HQ_OUT_EXPECT | HQ_OUT_TOPK)) == 0, HQ_ERR_ARG, "hq_batch_create: unknown output flags");
    HQ_CHECK(e->running.empty() && e->inflight.empty() && e->admit_q.empty(), HQ_ERR_ARG,
             "hq_batch_create: the engine is serving circuits");
    *out = nullptr;
    std::unique_ptr<hq_batch_s> b(new hq_batch_s());
    b->e = e;
    EngineOptions o;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        o = e->opt;
    }
    b->jobs.resize(count);
    for (int i = 0; i < count; i++) {
        HQ_CHECK(circuits[i] && circuits[i]->c, HQ_ERR_ARG, "hq_batch_create: null circuit");
        b->jobs[i].reset(new Job());
        b->jobs[i]->n = circuits[i]->c->numQubits;
        b->jobs[i]->outputs = outputs;
        b->jobs[i]->id = i;
    }
    // compile in parallel
    std::atomic<int> next(0);
    std::vector<std::string> errs(count);
    std::vector<hq_status> sts(count, HQ_OK);
    auto work = [&]() {
        for (;;) {
            int i = next++;
            if (i >= count) return;
            try {
                sts[i] = compile_job(circuits[i]->c->getGates(), b->jobs[i]->n, e->precision, o, &b->jobs[i]->plan,
                                     &e->kernels);
                if (sts[i] != HQ_OK) errs[i] = get_error();
            } catch (const std::exception& ex) {
                sts[i] = HQ_ERR_ARG;
                errs[i] = ex.what();
            }
        }
    };
    const int nt = std::max(1, std::min({16, o.compile_threads * 2, count}));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(work);
    for (auto& t : th) t.join();
    for (int i = 0; i < count; i++)
        if (sts[i] != HQ_OK) { set_error("hq_batch_create: circuit " + std::to_string(i) + ": " + errs[i]); return sts[i]; }
    // memory
    for (auto& j : b->jobs) {
        if (!job_alloc(j.get(), e->arena, e->precision)) {
            set_error("hq_batch_create: the batch does not fit the HBM budget");
            return HQ_ERR_OOM;
        }
    }
    // all rounds into one table
    std::vector<Job*> running;
    for (auto& j : b->jobs) running.push_back(j.get());
    HQ_HIP(hipMalloc((void**)&b->d_res, (size_t)count * RES_BYTES));
    HQ_HIP(hipHostMalloc((void**)&b->h_res, (size_t)count * RES_BYTES, hipHostMallocDefault));
    std::vector<RoundBuild> builds;
    std::vector<char> table;
    std::vector<size_t> round_off;
    std::vector<size_t> job_res(count);
    size_t res_used = 0;
    while (!running.empty()) {
        RoundBuild rb;
        build_round(running, (uint64_t)(uintptr_t)b->d_res + res_used * RES_BYTES, &rb);
        for (size_t i = 0; i < rb.finishing.size(); i++) job_res[rb.finishing[i]->id] = res_used + i;
        res_used += rb.finishing.size();
        size_t off = (table.size() + 255) & ~(size_t)255;
        table.resize(off + rb.table.size());
        memcpy(table.data() + off, rb.table.data(), rb.table.size());
        round_off.push_back(off);
        builds.push_back(std::move(rb));
        std::vector<Job*> still;
        for (Job* j : running) {
            j->next_group++;
            if (j->next_group < (int)j->plan.groups.size()) still.push_back(j);
        }
        running.swap(still);
    }
    HQ_HIP(hipMalloc((void**)&b->d_table, std::max<size_t>(table.size(), 256)));
    HQ_HIP(hipMemcpy(b->d_table, table.data(), table.size(), hipMemcpyHostToDevice));
    for (auto& j : b->jobs)
        HQ_HIP(hipMemcpy((void*)(uintptr_t)j->d_blob, j->plan.blobs.data(), j->plan.blobs.size(),
                         hipMemcpyHostToDevice));
    // the result of circuit i is at job_res[i]; keep the mapping in next_group
    for (int i = 0; i < count; i++) b->jobs[i]->next_group = (int)job_res[i];
    HQ_HIP(hipStreamSynchronize(e->stream));
    hq_status rc = e->side.create(e->opt.streams);
    if (rc != HQ_OK) return rc;
    HQ_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (size_t r = 0; r < builds.size() && rc == HQ_OK; r++) {
        rc = issue_round(builds[r], (uint64_t)(uintptr_t)b->d_table + round_off[r], e->precision, e->stream,
                         &e->side);
        b->launches += (int)builds[r].launches.size();
    }
    hipError_t ce = hipMemcpyAsync(b->h_res, b->d_res, (size_t)count * RES_BYTES, hipMemcpyDeviceToHost, e->stream);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(e->stream, &g);
    if (rc != HQ_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (ce != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error(std::string("hq_batch_create: graph capture failed: ") +
                  hipGetErrorString(ce != hipSuccess ? ce : ee));
        return HQ_ERR_HIP;
    }
    b->graph = g;
    HQ_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    HQ_HIP(hipEventCreate(&b->ev0));
    HQ_HIP(hipEventCreate(&b->ev1));
    b->rounds = (int)builds.size();
    e->batches.insert(b.get());
    for (auto& j : b->jobs) for (auto& x : j->blocks) e->batch_bytes += 1ull << x.order;
    *out = b.release();
    return HQ_OK;
}

hq_status hq_batch_run(hq_batch b, double* device_ms)
{
    HQ_CHECK(b, HQ_ERR_ARG, "hq_batch_run: null batch");
    hipStream_t st = b->e->stream;
    HQ_HIP(hipEventRecord(b->ev0, st));
    HQ_HIP(hipGraphLaunch(b->exec, st));
    HQ_HIP(hipEventRecord(b->ev1, st));
    HQ_HIP(hipEventSynchronize(b->ev1));
    if (device_ms) {
        float ms = 0;
        HQ_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *device_ms = ms;
    }
    return HQ_OK;
}

hq_status hq_batch_info(hq_batch b, int* count, int* rounds, int* launches, uint64_t* device_bytes)
{
    HQ_CHECK(b, HQ_ERR_ARG, "hq_batch_info: null batch");
    if (count) *count = (int)b->jobs.size();
    if (rounds) *rounds = b->rounds;
    if (launches) *launches = b->launches;
    if (device_bytes) {
        uint64_t t = 0;
        for (auto& j : b->jobs) for (auto& x : j->blocks) t += 1ull << x.order;
        *device_bytes = t;
    }
    return HQ_OK;
}

hq_status hq_batch_pass_bytes(hq_batch b, uint64_t* bytes)
{
    HQ_CHECK(b && bytes, HQ_ERR_ARG, "hq_batch_pass_bytes: null argument");
    const double amp = b->e->precision == HQ_C128 ? 16.0 : 8.0;
    double t = 0;
    for (auto& j : b->jobs)
        for (auto& g : j->plan.groups) t += 2.0 * amp * std::ldexp(1.0, j->plan.n) * g.live;
    *bytes = (uint64_t)t;
    return HQ_OK;
}

hq_status hq_batch_results(hq_batch b, double* expect3, uint64_t* topk_idx, double* topk_prob)
{
    HQ_CHECK(b, HQ_ERR_ARG, "hq_batch_results: null batch");
    const int count = (int)b->jobs.size();
    for (int i = 0; i < count; i++) {
        const char* r = (const char*)b->h_res + (size_t)b->jobs[i]->next_group * RES_BYTES;
        if (expect3) memcpy(expect3 + 3 * i, r, 3 * sizeof(double));
        if (topk_idx || topk_prob) {
            HQ_CHECK(b->jobs[i]->wants_topk(), HQ_ERR_ARG, "hq_batch_results: batch was created without HQ_OUT_TOPK");
            const Cand* c = (const Cand*)(r + 64);
            for (int k = 0; k < TOPK; k++) {
                if (topk_idx) topk_idx[(size_t)i * TOPK + k] = c[k].idx;
                if (topk_prob) topk_prob[(size_t)i * TOPK + k] = c[k].p;
            }
        }
    }
    return HQ_OK;
}

hq_status hq_batch_destroy(hq_batch b)
{
    delete b;
    return HQ_OK;
}

// ---- top-k of one state ----

hq_status hq_state_topk(hq_state s, int k, uint64_t* idx, double* prob)
{
    HQ_CHECK(s && (idx || prob), HQ_ERR_ARG, "hq_state_topk: null argument");
    HQ_CHECK(k >= 1 && k <= TOPK, HQ_ERR_ARG, "hq_state_topk: k must be in [1, 16]");
    auto lv = topk_levels(s->amps);
    size_t c0 = lv.size() > 1 ? lv[1] : 0, c1 = lv.size() > 2 ? lv[2] : 0;
    char* d = nullptr;
    const size_t bytes = (c0 + c1 + TOPK) * sizeof(Cand) + lv.size() * sizeof(TopkItem) + 1024;
    HQ_HIP(hipMalloc((void**)&d, bytes));
    Cand* cand[2] = {(Cand*)d, (Cand*)d + c0};
    Cand* fin = (Cand*)d + c0 + c1;
    TopkItem* items = (TopkItem*)(((uintptr_t)(fin + TOPK) + 63) & ~(uintptr_t)63);
    std::vector<TopkItem> h(lv.size());
    for (size_t L = 0; L < lv.size(); L++) {
        h[L].kind = L == 0 ? 0 : 1;
        h[L].pad = 0;
        h[L].src = L == 0 ? (uint64_t)(uintptr_t)s->data : (uint64_t)(uintptr_t)cand[(L - 1) & 1];
        h[L].count = lv[L];
        h[L].dst = L + 1 == lv.size() ? (uint64_t)(uintptr_t)fin : (uint64_t)(uintptr_t)cand[L & 1];
        h[L].first_block = 0;
    }
    hipError_t e = hipMemcpyAsync(items, h.data(), h.size() * sizeof(TopkItem), hipMemcpyHostToDevice, s->stream);
    for (size_t L = 0; L < lv.size() && e == hipSuccess; L++) {
        const unsigned blocks = (unsigned)((lv[L] + TOPK_CHUNK - 1) / TOPK_CHUNK);
        if (s->precision == HQ_C128)
            hipLaunchKernelGGL(k_topk<double>, dim3(blocks), dim3(256), 0, s->stream, items + L, 1u);
        else
            hipLaunchKernelGGL(k_topk<float>, dim3(blocks), dim3(256), 0, s->stream, items + L, 1u);
        e = hipGetLastError();
    }
    Cand res[TOPK];
    if (e == hipSuccess) e = hipMemcpyAsync(res, fin, sizeof(res), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e != hipSuccess) {
        set_error(std::string("hq_state_topk: ") + hipGetErrorString(e));
        return HQ_ERR_HIP;
    }
    for (int i = 0; i < k; i++) {
        if (idx) idx[i] = res[i].idx;
        if (prob) prob[i] = res[i].p;
    }
    return HQ_OK;
}

}  // extern "C"
