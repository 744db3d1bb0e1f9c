// jit.hip -- structure-specialised OShareMem kernels (hiprtc).
//
// The interpreter kernel (sharemem.hip) dispatches every op at run time; on
// gfx950 that costs about 2.6x the FP64 VALU work of the gates themselves
// (register copies at the dispatch merges, serialised scalar loads).  For big
// states each group is instead emitted as straight-line HIP: register
// indices, pair lists, controls, lane/thread/block bit positions and the LDS
// exchange addresses are literals; X/CNOT become register renaming; the
// angles of parametric gates stay runtime data (a coefficient array), so
// circuits that differ only in angles share compiled code.  Sources are
// compiled with hiprtc for gfx950 and cached in memory and on disk
// (HQ_JIT_CACHE, default ~/.cache/hyquas_amd/jit).  hiprtc compiles one
// program at a time per process, so concurrent compiles run as child
// processes of the helper hq_rtc next to the library (HQ_RTC_HELPER names
// another path, or "none" to compile in this process).

#include "hq_ops.h"
#include "jit_opts.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <mutex>
#include <spawn.h>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>
#include <sys/wait.h>
#include <thread>
#include <unistd.h>
#include <unordered_map>

extern char** environ;

namespace hq {

namespace {

std::string lit(double v, bool f32)
{
    char b[64];
    if (f32) snprintf(b, sizeof b, "%.9gf", (double)(float)v);
    else snprintf(b, sizeof b, "%.17g", v);
    std::string s = b;
    if (s.find_first_of(".eEn") == std::string::npos) s.insert(s.size() - (f32 ? 1 : 0), ".0");
    return "(R)(" + s + ")";
}

std::string hex64(uint64_t v)
{
    char b[32];
    snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
    return b;
}

std::string hex32(uint32_t v)
{
    char b[32];
    snprintf(b, sizeof b, "0x%xu", v);
    return b;
}

int ctz64(uint64_t x) { return __builtin_ctzll(x); }

uint32_t swz_of(uint32_t x, const ShmHeader& H, int K)
{
    uint32_t f = 0;
    for (int i = 3; i < K; i++)
        if ((x >> i) & 1) f ^= H.swz_col[i];
    return x ^ f;
}

struct Emitter {
    const ShmPlan& P;
    std::vector<double>* coef;
    std::ostringstream o;
    int NR, NT, K;
    bool f32;
    bool nt_ld, nt_st;       // non-temporal HBM loads / stores
    int bmap;                // block index -> tile: 0 natural, 1 reversed bits, 2 contiguous per XCD
    std::vector<int> name;   // register j is held in variable <gen><name[j]>
    char gen = 'r';          // variable set of the current registers (r or q)
    bool split;              // tile above 64 KiB: two-phase LDS exchanges over half the tile
    bool batch;              // batched form: blocks map to (circuit, tile) through an item table
    bool readout;            // batched form, last group: fold sum p, p Z, p ZZ into per-tile partials

    Emitter(const ShmPlan& p, std::vector<double>* c, bool batch_form = false, bool readout_fold = false)
        : P(p), coef(c), batch(batch_form), readout(readout_fold)
    {
        NR = 1 << p.rb;
        NT = 1 << p.tb;
        K = p.k;
        f32 = p.precision != HQ_C128;
        split = ((size_t)1 << K) * (f32 ? 8 : 16) > 64 * 1024;
        // non-temporal HBM accesses: chosen per group (they pay only where a
        // wave instruction covers whole 128-byte lines)
        nt_ld = p.nt_load;
        nt_st = p.nt_store;
        const char* bm = getenv("HQ_JIT_BMAP");
        bmap = bm ? atoi(bm) : 0;
        if ((bmap == 2 && P.hdr.num_block_bits < 3) || (bmap == 1 && P.hdr.num_block_bits == 0) || batch) bmap = 0;
        name.resize(NR);
        for (int j = 0; j < NR; j++) name[j] = j;
    }

    std::string r(int j) const { return std::string(1, gen) + std::to_string(name[j]); }

    // LDS index in a split exchange: the tile's swizzled index with tile bit
    // b removed (b is a register bit, fixed within a phase; the swizzle only
    // moves bits 0..2, so the lanes keep their bank-group spread)
    uint32_t swz_split(uint32_t x, int b) const
    {
        if (b < 3) {
            // the swizzle would move bit b: compact first, then swizzle the K-1 bits
            const uint32_t lo = x & ((1u << b) - 1), hi = (x >> (b + 1)) << b;
            return swz_of(lo | hi, P.hdr, K - 1);
        }
        const uint32_t y = swz_of(x, P.hdr, K);
        const uint32_t lo = y & ((1u << b) - 1), hi = (y >> (b + 1)) << b;
        return lo | hi;
    }

    void lane_lds_split(const ShmStage& st, int b, const char* var)
    {
        o << "    unsigned " << var << " = 0;\n";
        for (int i = 0; i < P.tb; i++) {
            uint32_t sw = swz_split(1u << st.thr_bits[i], b);
            o << "    " << var << " ^= (0u - ((t >> " << i << ") & 1u)) & " << hex32(sw) << ";\n";
        }
    }

    // exchange from stage A to stage B in two halves split on a register bit
    // common to both (the planner keeps one): write the half with bit = 0,
    // read it, then the other half; new values go to the other variable set
    void emit_stage_ops(const ShmStage& st)
    {
        for (uint32_t oi = st.op_begin; oi < st.op_end; oi++) emit_op(P.ops[oi], true);
    }

    void split_exchange(const ShmStage& A, const ShmStage& B)
    {
        int b = -1, ra = -1, rbb = -1;
        for (int i = P.rb - 1; i >= 0 && b < 0; i--)
            for (int k2 = 0; k2 < P.rb; k2++)
                if (A.reg_bits[i] == B.reg_bits[k2]) { b = (int)A.reg_bits[i]; ra = i; rbb = k2; break; }
        if (b < 0) throw std::runtime_error("split exchange without a common register bit");
        const char ng = gen == 'r' ? 'q' : 'r';
        o << "    {\n";
        lane_lds_split(A, b, "sw");
        lane_lds_split(B, b, "sr");
        for (int ph = 0; ph < 2; ph++) {
            if (ph == 1) o << "    __syncthreads();\n";
            for (int j = 0; j < NR; j++)
                if (((j >> ra) & 1) == ph)
                    o << "    lds[sw ^ " << hex32(swz_split(reg_tile(A, j), b)) << "] = " << r(j) << ";\n";
            o << "    __syncthreads();\n";
            for (int j = 0; j < NR; j++)
                if (((j >> rbb) & 1) == ph)
                    o << "    " << ng << j << " = lds[sr ^ " << hex32(swz_split(reg_tile(B, j), b)) << "];\n";
        }
        o << "    }\n";
        gen = ng;
        for (int j = 0; j < NR; j++) name[j] = j;
    }

    // complex64, whole-line variant: register-index bit holding the lowest
    // state qubit in stage st, if any; registers j and j | (1 << bit) are then
    // neighbours in HBM and move as one 16-byte access
    int pair_pos(const ShmStage& st) const
    {
        if (!f32 || !P.whole_line) return -1;
        for (int i = 0; i < P.rb; i++)
            if (P.hdr.tile_phys[st.reg_bits[i]] == 1) return i;
        return -1;
    }

    // physical offset (as a literal) of register index j in stage s
    uint64_t reg_off(const ShmStage& st, int j) const
    {
        uint64_t off = 0;
        for (int i = 0; i < P.rb; i++)
            if ((j >> i) & 1) off |= P.hdr.tile_phys[st.reg_bits[i]];
        return off;
    }

    uint32_t reg_tile(const ShmStage& st, int j) const
    {
        uint32_t e = 0;
        for (int i = 0; i < P.rb; i++)
            if ((j >> i) & 1) e |= 1u << st.reg_bits[i];
        return e;
    }

    // lane part of the physical address for stage st (runs of consecutive
    // bits are moved together)
    void lane_phys(const ShmStage& st, const char* var)
    {
        o << "    unsigned long long " << var << " = 0;\n";
        for (int i = 0; i < P.tb; i++) {
            int p = ctz64(P.hdr.tile_phys[st.thr_bits[i]]);
            int len = 1;
            while (i + len < P.tb && ctz64(P.hdr.tile_phys[st.thr_bits[i + len]]) == p + len) len++;
            o << "    " << var << " |= (unsigned long long)((t >> " << i << ") & " << hex32((1u << len) - 1)
              << ") << " << p << ";\n";
            i += len - 1;
        }
    }

    // swizzled LDS index of the lane part for stage st
    void lane_lds(const ShmStage& st, const char* var)
    {
        o << "    unsigned " << var << " = 0;\n";
        for (int i = 0; i < P.tb; i++) {
            uint32_t sw = swz_of(1u << st.thr_bits[i], P.hdr, K);
            o << "    " << var << " ^= (0u - ((t >> " << i << ") & 1u)) & " << hex32(sw) << ";\n";
        }
    }

    // coefficient expressions of op oi (constants or runtime array entries)
    std::vector<std::string> coeffs(const ShmOp& op)
    {
        std::vector<std::string> m(8);
        if (op.flags & SHM_FLAG_CONST) {
            for (int i = 0; i < 8; i++) m[i] = lit(op.m[i], f32);
        } else {
            size_t base = coef->size();
            for (int i = 0; i < 8; i++) coef->push_back(op.m[i]);
            for (int i = 0; i < 8; i++) m[i] = "cf[" + std::to_string(base + i) + "]";
        }
        return m;
    }

    void scale_stmt(int j, const std::string& fr, const std::string& fi)
    {
        o << "      { const C v = " << r(j) << "; " << r(j) << ".x = " << fr << " * v.x - " << fi << " * v.y; "
          << r(j) << ".y = " << fr << " * v.y + " << fi << " * v.x; }\n";
    }

    void emit_op(const ShmOp& op, bool identity_ok_renaming)
    {
        std::vector<std::string> m = coeffs(op);
        const bool const_op = op.flags & SHM_FLAG_CONST;
        o << "    {\n";
        bool blockc = op.cblk != 0;
        if (blockc) o << "     if ((b & " << hex64(op.cblk) << ") == " << hex64(op.cblk) << ") {\n";
        // per-lane identity where a thread-bit control is 0
        std::vector<std::string> mv(8);
        static const char* ident[8] = {"(R)1", "(R)0", "(R)0", "(R)0", "(R)0", "(R)0", "(R)1", "(R)0"};
        if (op.cthr) {
            o << "      const bool pc = (t & " << hex32(op.cthr) << ") == " << hex32(op.cthr) << ";\n";
            for (int i = 0; i < 8; i++) {
                o << "      const R m" << i << " = pc ? " << m[i] << " : " << ident[i] << ";\n";
                mv[i] = "m" + std::to_string(i);
            }
        } else {
            for (int i = 0; i < 8; i++) {
                if (const_op) {
                    mv[i] = m[i];
                } else {
                    o << "      const R m" << i << " = " << m[i] << ";\n";
                    mv[i] = "m" + std::to_string(i);
                }
            }
        }
        const bool rename = identity_ok_renaming && !blockc;
        // the shear forms come from uncontrolled gates only; a per-lane
        // identity select would not express the identity for them
        if (op.cthr && (op.cls == SHM_SHEAR || op.cls == SHM_SHEARX))
            throw std::runtime_error("shear op with a thread-bit control");
        if (op.cls == SHM_SCALE) {
            for (int j = 0; j < NR; j++) scale_stmt(j, mv[0], mv[1]);
        } else if (op.where == SHM_IN_REG) {
            const int Pb = op.tpos;
            for (int j = 0; j < NR; j++) {
                if (j & (1 << Pb)) continue;
                if ((j & op.creg) != op.creg) continue;
                const int jh = j | (1 << Pb);
                switch (op.cls) {
                case SHM_SWAP:
                    if (rename) std::swap(name[j], name[jh]);
                    else o << "      { const C tmp = " << r(j) << "; " << r(j) << " = " << r(jh) << "; " << r(jh) << " = tmp; }\n";
                    break;
                case SHM_Y:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; " << r(j) << ".x = hi.y; " << r(j)
                      << ".y = -hi.x; " << r(jh) << ".x = -lo.y; " << r(jh) << ".y = lo.x; }\n";
                    break;
                case SHM_REAL:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; "
                      << r(j) << ".x = " << mv[0] << " * lo.x + " << mv[2] << " * hi.x; "
                      << r(j) << ".y = " << mv[0] << " * lo.y + " << mv[2] << " * hi.y; "
                      << r(jh) << ".x = " << mv[4] << " * lo.x + " << mv[6] << " * hi.x; "
                      << r(jh) << ".y = " << mv[4] << " * lo.y + " << mv[6] << " * hi.y; }\n";
                    break;
                case SHM_RX:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; "
                      << r(j) << ".x = " << mv[0] << " * lo.x - " << mv[3] << " * hi.y; "
                      << r(j) << ".y = " << mv[0] << " * lo.y + " << mv[3] << " * hi.x; "
                      << r(jh) << ".x = " << mv[6] << " * hi.x - " << mv[5] << " * lo.y; "
                      << r(jh) << ".y = " << mv[6] << " * hi.y + " << mv[5] << " * lo.x; }\n";
                    break;
                case SHM_GEN:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; "
                      << r(j) << ".x = (" << mv[0] << " * lo.x - " << mv[1] << " * lo.y) + (" << mv[2] << " * hi.x - " << mv[3] << " * hi.y); "
                      << r(j) << ".y = (" << mv[0] << " * lo.y + " << mv[1] << " * lo.x) + (" << mv[2] << " * hi.y + " << mv[3] << " * hi.x); "
                      << r(jh) << ".x = (" << mv[4] << " * lo.x - " << mv[5] << " * lo.y) + (" << mv[6] << " * hi.x - " << mv[7] << " * hi.y); "
                      << r(jh) << ".y = (" << mv[4] << " * lo.y + " << mv[5] << " * lo.x) + (" << mv[6] << " * hi.y + " << mv[7] << " * hi.x); }\n";
                    break;
                case SHM_DIAG:
                    scale_stmt(j, mv[0], mv[1]);
                    scale_stmt(jh, mv[6], mv[7]);
                    break;
                case SHM_HADD:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; " << r(j) << ".x = lo.x + hi.x; "
                      << r(j) << ".y = lo.y + hi.y; " << r(jh) << ".x = lo.x - hi.x; " << r(jh) << ".y = lo.y - hi.y; }\n";
                    break;
                case SHM_ROTA:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; " << r(j) << ".x = lo.x - " << mv[0]
                      << " * hi.x; " << r(j) << ".y = lo.y - " << mv[0] << " * hi.y; " << r(jh) << ".x = hi.x + "
                      << mv[0] << " * lo.x; " << r(jh) << ".y = hi.y + " << mv[0] << " * lo.y; }\n";
                    break;
                case SHM_ROTB:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; " << r(j) << ".x = " << mv[0]
                      << " * lo.x - hi.x; " << r(j) << ".y = " << mv[0] << " * lo.y - hi.y; " << r(jh) << ".x = lo.x + "
                      << mv[0] << " * hi.x; " << r(jh) << ".y = lo.y + " << mv[0] << " * hi.y; }\n";
                    break;
                case SHM_RXA:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; " << r(j) << ".x = lo.x - " << mv[0]
                      << " * hi.y; " << r(j) << ".y = lo.y + " << mv[0] << " * hi.x; " << r(jh) << ".x = hi.x - "
                      << mv[0] << " * lo.y; " << r(jh) << ".y = hi.y + " << mv[0] << " * lo.x; }\n";
                    break;
                case SHM_RXB:
                    o << "      { const C lo = " << r(j) << ", hi = " << r(jh) << "; " << r(j) << ".x = " << mv[0]
                      << " * lo.x - hi.y; " << r(j) << ".y = " << mv[0] << " * lo.y + hi.x; " << r(jh) << ".x = "
                      << mv[0] << " * hi.x - lo.y; " << r(jh) << ".y = " << mv[0] << " * hi.y + lo.x; }\n";
                    break;
                case SHM_SHEAR:
                    o << "      { C lo = " << r(j) << ", hi = " << r(jh) << "; "
                      << "lo.x += " << mv[0] << " * hi.x; lo.y += " << mv[0] << " * hi.y; "
                      << "hi.x += " << mv[2] << " * lo.x; hi.y += " << mv[2] << " * lo.y; "
                      << "lo.x += " << mv[0] << " * hi.x; lo.y += " << mv[0] << " * hi.y; "
                      << r(j) << " = lo; " << r(jh) << " = hi; }\n";
                    break;
                case SHM_SHEARX:
                    o << "      { C lo = " << r(j) << ", hi = " << r(jh) << "; "
                      << "lo.x -= " << mv[0] << " * hi.y; lo.y += " << mv[0] << " * hi.x; "
                      << "hi.x += " << mv[2] << " * lo.y; hi.y -= " << mv[2] << " * lo.x; "
                      << "lo.x -= " << mv[0] << " * hi.y; lo.y += " << mv[0] << " * hi.x; "
                      << r(j) << " = lo; " << r(jh) << " = hi; }\n";
                    break;
                default:  // SHM_PHASE
                    scale_stmt(jh, mv[6], mv[7]);
                    break;
                }
            }
        } else {
            // diagonal on a thread bit (per lane) or a block bit (uniform)
            if (op.where == SHM_IN_THREAD)
                o << "      const bool hb = (t >> " << (int)op.tpos << ") & 1u;\n";
            else
                o << "      const bool hb = (b >> " << (int)op.tpos << ") & 1ull;\n";
            o << "      const R fr = hb ? " << mv[6] << " : " << mv[0] << ";\n";
            o << "      const R fi = hb ? " << mv[7] << " : " << mv[1] << ";\n";
            for (int j = 0; j < NR; j++) {
                if ((j & op.creg) != op.creg) continue;
                scale_stmt(j, "fr", "fi");
            }
        }
        if (blockc) o << "     }\n";
        o << "    }\n";
    }

    std::string emit()
    {
        const int nst = (int)P.stages.size();
        // HQ_G: global address space for pointers the batched form reads from
        // its item table (generic pointers would make every load per-lane)
        o << "#ifndef HQ_G\n#define HQ_G" << (batch ? " __attribute__((address_space(1)))" : "") << "\n#endif\n";
        o << "typedef " << (f32 ? "float" : "double") << " R;\n";
        o << "struct __attribute__((aligned(" << (f32 ? 8 : 16) << "))) C { R x, y; };\n";
        const int pp_ld = P.init ? -1 : pair_pos(P.stages[0]);
        const int pp_st = pair_pos(P.stages[nst - 1]);
        if (nt_ld || nt_st) o << "typedef R V2 __attribute__((ext_vector_type(2)));\n";
        if (pp_ld >= 0 || pp_st >= 0) o << "typedef float F4 __attribute__((ext_vector_type(4)));\n";
        if (nt_ld)
            o << "__device__ __forceinline__ C ld(const HQ_G C* p) { const V2 v = __builtin_nontemporal_load((const HQ_G V2*)p); "
                 "C c; c.x = v.x; c.y = v.y; return c; }\n";
        if (nt_st)
            o << "__device__ __forceinline__ void st(HQ_G C* p, const C c) { V2 v; v.x = c.x; v.y = c.y; "
                 "__builtin_nontemporal_store(v, (HQ_G V2*)p); }\n";
        // two waves per SIMD (two blocks per CU): a kernel that would need more
        // than 256 registers spills a few instead of halving the occupancy
        if (batch) {
            // same layout as the engine's BatchItem; `cf` is this group's
            // coefficient array, mode bit 1 = no store (the state is not needed)
            o << "struct BItem { unsigned long long psi, cf, part, first_block; unsigned n, mode; };\n";
            o << "extern \"C\" __global__ __launch_bounds__(" << NT << ") __attribute__((amdgpu_waves_per_eu(2)))"
              << " void hq_group_batch(const BItem* __restrict__ items, unsigned count)\n{\n";
        } else {
            o << "extern \"C\" __global__ __launch_bounds__(" << NT << ") __attribute__((amdgpu_waves_per_eu(2)))"
              << " void hq_group(HQ_G C* __restrict__ psi, const HQ_G R* __restrict__ cf)\n{\n";
        }
        if (nst > 1) o << "    __shared__ C lds[" << (1 << (split ? K - 1 : K)) << "];\n";
        o << "    const unsigned t = threadIdx.x;\n";
        if (batch) {
            o << "    unsigned lo_ = 0, hi_ = count;\n";
            o << "    const unsigned long long bid = blockIdx.x;\n";
            o << "    while (hi_ - lo_ > 1) { const unsigned mid = (lo_ + hi_) >> 1; "
                 "if (items[mid].first_block <= bid) lo_ = mid; else hi_ = mid; }\n";
            o << "    HQ_G C* __restrict__ psi = (HQ_G C*)items[lo_].psi;\n";
            o << "    const HQ_G R* __restrict__ cf = (const HQ_G R*)items[lo_].cf;\n";
            if (readout) o << "    HQ_G double* __restrict__ part = (HQ_G double*)items[lo_].part;\n";
            o << "    const unsigned mode = items[lo_].mode;\n";
            o << "    const unsigned long long b = bid - items[lo_].first_block;\n";
        } else if (bmap == 2)
            o << "    const unsigned long long b = (unsigned long long)(blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);\n";
        else if (bmap == 1)
            o << "    const unsigned long long b = __builtin_bitreverse64((unsigned long long)blockIdx.x) >> "
              << (64 - (int)P.hdr.num_block_bits) << ";\n";
        else
            o << "    const unsigned long long b = blockIdx.x;\n";
        o << "    unsigned long long base = 0;\n";
        for (uint32_t j = 0; j < P.hdr.num_block_bits; j++) {
            int p = ctz64(P.hdr.block_phys[j]);
            uint32_t len = 1;
            while (j + len < P.hdr.num_block_bits && ctz64(P.hdr.block_phys[j + len]) == p + (int)len) len++;
            o << "    base |= ((b >> " << j << ") & " << hex64((1ull << len) - 1) << ") << " << p << ";\n";
            j += len - 1;
        }
        // batched form, readout group: a tile that is zero contributes zero
        // partials (the per-circuit sum runs over every tile)
        const std::string zero_part = readout ? "    if (t == 0) { part[3 * b] = 0.0; part[3 * b + 1] = 0.0; part[3 * b + 2] = 0.0; }\n"
                                              : std::string();
        // batched form: mode bit 1 = the state is not needed afterwards (no store)
        const std::string store_if = batch ? "    if (!(mode & 2u)) {\n" : "    {\n";
        if (P.hdr.zero_mask) {
            // tiles known to be zero stay zero: nothing to load or store
            o << "    if (base & " << hex64(P.hdr.zero_mask) << ") {\n" << zero_part << "    return;\n    }\n";
        }
        if (P.init && P.hdr.num_block_bits > 0) {
            // first group of a run: every tile but the one holding |0...0>
            // starts and (the group's gates being linear) stays zero
            const ShmStage& st = P.stages[nst - 1];
            o << "    if (base != 0) {\n" << zero_part << store_if;
            lane_phys(st, "lz");
            o << "    HQ_G C* __restrict__ dz = psi + (base + lz);\n";
            o << "    C z; z.x = R(0); z.y = R(0);\n";
            for (int j = 0; j < NR; j++) {
                if (pp_st >= 0) {
                    if ((j >> pp_st) & 1) continue;
                    o << "    { F4 v; v.x = 0.f; v.y = 0.f; v.z = 0.f; v.w = 0.f; ";
                    if (nt_st) o << "__builtin_nontemporal_store(v, (HQ_G F4*)(dz + " << reg_off(st, j) << "ull)); }\n";
                    else o << "*(HQ_G F4*)(dz + " << reg_off(st, j) << "ull) = v; }\n";
                    continue;
                }
                if (nt_st) o << "    st(dz + " << reg_off(st, j) << "ull, z);\n";
                else o << "    dz[" << reg_off(st, j) << "ull] = z;\n";
            }
            o << "    }\n    return;\n    }\n";
        }
        o << "    C";
        for (int j = 0; j < NR; j++) o << (j ? ", " : " ") << "r" << j;
        if (split)
            for (int j = 0; j < NR; j++) o << ", q" << j;
        o << ";\n";
        {
            const ShmStage& st = P.stages[0];
            o << "    {\n";
            lane_phys(st, "lp");
            if (P.init) {
                // first group of a run: the tile of |0...0> is built in registers
                o << "    r0.x = (base + lp) == 0 ? R(1) : R(0); r0.y = R(0);\n";
                for (int j = 1; j < NR; j++) o << "    r" << j << ".x = R(0); r" << j << ".y = R(0);\n";
            } else {
                o << "    const HQ_G C* __restrict__ src = psi + (base + lp);\n";
                for (int j = 0; j < NR; j++) {
                    if (pp_ld >= 0) {
                        if ((j >> pp_ld) & 1) continue;
                        const int jh = j | (1 << pp_ld);
                        o << "    { const F4 v = ";
                        if (nt_ld) o << "__builtin_nontemporal_load((const HQ_G F4*)(src + " << reg_off(st, j) << "ull));";
                        else o << "*(const HQ_G F4*)(src + " << reg_off(st, j) << "ull);";
                        o << " r" << j << ".x = v.x; r" << j << ".y = v.y; r" << jh << ".x = v.z; r" << jh
                          << ".y = v.w; }\n";
                        continue;
                    }
                    if (nt_ld) o << "    r" << j << " = ld(src + " << reg_off(st, j) << "ull);\n";
                    else o << "    r" << j << " = src[" << reg_off(st, j) << "ull];\n";
                }
            }
            o << "    }\n";
        }
        for (int s = 0; s < nst; s++) {
            const ShmStage& st = P.stages[s];
            if (s > 0) {
                const ShmStage& prev = P.stages[s - 1];
                // every lane has read the previous exchange before it is overwritten
                if (s > 1) o << "    __syncthreads();\n";
                if (split) {
                    split_exchange(prev, st);
                    emit_stage_ops(st);
                    continue;
                }
                o << "    {\n";
                lane_lds(prev, "sw");
                for (int j = 0; j < NR; j++)
                    o << "    lds[sw ^ " << hex32(swz_of(reg_tile(prev, j), P.hdr, K)) << "] = " << r(j) << ";\n";
                o << "    }\n    __syncthreads();\n    {\n";
                lane_lds(st, "sr");
                for (int j = 0; j < NR; j++) name[j] = j;
                for (int j = 0; j < NR; j++)
                    o << "    r" << j << " = lds[sr ^ " << hex32(swz_of(reg_tile(st, j), P.hdr, K)) << "];\n";
                o << "    }\n";
            }
            emit_stage_ops(st);
        }
        const ShmStage& last = P.stages[nst - 1];
        if (readout) emit_readout(last);
        {
            o << store_if;
            lane_phys(last, "lq");
            o << "    HQ_G C* __restrict__ dst = psi + (base + lq);\n";
            for (int j = 0; j < NR; j++) {
                if (pp_st >= 0) {
                    if ((j >> pp_st) & 1) continue;
                    const int jh = j | (1 << pp_st);
                    o << "    { F4 v; v.x = " << r(j) << ".x; v.y = " << r(j) << ".y; v.z = " << r(jh) << ".x; v.w = "
                      << r(jh) << ".y; ";
                    if (nt_st) o << "__builtin_nontemporal_store(v, (HQ_G F4*)(dst + " << reg_off(last, j) << "ull)); }\n";
                    else o << "*(HQ_G F4*)(dst + " << reg_off(last, j) << "ull) = v; }\n";
                    continue;
                }
                if (nt_st) o << "    st(dst + " << reg_off(last, j) << "ull, " << r(j) << ");\n";
                else o << "    dst[" << reg_off(last, j) << "ull] = " << r(j) << ";\n";
            }
            o << "    }\n";
        }
        o << "}\n";
        return o.str();
    }

    // Readout of the batched form: per tile, sum p, sum p (number of |0>
    // qubits minus |1> qubits) and sum p (equal minus unequal neighbour
    // pairs) over the tile's amplitudes (p = |a|^2, FP64), then a fixed-order
    // tree over the block's threads; thread 0 writes the tile's three partials.
    void emit_readout(const ShmStage& st)
    {
        const int n = P.n;
        const uint64_t nn1 = n >= 2 ? ((1ull << (n - 1)) - 1) : 0;
        o << "    {\n";
        lane_phys(st, "lr");
        o << "    double s0 = 0.0, s1 = 0.0, s2 = 0.0;\n";
        for (int j = 0; j < NR; j++) {
            o << "    { const unsigned long long ix = base + lr + " << reg_off(st, j) << "ull; const double re = (double)"
              << r(j) << ".x, im = (double)" << r(j) << ".y; const double p = re * re + im * im; "
              << "s0 += p; s1 += p * (double)(" << n << " - 2 * (int)__builtin_popcountll(ix)); "
              << "s2 += p * (double)(" << (n - 1) << " - 2 * (int)__builtin_popcountll((ix ^ (ix >> 1)) & "
              << hex64(nn1) << ")); }\n";
        }
        o << "    __shared__ double red[" << 3 * NT << "];\n";
        o << "    red[t] = s0; red[" << NT << " + t] = s1; red[" << 2 * NT << " + t] = s2;\n";
        o << "    __syncthreads();\n";
        o << "    for (unsigned h = " << NT / 2 << "u; h > 0u; h >>= 1) {\n";
        o << "      if (t < h) { red[t] += red[t + h]; red[" << NT << " + t] += red[" << NT << " + t + h]; red[" << 2 * NT
          << " + t] += red[" << 2 * NT << " + t + h]; }\n";
        o << "      __syncthreads();\n    }\n";
        o << "    if (t == 0) { part[3 * b] = red[0]; part[3 * b + 1] = red[" << NT << "]; part[3 * b + 2] = red["
          << 2 * NT << "]; }\n";
        o << "    }\n";
    }
};

// ---- compile cache ---------------------------------------------------------------

uint64_t fnv1a(const std::string& s)
{
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
    return h;
}

std::mutex g_cache_mu;
std::unordered_map<std::string, std::vector<char>>& mem_cache()
{
    static std::unordered_map<std::string, std::vector<char>> m;
    return m;
}

std::string cache_dir()
{
    const char* e = getenv("HQ_JIT_CACHE");
    if (e && *e) return e;
    const char* x = getenv("XDG_CACHE_HOME");
    if (x && *x) return std::string(x) + "/hyquas_amd/jit";
    const char* h = getenv("HOME");
    if (h && *h) return std::string(h) + "/.cache/hyquas_amd/jit";
    return "";
}

void mkdirs(const std::string& path)
{
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

bool read_file(const std::string& p, std::string* out)
{
    std::ifstream f(p, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    *out = ss.str();
    return true;
}

const char* kSrcHeader = "// hyquas_amd group kernel v1\n";

// What besides the source decides the code object: the options and the
// hiprtc version.  Part of the disk cache key, so a flag or toolchain change
// never serves a code object built under other settings.
const std::string& build_tag()
{
    static const std::string tag = [] {
        std::string t;
        for (const char* o : kJitOpts) t += std::string(o) + " ";
        int major = 0, minor = 0;
        if (hiprtcVersion(&major, &minor) == HIPRTC_SUCCESS)
            t += "hiprtc " + std::to_string(major) + "." + std::to_string(minor);
        return t;
    }();
    return tag;
}

bool write_file(const std::string& p, const char* data, size_t n)
{
    std::ofstream f(p, std::ios::binary);
    if (!f) return false;
    f.write(data, (std::streamsize)n);
    f.close();
    return !f.fail();
}

// The compile helper: $HQ_RTC_HELPER, else hq_rtc in the library's folder;
// empty when there is none ("none" asks for in-process compiles).
const std::string& helper_path()
{
    static const std::string path = [] {
        const char* e = getenv("HQ_RTC_HELPER");
        if (e && *e) return strcmp(e, "none") == 0 ? std::string() : std::string(e);
        Dl_info info;
        if (dladdr((void*)&helper_path, &info) && info.dli_fname) {
            const std::string lib = info.dli_fname;
            const size_t slash = lib.rfind('/');
            const std::string h = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/hq_rtc";
            if (access(h.c_str(), X_OK) == 0) return h;
        }
        return std::string();
    }();
    return path;
}

std::atomic<bool> g_helper_failed{false};

// Compile `src` in a child process of the helper.  1: compiled; -1: the
// source does not compile (err holds the log); 0: no helper or it could not
// run (the caller compiles in this process).
int helper_compile(const std::string& src, std::vector<char>* code, std::string* err)
{
    const std::string& h = helper_path();
    if (h.empty() || g_helper_failed.load()) return 0;
    // sources and code objects go through a private directory (mode 0700,
    // made once per process under TMPDIR), removed at exit
    static const std::string dir = [] {
        const char* t = getenv("TMPDIR");
        std::string tmpl = std::string(t && *t ? t : "/tmp") + "/hq_rtc_XXXXXX";
        std::vector<char> buf(tmpl.begin(), tmpl.end());
        buf.push_back('\0');
        if (!mkdtemp(buf.data())) return std::string();
        static std::string made;
        made = buf.data();
        atexit([] { (void)rmdir(made.c_str()); });
        return made;
    }();
    if (dir.empty()) return 0;
    static std::atomic<unsigned long long> counter{0};
    const std::string base = dir + "/k" + std::to_string(counter++);
    const std::string fin = base + ".hip", fout = base + ".co", flog = base + ".log";
    if (!write_file(fin, src.data(), src.size())) return 0;
    auto cleanup = [&] {
        (void)remove(fin.c_str());
        (void)remove(fout.c_str());
        (void)remove(flog.c_str());
    };
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_addclosefrom_np(&fa, 3);   // the child inherits no device or file handles
    std::vector<char*> argv = {const_cast<char*>(h.c_str()), const_cast<char*>(fin.c_str()),
                               const_cast<char*>(fout.c_str()), const_cast<char*>(flog.c_str()), nullptr};
    pid_t pid = -1;
    const int rc = posix_spawn(&pid, h.c_str(), &fa, nullptr, argv.data(), environ);
    posix_spawn_file_actions_destroy(&fa);
    if (rc != 0) {
        cleanup();
        if (!g_helper_failed.exchange(true))
            fprintf(stderr, "hyquas_amd: compile helper %s could not start (%s); compiling in-process\n", h.c_str(),
                    strerror(rc));
        return 0;
    }
    int status = 0;
    pid_t w;
    do { w = waitpid(pid, &status, 0); } while (w < 0 && errno == EINTR);
    const int code_rc = (w == pid && WIFEXITED(status)) ? WEXITSTATUS(status) : -1;
    if (code_rc == 0) {
        std::string c;
        if (read_file(fout, &c) && !c.empty()) {
            code->assign(c.begin(), c.end());
            cleanup();
            return 1;
        }
    } else if (code_rc == 1) {
        std::string log;
        (void)read_file(flog, &log);
        *err = "hiprtc: " + log.substr(0, 2000);
        cleanup();
        return -1;
    }
    cleanup();
    if (!g_helper_failed.exchange(true))
        fprintf(stderr, "hyquas_amd: compile helper %s failed (status %d); compiling in-process\n", h.c_str(),
                code_rc);
    return 0;
}

hq_status compile_source(const std::string& src, std::vector<char>* code, std::string* err)
{
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = mem_cache().find(src);
        if (it != mem_cache().end()) { *code = it->second; return HQ_OK; }
    }
    const std::string dir = cache_dir();
    char key[64];
    snprintf(key, sizeof key, "%016llx_%016llx_%zu", (unsigned long long)fnv1a(build_tag()),
             (unsigned long long)fnv1a(src), src.size());
    if (!dir.empty()) {
        std::string s2, c2;
        if (read_file(dir + "/" + key + ".hip", &s2) && s2 == src && read_file(dir + "/" + key + ".co", &c2) &&
            !c2.empty()) {
            code->assign(c2.begin(), c2.end());
            std::lock_guard<std::mutex> lk(g_cache_mu);
            mem_cache()[src] = *code;
            return HQ_OK;
        }
    }
    const int via = helper_compile(src, code, err);
    if (via < 0) return HQ_ERR_HIP;
    if (via == 0) {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), "hq_group.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
            *err = "hiprtcCreateProgram failed";
            return HQ_ERR_HIP;
        }
        hiprtcResult rc = hiprtcCompileProgram(prog, kNumJitOpts, kJitOpts);
        if (rc != HIPRTC_SUCCESS) {
            size_t n = 0;
            hiprtcGetProgramLogSize(prog, &n);
            std::string log(n, '\0');
            if (n) hiprtcGetProgramLog(prog, &log[0]);
            *err = std::string("hiprtc: ") + hiprtcGetErrorString(rc) + "\n" + log.substr(0, 2000);
            hiprtcDestroyProgram(&prog);
            return HQ_ERR_HIP;
        }
        size_t sz = 0;
        hiprtcGetCodeSize(prog, &sz);
        code->resize(sz);
        hiprtcGetCode(prog, code->data());
        hiprtcDestroyProgram(&prog);
    }
    if (!dir.empty()) {
        // write both files under temporary names, then rename; a failed or
        // short write leaves nothing behind (the disk cache is best effort)
        mkdirs(dir);
        std::string tmp = dir + "/" + key + ".tmp" + std::to_string((unsigned long long)(uintptr_t)code);
        const bool ok = write_file(tmp + ".co", code->data(), code->size()) &&
                        write_file(tmp + ".hip", src.data(), src.size());
        if (ok && rename((tmp + ".co").c_str(), (dir + "/" + key + ".co").c_str()) == 0) {
            if (rename((tmp + ".hip").c_str(), (dir + "/" + key + ".hip").c_str()) != 0)
                (void)remove((tmp + ".hip").c_str());
        } else {
            (void)remove((tmp + ".co").c_str());
            (void)remove((tmp + ".hip").c_str());
        }
    }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    mem_cache()[src] = *code;
    return HQ_OK;
}

}  // namespace

// Source of the specialised kernel; empty (and the error set) if the plan
// cannot be emitted.
std::string jit_source(const ShmPlan& plan, std::vector<double>* coef, bool batch, bool readout)
{
    try {
        Emitter e(plan, coef, batch, readout);
        return std::string(kSrcHeader) + e.emit();
    } catch (const std::exception& ex) {
        set_error(std::string("specialised kernel: ") + ex.what());
        return std::string();
    }
}

static hq_status load_program(const ShmPlan& plan, const std::vector<char>& code, const std::vector<double>& coef,
                              hq_program p)
{
    // a program loaded before (a retried batch) gives its module and coefficients back first
    if (p->jit_module) { (void)hipModuleUnload(p->jit_module); p->jit_module = nullptr; p->jit_fn = nullptr; }
    if (p->d_coef) { (void)hipFree(p->d_coef); p->d_coef = nullptr; }
    HQ_HIP(hipModuleLoadData(&p->jit_module, code.data()));
    HQ_HIP(hipModuleGetFunction(&p->jit_fn, p->jit_module, "hq_group"));
    size_t rbytes = plan.precision == HQ_C128 ? 8 : 4;
    size_t n = std::max<size_t>(coef.size(), 1);
    HQ_HIP(hipMalloc(&p->d_coef, n * rbytes));
    if (!coef.empty()) {
        if (rbytes == 8) {
            HQ_HIP(hipMemcpy(p->d_coef, coef.data(), coef.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(coef.begin(), coef.end());
            HQ_HIP(hipMemcpy(p->d_coef, f.data(), f.size() * 4, hipMemcpyHostToDevice));
        }
    }
    return HQ_OK;
}

hq_status jit_compile_load(const std::string& src, const char* name, hipModule_t* mod, hipFunction_t* fn)
{
    *mod = nullptr;
    *fn = nullptr;
    std::vector<char> code;
    std::string err;
    hq_status st = compile_source(src, &code, &err);
    if (st != HQ_OK) { set_error(err); return st; }
    HQ_HIP(hipModuleLoadData(mod, code.data()));
    hipError_t e = hipModuleGetFunction(fn, *mod, name);
    if (e != hipSuccess) {
        (void)hipModuleUnload(*mod);
        *mod = nullptr;
        set_error(std::string("hipModuleGetFunction(") + name + "): " + hipGetErrorString(e));
        return HQ_ERR_HIP;
    }
    return HQ_OK;
}

hq_status jit_sharemem(const ShmPlan& plan, hq_program p)
{
    std::vector<double> coef;
    std::string src = jit_source(plan, &coef);
    if (src.empty()) return HQ_ERR_UNSUPPORTED;
    std::vector<char> code;
    std::string err;
    hq_status st = compile_source(src, &code, &err);
    if (st != HQ_OK) { set_error(err); return st; }
    return load_program(plan, code, coef, p);
}

// Compile many plans in parallel (child processes of the compile helper;
// in this process hiprtc runs them one at a time).
hq_status jit_sharemem_many(const std::vector<const ShmPlan*>& plans, const std::vector<hq_program>& progs,
                            int threads)
{
    const size_t n = plans.size();
    std::vector<std::string> srcs(n);
    std::vector<std::vector<double>> coefs(n);
    for (size_t i = 0; i < n; i++) {
        srcs[i] = jit_source(*plans[i], &coefs[i]);
        if (srcs[i].empty()) return HQ_ERR_UNSUPPORTED;
    }
    std::vector<std::vector<char>> codes(n);
    std::vector<hq_status> sts(n, HQ_OK);
    std::vector<std::string> errs(n);
    // in-process compiles: warm the compiler once before fanning out
    const bool warm = helper_path().empty();
    if (n > 0 && warm) sts[0] = compile_source(srcs[0], &codes[0], &errs[0]);
    std::mutex mu;
    size_t next = warm ? 1 : 0;
    auto worker = [&]() {
        while (true) {
            size_t i;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (next >= n) return;
                i = next++;
            }
            sts[i] = compile_source(srcs[i], &codes[i], &errs[i]);
        }
    };
    const int nt = std::max(1, std::min<int>(threads, (int)(n - next)));
    std::vector<std::thread> pool;
    for (int i = 0; i < nt && next < n; i++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    for (size_t i = 0; i < n; i++) {
        if (sts[i] != HQ_OK) { set_error(errs[i]); return sts[i]; }
        hq_status st = load_program(*plans[i], codes[i], coefs[i], progs[i]);
        if (st != HQ_OK) return st;
    }
    return HQ_OK;
}

hq_status launch_sharemem_jit(hq_state s, hq_program p)
{
    void* psi = s->data;
    void* cf = p->d_coef;
    void* args[] = {&psi, &cf};
    unsigned blocks = (unsigned)((uint64_t)1 << (p->n - p->tile));
    HQ_HIP(hipModuleLaunchKernel(p->jit_fn, blocks, 1, 1, 1u << p->tb, 1, 1, 0, s->stream, args, nullptr));
    return HQ_OK;
}

}  // namespace hq

extern "C" hq_status hq_sharemem_source(int n, int precision, const hq_gate* gates, int num_gates,
                                        uint64_t shared_mask, int reg_qubits, int variant, char* buf, size_t cap,
                                        size_t* len, double* coef, size_t coef_cap, size_t* num_coef, int* shape)
{
    HQ_CHECK(n >= 1 && n <= 36, HQ_ERR_ARG, "hq_sharemem_source: num_qubits must be in [1, 36]");
    HQ_CHECK(precision == HQ_C128 || precision == HQ_C64, HQ_ERR_ARG, "hq_sharemem_source: bad precision");
    HQ_CHECK(variant >= 0 && variant <= 63, HQ_ERR_ARG, "hq_sharemem_source: variant must be in [0, 63]");
    const bool whole_line = variant & 1, init = (variant & 2) != 0, batch = (variant & 4) != 0,
               readout = (variant & 8) != 0, shared_code = (variant & 16) != 0, nt_store = (variant & 32) != 0;
    HQ_CHECK(batch || !readout, HQ_ERR_ARG, "hq_sharemem_source: the readout fold needs the batched form");
    hq::ShmPlan plan;
    hq_status st = hq::plan_sharemem(n, precision, gates, num_gates, shared_mask, reg_qubits, &plan, init,
                                     whole_line ? 3 : 0, shared_code);
    if (st != HQ_OK) return st;
    // bit 0 alone: whole-line stages with non-temporal loads and stores; with
    // bit 5: whole-line stages, non-temporal stores only; bit 5 alone: the
    // plain stages with non-temporal stores
    plan.whole_line = whole_line;
    plan.nt_load = whole_line && !nt_store;
    plan.nt_store = whole_line || nt_store;
    plan.init = init;
    std::vector<double> cf;
    std::string src = hq::jit_source(plan, &cf, batch, readout);
    if (src.empty()) return HQ_ERR_UNSUPPORTED;
    if (len) *len = src.size();
    if (buf && cap) {
        size_t k = std::min(src.size(), cap - 1);
        memcpy(buf, src.data(), k);
        buf[k] = 0;
    }
    if (num_coef) *num_coef = cf.size();
    if (coef)
        for (size_t i = 0; i < cf.size() && i < coef_cap; i++) coef[i] = cf[i];
    if (shape) { shape[0] = plan.k; shape[1] = plan.rb; shape[2] = plan.tb; }
    return HQ_OK;
}
