// sharemem.hip -- OShareMem: one HBM pass applies a whole gate group.
//
// Reference behaviour (src/kernelOpt.cu:388-433, src/executor.cpp:462-493,
// 585-632): each block owns the 2^k amplitudes that differ only in the k
// "shared" qubits, loads them, applies every gate of the group, stores them.
// Gates whose target is outside the tile must be diagonal; controls and
// diagonal targets may sit on block bits.
//
// MI355X design (not the reference's): the tile lives in REGISTERS, not in
// shared memory.  Each thread holds 2^RB amplitudes; the tile bits are split
// per stage into RB register bits and TB thread bits (k = RB + TB).  Every
// non-diagonal gate of a stage targets a register bit, so it is a pure
// register update; diagonal gates and controls on thread or block bits are
// per-lane / per-block predicates.  Between stages the tile is re-split by
// one LDS exchange (write, barrier, read).  The first stage reads HBM directly
// and the last writes HBM directly, so a group with s stages costs one HBM
// read + write and s-1 LDS round trips instead of one LDS round trip per gate.
// LDS addresses are XOR-swizzled (tile bits >= 3 fold into bits 0..2) and
// the host picks lane bits 0..2 so that 8-lane groups hit distinct 16-byte
// bank groups.

#include "sharemem_dev.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>

namespace hq {

// ---- device side ----------------------------------------------------------------------

template <typename R, int RB, int TB>
__global__ __launch_bounds__(1 << TB) void k_sharemem(cplx<R>* __restrict__ psi, const char* __restrict__ prog,
                                                     uint32_t mode)
{
    __shared__ cplx<R> lds[1 << (RB + TB)];
    tile_pass<R, RB, TB>(psi, prog, blockIdx.x, lds, mode, 0, nullptr);
}

// ---- launch table -------------------------------------------------------------------------

template <typename R, int RB, int TB>
static hq_status launch_variant(hq_state s, hq_program p)
{
    uint64_t blocks = (uint64_t)1 << (p->n - p->tile);
    const uint32_t mode = p->plan && p->plan->init ? TILE_INIT : 0u;
    hipLaunchKernelGGL((k_sharemem<R, RB, TB>), dim3((unsigned)blocks), dim3(1 << TB), 0, s->stream,
                       (cplx<R>*)s->data, (const char*)p->d_prog, mode);
    HQ_HIP(hipGetLastError());
    return HQ_OK;
}

template <typename R>
static hq_status launch_by_shape(hq_state s, hq_program p)
{
    const int rb = p->rb, tb = p->tb;
#define HQ_SHM_CASE(RBV, TBV) \
    if (rb == RBV && tb == TBV) return launch_variant<R, RBV, TBV>(s, p);
    HQ_SHM_CASE(1, 6) HQ_SHM_CASE(2, 6) HQ_SHM_CASE(3, 6) HQ_SHM_CASE(4, 6) HQ_SHM_CASE(5, 6)
    HQ_SHM_CASE(2, 7) HQ_SHM_CASE(3, 7) HQ_SHM_CASE(4, 7) HQ_SHM_CASE(5, 7)
    HQ_SHM_CASE(3, 8) HQ_SHM_CASE(4, 8) HQ_SHM_CASE(5, 8) HQ_SHM_CASE(3, 9) HQ_SHM_CASE(4, 9)
    if constexpr (sizeof(R) == 4) { HQ_SHM_CASE(6, 8) }   // 14-qubit complex64 tiles
#undef HQ_SHM_CASE
    set_error("sharemem: no kernel for this register/thread split");
    return HQ_ERR_UNSUPPORTED;
}

static bool shape_supported(int rb, int tb, int precision)
{
    static const int shapes[][2] = {{1, 6}, {2, 6}, {3, 6}, {4, 6}, {5, 6}, {2, 7}, {3, 7},
                                    {4, 7}, {5, 7}, {3, 8}, {4, 8}, {5, 8}, {3, 9}, {4, 9}};
    for (auto& sh : shapes)
        if (sh[0] == rb && sh[1] == tb) return true;
    return precision == HQ_C64 && rb == 6 && tb == 8;
}

hq_status launch_sharemem(hq_state s, hq_program p)
{
    return s->precision == HQ_C128 ? launch_by_shape<double>(s, p) : launch_by_shape<float>(s, p);
}

int sharemem_default_rb(int precision, int tile)
{
    // tile = RB + TB; 256-thread blocks with 16 (complex128) or 32 (complex64)
    // amplitudes per thread at the default 12/13-qubit tiles
    if (tile <= 10) return std::max(1, tile - 6);
    if (precision == HQ_C128) return tile >= 13 ? 5 : 4;
    // complex64: 256 threads x 16 amplitudes at 12 qubits measured fastest;
    // 14-qubit tiles hold 64 amplitudes per thread (the register bytes of a
    // 13-qubit complex128 tile)
    return tile >= 14 ? 6 : tile == 13 ? 5 : tile == 12 ? 4 : tile - 6;
}

// ---- host side: lowering + register-stage planning -------------------------------------------

namespace {

struct Lowered {
    int cls;
    bool needs_reg;       // non-diagonal update: target must be a register bit
    int t_tile, t_block;  // target as tile bit or block bit (-1 if not)
    int c_tile[2], c_block[2];
    double m[8];
};

// Rotation [[c, -s], [s, c]] (c >= 0 after the sign is moved into the
// scalar) as three shears: [[1, a], [0, 1]] [[1, 0], [b, 1]] [[1, a], [0, 1]]
// with a = -tan(phi / 2) = -s / (1 + c), b = s = sin(phi); |a|, |b| <= 1 and
// no division by a small number, for every angle.
void shear_form(double c, double s, double* a, double* b, std::complex<double>* S)
{
    const double r = std::hypot(c, s);
    c /= r;
    s /= r;
    double sigma = 1.0;
    if (c < 0) { c = -c; s = -s; sigma = -1.0; }
    *S *= sigma * r;
    *a = -s / (1.0 + c);
    *b = s;
}

// Factor a scalar out of an uncontrolled gate so that its pair update needs
// half the multiplies: H = r [[1, 1], [1, -1]]; a rotation [[c, -s], [s, c]]
// = c [[1, -t], [t, 1]] (t = s/c) or s [[u, -1], [1, u]] (u = c/s), whichever
// keeps |t|, |u| <= 1; RX-type [[a, ib], [ib, a]] likewise; a unit diagonal
// diag(d0, d1) = d0 diag(1, d1/d0).  Other gates are left as they are.
// True when the gate was rewritten (its scalar multiplied into S).
bool factor_scalar(Lowered& L, std::complex<double>* S, bool shared_code)
{
    double* m = L.m;
    // shared_code: only factorings whose op class does not depend on the
    // angle, so groups that differ only in angles emit the same specialised
    // kernel source: H, unit diagonals, and rotations as three shears
    if (shared_code && L.cls == SHM_REAL && m[0] == m[6] && m[2] == -m[4] && (m[0] != 0 || m[4] != 0)) {
        double a, b;
        shear_form(m[0], m[4], &a, &b, S);
        L.cls = SHM_SHEAR;
        memset(m, 0, sizeof(L.m));
        m[0] = a;
        m[2] = b;
        return true;
    }
    if (shared_code && L.cls == SHM_RX && m[0] == m[6] && m[3] == m[5] && (m[0] != 0 || m[3] != 0)) {
        // [[c, -i s], [-i s, c]] = diag(1, -i) [[c, -s], [s, c]] diag(1, i)
        double a, b;
        shear_form(m[0], -m[3], &a, &b, S);
        L.cls = SHM_SHEARX;
        memset(m, 0, sizeof(L.m));
        m[0] = a;
        m[2] = b;
        return true;
    }
    if (shared_code && (L.cls == SHM_RX || (L.cls == SHM_REAL && !(m[0] == m[2] && m[0] == m[4] && m[0] == -m[6]))))
        return false;
    const int cls_in = L.cls;
    if (L.cls == SHM_REAL) {
        if (m[0] == m[2] && m[0] == m[4] && m[0] == -m[6] && m[0] != 0) {
            *S *= m[0];
            L.cls = SHM_HADD;
            memset(m, 0, sizeof(L.m));
        } else if (m[0] == m[6] && m[2] == -m[4] && (m[0] != 0 || m[4] != 0)) {
            const double c = m[0], s = m[4];
            if (std::fabs(c) >= std::fabs(s)) { *S *= c; L.cls = SHM_ROTA; memset(m, 0, sizeof(L.m)); m[0] = s / c; }
            else { *S *= s; L.cls = SHM_ROTB; memset(m, 0, sizeof(L.m)); m[0] = c / s; }
        }
    } else if (L.cls == SHM_RX) {
        // lo' = a lo + i b hi, hi' = i c lo + d hi with a = d, b = c
        const double a = m[0], b = m[3];
        if (a == m[6] && b == m[5] && (a != 0 || b != 0)) {
            if (std::fabs(a) >= std::fabs(b)) { *S *= a; L.cls = SHM_RXA; memset(m, 0, sizeof(L.m)); m[0] = b / a; }
            else { *S *= b; L.cls = SHM_RXB; memset(m, 0, sizeof(L.m)); m[0] = a / b; }
        }
    } else if (L.cls == SHM_DIAG) {
        const std::complex<double> d0(m[0], m[1]), d1(m[6], m[7]);
        if (std::fabs(std::abs(d0) - 1.0) < 1e-12) {
            const std::complex<double> q = d1 / d0;
            *S *= d0;
            L.cls = SHM_PHASE;
            memset(m, 0, sizeof(L.m));
            m[0] = 1.0;
            m[6] = q.real();
            m[7] = q.imag();
        }
    }
    return L.cls != cls_in;
}

// swizzle column of tile bit i (bits >= 3 fold into the low 3 bits)
uint32_t swz_column(int i)
{
    static const uint32_t cyc[7] = {3, 5, 6, 7, 1, 2, 4};
    return i < 3 ? (1u << i) : cyc[(i - 3) % 7];
}

// GF(2) rank of a set of 3-bit vectors
int rank3(const std::vector<uint32_t>& v)
{
    uint32_t basis[3] = {0, 0, 0};
    int r = 0;
    for (uint32_t x : v) {
        for (int b = 2; b >= 0; b--) {
            if (!((x >> b) & 1)) continue;
            if (!basis[b]) { basis[b] = x; r++; break; }
            x ^= basis[b];
        }
    }
    return r;
}

}  // namespace

hq_status plan_sharemem(int n, int precision, const hq_gate* gates, int num_gates, uint64_t shared_mask, int rb_req,
                        ShmPlan* plan, bool init, int io, bool shared_code)
{
    HQ_CHECK(num_gates >= 0 && (num_gates == 0 || gates), HQ_ERR_ARG, "sharemem: bad gate array");
    const size_t amp_bytes = precision == HQ_C128 ? 16 : 8;
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    HQ_CHECK((shared_mask & ~all) == 0, HQ_ERR_ARG, "sharemem: shared mask has bits beyond the state");
    const int k = __builtin_popcountll(shared_mask);
    HQ_CHECK(k >= 7 && k <= SHM_MAX_TILE, HQ_ERR_ARG, "sharemem: tile must have 7..14 qubits");
    HQ_CHECK(((size_t)1 << k) * amp_bytes <= 160 * 1024, HQ_ERR_ARG, "sharemem: tile exceeds the 160 KiB LDS");
    const int rb = rb_req > 0 ? rb_req : sharemem_default_rb(precision, k);
    const int tb = k - rb;
    HQ_CHECK(shape_supported(rb, tb, precision), HQ_ERR_UNSUPPORTED, "sharemem: unsupported register/thread split");

    int tile_of[64], block_of[64];
    std::vector<int> tile_pos, block_pos;
    for (int q = 0; q < n; q++) {
        if ((shared_mask >> q) & 1) { tile_of[q] = (int)tile_pos.size(); block_of[q] = -1; tile_pos.push_back(q); }
        else { block_of[q] = (int)block_pos.size(); tile_of[q] = -1; block_pos.push_back(q); }
    }

    std::vector<Lowered> low;
    std::vector<int> low_type;
    low.reserve(num_gates);
    // uncontrolled H, RX, RY and unit diagonals lose a scalar factor; the
    // product S is applied once at the end of the pass (HQ_SHM_FACTOR=0: off)
    static const bool factor = !getenv("HQ_SHM_FACTOR") || atoi(getenv("HQ_SHM_FACTOR")) != 0;
    std::complex<double> S(1.0, 0.0);
    bool factored = false;
    for (int i = 0; i < num_gates; i++) {
        const hq_gate& g = gates[i];
        hq_status st = validate_gate(g, n);
        if (st != HQ_OK) return st;
        if (g.type == HQ_ID) continue;
        Lowered L;
        L.cls = classify_gate(g);
        L.needs_reg = !(L.cls == SHM_DIAG || L.cls == SHM_PHASE);
        L.t_tile = tile_of[g.target];
        L.t_block = block_of[g.target];
        if (L.needs_reg && L.t_tile < 0) {
            set_error("sharemem: gate " + std::to_string(i) + " is not diagonal but its target is outside the tile");
            return HQ_ERR_ARG;
        }
        int cs[2] = {g.control, g.control2};
        for (int c = 0; c < 2; c++) {
            L.c_tile[c] = cs[c] >= 0 ? tile_of[cs[c]] : -1;
            L.c_block[c] = cs[c] >= 0 ? block_of[cs[c]] : -1;
        }
        memcpy(L.m, g.mat, sizeof(L.m));
        if (factor && g.control < 0 && g.control2 < 0 && factor_scalar(L, &S, shared_code)) factored = true;
        low.push_back(L);
        low_type.push_back(g.type);
    }

    // greedy stage planning: take gates in order; a stage has at most RB
    // distinct register targets; a gate that cannot join blocks its qubits
    struct Stage {
        std::vector<int> reg_bits, thr_bits;
        std::vector<int> ops;
    };
    std::vector<Stage> stages;
    // thread bits: lanes 0..2 get three bits with independent swizzle
    // columns (lowest first), the rest follow in ascending order
    auto assign_lanes = [&](Stage& S, uint32_t regset) {
        std::vector<int> free_bits;
        for (int b = 0; b < k; b++)
            if (!((regset >> b) & 1)) free_bits.push_back(b);
        std::vector<int> lane3;
        std::vector<uint32_t> cols;
        for (int b : free_bits) {
            if (lane3.size() == 3) break;
            cols.push_back(swz_column(b));
            if (rank3(cols) == (int)cols.size()) lane3.push_back(b);
            else cols.pop_back();
        }
        for (int b : lane3) S.thr_bits.push_back(b);
        for (int b : free_bits)
            if (std::find(lane3.begin(), lane3.end(), b) == lane3.end()) S.thr_bits.push_back(b);
    };
    std::vector<int> pending(low.size());
    for (size_t i = 0; i < low.size(); i++) pending[i] = (int)i;
    // a tile above 64 KiB is exchanged in two LDS halves by the specialised
    // kernel, split on a register bit that consecutive stages share: keep one
    const bool split_lds = ((size_t)1 << k) * amp_bytes > 64 * 1024;
    uint32_t prev_reg = 0;
    do {
        Stage S;
        uint32_t regset = 0;
        int cnt = 0;
        uint32_t blk_tile = 0;
        uint64_t blk_block = 0;
        std::vector<int> rest;
        for (int idx : pending) {
            const Lowered& L = low[idx];
            uint32_t tt = L.t_tile >= 0 ? 1u << L.t_tile : 0;
            uint64_t tbk = L.t_block >= 0 ? 1ull << L.t_block : 0;
            for (int c = 0; c < 2; c++) {
                if (L.c_tile[c] >= 0) tt |= 1u << L.c_tile[c];
                if (L.c_block[c] >= 0) tbk |= 1ull << L.c_block[c];
            }
            bool ok = !(tt & blk_tile) && !(tbk & blk_block);
            if (ok && L.needs_reg) {
                uint32_t tb1 = 1u << L.t_tile;
                if (!(regset & tb1)) {
                    const bool keep = split_lds && !stages.empty() && !((regset | tb1) & prev_reg);
                    if (cnt < (keep ? rb - 1 : rb)) { regset |= tb1; cnt++; }
                    else ok = false;
                }
            }
            if (ok) S.ops.push_back(idx);
            else { blk_tile |= tt; blk_block |= tbk; rest.push_back(idx); }
        }
        if (split_lds && !stages.empty() && !(regset & prev_reg) && cnt < rb) {
            for (int b = k - 1; b >= 0; b--)
                if ((prev_reg >> b) & 1) { regset |= 1u << b; cnt++; break; }
        }
        // fill the register set with the highest remaining tile bits, so the
        // lanes cover the low (contiguous) tile bits
        for (int b = k - 1; b >= 0 && cnt < rb; b--)
            if (!((regset >> b) & 1)) { regset |= 1u << b; cnt++; }
        for (int b = 0; b < k; b++)
            if ((regset >> b) & 1) S.reg_bits.push_back(b);
        assign_lanes(S, regset);
        prev_reg = regset;
        stages.push_back(std::move(S));
        pending.swap(rest);
    } while (!pending.empty());

    // HBM-facing stages: the first stage's lanes load the tile and the last
    // stage's lanes store it.  When one of them holds one of the lowest tile
    // bits of a 128-byte line (3 for complex128, 4 for complex64) in
    // registers, a wave's access covers each line in pieces; an added stage
    // without gates (register bits = highest tile bits, plus the neighbour's
    // highest register bit for split exchanges) moves the HBM access to
    // whole-line lanes at the cost of one LDS exchange.  complex64 tiles that
    // hold the lowest state qubit keep it as a register bit in such a stage:
    // a lane then moves two neighbouring amplitudes as one 16-byte access and
    // 8 lanes cover a 128-byte line.
    const bool pair64 = amp_bytes == 8 && tile_pos[0] == 0;
    auto short_runs = [&](const Stage& st) {
        uint32_t regs = 0;
        for (int b : st.reg_bits) regs |= 1u << b;
        return (regs & ((1u << (amp_bytes == 16 ? 3 : 4)) - 1)) != 0;
    };
    auto make_io = [&](const Stage& nb) {
        Stage S;
        uint32_t regset = 0;
        int cnt = 0;
        if (pair64) { regset |= 1u; cnt++; }
        if (split_lds && !((regset >> nb.reg_bits.back()) & 1)) { regset |= 1u << nb.reg_bits.back(); cnt++; }
        for (int b = k - 1; b >= 0 && cnt < rb; b--)
            if (!((regset >> b) & 1)) { regset |= 1u << b; cnt++; }
        for (int b = 0; b < k; b++)
            if ((regset >> b) & 1) S.reg_bits.push_back(b);
        assign_lanes(S, regset);
        return S;
    };
    if ((io & 1) && !init && short_runs(stages.front())) stages.insert(stages.begin(), make_io(stages.front()));
    if ((io & 2) && short_runs(stages.back())) stages.push_back(make_io(stages.back()));

    // serialize
    std::vector<ShmOp> ops;
    std::vector<ShmStage> dst;
    for (auto& S : stages) {
        ShmStage d;
        memset(&d, 0, sizeof(d));
        int reg_index[SHM_MAX_TILE], thr_index[SHM_MAX_TILE];
        for (int i = 0; i < SHM_MAX_TILE; i++) reg_index[i] = thr_index[i] = -1;
        for (int i = 0; i < rb; i++) { d.reg_bits[i] = S.reg_bits[i]; reg_index[S.reg_bits[i]] = i; }
        for (int i = 0; i < tb; i++) { d.thr_bits[i] = S.thr_bits[i]; thr_index[S.thr_bits[i]] = i; }
        d.op_begin = (uint32_t)ops.size();
        for (int idx : S.ops) {
            const Lowered& L = low[idx];
            ShmOp o;
            memset(&o, 0, sizeof(o));
            o.cls = (uint8_t)L.cls;
            if (L.t_tile >= 0 && reg_index[L.t_tile] >= 0) { o.where = SHM_IN_REG; o.tpos = (uint8_t)reg_index[L.t_tile]; }
            else if (L.t_tile >= 0) { o.where = SHM_IN_THREAD; o.tpos = (uint8_t)thr_index[L.t_tile]; }
            else { o.where = SHM_IN_BLOCK; o.tpos = (uint8_t)L.t_block; }
            for (int c = 0; c < 2; c++) {
                if (L.c_tile[c] >= 0) {
                    if (reg_index[L.c_tile[c]] >= 0) o.creg |= 1u << reg_index[L.c_tile[c]];
                    else o.cthr |= 1u << thr_index[L.c_tile[c]];
                } else if (L.c_block[c] >= 0) {
                    o.cblk |= 1ull << L.c_block[c];
                }
            }
            memcpy(o.m, L.m, sizeof(o.m));
            // a thread-bit control makes the op the identity on some lanes:
            // use an update form that can express the identity
            if (o.cthr && o.cls == SHM_SWAP) o.cls = SHM_REAL;
            if (o.cthr && o.cls == SHM_Y) o.cls = SHM_RX;
            switch (low_type[idx]) {
            case HQ_CCX: case HQ_CNOT: case HQ_CY: case HQ_CZ: case HQ_H: case HQ_X: case HQ_Y: case HQ_Z:
            case HQ_S: case HQ_SDG: case HQ_T: case HQ_TDG: case HQ_ID: case HQ_GII: case HQ_GZZ:
                o.flags |= SHM_FLAG_CONST;
                break;
            default:
                break;
            }
            ops.push_back(o);
        }
        d.op_end = (uint32_t)ops.size();
        dst.push_back(d);
    }
    // shared code: the scale op whenever a gate was factored (a structural
    // fact), even where the angles make the product exactly 1
    if (S != std::complex<double>(1.0, 0.0) || (shared_code && factored)) {
        ShmOp o;
        memset(&o, 0, sizeof(o));
        o.cls = SHM_SCALE;
        o.where = SHM_IN_REG;
        o.m[0] = S.real();
        o.m[1] = S.imag();
        ops.push_back(o);
        dst.back().op_end = (uint32_t)ops.size();
    }

    ShmHeader hdr;
    memset(&hdr, 0, sizeof(hdr));
    for (size_t j = 0; j < block_pos.size(); j++) hdr.block_phys[j] = 1ull << block_pos[j];
    for (int i = 0; i < k; i++) { hdr.tile_phys[i] = 1ull << tile_pos[i]; hdr.swz_col[i] = swz_column(i); }
    hdr.num_block_bits = (uint32_t)block_pos.size();
    hdr.num_stages = (uint32_t)dst.size();
    hdr.num_ops = (uint32_t)ops.size();

    plan->n = n;
    plan->precision = precision;
    plan->k = k;
    plan->rb = rb;
    plan->tb = tb;
    plan->shared_mask = shared_mask;
    plan->hdr = hdr;
    plan->stages = std::move(dst);
    plan->ops = std::move(ops);
    return HQ_OK;
}

hq_status build_sharemem_program(hq_state s, const hq_gate* gates, int num_gates, uint64_t shared_mask, int rb_req,
                                 hq_program* out, bool init, int variant, uint64_t zero_mask)
{
    HQ_CHECK(s && out, HQ_ERR_ARG, "sharemem: null argument");
    *out = nullptr;
    ShmPlan plan;
    // variant bit 0: whole-line stages; bit 1: non-temporal stores; bit 2:
    // non-temporal loads
    const bool whole_line = variant & 1;
    hq_status st = plan_sharemem(s->n, s->precision, gates, num_gates, shared_mask, rb_req, &plan, init,
                                 whole_line ? 3 : 0);
    if (st != HQ_OK) return st;
    plan.init = init;
    plan.whole_line = whole_line;
    plan.nt_store = (variant & 2) != 0;
    plan.nt_load = (variant & 4) != 0;
    plan.hdr.zero_mask = init ? 0 : zero_mask;
    const auto& dst = plan.stages;
    const auto& ops = plan.ops;
    size_t bytes = sizeof(ShmHeader) + dst.size() * sizeof(ShmStage) + ops.size() * sizeof(ShmOp);
    std::vector<char> blob(bytes);
    memcpy(blob.data(), &plan.hdr, sizeof(ShmHeader));
    memcpy(blob.data() + sizeof(ShmHeader), dst.data(), dst.size() * sizeof(ShmStage));
    if (!ops.empty())
        memcpy(blob.data() + sizeof(ShmHeader) + dst.size() * sizeof(ShmStage), ops.data(), ops.size() * sizeof(ShmOp));

    hq_program p = new hq_program_s();
    p->kind = 1;
    p->n = plan.n;
    p->precision = s->precision;
    p->tile = plan.k;
    p->rb = plan.rb;
    p->tb = plan.tb;
    p->num_stages = (int)dst.size();
    p->num_ops = (int)ops.size();
    p->shared_mask = shared_mask;
    p->prog_bytes = bytes;
    p->plan = new ShmPlan(std::move(plan));
    hipError_t e = hipMalloc(&p->d_prog, bytes);
    if (e != hipSuccess) {
        set_error(std::string("sharemem: hipMalloc: ") + hipGetErrorString(e));
        delete p->plan;
        delete p;
        return HQ_ERR_OOM;
    }
    e = hipMemcpy(p->d_prog, blob.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("sharemem: hipMemcpy: ") + hipGetErrorString(e));
        (void)hipFree(p->d_prog);
        delete p->plan;
        delete p;
        return HQ_ERR_HIP;
    }
    *out = p;
    return HQ_OK;
}

}  // namespace hq
