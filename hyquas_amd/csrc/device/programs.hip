// programs.hip -- device-resident group programs and their launches.
//
// A program is what Circuit::compile leaves behind for one group: the
// lowered OShareMem op table (src/executor.cpp:462-493 rebuilt it and copied
// it to a global __constant__ table on every launch; here it is built once
// and owned by the program, so launches are re-entrant), or the TransMM
// transpose plan + fused matrix (src/schedule.cpp:704-783).

#include "hq_device.h"

#include <cmath>
#include <cstring>
#include <thread>

using namespace hq;

extern "C" {

hq_status hq_program_create_sharemem(hq_state s, const hq_gate* gates, int num_gates, uint64_t shared_mask,
                                     hq_program* out)
{
    return build_sharemem_program(s, gates, num_gates, shared_mask, 0, out);
}

// U as split planes (re | im) in the state's precision: FP64 fusion, rounded
// once to FP32 for complex64 states; uploaded to p->d_mat
static hq_status upload_matrix(hq_state s, hq_program p, const double* mat, int mat_qubits, const char* who)
{
    const size_t K = (size_t)1 << mat_qubits;
    const size_t rbytes = s->precision == HQ_C128 ? 8 : 4;
    std::vector<char> planes(2 * K * K * rbytes);
    for (size_t i = 0; i < K * K; i++) {
        double re = mat[2 * i], im = mat[2 * i + 1];
        if (!std::isfinite(re) || !std::isfinite(im)) {
            set_error(std::string(who) + ": matrix entry is not finite");
            return HQ_ERR_ARG;
        }
        if (rbytes == 8) {
            ((double*)planes.data())[i] = re;
            ((double*)planes.data())[K * K + i] = im;
        } else {
            ((float*)planes.data())[i] = (float)re;
            ((float*)planes.data())[K * K + i] = (float)im;
        }
    }
    hipError_t e = hipMalloc(&p->d_mat, planes.size());
    if (e != hipSuccess) {
        set_error(std::string(who) + ": hipMalloc: " + hipGetErrorString(e));
        return HQ_ERR_OOM;
    }
    e = hipMemcpy(p->d_mat, planes.data(), planes.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string(who) + ": hipMemcpy: " + hipGetErrorString(e));
        return HQ_ERR_HIP;
    }
    return HQ_OK;
}

hq_status hq_program_create_dense(hq_state s, const int* positions, const double* mat, int mat_qubits,
                                  hq_program* out)
{
    HQ_CHECK(s && out && positions && mat, HQ_ERR_ARG, "hq_program_create_dense: null argument");
    *out = nullptr;
    HQ_CHECK(mat_qubits >= 4 && mat_qubits <= 7, HQ_ERR_ARG, "hq_program_create_dense: mat_qubits must be 4..7");
    HQ_CHECK(s->n >= dense_at_min_qubits(s->precision, mat_qubits), HQ_ERR_ARG,
             "hq_program_create_dense: the state is smaller than one tile of the dense kernel");
    const int m = mat_qubits;
    uint64_t rows = 0;
    for (int j = 0; j < m; j++) {
        HQ_CHECK(positions[j] >= 0 && positions[j] < s->n && !((rows >> positions[j]) & 1), HQ_ERR_ARG,
                 "hq_program_create_dense: positions must be distinct qubits of the state");
        rows |= 1ull << positions[j];
    }
    // the kernel numbers the matrix rows by increasing qubit position: row
    // index bit jj of the kernel is matrix bit sigma[jj]
    int sigma[7];
    for (int jj = 0, b = 0; b < 64 && jj < m; b++) {
        if (!((rows >> b) & 1)) continue;
        for (int j = 0; j < m; j++)
            if (positions[j] == b) sigma[jj] = j;
        jj++;
    }
    const size_t K = (size_t)1 << m;
    std::vector<size_t> orig(K);
    for (size_t x = 0; x < K; x++) {
        size_t o = 0;
        for (int jj = 0; jj < m; jj++)
            if ((x >> jj) & 1) o |= (size_t)1 << sigma[jj];
        orig[x] = o;
    }
    std::vector<double> pm(2 * K * K);
    for (size_t c = 0; c < K; c++)
        for (size_t r = 0; r < K; r++) {
            const size_t from = orig[c] * K + orig[r], to = c * K + r;
            pm[2 * to] = mat[2 * from];
            pm[2 * to + 1] = mat[2 * from + 1];
        }
    hq_program p = new hq_program_s();
    p->kind = 3;
    p->n = s->n;
    p->precision = s->precision;
    p->mat_qubits = m;
    p->rows_mask = rows;
    hq_status st = upload_matrix(s, p, pm.data(), m, "hq_program_create_dense");
    if (st != HQ_OK) { hq_program_destroy(p); return st; }
    *out = p;
    return HQ_OK;
}

hq_status hq_program_create_blas(hq_state s, const int* perm, const double* mat, int mat_qubits, hq_program* out)
{
    HQ_CHECK(s && out, HQ_ERR_ARG, "hq_program_create_blas: null argument");
    *out = nullptr;
    HQ_CHECK(mat_qubits >= 0 && mat_qubits <= 7 && mat_qubits <= s->n, HQ_ERR_ARG,
             "hq_program_create_blas: mat_qubits must be in [0, min(7, n)] (0: transpose only)");
    HQ_CHECK(mat || !mat_qubits, HQ_ERR_ARG, "hq_program_create_blas: null matrix");
    const int n = s->n;
    hq_program p = new hq_program_s();
    p->kind = 2;
    p->n = n;
    p->precision = s->precision;
    p->mat_qubits = mat_qubits;
    p->identity_perm = true;
    if (perm) {
        uint64_t seen = 0;
        for (int i = 0; i < n; i++) {
            if (perm[i] < 0 || perm[i] >= n || ((seen >> perm[i]) & 1)) {
                delete p;
                set_error("hq_program_create_blas: perm is not a permutation of 0..n-1");
                return HQ_ERR_ARG;
            }
            seen |= 1ull << perm[i];
            if (perm[i] != i) p->identity_perm = false;
        }
        p->perm.assign(perm, perm + n);
    }
    hq_status st = HQ_OK;
    if (!p->identity_perm) {
        st = transpose_prepare(s, p->perm.data(), &p->d_targs);
        if (st != HQ_OK) { hq_program_destroy(p); return st; }
    }
    if (mat_qubits == 0) { *out = p; return HQ_OK; }
    st = upload_matrix(s, p, mat, mat_qubits, "hq_program_create_blas");
    if (st != HQ_OK) { hq_program_destroy(p); return st; }
    *out = p;
    return HQ_OK;
}

hq_status hq_program_launch(hq_state s, hq_program p)
{
    HQ_CHECK(s && p, HQ_ERR_ARG, "hq_program_launch: null argument");
    HQ_CHECK(p->n == s->n && p->precision == s->precision, HQ_ERR_ARG,
             "hq_program_launch: program was built for another qubit count or precision");
    int slot;
    hq_status st;
    if (p->kind == 1) {
        st = prof_begin(s, 1, &slot);
        if (st != HQ_OK) return st;
        st = p->jit_fn ? launch_sharemem_jit(s, p) : launch_sharemem(s, p);
        if (st != HQ_OK) return st;
        return prof_end(s, slot);
    }
    if (p->kind == 3) {
        st = prof_begin(s, 3, &slot);
        if (st != HQ_OK) return st;
        st = launch_dense_at(s, p->d_mat, p->mat_qubits, p->rows_mask);
        if (st != HQ_OK) return st;
        return prof_end(s, slot);
    }
    HQ_CHECK(p->kind == 2, HQ_ERR_ARG, "hq_program_launch: unknown program kind");
    HQ_CHECK(s->buf, HQ_ERR_STATE, "hq_program_launch: TransMM needs a state created with a work buffer");
    if (!p->identity_perm) {
        st = prof_begin(s, 2, &slot);
        if (st != HQ_OK) return st;
        st = transpose_launch(s, p->d_targs, s->data, s->buf);
        if (st != HQ_OK) return st;
        std::swap(s->data, s->buf);
        st = prof_end(s, slot);
        if (st != HQ_OK) return st;
    }
    if (p->mat_qubits > 0) {
        st = prof_begin(s, 3, &slot);
        if (st != HQ_OK) return st;
        st = launch_dense(s, p->d_mat, p->mat_qubits, s->data, s->buf);
        if (st != HQ_OK) return st;
        std::swap(s->data, s->buf);
        st = prof_end(s, slot);
        if (st != HQ_OK) return st;
    }
    return HQ_OK;
}

hq_status hq_program_destroy(hq_program p)
{
    if (!p) return HQ_OK;
    if (p->d_prog) (void)hipFree(p->d_prog);
    if (p->d_targs) (void)hipFree(p->d_targs);
    if (p->d_mat) (void)hipFree(p->d_mat);
    if (p->d_coef) (void)hipFree(p->d_coef);
    if (p->jit_module) (void)hipModuleUnload(p->jit_module);
    delete p->plan;
    delete p;
    return HQ_OK;
}

hq_status hq_program_specialize(hq_program p)
{
    HQ_CHECK(p, HQ_ERR_ARG, "hq_program_specialize: null program");
    if (p->kind != 1) return HQ_OK;   // TransMM programs have nothing to specialise
    if (p->jit_fn) return HQ_OK;
    HQ_CHECK(p->plan, HQ_ERR_STATE, "hq_program_specialize: program has no plan");
    return jit_sharemem(*p->plan, p);
}

hq_status hq_program_is_specialized(hq_program p, int* yes)
{
    HQ_CHECK(p && yes, HQ_ERR_ARG, "hq_program_is_specialized: null argument");
    *yes = p->jit_fn != nullptr;
    return HQ_OK;
}

hq_status hq_program_info(hq_program p, int* tile_qubits, int* num_stages, int* num_ops)
{
    HQ_CHECK(p, HQ_ERR_ARG, "hq_program_info: null program");
    if (tile_qubits) *tile_qubits = p->kind == 1 ? p->tile : p->mat_qubits;
    if (num_stages) *num_stages = p->kind == 1 ? p->num_stages : 0;
    if (num_ops) *num_ops = p->kind == 1 ? p->num_ops : 0;
    return HQ_OK;
}

hq_status hq_apply_sharemem(hq_state s, const hq_gate* gates, int num_gates, uint64_t shared_mask)
{
    hq_program p = nullptr;
    hq_status st = hq_program_create_sharemem(s, gates, num_gates, shared_mask, &p);
    if (st != HQ_OK) return st;
    st = hq_program_launch(s, p);
    hipError_t e = hipStreamSynchronize(s->stream);
    hq_program_destroy(p);
    if (st == HQ_OK && e != hipSuccess) { set_error(hipGetErrorString(e)); return HQ_ERR_HIP; }
    return st;
}

hq_status hq_apply_transpose(hq_state s, const int* perm)
{
    HQ_CHECK(s && perm, HQ_ERR_ARG, "hq_apply_transpose: null argument");
    hq_program p = nullptr;
    hq_status st = hq_program_create_blas(s, perm, nullptr, 0, &p);
    if (st != HQ_OK) return st;
    st = hq_program_launch(s, p);
    hipError_t e = hipStreamSynchronize(s->stream);
    hq_program_destroy(p);
    if (st == HQ_OK && e != hipSuccess) { set_error(hipGetErrorString(e)); return HQ_ERR_HIP; }
    return st;
}

hq_status hq_apply_dense(hq_state s, const double* mat, int mat_qubits)
{
    hq_program p = nullptr;
    hq_status st = hq_program_create_blas(s, nullptr, mat, mat_qubits, &p);
    if (st != HQ_OK) return st;
    st = hq_program_launch(s, p);
    hipError_t e = hipStreamSynchronize(s->stream);
    hq_program_destroy(p);
    if (st == HQ_OK && e != hipSuccess) { set_error(hipGetErrorString(e)); return HQ_ERR_HIP; }
    return st;
}

}  // extern "C"

namespace hq {

// each variant of every group is launched `rounds` times, interleaved with
// the other variants; its fastest launch counts (t[v][i])
static hq_status time_variants(hq_state s, const std::vector<std::vector<hq_program>>& built, int rounds,
                               std::vector<std::vector<float>>* t)
{
    const size_t nv = built.size(), n = nv ? built[0].size() : 0;
    t->assign(nv, std::vector<float>(n, 1e30f));
    hipEvent_t e0, e1;
    HQ_HIP(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error(std::string("tune: ") + hipGetErrorString(he));
        return HQ_ERR_HIP;
    }
    const bool prof = s->profiling;
    s->profiling = false;
    hq_status st = HQ_OK;
    for (int r = 0; r < rounds && st == HQ_OK; r++) {
        for (size_t i = 0; i < n && st == HQ_OK; i++) {
            for (size_t v = 0; v < nv && st == HQ_OK; v++) {
                if ((he = hipEventRecord(e0, s->stream)) != hipSuccess) break;
                st = hq_program_launch(s, built[v][i]);
                if (st != HQ_OK) break;
                if ((he = hipEventRecord(e1, s->stream)) != hipSuccess) break;
                if ((he = hipEventSynchronize(e1)) != hipSuccess) break;
                float ms = 0;
                if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
                (*t)[v][i] = std::min((*t)[v][i], ms);
            }
        }
        if (he != hipSuccess) {
            set_error(std::string("tune: ") + hipGetErrorString(he));
            st = HQ_ERR_HIP;
        }
    }
    s->profiling = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

hq_status build_sharemem_programs(hq_state s, const std::vector<SharememRequest>& reqs, int reg_qubits, bool jit,
                                  std::vector<hq_program>* out, int tune_mode)
{
    // tune_mode 1: two variants of build_sharemem_program, timed: the plain
    // stages, and the whole-line stages with non-temporal loads and stores.
    // (Four candidates, adding each with non-temporal stores only, averaged
    // 0.1575 s per 30-qubit simulation over five boxes against 0.1567 s for
    // two; all eight combinations 0.1579 s: with two timed launches per
    // candidate, more close candidates let the noise pick losers.)  2: the
    // whole-line variant only
    bool tune = tune_mode == 1 && jit && !reqs.empty();
    const int only = tune_mode == 2 && jit ? 7 : 0;
    static constexpr int kTuned[] = {0, 7};
    const int variants = tune ? 2 : 1;
    std::vector<std::vector<hq_program>> built(variants, std::vector<hq_program>(reqs.size(), nullptr));
    auto destroy_all = [&]() {
        for (auto& v : built)
            for (auto p : v) hq_program_destroy(p);
    };
    out->assign(reqs.size(), nullptr);
    for (int v = 0; v < variants; v++) {
        for (size_t i = 0; i < reqs.size(); i++) {
            hq_status st = build_sharemem_program(s, reqs[i].gates.data(), (int)reqs[i].gates.size(), reqs[i].mask,
                                                  reg_qubits, &built[v][i], reqs[i].init, tune ? kTuned[v] : only,
                                                  reqs[i].zero_mask);
            if (st != HQ_OK) { destroy_all(); return st; }
        }
    }
    if (jit) {
        std::vector<const ShmPlan*> plans;
        std::vector<hq_program> progs;
        for (auto& v : built)
            for (auto p : v) { plans.push_back(p->plan); progs.push_back(p); }
        unsigned hw = std::thread::hardware_concurrency();
        const char* e = getenv("OMP_NUM_THREADS");
        int threads = e && atoi(e) > 0 ? atoi(e) : (int)(hw ? hw : 4);
        hq_status st = jit_sharemem_many(plans, progs, std::min(threads, 16));
        if (st != HQ_OK && tune) {
            // the plain kernels alone (the other variants are options, not requirements)
            for (size_t v = 1; v < built.size(); v++)
                for (auto p : built[v]) hq_program_destroy(p);
            built.resize(1);
            tune = false;
            plans.resize(reqs.size());
            progs.resize(reqs.size());
            st = jit_sharemem_many(plans, progs, std::min(threads, 16));
        }
        if (st != HQ_OK) { destroy_all(); return st; }
    }
    if (tune) {
        std::vector<std::vector<float>> t;
        hq_status st = time_variants(s, built, 2, &t);
        if (st != HQ_OK) { destroy_all(); return st; }
        for (size_t i = 0; i < reqs.size(); i++) {
            size_t keep = 0;
            for (size_t v = 1; v < built.size(); v++)
                if (t[v][i] < t[keep][i]) keep = v;
            (*out)[i] = built[keep][i];
            for (size_t v = 0; v < built.size(); v++)
                if (v != keep) hq_program_destroy(built[v][i]);
        }
        return HQ_OK;
    }
    *out = built[0];
    return HQ_OK;
}

}  // namespace hq
