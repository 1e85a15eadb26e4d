// Device engine: context, resident problem and the evaluation sequence behind the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xmvb_amd.h"
#include "common.hpp"
#include "kernels.hpp"

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StatusError : std::runtime_error {
    int status;
    StatusError(int s, const std::string& w) : std::runtime_error(w), status(s) {}
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

enum Phase {
    PH_START, PH_PREP, PH_ONE_INDEX, PH_TRANSFORM, PH_PAIR_FWD, PH_STRUCT, PH_EIGEN, PH_BWD_BEGIN, PH_BWD_KERNEL,
    PH_PAIR_BWD, PH_TAIL, PH_COUNT
};

}  // namespace

struct vbe_context {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    std::string error;
};

struct vbe_problem {
    vbe_context* ctx = nullptr;
    vbe_dims dims{};
    int n_str = 0, n_det = 0, n_astr = 0, n_bstr = 0, np = 0, ng = 0;
    long long n_terms = 0;
    double e_nuc = 0.0;
    int pair_blocks_fwd = 0, pair_blocks_bwd = 0;
    std::vector<void*> allocations;
    uint64_t bytes = 0;
    // static device data
    double *S = nullptr, *h = nullptr, *B = nullptr;
    uint64_t *amask = nullptr, *bmask = nullptr;
    int *det_a = nullptr, *det_b = nullptr;
    long long* off = nullptr;
    int* ts = nullptr;
    double* tc = nullptr;
    long long* soff = nullptr;
    int* sdet = nullptr;
    double* scoef = nullptr;
    // per-evaluation workspace
    double *C = nullptr, *grad = nullptr;
    const double* cur_C = nullptr;  // orbitals of the current evaluation (problem buffer or caller's device array)
    uint8_t* support = nullptr;
    double *Cn = nullptr, *nu2 = nullptr, *SC = nullptr, *HC = nullptr, *sact = nullptr, *hact = nullptr;
    double *T = nullptr, *L = nullptr, *G = nullptr;
    double *Hdet = nullptr, *Sdet = nullptr, *Yh = nullptr, *Ys = nullptr, *Hs = nullptr, *Ss = nullptr;
    double *gep_work = nullptr, *coef = nullptr, *energy = nullptr, *dcoef = nullptr;
    double *partials = nullptr, *dsum = nullptr, *dL = nullptr, *gcn = nullptr;
    int* flags = nullptr;
    hipEvent_t ev[PH_COUNT] = {};
    // state of the last evaluation
    bool forward_valid = false;
    const uint8_t* last_support = nullptr;
    double last_energy = 0.0;
    int64_t last_singular = 0;
    vbe_phase_times times{};

    template <typename T>
    T* alloc(size_t count) {
        const size_t nbytes = std::max<size_t>(count, 1) * sizeof(T);
        if (ctx->budget && ctx->in_use + nbytes > ctx->budget)
            throw StatusError(VBE_ERR_MEMORY, "HBM budget exceeded while allocating problem storage");
        void* p = nullptr;
        check(hipMalloc(&p, nbytes), "hipMalloc");
        allocations.push_back(p);
        bytes += nbytes;
        ctx->in_use += nbytes;
        return static_cast<T*>(p);
    }
    ~vbe_problem() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : allocations) (void)hipFree(p);
        if (ctx) ctx->in_use -= std::min(ctx->in_use, bytes);
    }
};

namespace {

int fail(vbe_context* ctx, int status, const std::string& what) {
    if (ctx) ctx->error = what;
    xmvb_amd::set_global_error(what);
    return status;
}

template <typename F>
int guarded(vbe_context* ctx, F&& body) {
    try {
        return body();
    } catch (const StatusError& e) {
        return fail(ctx, e.status, e.what());
    } catch (const HipError& e) {
        return fail(ctx, VBE_ERR_DEVICE, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, VBE_ERR_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, VBE_ERR_ARGUMENT, e.what());
    }
}

void validate(const vbe_dims& d, const vbe_topology_view& t) {
    if (d.n_bf <= 0 || d.n_act <= 0 || d.n_aux <= 0 || d.n_alpha < 0 || d.n_beta < 0)
        throw StatusError(VBE_ERR_ARGUMENT, "dimensions must be positive");
    if (d.n_act > xmvb_amd::MAX_ACT)
        throw StatusError(VBE_ERR_UNSUPPORTED, "n_act above " + std::to_string(xmvb_amd::MAX_ACT) +
                                                   " is not supported by the LDS-resident determinant-pair kernels");
    if (d.n_alpha > xmvb_amd::MAX_SPIN_BLOCK || d.n_beta > xmvb_amd::MAX_SPIN_BLOCK)
        throw StatusError(VBE_ERR_UNSUPPORTED, "more than 8 electrons of one spin are not supported");
    if (d.n_alpha > d.n_act || d.n_beta > d.n_act) throw StatusError(VBE_ERR_ARGUMENT, "more electrons than orbitals");
    if (t.n_structures <= 0 || t.n_determinants <= 0 || t.n_alpha_strings <= 0 || t.n_beta_strings <= 0)
        throw StatusError(VBE_ERR_ARGUMENT, "empty topology");
    if (!t.alpha_masks || !t.beta_masks || !t.det_alpha || !t.det_beta || !t.term_offsets || !t.term_structure ||
        !t.term_coefficient)
        throw StatusError(VBE_ERR_ARGUMENT, "null topology array");
    const uint64_t valid = d.n_act == 64 ? ~0ull : ((1ull << d.n_act) - 1);
    for (int i = 0; i < t.n_alpha_strings; ++i)
        if ((t.alpha_masks[i] & ~valid) || __builtin_popcountll(t.alpha_masks[i]) != d.n_alpha)
            throw StatusError(VBE_ERR_ARGUMENT, "alpha string inconsistent with n_act/n_alpha");
    for (int i = 0; i < t.n_beta_strings; ++i)
        if ((t.beta_masks[i] & ~valid) || __builtin_popcountll(t.beta_masks[i]) != d.n_beta)
            throw StatusError(VBE_ERR_ARGUMENT, "beta string inconsistent with n_act/n_beta");
    for (int i = 0; i < t.n_determinants; ++i)
        if (t.det_alpha[i] < 0 || t.det_alpha[i] >= t.n_alpha_strings || t.det_beta[i] < 0 ||
            t.det_beta[i] >= t.n_beta_strings)
            throw StatusError(VBE_ERR_ARGUMENT, "determinant string index out of range");
    if (t.term_offsets[0] != 0) throw StatusError(VBE_ERR_ARGUMENT, "term offsets must start at 0");
    for (int i = 0; i < t.n_determinants; ++i)
        if (t.term_offsets[i + 1] < t.term_offsets[i]) throw StatusError(VBE_ERR_ARGUMENT, "term offsets decrease");
    const long long nt = t.term_offsets[t.n_determinants];
    for (long long k = 0; k < nt; ++k)
        if (t.term_structure[k] < 0 || t.term_structure[k] >= t.n_structures)
            throw StatusError(VBE_ERR_ARGUMENT, "term structure index out of range");
}

template <typename T>
void upload(T* dst, const T* src, size_t count, hipStream_t st) {
    check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D");
}

// Reads the device flags and energy with one synchronisation, then raises on any flagged failure.
void finish(vbe_problem* p) {
    int flags[xmvb_amd::FLAG_COUNT];
    double e[2];
    check(hipMemcpyAsync(flags, p->flags, sizeof(flags), hipMemcpyDeviceToHost, p->ctx->stream), "flags D2H");
    check(hipMemcpyAsync(e, p->energy, sizeof(e), hipMemcpyDeviceToHost, p->ctx->stream), "energy D2H");
    check(hipStreamSynchronize(p->ctx->stream), "hipStreamSynchronize");
    p->last_energy = e[1];
    p->last_singular = flags[xmvb_amd::FLAG_SINGULAR_PAIRS];
    if (flags[xmvb_amd::FLAG_NORMALIZATION])
        throw StatusError(VBE_ERR_NUMERIC, "orbital normalization failed for orbital " +
                                               std::to_string(flags[xmvb_amd::FLAG_NORMALIZATION] - 1));
    if (flags[xmvb_amd::FLAG_ACTIVE_OVERLAP])
        throw StatusError(VBE_ERR_NUMERIC, "active orbital overlap matrix is not positive definite");
    if (flags[xmvb_amd::FLAG_SINGULAR_PAIRS])
        throw StatusError(VBE_ERR_UNSUPPORTED,
                          std::to_string(flags[xmvb_amd::FLAG_SINGULAR_PAIRS]) +
                              " determinant pairs have rank-deficient overlap blocks; the singular-pair device path "
                              "is not implemented");
    if (flags[xmvb_amd::FLAG_METRIC])
        throw StatusError(VBE_ERR_METRIC, "structure metric is not positive definite");
}

// Forward phases on the context stream: orbitals -> integrals -> pairs -> structures -> eigensolve.
void enqueue_forward(vbe_problem* p) {
    hipStream_t st = p->ctx->stream;
    const auto& d = p->dims;
    using namespace xmvb_amd;
    check(hipMemsetAsync(p->flags, 0, sizeof(int) * FLAG_COUNT, st), "hipMemsetAsync flags");
    check(hipEventRecord(p->ev[PH_START], st), "hipEventRecord");
    launch_orbital_prep(st, p->cur_C, p->last_support, p->S, p->h, d.n_bf, d.n_act, p->Cn, p->SC, p->HC, p->nu2, p->flags);
    launch_active_matrices(st, p->Cn, p->SC, p->HC, d.n_bf, d.n_act, p->sact, p->hact, p->flags);
    check(hipEventRecord(p->ev[PH_PREP], st), "hipEventRecord");
    launch_ri_one_index(st, p->B, p->Cn, d.n_bf, d.n_act, d.n_aux, p->T);
    check(hipEventRecord(p->ev[PH_ONE_INDEX], st), "hipEventRecord");
    launch_ri_pair_stage(st, p->T, p->Cn, d.n_bf, d.n_act, d.n_aux, p->L, p->G);
    check(hipEventRecord(p->ev[PH_TRANSFORM], st), "hipEventRecord");
    launch_pairs(st, false, p->pair_blocks_fwd, p->G, p->sact, p->hact, d.n_act, p->ng, p->amask, p->bmask, p->det_a,
                 p->det_b, p->n_det, p->Hdet, p->Sdet, nullptr, nullptr, nullptr, p->flags);
    check(hipGetLastError(), "pair forward launch");
    check(hipEventRecord(p->ev[PH_PAIR_FWD], st), "hipEventRecord");
    launch_structure(st, p->Hdet, p->Sdet, p->n_det, p->n_str, p->soff, p->sdet, p->scoef, p->Yh, p->Ys, p->Hs, p->Ss);
    check(hipEventRecord(p->ev[PH_STRUCT], st), "hipEventRecord");
    launch_gep(st, p->Hs, p->Ss, p->n_str, p->e_nuc, p->gep_work, p->coef, p->energy, p->flags, p->off, p->ts, p->tc,
               p->n_det, p->dcoef);
    check(hipGetLastError(), "forward launch");
    check(hipEventRecord(p->ev[PH_EIGEN], st), "hipEventRecord");
}

void enqueue_gradient(vbe_problem* p, double* grad_out) {
    hipStream_t st = p->ctx->stream;
    const auto& d = p->dims;
    using namespace xmvb_amd;
    check(hipEventRecord(p->ev[PH_BWD_BEGIN], st), "hipEventRecord");
    launch_pairs(st, true, p->pair_blocks_bwd, p->G, p->sact, p->hact, d.n_act, p->ng, p->amask, p->bmask, p->det_a,
                 p->det_b, p->n_det, nullptr, nullptr, p->dcoef, p->energy, p->partials, p->flags);
    check(hipGetLastError(), "pair backward launch");
    check(hipEventRecord(p->ev[PH_BWD_KERNEL], st), "hipEventRecord");
    const int len = p->ng + 2 * d.n_act * d.n_act;
    launch_reduce_partials(st, p->partials, p->pair_blocks_bwd, len, p->dsum);
    check(hipEventRecord(p->ev[PH_PAIR_BWD], st), "hipEventRecord");
    launch_gradient_tail(st, p->T, p->L, p->dsum, p->SC, p->HC, p->Cn, p->nu2, p->last_support, d.n_bf, d.n_act,
                         d.n_aux, p->dL, p->gcn, grad_out);
    check(hipGetLastError(), "gradient launch");
    check(hipEventRecord(p->ev[PH_TAIL], st), "hipEventRecord");
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

void record_times(vbe_problem* p, bool with_gradient) {
    auto& t = p->times;
    t.orbital_prep_ms = elapsed(p->ev[PH_START], p->ev[PH_PREP]);
    t.transform_ms = elapsed(p->ev[PH_PREP], p->ev[PH_TRANSFORM]);
    t.ri_one_index_ms = elapsed(p->ev[PH_PREP], p->ev[PH_ONE_INDEX]);
    t.pair_forward_ms = elapsed(p->ev[PH_TRANSFORM], p->ev[PH_PAIR_FWD]);
    t.structure_ms = elapsed(p->ev[PH_PAIR_FWD], p->ev[PH_STRUCT]);
    t.eigensolve_ms = elapsed(p->ev[PH_STRUCT], p->ev[PH_EIGEN]);
    t.pair_backward_ms = with_gradient ? elapsed(p->ev[PH_EIGEN], p->ev[PH_PAIR_BWD]) : 0.f;
    t.pair_backward_kernel_ms = with_gradient ? elapsed(p->ev[PH_BWD_BEGIN], p->ev[PH_BWD_KERNEL]) : 0.f;
    t.gradient_tail_ms = with_gradient ? elapsed(p->ev[PH_PAIR_BWD], p->ev[PH_TAIL]) : 0.f;
    t.total_ms = elapsed(p->ev[PH_START], with_gradient ? p->ev[PH_TAIL] : p->ev[PH_EIGEN]);
}

int eval_core(vbe_context* ctx, vbe_problem* p, const double* d_C, const uint8_t* d_support, bool want_grad,
              double* E_total, double* d_grad) {
    if (!ctx || !p || p->ctx != ctx) throw StatusError(VBE_ERR_ARGUMENT, "problem does not belong to this context");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    p->cur_C = d_C;
    p->last_support = d_support;
    p->forward_valid = false;
    // everything is enqueued back to back; a failure flagged by the forward (singular pairs,
    // indefinite metric) makes the backward's output meaningless but never out of bounds, and
    // is reported before any result is returned
    enqueue_forward(p);
    if (want_grad) enqueue_gradient(p, d_grad ? d_grad : p->grad);
    finish(p);
    record_times(p, want_grad);
    p->forward_valid = true;
    if (E_total) *E_total = p->last_energy;
    return VBE_OK;
}

}  // namespace

extern "C" {

int vbe_context_create(int32_t device, uint64_t hbm_budget_bytes, vbe_context** out) {
    return guarded(nullptr, [&]() -> int {
        if (!out) throw StatusError(VBE_ERR_ARGUMENT, "null output");
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) throw StatusError(VBE_ERR_ARGUMENT, "invalid HIP device index");
        auto ctx = std::make_unique<vbe_context>();
        ctx->device = device;
        check(hipSetDevice(device), "hipSetDevice");
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        if (hbm_budget_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            ctx->budget = static_cast<uint64_t>(0.9 * static_cast<double>(free_b));
        } else {
            ctx->budget = hbm_budget_bytes;
        }
        *out = ctx.release();
        return VBE_OK;
    });
}

void vbe_context_destroy(vbe_context* ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int vbe_last_error(const vbe_context* ctx, char* buf, size_t len) {
    if (!buf || len == 0) return VBE_ERR_ARGUMENT;
    const std::string& e = ctx ? ctx->error : std::string(vbe_global_error());
    const size_t n = std::min(len - 1, e.size());
    std::memcpy(buf, e.data(), n);
    buf[n] = '\0';
    return VBE_OK;
}

int vbe_problem_upload(vbe_context* ctx, const vbe_dims* dims, const double* S_ao, const double* h_ao,
                       const double* B_packed, const vbe_topology_view* topo, double e_nuc, vbe_problem** out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !dims || !S_ao || !h_ao || !B_packed || !topo || !out)
            throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        validate(*dims, *topo);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        auto p = std::make_unique<vbe_problem>();
        p->ctx = ctx;
        p->dims = *dims;
        p->e_nuc = e_nuc;
        p->n_str = topo->n_structures;
        p->n_det = topo->n_determinants;
        p->n_astr = topo->n_alpha_strings;
        p->n_bstr = topo->n_beta_strings;
        p->n_terms = topo->term_offsets[topo->n_determinants];
        const int n = dims->n_act, nbf = dims->n_bf, naux = dims->n_aux;
        p->np = n * (n + 1) / 2;
        p->ng = p->np * (p->np + 1) / 2;
        const size_t nbp = (size_t)nbf * (nbf + 1) / 2;
        hipStream_t st = ctx->stream;
        // reverse incidence: structure -> (determinant, coefficient)
        std::vector<long long> soff(p->n_str + 1, 0);
        for (long long k = 0; k < p->n_terms; ++k) ++soff[topo->term_structure[k] + 1];
        for (int s = 0; s < p->n_str; ++s) soff[s + 1] += soff[s];
        std::vector<int> sdet(p->n_terms);
        std::vector<double> scoef(p->n_terms);
        {
            auto cursor = soff;
            for (int dd = 0; dd < p->n_det; ++dd)
                for (long long k = topo->term_offsets[dd]; k < topo->term_offsets[dd + 1]; ++k) {
                    const long long at = cursor[topo->term_structure[k]]++;
                    sdet[at] = dd;
                    scoef[at] = topo->term_coefficient[k];
                }
        }
        std::vector<long long> off(topo->term_offsets, topo->term_offsets + p->n_det + 1);
        p->S = p->alloc<double>((size_t)nbf * nbf);
        p->h = p->alloc<double>((size_t)nbf * nbf);
        p->B = p->alloc<double>((size_t)naux * nbp);
        p->amask = p->alloc<uint64_t>(p->n_astr);
        p->bmask = p->alloc<uint64_t>(p->n_bstr);
        p->det_a = p->alloc<int>(p->n_det);
        p->det_b = p->alloc<int>(p->n_det);
        p->off = p->alloc<long long>(p->n_det + 1);
        p->ts = p->alloc<int>(p->n_terms);
        p->tc = p->alloc<double>(p->n_terms);
        p->soff = p->alloc<long long>(p->n_str + 1);
        p->sdet = p->alloc<int>(p->n_terms);
        p->scoef = p->alloc<double>(p->n_terms);
        upload(p->S, S_ao, (size_t)nbf * nbf, st);
        upload(p->h, h_ao, (size_t)nbf * nbf, st);
        upload(p->B, B_packed, (size_t)naux * nbp, st);
        upload(p->amask, topo->alpha_masks, p->n_astr, st);
        upload(p->bmask, topo->beta_masks, p->n_bstr, st);
        upload(p->det_a, topo->det_alpha, p->n_det, st);
        upload(p->det_b, topo->det_beta, p->n_det, st);
        upload(p->off, off.data(), off.size(), st);
        upload(p->ts, topo->term_structure, p->n_terms, st);
        upload(p->tc, topo->term_coefficient, p->n_terms, st);
        upload(p->soff, soff.data(), soff.size(), st);
        upload(p->sdet, sdet.data(), sdet.size(), st);
        upload(p->scoef, scoef.data(), scoef.size(), st);
        // workspace
        const size_t nd = p->n_det, ns = p->n_str;
        const long long n_pairs = (long long)nd * (nd + 1) / 2;
        // forward: about two pairs per wavefront; backward: fewer blocks so that the per-block
        // adjoint partials stay small
        p->pair_blocks_fwd = (int)std::min<long long>(std::max<long long>(1, (n_pairs + 15) / 16), 4096);
        p->pair_blocks_bwd = (int)std::min<long long>(std::max<long long>(1, (n_pairs + 7) / 8), 512);
        p->C = p->alloc<double>((size_t)n * nbf);
        p->grad = p->alloc<double>((size_t)n * nbf);
        p->Cn = p->alloc<double>((size_t)n * nbf);
        p->nu2 = p->alloc<double>(n);
        p->SC = p->alloc<double>((size_t)n * nbf);
        p->HC = p->alloc<double>((size_t)n * nbf);
        p->sact = p->alloc<double>((size_t)n * n);
        p->hact = p->alloc<double>((size_t)n * n);
        p->T = p->alloc<double>((size_t)naux * n * nbf);
        p->L = p->alloc<double>((size_t)naux * p->np);
        p->G = p->alloc<double>(p->ng);
        p->Hdet = p->alloc<double>(nd * nd);
        p->Sdet = p->alloc<double>(nd * nd);
        p->Yh = p->alloc<double>(nd * ns);
        p->Ys = p->alloc<double>(nd * ns);
        p->Hs = p->alloc<double>(ns * ns);
        p->Ss = p->alloc<double>(ns * ns);
        p->gep_work = p->alloc<double>(xmvb_amd::gep_work_doubles((int)ns));
        p->coef = p->alloc<double>(ns);
        p->energy = p->alloc<double>(2);
        p->dcoef = p->alloc<double>(nd);
        const size_t len = (size_t)p->ng + 2 * (size_t)n * n;
        p->partials = p->alloc<double>((size_t)p->pair_blocks_bwd * len);
        p->dsum = p->alloc<double>(len);
        p->dL = p->alloc<double>((size_t)naux * p->np);
        p->gcn = p->alloc<double>((size_t)n * nbf);
        p->flags = p->alloc<int>(xmvb_amd::FLAG_COUNT);
        for (auto& e : p->ev) check(hipEventCreate(&e), "hipEventCreate");
        check(xmvb_amd::configure_kernels(n, p->ng, p->n_str), "hipFuncSetAttribute");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        *out = p.release();
        return VBE_OK;
    });
}

void vbe_problem_release(vbe_problem* prob) {
    if (!prob) return;
    if (prob->ctx && prob->ctx->stream) (void)hipStreamSynchronize(prob->ctx->stream);
    delete prob;
}

int vbe_problem_device_bytes(const vbe_problem* prob, uint64_t* bytes) {
    if (!prob || !bytes) return VBE_ERR_ARGUMENT;
    *bytes = prob->bytes;
    return VBE_OK;
}

int vbe_eval(vbe_context* ctx, vbe_problem* prob, const double* C_raw, const uint8_t* support, vbe_flags flags,
             double* E_total, double* E_states, double* grad_raw) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !C_raw || !E_total) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (flags.precision != VBE_PRECISION_FP64)
            throw StatusError(VBE_ERR_UNSUPPORTED, "only the FP64 precision path is implemented");
        if (flags.eigensolver != VBE_EIGENSOLVER_DENSE)
            throw StatusError(VBE_ERR_UNSUPPORTED, "only EIGENSOLVER=DENSE is implemented");
        if (flags.want_gradient && !grad_raw) throw StatusError(VBE_ERR_ARGUMENT, "null gradient buffer");
        const size_t nc = (size_t)prob->dims.n_act * prob->dims.n_bf;
        check(hipSetDevice(ctx->device), "hipSetDevice");
        upload(prob->C, C_raw, nc, ctx->stream);
        const uint8_t* dsup = nullptr;
        if (support) {
            if (!prob->support) prob->support = prob->alloc<uint8_t>(nc);
            upload(prob->support, support, nc, ctx->stream);
            dsup = prob->support;
        }
        const int rc = eval_core(ctx, prob, prob->C, dsup, flags.want_gradient != 0, E_total, prob->grad);
        if (rc != VBE_OK) return rc;
        if (E_states) E_states[0] = *E_total;
        if (flags.want_gradient) {
            check(hipMemcpyAsync(grad_raw, prob->grad, nc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                  "gradient D2H");
            check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        }
        return VBE_OK;
    });
}

int vbe_eval_device(vbe_context* ctx, vbe_problem* prob, const double* d_C_raw, const uint8_t* d_support,
                    vbe_flags flags, double* E_total, double* d_grad_raw) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !d_C_raw || !E_total) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (flags.precision != VBE_PRECISION_FP64)
            throw StatusError(VBE_ERR_UNSUPPORTED, "only the FP64 precision path is implemented");
        if (flags.eigensolver != VBE_EIGENSOLVER_DENSE)
            throw StatusError(VBE_ERR_UNSUPPORTED, "only EIGENSOLVER=DENSE is implemented");
        if (flags.want_gradient && !d_grad_raw) throw StatusError(VBE_ERR_ARGUMENT, "null gradient buffer");
        return eval_core(ctx, prob, d_C_raw, d_support, flags.want_gradient != 0, E_total, d_grad_raw);
    });
}

int vbe_eval_forward(vbe_context* ctx, vbe_problem* prob, const double* C_raw, const uint8_t* support,
                     double* E_total) {
    vbe_flags f{VBE_PRECISION_FP64, VBE_EIGENSOLVER_DENSE, 0};
    return vbe_eval(ctx, prob, C_raw, support, f, E_total, nullptr, nullptr);
}

int vbe_complete_gradient(vbe_context* ctx, vbe_problem* prob, double* grad_raw) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !grad_raw) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (!prob->forward_valid) throw StatusError(VBE_ERR_ARGUMENT, "no retained forward evaluation");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        enqueue_gradient(prob, prob->grad);
        finish(prob);
        const size_t nc = (size_t)prob->dims.n_act * prob->dims.n_bf;
        check(hipMemcpyAsync(grad_raw, prob->grad, nc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
              "gradient D2H");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        return VBE_OK;
    });
}

int vbe_last_intermediates(vbe_context* ctx, vbe_problem* prob, double* H, double* S, double* coef, double* S_act,
                           double* h_act, double* g_packed) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (!prob->forward_valid) throw StatusError(VBE_ERR_ARGUMENT, "no retained forward evaluation");
        hipStream_t st = ctx->stream;
        const size_t ns = prob->n_str, n = prob->dims.n_act;
        auto get = [&](double* dst, const double* src, size_t count) {
            if (dst) check(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
        };
        get(H, prob->Hs, ns * ns);
        get(S, prob->Ss, ns * ns);
        get(coef, prob->coef, ns);
        get(S_act, prob->sact, n * n);
        get(h_act, prob->hact, n * n);
        get(g_packed, prob->G, prob->ng);
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        return VBE_OK;
    });
}

int vbe_last_phase_times(const vbe_problem* prob, vbe_phase_times* times) {
    if (!prob || !times) return VBE_ERR_ARGUMENT;
    *times = prob->times;
    return VBE_OK;
}

int vbe_last_singular_pairs(const vbe_problem* prob, int64_t* count) {
    if (!prob || !count) return VBE_ERR_ARGUMENT;
    *count = prob->last_singular;
    return VBE_OK;
}

}  // extern "C"
