// Device engine behind the C ABI: contexts, resident problem pools and the evaluation sequence.
//
// A pool holds `capacity` resident problem slots that share one structure topology (same active
// space and structure set); each slot has its own AO dimensions, static data (S, h, B) and
// workspace, all allocated when the slot is loaded, so evaluations allocate nothing. One
// evaluation launch processes any subset of slots with one kernel per phase (blockIdx.z = slot).
// A single problem (vbe_problem) is a pool of one slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/xmvb_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"
#include "optimizer.hpp"
#include "server.hpp"

namespace {

using xmvb_amd::FLAG_COUNT;
using xmvb_amd::LaunchShape;
using xmvb_amd::SlotDesc;
using xmvb_amd::TopoDesc;

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

constexpr int kLanes = 4;              // launch lanes of a pool (in-flight submissions)
constexpr int kContextStreams = 4;     // HIP streams of a context, shared by all of its pools
constexpr int kMaxBwdBlocksPerSlot = 512;
constexpr int kLaneBlocks = 2048;      // default one-pair-per-lane blocks per launch (all slots)
constexpr int kAutotuneTargets[VBE_AUTOTUNE_CANDIDATES] = {256, 512, 1024, 2048, 4096, 8192};

}  // namespace

struct vbe_context {
    int device = 0;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    double pair_condition_limit = xmvb_amd::kLaneConditionDefault;
    int ri_precision = VBE_PRECISION_FP64;
    // structure eigensolver of pools created afterwards; Davidson accuracy (StructureSolveAccuracy
    // defaults, src/vbscf/core/contracts/eigensolver.hpp:23-45) and root recycling
    int eigensolver = VBE_EIGENSOLVER_DENSE;
    double dav_energy_tol = 1.0e-7, dav_residual_tol = 1.0e-3, dav_stationarity_tol = 1.0;
    bool dav_reuse_roots = true;
    // selected structure states of problems and pools created afterwards (normalised weights)
    int n_states = 1;
    int state_idx[xmvb_amd::kMaxStates] = {0};
    double state_w[xmvb_amd::kMaxStates] = {1.0};
    std::string error;
    // guards in_use, error and the stream cursor: pools of one context may be driven from several
    // host threads (for example the request scheduler beside a convergence worker)
    std::mutex mu;
    // The context's streams (created with its first pool). Every pool's launch lanes take their
    // streams from this set, so however many pools are resident, the device sees at most
    // kContextStreams queues of this context, one per hardware queue; launches of different pools
    // on one stream run in submission order. Completion is tracked per lane with events, never by
    // synchronizing a shared stream.
    std::vector<hipStream_t> streams;
    int next_stream = 0;
    hipStream_t take_stream() {
        std::lock_guard<std::mutex> g(mu);
        if (streams.empty()) {
            for (int k = 0; k < kContextStreams; ++k) {
                hipStream_t st = nullptr;
                check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
                streams.push_back(st);
            }
        }
        return streams[next_stream++ % streams.size()];
    }
    // the stream after `main` in the set: a lane's side stream for launches of few problems
    hipStream_t side_stream(hipStream_t main) {
        std::lock_guard<std::mutex> g(mu);
        for (size_t k = 0; k < streams.size(); ++k)
            if (streams[k] == main) return streams[(k + 1) % streams.size()];
        return main;
    }
    ~vbe_context() {
        for (hipStream_t st : streams) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
};

namespace {

// Device allocations charged to a context budget.
struct DeviceArena {
    vbe_context* ctx = nullptr;
    std::vector<void*> ptrs;
    uint64_t bytes = 0;
    template <typename T>
    T* alloc(size_t count) {
        const size_t nbytes = std::max<size_t>(count, 1) * sizeof(T);
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            if (ctx->budget && ctx->in_use + nbytes > ctx->budget)
                throw StatusError(VBE_ERR_MEMORY, "HBM budget exceeded while allocating problem storage");
            ctx->in_use += nbytes;  // reserved before the allocation, returned if it fails
        }
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, nbytes);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> g(ctx->mu);
            ctx->in_use -= std::min(ctx->in_use, (uint64_t)nbytes);
        }
        check(e, "hipMalloc");
        ptrs.push_back(p);
        bytes += nbytes;
        return static_cast<T*>(p);
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
        if (ctx) {
            std::lock_guard<std::mutex> g(ctx->mu);
            ctx->in_use -= std::min(ctx->in_use, bytes);
        }
        bytes = 0;
    }
    ~DeviceArena() { release(); }
};

struct Slot {
    bool loaded = false;
    int n_bf = 0, n_aux = 0;
    double e_nuc = 0.0;
    std::unique_ptr<DeviceArena> arena;
    SlotDesc desc{};           // workspace + static pointers; per-launch fields filled at launch
    std::vector<double> S_host;  // AO overlap (the optimiser's chart metric)
    bool has_root = false;     // desc.coef holds this slot's last structure eigenvector
    double* C_buf = nullptr;   // slot-owned orbitals / gradient / support for host-buffer evaluations
    double* grad_buf = nullptr;
    uint8_t* support_buf = nullptr;
};

// One stream with its descriptor table and compact per-launch outputs.
struct Lane {
    hipStream_t stream = nullptr;
    SlotDesc* d_desc = nullptr;
    SlotDesc* h_desc = nullptr;  // pinned
    double* d_energy = nullptr;
    int* d_flags = nullptr;
    double* h_energy = nullptr;  // pinned
    int* h_flags = nullptr;      // pinned
    double* d_C = nullptr;       // compact orbitals / gradients of submissions
    double* d_grad = nullptr;
    double* h_C = nullptr;       // pinned staging
    double* h_grad = nullptr;
    size_t stage_doubles = 0;
    hipEvent_t ev[xmvb_amd::EV_COUNT] = {};
    hipEvent_t done = nullptr;
    xmvb_amd::SideStream side;  // concurrent kernels of launches of few problems
    bool busy = false;
    bool with_grad = false;
    bool events_valid = false;
    bool scratch = false;  // the launch ran on a workspace outside the slots (no-allocation entry point)
    int ticket = -1;
    std::vector<int> slots;
    std::vector<size_t> grad_offsets;
};

}  // namespace

struct vbe_pool {
    vbe_context* ctx = nullptr;
    int n_act = 0, n_alpha = 0, n_beta = 0, capacity = 0;
    int ri_precision = VBE_PRECISION_FP64;
    int eigensolver = VBE_EIGENSOLVER_DENSE;
    TopoDesc topo{};
    DeviceArena topo_arena;
    DeviceArena ri_stage_arena;  // staging of flat RI factors for re-layout at load
    // the optimizer's device state, kept between calls and grown when a call needs more: a
    // convergence request then neither allocates nor frees (hipFree waits for the whole device)
    DeviceArena opt_arena;
    void* opt_block = nullptr;
    size_t opt_block_bytes = 0;
    void* opt_pinned = nullptr;  // and its pinned read-back buffer
    size_t opt_pinned_bytes = 0;
    double* ri_stage = nullptr;
    size_t ri_stage_doubles = 0;
    std::vector<Slot> slots;
    Lane lanes[kLanes];
    int next_ticket = 0;
    int last_lane = 0;
    int64_t last_singular = 0;
    vbe_phase_times times{};
    // pair-kernel blocks per launch (all listed slots together), chosen by vbe_pool_autotune:
    // forward one-pair-per-lane blocks (the lane-group kernel takes twice this), backward blocks
    int fwd_target = kLaneBlocks;
    int bwd_target = kLaneBlocks;

    // serializes host access to the pool's lanes and slots between the caller's thread and the
    // request scheduler's dispatcher thread
    std::recursive_mutex mu;

    ~vbe_pool() {
        for (auto& l : lanes) {
            if (l.done) (void)hipEventSynchronize(l.done);  // the lane's last launch (streams are shared)
            for (auto& e : l.ev)
                if (e) (void)hipEventDestroy(e);
            if (l.done) (void)hipEventDestroy(l.done);
            if (l.side.fork) (void)hipEventDestroy(l.side.fork);
            if (l.side.join) (void)hipEventDestroy(l.side.join);
            if (l.h_desc) (void)hipHostFree(l.h_desc);
            if (l.h_energy) (void)hipHostFree(l.h_energy);
            if (l.h_flags) (void)hipHostFree(l.h_flags);
            if (l.h_C) (void)hipHostFree(l.h_C);
            if (l.h_grad) (void)hipHostFree(l.h_grad);
        }
        if (opt_pinned) (void)hipHostFree(opt_pinned);
    }
};

struct vbe_problem {
    std::unique_ptr<vbe_pool> pool;
    bool forward_valid = false;
};

namespace {

int fail(vbe_context* ctx, int status, const std::string& what) {
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->error = what;
    }
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

void validate_topology(int n_act, int n_alpha, int n_beta, const vbe_topology_view& t) {
    if (n_act <= 0 || n_alpha < 0 || n_beta < 0) throw StatusError(VBE_ERR_ARGUMENT, "dimensions must be positive");
    if (n_act > xmvb_amd::MAX_ACT)
        throw StatusError(VBE_ERR_UNSUPPORTED, "n_act above " + std::to_string(xmvb_amd::MAX_ACT) +
                                                   " is not supported by the LDS-resident determinant-pair kernels");
    if (n_alpha > xmvb_amd::MAX_SPIN_BLOCK || n_beta > xmvb_amd::MAX_SPIN_BLOCK)
        throw StatusError(VBE_ERR_UNSUPPORTED, "more than 8 electrons of one spin are not supported");
    if (n_alpha > n_act || n_beta > n_act) throw StatusError(VBE_ERR_ARGUMENT, "more electrons than orbitals");
    if (t.n_structures <= 0 || t.n_determinants <= 0 || t.n_alpha_strings <= 0 || t.n_beta_strings <= 0)
        throw StatusError(VBE_ERR_ARGUMENT, "empty topology");
    if ((long long)t.n_determinants * (t.n_determinants + 1) / 2 > 0x7fffffffLL)
        throw StatusError(VBE_ERR_UNSUPPORTED, "more than 2^31 determinant pairs are not supported");
    if (!t.alpha_masks || !t.beta_masks || !t.det_alpha || !t.det_beta || !t.term_offsets || !t.term_structure ||
        !t.term_coefficient)
        throw StatusError(VBE_ERR_ARGUMENT, "null topology array");
    const uint64_t valid = n_act == 64 ? ~0ull : ((1ull << n_act) - 1);
    for (int i = 0; i < t.n_alpha_strings; ++i)
        if ((t.alpha_masks[i] & ~valid) || __builtin_popcountll(t.alpha_masks[i]) != n_alpha)
            throw StatusError(VBE_ERR_ARGUMENT, "alpha string inconsistent with n_act/n_alpha");
    for (int i = 0; i < t.n_beta_strings; ++i)
        if ((t.beta_masks[i] & ~valid) || __builtin_popcountll(t.beta_masks[i]) != n_beta)
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

long long pair_count(const vbe_pool* p) { return (long long)p->topo.n_det * (p->topo.n_det + 1) / 2; }

int bwd_blocks_cap(const vbe_pool* p) {
    return (int)std::min<long long>(kMaxBwdBlocksPerSlot, std::max<long long>(1, (pair_count(p) + 7) / 8));
}

std::unique_ptr<vbe_pool> make_pool(vbe_context* ctx, const vbe_topology_view& t, int n_act, int n_alpha, int n_beta,
                                    int capacity) {
    validate_topology(n_act, n_alpha, n_beta, t);
    if (capacity <= 0) throw StatusError(VBE_ERR_ARGUMENT, "pool capacity must be positive");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto p = std::make_unique<vbe_pool>();
    p->ctx = ctx;
    p->ri_precision = ctx->ri_precision;
    p->eigensolver = ctx->eigensolver;
    p->n_act = n_act;
    p->n_alpha = n_alpha;
    p->n_beta = n_beta;
    p->capacity = capacity;
    p->topo_arena.ctx = ctx;
    p->opt_arena.ctx = ctx;
    for (auto& l : p->lanes) {
        l.stream = ctx->take_stream();
        for (auto& e : l.ev) check(hipEventCreate(&e), "hipEventCreate");
        check(hipEventCreateWithFlags(&l.done, hipEventDisableTiming), "hipEventCreate");
        l.side.st = ctx->side_stream(l.stream);
        check(hipEventCreateWithFlags(&l.side.fork, hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&l.side.join, hipEventDisableTiming), "hipEventCreate");
        check(hipHostMalloc(reinterpret_cast<void**>(&l.h_desc), sizeof(SlotDesc) * capacity, hipHostMallocDefault),
              "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&l.h_energy), sizeof(double) * 2 * capacity, hipHostMallocDefault),
              "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&l.h_flags), sizeof(int) * FLAG_COUNT * capacity,
                            hipHostMallocDefault),
              "hipHostMalloc");
        l.d_desc = p->topo_arena.alloc<SlotDesc>(capacity);
        l.d_energy = p->topo_arena.alloc<double>(2 * (size_t)capacity);
        l.d_flags = p->topo_arena.alloc<int>((size_t)FLAG_COUNT * capacity);
    }
    hipStream_t st = p->lanes[0].stream;
    const int nd = t.n_determinants, ns = t.n_structures;
    const long long n_terms = t.term_offsets[nd];
    // reverse incidence: structure -> (determinant, coefficient)
    std::vector<long long> soff(ns + 1, 0);
    for (long long k = 0; k < n_terms; ++k) ++soff[t.term_structure[k] + 1];
    for (int s = 0; s < ns; ++s) soff[s + 1] += soff[s];
    std::vector<int> sdet(n_terms);
    std::vector<double> scoef(n_terms);
    {
        auto cursor = soff;
        for (int d = 0; d < nd; ++d)
            for (long long k = t.term_offsets[d]; k < t.term_offsets[d + 1]; ++k) {
                const long long at = cursor[t.term_structure[k]]++;
                sdet[at] = d;
                scoef[at] = t.term_coefficient[k];
            }
    }
    std::vector<long long> off(t.term_offsets, t.term_offsets + nd + 1);
    auto& A = p->topo_arena;
    TopoDesc& T = p->topo;
    T.n = n_act;
    T.np = n_act * (n_act + 1) / 2;
    T.ng = T.np * (T.np + 1) / 2;
    T.n_det = nd;
    T.n_str = ns;
    const long long pairs = (long long)nd * (nd + 1) / 2;
    // spin-flip partners: determinant (a, b) <-> the determinant with strings (b, a), when every
    // determinant has one (equal electron counts and a flip-closed determinant set)
    std::vector<int> flip, lane_pairs;
    if (n_alpha == n_beta && pairs < (1LL << 31)) {
        std::vector<std::pair<std::pair<uint64_t, uint64_t>, int>> key(nd);
        for (int d = 0; d < nd; ++d) key[d] = {{t.alpha_masks[t.det_alpha[d]], t.beta_masks[t.det_beta[d]]}, d};
        std::sort(key.begin(), key.end());
        flip.resize(nd);
        for (int d = 0; d < nd && !flip.empty(); ++d) {
            const std::pair<uint64_t, uint64_t> want{t.beta_masks[t.det_beta[d]], t.alpha_masks[t.det_alpha[d]]};
            auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(want, -1));
            if (it == key.end() || it->first != want) flip.clear();
            else flip[d] = it->second;
        }
        if (!flip.empty()) {
            lane_pairs.reserve((size_t)(pairs / 2 + nd));
            for (long long e = 0; e < nd; ++e)
                for (long long d = 0; d <= e; ++d) {
                    const long long lo = std::min(flip[d], flip[e]), hi = std::max(flip[d], flip[e]);
                    const long long tt = e * (e + 1) / 2 + d;
                    if (tt <= hi * (hi + 1) / 2 + lo) lane_pairs.push_back((int)tt);
                }
        }
    }
    T.n_lane_pairs = flip.empty() ? pairs : (long long)lane_pairs.size();
    {
        // a multiplier coprime with the work count near count / golden ratio: lanes of one wavefront
        // then take pairs of different determinants, which spreads the backward kernels' LDS atomic
        // targets (consecutive pairs share their ket determinant and hit the same rows)
        const long long work = T.n_lane_pairs;
        long long m = std::max<long long>(1, (long long)(0.6180339887498949 * (double)work)) | 1;
        while (std::gcd(m, work) != 1) m += 2;
        T.pair_mul = work > 1 && work < (1LL << 32) ? m : 1;  // the kernels form i * pair_mul in 64 bits
    }
    auto* amask = A.alloc<uint64_t>(t.n_alpha_strings);
    auto* bmask = A.alloc<uint64_t>(t.n_beta_strings);
    auto* det_a = A.alloc<int>(nd);
    auto* det_b = A.alloc<int>(nd);
    auto* d_off = A.alloc<long long>(nd + 1);
    auto* ts = A.alloc<int>(n_terms);
    auto* tc = A.alloc<double>(n_terms);
    auto* d_soff = A.alloc<long long>(ns + 1);
    auto* d_sdet = A.alloc<int>(n_terms);
    auto* d_scoef = A.alloc<double>(n_terms);
    int* d_flip = flip.empty() ? nullptr : A.alloc<int>(nd);
    int* d_lane_pairs = flip.empty() ? nullptr : A.alloc<int>(lane_pairs.size());
    upload(amask, t.alpha_masks, t.n_alpha_strings, st);
    upload(bmask, t.beta_masks, t.n_beta_strings, st);
    upload(det_a, t.det_alpha, nd, st);
    upload(det_b, t.det_beta, nd, st);
    upload(d_off, off.data(), off.size(), st);
    upload(ts, t.term_structure, n_terms, st);
    upload(tc, t.term_coefficient, n_terms, st);
    upload(d_soff, soff.data(), soff.size(), st);
    upload(d_sdet, sdet.data(), sdet.size(), st);
    upload(d_scoef, scoef.data(), scoef.size(), st);
    if (d_flip) {
        upload(d_flip, flip.data(), flip.size(), st);
        upload(d_lane_pairs, lane_pairs.data(), lane_pairs.size(), st);
    }
    T.flip = d_flip;
    T.lane_pairs = d_lane_pairs;
    T.amask = amask;
    T.bmask = bmask;
    T.det_a = det_a;
    T.det_b = det_b;
    T.off = d_off;
    T.ts = ts;
    T.tc = tc;
    T.soff = d_soff;
    T.sdet = d_sdet;
    T.scoef = d_scoef;
    T.n_states = ctx->n_states;
    for (int k = 0; k < xmvb_amd::kMaxStates; ++k) {
        T.state_idx[k] = k < ctx->n_states ? ctx->state_idx[k] : 0;
        T.state_w[k] = k < ctx->n_states ? ctx->state_w[k] : 0.0;
    }
    for (int k = 0; k < T.n_states; ++k)
        if (T.state_idx[k] >= ns)
            throw StatusError(VBE_ERR_ARGUMENT, "selected state " + std::to_string(T.state_idx[k]) +
                                                    " is out of range for " + std::to_string(ns) + " structures");
    check(xmvb_amd::configure_kernels(n_act, T.ng, ns, std::max(n_alpha, n_beta)), "hipFuncSetAttribute");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    p->slots.resize(capacity);
    return p;
}

// Every path that starts work on slots (evaluation, submission, load, unload) rejects a slot that a
// pending submission on either lane still holds: the two launches would share its workspace.
void check_slots_free(const vbe_pool* p, const int* slots, int count) {
    for (const auto& l : p->lanes)
        if (l.busy)
            for (int i = 0; i < count; ++i)
                if (std::find(l.slots.begin(), l.slots.end(), slots[i]) != l.slots.end())
                    throw StatusError(VBE_ERR_ARGUMENT, "slot " + std::to_string(slots[i]) +
                                                            " is in use by a pending submission");
}

// Per-evaluation workspace of one slot of a pool, carved from one block of device memory (arrays
// aligned to 256 bytes). With base == nullptr nothing is assigned and only the size is returned.
// The slot's own arena and the no-allocation entry point's static scratch use this same layout.
size_t carve_workspace(const vbe_pool* p, int n_bf, int n_aux, unsigned char* base, SlotDesc& d) {
    const TopoDesc& T = p->topo;
    const size_t n = T.n, np = T.np, nd = T.n_det, ns = T.n_str;
    size_t off = 0;
    auto take = [&](auto*& ptr, size_t count) {
        using E = std::remove_pointer_t<std::remove_reference_t<decltype(ptr)>>;
        off = (off + 255) & ~size_t(255);
        if (base) ptr = reinterpret_cast<E*>(base + off);
        off += std::max<size_t>(count, 1) * sizeof(E);
    };
    take(d.Cn, n * n_bf);
    take(d.CnT, (size_t)(n_bf + 31) / 32 * 32 * xmvb_amd::kCnStride);
    take(d.SC, n * n_bf);
    take(d.HC, n * n_bf);
    take(d.nu2, n);
    take(d.sact, n * n);
    take(d.hact, n * n);
    take(d.T, (size_t)n_aux * n * n_bf);
    take(d.L, (size_t)n_aux * np);
    take(d.G, (size_t)T.ng);
    take(d.Hdet, nd * nd);
    take(d.Sdet, nd * nd);
    take(d.Yh, nd * ns);
    take(d.Ys, nd * ns);
    take(d.Hs, ns * ns);
    take(d.Ss, ns * ns);
    take(d.gep, xmvb_amd::gep_work_doubles((int)ns));
    take(d.coef, ns * (size_t)T.n_states);
    take(d.dcoef, nd * (size_t)T.n_states);
    take(d.state_energy, (size_t)xmvb_amd::kMaxStates);
    const size_t len = (size_t)T.ng + 2 * n * n;
    take(d.partials, (size_t)(bwd_blocks_cap(p) + xmvb_amd::kListBlocks + xmvb_amd::kSingularBlocks) * len);
    take(d.dsum, len);
    take(d.dL, (size_t)n_aux * np);
    take(d.gcn, n * n_bf);
    take(d.defer, (size_t)pair_count(p));
    take(d.singular, (size_t)pair_count(p));
    take(d.defer_flag, (size_t)T.n_lane_pairs);
    take(d.singular_flag, (size_t)pair_count(p));
    take(d.chunk_count,
         (size_t)(std::max<long long>(T.n_lane_pairs, pair_count(p)) / xmvb_amd::kCompactChunk + 2));
    return (off + 255) & ~size_t(255);
}

// aux rows of the tiled RI layout (padded to whole tiles)
size_t ri_tiled_aux(int n_aux) {
    return (size_t)(n_aux + xmvb_amd::kRiTileAux - 1) / xmvb_amd::kRiTileAux * xmvb_amd::kRiTileAux;
}

// The pool's device staging buffer for flat RI factors on their way to a slot's layout; it only
// grows (a larger problem replaces it) and is charged to the context budget.
double* ensure_ri_stage(vbe_pool* p, size_t doubles) {
    if (p->ri_stage_doubles >= doubles) return p->ri_stage;
    p->ri_stage_arena.release();
    p->ri_stage_arena.ctx = p->ctx;
    p->ri_stage = nullptr;
    p->ri_stage_doubles = 0;
    p->ri_stage = p->ri_stage_arena.alloc<double>(doubles);
    p->ri_stage_doubles = doubles;
    return p->ri_stage;
}

void load_slot(vbe_pool* p, int slot, int n_bf, int n_aux, const double* S, const double* h, const double* B,
               double e_nuc) {
    if (slot < 0 || slot >= p->capacity) throw StatusError(VBE_ERR_ARGUMENT, "slot index out of range");
    if (n_bf <= 0 || n_aux <= 0 || !S || !h || !B) throw StatusError(VBE_ERR_ARGUMENT, "invalid slot data");
    check_slots_free(p, &slot, 1);
    check(hipSetDevice(p->ctx->device), "hipSetDevice");
    Slot& s = p->slots[slot];
    s.S_host.assign(S, S + (size_t)n_bf * n_bf);
    s.has_root = false;
    const size_t n = p->topo.n;
    const size_t nbp = (size_t)n_bf * (n_bf + 1) / 2;
    const bool realloc = !s.arena || s.n_bf != n_bf || s.n_aux != n_aux;
    // the slot is unusable until the upload below has completed; a failed (re)allocation, e.g.
    // over the HBM budget, leaves it unloaded with no storage rather than with stale pointers
    s.loaded = false;
    if (realloc) {
        s.arena.reset();
        s.desc = SlotDesc{};
        s.C_buf = s.grad_buf = nullptr;
        s.support_buf = nullptr;
        s.n_bf = s.n_aux = 0;
        auto arena = std::make_unique<DeviceArena>();
        arena->ctx = p->ctx;
        auto& A = *arena;
        SlotDesc d{};
        d.S = A.alloc<double>((size_t)n_bf * n_bf);
        d.h = A.alloc<double>((size_t)n_bf * n_bf);
        // aux-tiled RI factors for the one-read transform where its bands cover n_bf (plain FP32
        // storage keeps the flat layout); tiles are padded with zero aux rows
        const bool tiled = p->ri_precision != VBE_PRECISION_FP32_RI && n_bf <= xmvb_amd::kRiTiledMaxBf;
        const size_t nb_elems = (tiled ? ri_tiled_aux(n_aux) : (size_t)n_aux) * nbp;
        if (p->ri_precision == VBE_PRECISION_FP64) {
            d.B = A.alloc<double>(nb_elems);
        } else {
            d.B32 = A.alloc<float>(nb_elems);
            if (p->ri_precision == VBE_PRECISION_FP32C_RI) d.Blo = A.alloc<signed char>(nb_elems);
        }
        d.b_tiled = tiled ? 1 : 0;
        // the per-evaluation workspace as one block, in the layout the no-allocation entry point
        // reproduces in its static scratch
        unsigned char* ws = A.alloc<unsigned char>(carve_workspace(p, n_bf, n_aux, nullptr, d));
        carve_workspace(p, n_bf, n_aux, ws, d);
        double* C_buf = A.alloc<double>(n * n_bf);
        double* grad_buf = A.alloc<double>(n * n_bf);
        uint8_t* support_buf = A.alloc<uint8_t>(n * n_bf);
        s.arena = std::move(arena);
        s.C_buf = C_buf;
        s.grad_buf = grad_buf;
        s.support_buf = support_buf;
        s.desc = d;
    }
    s.n_bf = n_bf;
    s.n_aux = n_aux;
    s.e_nuc = e_nuc;
    s.desc.n_bf = n_bf;
    s.desc.n_aux = n_aux;
    s.desc.e_nuc = e_nuc;
    hipStream_t st = p->lanes[0].stream;
    upload(const_cast<double*>(s.desc.S), S, (size_t)n_bf * n_bf, st);
    upload(const_cast<double*>(s.desc.h), h, (size_t)n_bf * n_bf, st);
    const size_t flat = (size_t)n_aux * nbp;
    if (p->ri_precision == VBE_PRECISION_FP32_RI) {
        std::vector<float> b32(flat);
        for (size_t k = 0; k < b32.size(); ++k) b32[k] = (float)B[k];
        upload(const_cast<float*>(s.desc.B32), b32.data(), b32.size(), st);
        check(hipStreamSynchronize(st), "hipStreamSynchronize");  // b32 is a host temporary
    } else if (p->ri_precision == VBE_PRECISION_FP64 && !s.desc.b_tiled) {
        upload(const_cast<double*>(s.desc.B), B, flat, st);
    } else {
        // the flat factors go to the pool's staging buffer, a kernel writes the slot's layout
        // (tiling and, for compensated storage, the FP32 + correction encoding)
        double* stage = ensure_ri_stage(p, flat);
        upload(stage, B, flat, st);
        xmvb_amd::launch_ri_relayout(st, stage, n_aux, (long long)nbp, s.desc.b_tiled != 0,
                                     const_cast<double*>(s.desc.B), const_cast<float*>(s.desc.B32),
                                     const_cast<signed char*>(s.desc.Blo));
        check(hipGetLastError(), "RI factor layout kernel");
    }
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    s.loaded = true;
}

// Per-slot inputs of one launch.
struct SlotInput {
    int slot;
    const double* C;         // device orbitals
    const uint8_t* support;  // device support mask or null
    double* grad;            // device gradient output
};

LaunchShape shape_for(const vbe_pool* p, const std::vector<SlotInput>& in, bool pinned) {
    LaunchShape sh{};
    sh.pinned = pinned;
    for (const auto& x : in) {
        sh.max_n_bf = std::max(sh.max_n_bf, p->slots[x.slot].n_bf);
        sh.max_n_aux = std::max(sh.max_n_aux, p->slots[x.slot].n_aux);
    }
    sh.mmax = std::max(p->n_alpha, p->n_beta);
    const long long pairs = pair_count(p);
    const long long count = (long long)in.size();
    // at least ~4 sweeps of 4 wavefronts' pair groups per forward block
    const long long per_sweep = 4LL * (64 / xmvb_amd::pair_group(sh.mmax));
    const long long fwd_group = 2LL * p->fwd_target, bwd_group = p->bwd_target;
    sh.pair_blocks_fwd = (int)std::clamp<long long>((fwd_group + count - 1) / count, 1,
                                                    std::max<long long>(1, (pairs + 4 * per_sweep - 1) / (4 * per_sweep)));
    sh.pair_blocks_bwd = pinned ? bwd_blocks_cap(p)
                                : (int)std::clamp<long long>((bwd_group + count - 1) / count, 1, bwd_blocks_cap(p));
    // one pair per lane: 256 pairs per block sweep
    sh.n_alpha = p->n_alpha;
    sh.n_beta = p->n_beta;
    sh.lane_fwd = xmvb_amd::pairs_lane_supported(p->n_alpha, p->n_beta, false);
    sh.lane_bwd = xmvb_amd::pairs_lane_supported(p->n_alpha, p->n_beta, true);
    const long long lane_sweep = std::max<long long>(1, (p->topo.n_lane_pairs + 255) / 256);
    sh.lane_blocks_fwd = (int)std::clamp<long long>((p->fwd_target + count - 1) / count, 1, lane_sweep);
    sh.lane_blocks_bwd = pinned ? (int)std::min<long long>(lane_sweep, bwd_blocks_cap(p))
                                : (int)std::clamp<long long>((p->bwd_target + count - 1) / count, 1,
                                                             std::min<long long>(lane_sweep, bwd_blocks_cap(p)));
    // pinned shapes keep the full count so that a problem's partial rows do not depend on its batch
    sh.list_blocks = pinned ? xmvb_amd::kListBlocks
                            : (int)std::clamp<long long>(256 / std::max(count, 1LL), xmvb_amd::kListBlocksBatched,
                                                         xmvb_amd::kListBlocks);
    sh.lane_cond_limit = p->ctx->pair_condition_limit;
    sh.b32 = p->ri_precision == VBE_PRECISION_FP32_RI;
    sh.b32c = p->ri_precision == VBE_PRECISION_FP32C_RI;
    for (const auto& x : in) {
        const Slot& sl = p->slots[x.slot];
        if (sl.desc.b_tiled) {
            sh.any_tiled = true;
            sh.max_n_bf_tiled = std::max(sh.max_n_bf_tiled, sl.n_bf);
        } else {
            sh.any_flat = true;
        }
    }
    sh.davidson = p->eigensolver == VBE_EIGENSOLVER_DAVIDSON;
    if (sh.davidson && (p->topo.n_states > 1 || p->topo.state_idx[0] != 0))
        throw StatusError(VBE_ERR_UNSUPPORTED, "the Davidson eigensolver finds the lowest structure root only; "
                                               "state-averaged evaluations take the dense eigensolver");
    sh.dav_energy_tol = p->ctx->dav_energy_tol;
    sh.dav_residual_tol = p->ctx->dav_residual_tol;
    sh.dav_stationarity_tol = p->ctx->dav_stationarity_tol;
    return sh;
}

// Enqueue forward and/or backward for a set of slots on a lane, then the D2H of energies and flags.
// ws (optional, single-slot launches): a descriptor whose workspace replaces the slot's own.
// timed: record the phase events (read by finish into the pool's phase times); the optimiser's
// inner evaluations skip them, a dozen event records per launch being host time on its critical path.
void enqueue(vbe_pool* p, Lane& lane, const std::vector<SlotInput>& in, bool with_grad, bool forward,
             const SlotDesc* ws = nullptr, bool pinned = false, bool timed = true) {
    const int count = (int)in.size();
    if (count <= 0 || count > p->capacity) throw StatusError(VBE_ERR_ARGUMENT, "invalid number of slots");
    if (ws && count != 1) throw StatusError(VBE_ERR_ARGUMENT, "a replacement workspace serves one slot");
    hipStream_t st = lane.stream;
    const LaunchShape sh = shape_for(p, in, pinned);
    // the side stream only when it is idle: queued behind another lane's launch it would delay this one
    const xmvb_amd::SideStream* side =
        lane.side.st != st && hipStreamQuery(lane.side.st) == hipSuccess ? &lane.side : nullptr;
    if (forward) {
        std::vector<char> seen(p->capacity, 0);
        for (int i = 0; i < count; ++i) {
            const int s = in[i].slot;
            if (s < 0 || s >= p->capacity || !p->slots[s].loaded)
                throw StatusError(VBE_ERR_ARGUMENT, "slot not loaded");
            if (seen[s]) throw StatusError(VBE_ERR_ARGUMENT, "a slot appears twice in one launch");
            seen[s] = 1;
            SlotDesc d = ws ? *ws : p->slots[s].desc;
            d.C = in[i].C;
            d.support = in[i].support;
            d.grad = in[i].grad;
            d.energy = lane.d_energy + 2 * i;
            d.flags = lane.d_flags + FLAG_COUNT * i;
            // a foreign workspace holds no previous eigenvector of this slot
            d.seeded = !ws && p->slots[s].has_root && p->ctx->dav_reuse_roots ? 1 : 0;
            lane.h_desc[i] = d;
        }
        upload(lane.d_desc, lane.h_desc, count, st);
        check(hipMemsetAsync(lane.d_flags, 0, sizeof(int) * FLAG_COUNT * count, st), "hipMemsetAsync");
        xmvb_amd::launch_forward(st, lane.d_desc, count, p->topo, sh, timed ? lane.ev : nullptr, side);
        check(hipGetLastError(), "forward launch");
    }
    if (with_grad) {
        xmvb_amd::launch_backward(st, lane.d_desc, count, p->topo, sh, timed ? lane.ev : nullptr, side);
        check(hipGetLastError(), "backward launch");
    }
    check(hipMemcpyAsync(lane.h_energy, lane.d_energy, sizeof(double) * 2 * count, hipMemcpyDeviceToHost, st),
          "energy D2H");
    check(hipMemcpyAsync(lane.h_flags, lane.d_flags, sizeof(int) * FLAG_COUNT * count, hipMemcpyDeviceToHost, st),
          "flags D2H");
    // the lane's completion point (the stream may carry other pools' launches after it)
    check(hipEventRecord(lane.done, st), "hipEventRecord");
    lane.slots.clear();
    for (const auto& x : in) lane.slots.push_back(x.slot);
    lane.with_grad = with_grad;
    lane.events_valid = forward && timed;
    lane.scratch = ws != nullptr;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// Wait for a lane, record phase times, raise on device flags (or, with slot_status, report each
// listed slot's status instead of raising).
void finish(vbe_pool* p, Lane& lane, std::vector<int>* slot_status = nullptr) {
    check(hipEventSynchronize(lane.done), "hipEventSynchronize");
    lane.busy = false;
    using namespace xmvb_amd;
    if (lane.events_valid) {
        auto& t = p->times;
        t.orbital_prep_ms = elapsed(lane.ev[EV_START], lane.ev[EV_PREP]);
        t.transform_ms = elapsed(lane.ev[EV_PREP], lane.ev[EV_TRANSFORM]);
        t.ri_one_index_ms = elapsed(lane.ev[EV_PREP], lane.ev[EV_ONE_INDEX]);
        t.pair_forward_ms = elapsed(lane.ev[EV_TRANSFORM], lane.ev[EV_PAIR_FWD]);
        t.structure_ms = elapsed(lane.ev[EV_PAIR_FWD], lane.ev[EV_STRUCT]);
        t.eigensolve_ms = elapsed(lane.ev[EV_STRUCT], lane.ev[EV_EIGEN]);
        t.pair_backward_ms = lane.with_grad ? elapsed(lane.ev[EV_EIGEN], lane.ev[EV_PAIR_BWD]) : 0.f;
        t.pair_backward_kernel_ms = lane.with_grad ? elapsed(lane.ev[EV_BWD_BEGIN], lane.ev[EV_BWD_KERNEL]) : 0.f;
        t.gradient_tail_ms = lane.with_grad ? elapsed(lane.ev[EV_PAIR_BWD], lane.ev[EV_TAIL]) : 0.f;
        t.total_ms = elapsed(lane.ev[EV_START], lane.with_grad ? lane.ev[EV_TAIL] : lane.ev[EV_EIGEN]);
        t.pair_forward_kernel_ms = elapsed(lane.ev[EV_TRANSFORM], lane.ev[EV_PAIR_FWD_MAIN]);
        t.pair_backward_field_ms = lane.with_grad ? elapsed(lane.ev[EV_BWD_BEGIN], lane.ev[EV_BWD_FIELD]) : 0.f;
        t.pair_backward_adjoint_ms = lane.with_grad ? elapsed(lane.ev[EV_BWD_FIELD], lane.ev[EV_BWD_ADJOINT]) : 0.f;
    }
    int64_t singular = 0;
    std::string err;
    int status = VBE_OK;
    if (slot_status) slot_status->assign(lane.slots.size(), VBE_OK);
    for (size_t i = 0; i < lane.slots.size(); ++i) {
        const int* f = lane.h_flags + FLAG_COUNT * i;
        singular += f[FLAG_SINGULAR_PAIRS];
        const int si = f[FLAG_NORMALIZATION] || f[FLAG_ACTIVE_OVERLAP] || f[xmvb_amd::FLAG_EIGEN] ? VBE_ERR_NUMERIC
                       : f[FLAG_METRIC]                                                        ? VBE_ERR_METRIC
                                                                                               : VBE_OK;
        if (slot_status) (*slot_status)[i] = si;
        if (lane.events_valid && !lane.scratch) p->slots[lane.slots[i]].has_root = si == VBE_OK;
        if (status != VBE_OK || si == VBE_OK) continue;
        status = si;
        const std::string where = lane.slots.size() > 1 ? " (slot " + std::to_string(lane.slots[i]) + ")" : "";
        if (f[FLAG_NORMALIZATION])
            err = "orbital normalization failed for orbital " + std::to_string(f[FLAG_NORMALIZATION] - 1) + where;
        else if (f[FLAG_ACTIVE_OVERLAP])
            err = "active orbital overlap matrix is not positive definite" + where;
        else if (f[xmvb_amd::FLAG_EIGEN])
            err = std::string("Davidson stopped without meeting its tolerances (") +
                  (f[xmvb_amd::FLAG_EIGEN] == 1   ? "initial vectors dependent in the overlap metric"
                   : f[xmvb_amd::FLAG_EIGEN] == 2 ? "complete basis cannot meet the residual"
                   : f[xmvb_amd::FLAG_EIGEN] == 3 ? "correction space became linearly dependent"
                                                  : "iteration limit") +
                  ")" + where;
        else
            err = "structure metric is not positive definite" + where;
    }
    p->last_singular = singular;
    if (status != VBE_OK && !slot_status) throw StatusError(status, err);
}

Lane& acquire_lane(vbe_pool* p) {
    for (int k = 0; k < kLanes; ++k) {
        Lane& l = p->lanes[(p->last_lane + 1 + k) % kLanes];
        if (!l.busy) return l;
    }
    throw StatusError(VBE_ERR_ARGUMENT, "both submission lanes are busy; wait for a ticket first");
}

void ensure_stage(vbe_pool* p, Lane& l, size_t doubles) {
    if (l.stage_doubles >= doubles) return;
    if (l.h_C) (void)hipHostFree(l.h_C);
    if (l.h_grad) (void)hipHostFree(l.h_grad);
    l.h_C = l.h_grad = nullptr;
    check(hipHostMalloc(reinterpret_cast<void**>(&l.h_C), sizeof(double) * doubles, hipHostMallocDefault),
          "hipHostMalloc");
    check(hipHostMalloc(reinterpret_cast<void**>(&l.h_grad), sizeof(double) * doubles, hipHostMallocDefault),
          "hipHostMalloc");
    // compact device buffers live as long as the pool
    l.d_C = p->topo_arena.alloc<double>(doubles);
    l.d_grad = p->topo_arena.alloc<double>(doubles);
    l.stage_doubles = doubles;
}

void check_flags_args(const vbe_flags& flags, const vbe_pool* p) {
    if (flags.precision != VBE_PRECISION_FP64 && flags.precision != VBE_PRECISION_FP32_RI &&
        flags.precision != VBE_PRECISION_FP32C_RI)
        throw StatusError(VBE_ERR_ARGUMENT, "unknown precision");
    if (flags.precision != p->ri_precision)
        throw StatusError(VBE_ERR_ARGUMENT, "precision differs from the one the problem was uploaded with "
                                            "(vbe_set_ri_precision)");
    if (flags.eigensolver != VBE_EIGENSOLVER_DENSE && flags.eigensolver != VBE_EIGENSOLVER_DAVIDSON)
        throw StatusError(VBE_ERR_ARGUMENT, "unknown eigensolver");
}

// Evaluate slot 0 of a single problem synchronously; C/support/grad are device pointers.
double eval_problem(vbe_problem* prob, const double* d_C, const uint8_t* d_support, bool want_grad, double* d_grad,
                    bool forward) {
    vbe_pool* p = prob->pool.get();
    check(hipSetDevice(p->ctx->device), "hipSetDevice");
    Lane& lane = p->lanes[0];
    std::vector<SlotInput> in{{0, d_C, d_support, d_grad}};
    if (forward) prob->forward_valid = false;
    lane.busy = true;
    try {
        // a single problem's launch shape is the pinned one (the same bits as in any pinned batch)
        enqueue(p, lane, in, want_grad, forward, nullptr, true);
    } catch (...) {
        lane.busy = false;
        throw;
    }
    finish(p, lane);
    prob->forward_valid = true;
    return lane.h_energy[1];
}

// P-LBFGS over the listed slots of a pool (a single problem is a pool of one, passed as prob so
// that energy-only trials can keep their forward state for the gradient).
void optimize_slots(vbe_pool* p, vbe_problem* prob, const std::vector<int>& slots,
                    std::vector<xmvb_amd::OptProblem>& batch, const vbe_optimizer_options& options) {
    check(hipSetDevice(p->ctx->device), "hipSetDevice");
    Lane& lane = p->lanes[0];
    xmvb_amd::OptimizerHooks hooks;
    hooks.stream = lane.stream;
    hooks.n_act = p->n_act;
    hooks.evaluate = [&](const std::vector<int>& which, const std::vector<const double*>& C,
                         const std::vector<const uint8_t*>& sup, const std::vector<double*>& G, bool with_grad,
                         std::vector<double>& E) {
        std::vector<SlotInput> in;
        for (size_t j = 0; j < which.size(); ++j) in.push_back({slots[which[j]], C[j], sup[j], G[j]});
        if (prob) prob->forward_valid = false;
        lane.busy = true;
        try {
            // pinned: each slot follows the iterates it would follow alone, whatever the batch
            enqueue(p, lane, in, with_grad, true, nullptr, true, false);
        } catch (...) {
            lane.busy = false;
            throw;
        }
        std::vector<int> status;
        finish(p, lane, &status);
        if (prob) prob->forward_valid = true;
        for (size_t j = 0; j < which.size(); ++j)
            E[j] = status[j] == VBE_OK ? lane.h_energy[2 * j + 1] : std::numeric_limits<double>::quiet_NaN();
    };
    hooks.energy_only_trials = prob != nullptr;
    hooks.device_block = [p](size_t bytes) -> void* {
        if (p->opt_block_bytes < bytes) {
            p->opt_arena.release();
            p->opt_block = nullptr;
            p->opt_block_bytes = 0;
            p->opt_block = p->opt_arena.alloc<unsigned char>(bytes);
            p->opt_block_bytes = bytes;
        }
        return p->opt_block;
    };
    hooks.host_block = [p](size_t bytes) -> void* {
        if (p->opt_pinned_bytes < bytes) {
            if (p->opt_pinned) (void)hipHostFree(p->opt_pinned);
            p->opt_pinned = nullptr;
            p->opt_pinned_bytes = 0;
            check(hipHostMalloc(&p->opt_pinned, bytes, hipHostMallocDefault), "hipHostMalloc");
            p->opt_pinned_bytes = bytes;
        }
        return p->opt_pinned;
    };
    hooks.complete = [&](int, const double* C, double* G) {
        if (!prob) throw StatusError(VBE_ERR_ARGUMENT, "energy-only trials need a single problem");
        eval_problem(prob, C, lane.h_desc[0].support, true, G, false);
    };
    try {
        xmvb_amd::run_plbfgs(hooks, batch, options);
    } catch (const std::invalid_argument& e) {
        throw StatusError(VBE_ERR_ARGUMENT, e.what());
    } catch (const StatusError&) {
        throw;
    } catch (const HipError&) {
        throw;
    } catch (const std::runtime_error& e) {
        throw StatusError(VBE_ERR_NUMERIC, e.what());
    }
}

}  // namespace

namespace xmvb_amd {

int context_device(const vbe_context* ctx) { return ctx ? ctx->device : -1; }

int report_error(vbe_context* ctx, int status, const char* what) { return fail(ctx, status, what); }

int eval_device_in_scratch(vbe_context* ctx, vbe_problem* prob, const double* d_C, const uint8_t* d_support,
                           vbe_flags flags, double* E_total, double* d_grad, unsigned char* scratch,
                           size_t scratch_bytes) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !d_C || !E_total || !scratch) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        vbe_pool* p = prob->pool.get();
        if (p->ctx != ctx) throw StatusError(VBE_ERR_ARGUMENT, "problem does not belong to this context");
        check_flags_args(flags, p);
        if (flags.want_gradient && !d_grad) throw StatusError(VBE_ERR_ARGUMENT, "null gradient buffer");
        const Slot& s = p->slots[0];
        if (!s.loaded) throw StatusError(VBE_ERR_ARGUMENT, "problem not loaded");
        SlotDesc ws = s.desc;
        const size_t need = carve_workspace(p, s.n_bf, s.n_aux, nullptr, ws);
        if (need > scratch_bytes)
            throw StatusError(VBE_ERR_MEMORY, "evaluation workspace of " + std::to_string(need) +
                                                  " bytes exceeds the static scratch maximum of " +
                                                  std::to_string(scratch_bytes) + " bytes");
        carve_workspace(p, s.n_bf, s.n_aux, scratch, ws);
        p->eigensolver = flags.eigensolver;
        check(hipSetDevice(ctx->device), "hipSetDevice");
        Lane& lane = p->lanes[0];
        std::vector<SlotInput> in{{0, d_C, d_support, flags.want_gradient ? d_grad : nullptr}};
        // the problem's own workspace no longer holds the last forward
        prob->forward_valid = false;
        lane.busy = true;
        try {
            enqueue(p, lane, in, flags.want_gradient != 0, true, &ws);
        } catch (...) {
            lane.busy = false;
            throw;
        }
        finish(p, lane);
        *E_total = lane.h_energy[1];
        return VBE_OK;
    });
}

// ---- request-scheduler hooks ------------------------------------------------------------------

int pool_device(const vbe_pool* p) { return p->ctx->device; }

int context_streams(vbe_context* ctx, void** out, int max_streams) {
    (void)ctx->take_stream();  // creates the set on first use
    std::lock_guard<std::mutex> g(ctx->mu);
    const int n = std::min<int>(max_streams, (int)ctx->streams.size());
    for (int k = 0; k < n; ++k) out[k] = ctx->streams[k];
    return n;
}

int pool_slot_dims(vbe_pool* p, int slot, int* n_act, int* n_bf, std::string* err) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    if (slot < 0 || slot >= p->capacity || !p->slots[slot].loaded) {
        *err = "slot " + std::to_string(slot) + " is not loaded";
        return VBE_ERR_ARGUMENT;
    }
    *n_act = p->n_act;
    *n_bf = p->slots[slot].n_bf;
    return VBE_OK;
}

bool pool_slot_busy(vbe_pool* p, int slot) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    for (const auto& l : p->lanes)
        if (l.busy && std::find(l.slots.begin(), l.slots.end(), slot) != l.slots.end()) return true;
    return false;
}

int pool_launch(vbe_pool* p, int count, const int* slots, const double* const* C, bool gradient, void* stream,
                void* side, int* lane_out, std::string* err) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    *lane_out = -1;
    Lane* lane = nullptr;
    for (int k = 0; k < kLanes && !lane; ++k)
        if (!p->lanes[(p->last_lane + 1 + k) % kLanes].busy) lane = &p->lanes[(p->last_lane + 1 + k) % kLanes];
    if (!lane) return VBE_OK;
    try {
        check(hipSetDevice(p->ctx->device), "hipSetDevice");
        check_slots_free(p, slots, count);
        // an idle lane may change streams: its events are recorded afresh on the new one; with no
        // side stream the side work is ordered on the launch's own stream
        if (stream) {
            lane->stream = static_cast<hipStream_t>(stream);
            lane->side.st = side ? static_cast<hipStream_t>(side) : lane->stream;
        }
        size_t total = 0;
        for (int i = 0; i < count; ++i) {
            if (slots[i] < 0 || slots[i] >= p->capacity || !p->slots[slots[i]].loaded)
                throw StatusError(VBE_ERR_ARGUMENT, "slot not loaded");
            total += (size_t)p->n_act * p->slots[slots[i]].n_bf;
        }
        ensure_stage(p, *lane, total);
        std::vector<SlotInput> in;
        lane->grad_offsets.clear();
        size_t off = 0;
        for (int i = 0; i < count; ++i) {
            const size_t len = (size_t)p->n_act * p->slots[slots[i]].n_bf;
            std::memcpy(lane->h_C + off, C[i], len * sizeof(double));
            in.push_back({slots[i], lane->d_C + off, nullptr, gradient ? lane->d_grad + off : nullptr});
            lane->grad_offsets.push_back(off);
            off += len;
        }
        lane->grad_offsets.push_back(total);
        upload(lane->d_C, lane->h_C, total, lane->stream);
        lane->busy = true;
        try {
            enqueue(p, *lane, in, gradient, true);
            if (gradient) {
                check(hipMemcpyAsync(lane->h_grad, lane->d_grad, total * sizeof(double), hipMemcpyDeviceToHost,
                                     lane->stream),
                      "gradient D2H");
                check(hipEventRecord(lane->done, lane->stream), "hipEventRecord");
            }
        } catch (...) {
            (void)hipStreamSynchronize(lane->stream);
            lane->busy = false;
            throw;
        }
        lane->ticket = -1;  // owned by the scheduler, not a caller ticket
        p->last_lane = int(lane - p->lanes);
        *lane_out = p->last_lane;
        return VBE_OK;
    } catch (const StatusError& e) {
        *err = e.what();
        return e.status;
    } catch (const std::exception& e) {
        *err = e.what();
        return VBE_ERR_DEVICE;
    }
}

int pool_launch_query(vbe_pool* p, int lane, std::string* err) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    const hipError_t q = hipEventQuery(p->lanes[lane].done);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) {
        *err = std::string("hipEventQuery: ") + hipGetErrorString(q);
        return -VBE_ERR_DEVICE;
    }
    return 1;
}

int pool_launch_collect(vbe_pool* p, int lane_index, double* energies, double* const* grads, int* status,
                        float* device_ms, std::string* err) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    Lane& lane = p->lanes[lane_index];
    try {
        std::vector<int> st;
        const size_t count = lane.slots.size();
        finish(p, lane, &st);
        *device_ms = p->times.total_ms;
        for (size_t i = 0; i < count; ++i) {
            energies[i] = lane.h_energy[2 * i + 1];
            status[i] = st[i];
            if (lane.with_grad && grads && grads[i])
                std::memcpy(grads[i], lane.h_grad + lane.grad_offsets[i],
                            (lane.grad_offsets[i + 1] - lane.grad_offsets[i]) * sizeof(double));
        }
        return VBE_OK;
    } catch (const std::exception& e) {
        lane.busy = false;
        *err = e.what();
        return VBE_ERR_DEVICE;
    }
}

}  // namespace xmvb_amd

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

void vbe_context_destroy(vbe_context* ctx) { delete ctx; }

int vbe_set_ri_precision(vbe_context* ctx, int32_t precision) {
    if (!ctx) return fail(nullptr, VBE_ERR_ARGUMENT, "null context");
    if (precision != VBE_PRECISION_FP64 && precision != VBE_PRECISION_FP32_RI && precision != VBE_PRECISION_FP32C_RI)
        return fail(ctx, VBE_ERR_ARGUMENT, "unknown precision");
    ctx->ri_precision = precision;
    return VBE_OK;
}

int vbe_set_eigensolver(vbe_context* ctx, int32_t eigensolver, double energy_tolerance, double residual_tolerance,
                        double stationarity_tolerance, int32_t reuse_roots) {
    if (!ctx) return VBE_ERR_ARGUMENT;
    if (eigensolver != VBE_EIGENSOLVER_DENSE && eigensolver != VBE_EIGENSOLVER_DAVIDSON)
        return fail(ctx, VBE_ERR_ARGUMENT, "unknown eigensolver");
    if (!(energy_tolerance > 0.0) || !(residual_tolerance > 0.0) || !(stationarity_tolerance > 0.0) ||
        !std::isfinite(energy_tolerance) || !std::isfinite(residual_tolerance) || !std::isfinite(stationarity_tolerance))
        return fail(ctx, VBE_ERR_ARGUMENT, "Davidson accuracy tolerances must be positive");
    ctx->eigensolver = eigensolver;
    ctx->dav_energy_tol = energy_tolerance;
    ctx->dav_residual_tol = residual_tolerance;
    ctx->dav_stationarity_tol = stationarity_tolerance;
    ctx->dav_reuse_roots = reuse_roots != 0;
    return VBE_OK;
}

int vbe_set_states(vbe_context* ctx, int32_t n_states, const int32_t* state_indices, const double* weights) {
    if (!ctx) return fail(nullptr, VBE_ERR_ARGUMENT, "null context");
    if (n_states < 1 || n_states > xmvb_amd::kMaxStates)
        return fail(ctx, VBE_ERR_ARGUMENT, "n_states must lie in [1, " + std::to_string(xmvb_amd::kMaxStates) + "]");
    // mcscf::normalize_weights (src/mcscf/states/weights.cpp:7-33): non-negative, positive sum
    double sum = 0.0;
    for (int k = 0; k < n_states; ++k) {
        const double w = weights ? weights[k] : 1.0;
        if (!(w >= 0.0) || !std::isfinite(w)) return fail(ctx, VBE_ERR_ARGUMENT, "state_average_weights must be non-negative");
        sum += w;
    }
    if (!(sum > 0.0)) return fail(ctx, VBE_ERR_ARGUMENT, "state_average_weights must sum to a positive value");
    for (int k = 0; k < n_states; ++k) {
        const int idx = state_indices ? state_indices[k] : k;
        if (idx < 0) return fail(ctx, VBE_ERR_ARGUMENT, "selected state index is out of range");
        for (int j = 0; j < k; ++j)
            if ((state_indices ? state_indices[j] : j) == idx)
                return fail(ctx, VBE_ERR_ARGUMENT, "selected states must be distinct");
    }
    ctx->n_states = n_states;
    for (int k = 0; k < xmvb_amd::kMaxStates; ++k) {
        ctx->state_idx[k] = k < n_states ? (state_indices ? state_indices[k] : k) : 0;
        ctx->state_w[k] = k < n_states ? (weights ? weights[k] : 1.0) / sum : 0.0;
    }
    return VBE_OK;
}

int vbe_set_pair_condition_limit(vbe_context* ctx, double limit) {
    if (!ctx) return fail(nullptr, VBE_ERR_ARGUMENT, "null context");
    if (!(limit >= 0.0 && limit <= xmvb_amd::kLaneConditionMax))
        return fail(ctx, VBE_ERR_ARGUMENT, "pair condition limit must lie in [0, 1e9]");
    ctx->pair_condition_limit = limit;
    return VBE_OK;
}

int vbe_last_error(const vbe_context* ctx, char* buf, size_t len) {
    if (!buf || len == 0) return VBE_ERR_ARGUMENT;
    std::string e;
    if (ctx) {
        std::lock_guard<std::mutex> g(const_cast<vbe_context*>(ctx)->mu);
        e = ctx->error;
    } else {
        e = vbe_global_error();
    }
    const size_t n = std::min(len - 1, e.size());
    std::memcpy(buf, e.data(), n);
    buf[n] = '\0';
    return VBE_OK;
}

// ---- pools ----------------------------------------------------------------------------------

int vbe_pool_create(vbe_context* ctx, const vbe_topology_view* topo, int32_t n_act, int32_t n_alpha, int32_t n_beta,
                    int32_t capacity, vbe_pool** out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !topo || !out) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        *out = make_pool(ctx, *topo, n_act, n_alpha, n_beta, capacity).release();
        return VBE_OK;
    });
}

void vbe_pool_destroy(vbe_pool* pool) { delete pool; }

int vbe_pool_load(vbe_pool* pool, int32_t slot, int32_t n_bf, int32_t n_aux, const double* S_ao, const double* h_ao,
                  const double* B_packed, double e_nuc) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool) throw StatusError(VBE_ERR_ARGUMENT, "null pool");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        load_slot(pool, slot, n_bf, n_aux, S_ao, h_ao, B_packed, e_nuc);
        return VBE_OK;
    });
}

int vbe_pool_unload(vbe_pool* pool, int32_t slot) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool) throw StatusError(VBE_ERR_ARGUMENT, "null pool");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        if (slot < 0 || slot >= pool->capacity) throw StatusError(VBE_ERR_ARGUMENT, "slot index out of range");
        check_slots_free(pool, &slot, 1);
        check(hipSetDevice(pool->ctx->device), "hipSetDevice");
        // launches that read the slot were ordered on the lanes' streams and have completed
        // (no pending submission holds it); a synchronous evaluation returns only when done
        Slot& s = pool->slots[slot];
        s.arena.reset();
        s.desc = SlotDesc{};
        s.C_buf = s.grad_buf = nullptr;
        s.support_buf = nullptr;
        s.S_host.clear();
        s.S_host.shrink_to_fit();
        s.n_bf = s.n_aux = 0;
        s.has_root = false;
        s.loaded = false;
        return VBE_OK;
    });
}

int vbe_context_memory(const vbe_context* ctx, uint64_t* in_use, uint64_t* budget) {
    if (!ctx || !in_use || !budget) return VBE_ERR_ARGUMENT;
    std::lock_guard<std::mutex> g(const_cast<vbe_context*>(ctx)->mu);
    *in_use = ctx->in_use;
    *budget = ctx->budget;
    return VBE_OK;
}

int vbe_pool_device_bytes(const vbe_pool* pool, uint64_t* bytes) {
    if (!pool || !bytes) return VBE_ERR_ARGUMENT;
    uint64_t b = pool->topo_arena.bytes + pool->ri_stage_arena.bytes + pool->opt_arena.bytes;
    for (const auto& s : pool->slots)
        if (s.arena) b += s.arena->bytes;
    *bytes = b;
    return VBE_OK;
}

int vbe_pool_eval_device(vbe_pool* pool, int32_t count, const int32_t* slots, const double* d_C,
                         int32_t want_gradient, double* energies, double* d_grad) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool || !slots || !d_C || !energies || count <= 0) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        if (want_gradient && !d_grad) throw StatusError(VBE_ERR_ARGUMENT, "null gradient buffer");
        check(hipSetDevice(pool->ctx->device), "hipSetDevice");
        Lane& lane = acquire_lane(pool);
        check_slots_free(pool, slots, count);
        std::vector<SlotInput> in;
        size_t off = 0;
        for (int i = 0; i < count; ++i) {
            const int s = slots[i];
            if (s < 0 || s >= pool->capacity || !pool->slots[s].loaded)
                throw StatusError(VBE_ERR_ARGUMENT, "slot not loaded");
            in.push_back({s, d_C + off, nullptr, want_gradient ? d_grad + off : nullptr});
            off += (size_t)pool->n_act * pool->slots[s].n_bf;
        }
        lane.busy = true;
        try {
            enqueue(pool, lane, in, want_gradient != 0, true);
        } catch (...) {
            lane.busy = false;
            throw;
        }
        pool->last_lane = int(&lane - pool->lanes);
        finish(pool, lane);
        for (int i = 0; i < count; ++i) energies[i] = lane.h_energy[2 * i + 1];
        return VBE_OK;
    });
}

int vbe_pool_submit(vbe_pool* pool, int32_t count, const int32_t* slots, const double* C, int32_t want_gradient,
                    int32_t* ticket) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool || !slots || !C || !ticket || count <= 0) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        if (count > pool->capacity) throw StatusError(VBE_ERR_ARGUMENT, "more slots than the pool capacity");
        check(hipSetDevice(pool->ctx->device), "hipSetDevice");
        Lane& lane = acquire_lane(pool);
        check_slots_free(pool, slots, count);
        size_t total = 0;
        for (int i = 0; i < count; ++i) {
            if (slots[i] < 0 || slots[i] >= pool->capacity || !pool->slots[slots[i]].loaded)
                throw StatusError(VBE_ERR_ARGUMENT, "slot not loaded");
            total += (size_t)pool->n_act * pool->slots[slots[i]].n_bf;
        }
        ensure_stage(pool, lane, total);
        std::memcpy(lane.h_C, C, total * sizeof(double));
        upload(lane.d_C, lane.h_C, total, lane.stream);
        std::vector<SlotInput> in;
        lane.grad_offsets.clear();
        size_t off = 0;
        for (int i = 0; i < count; ++i) {
            in.push_back({slots[i], lane.d_C + off, nullptr, want_gradient ? lane.d_grad + off : nullptr});
            lane.grad_offsets.push_back(off);
            off += (size_t)pool->n_act * pool->slots[slots[i]].n_bf;
        }
        lane.grad_offsets.push_back(total);
        lane.busy = true;
        try {
            enqueue(pool, lane, in, want_gradient != 0, true);
            if (want_gradient)
                check(hipMemcpyAsync(lane.h_grad, lane.d_grad, total * sizeof(double), hipMemcpyDeviceToHost,
                                     lane.stream),
                      "gradient D2H");
            check(hipEventRecord(lane.done, lane.stream), "hipEventRecord");
        } catch (...) {
            (void)hipStreamSynchronize(lane.stream);
            lane.busy = false;
            throw;
        }
        lane.ticket = pool->next_ticket++;
        pool->last_lane = int(&lane - pool->lanes);
        *ticket = lane.ticket;
        return VBE_OK;
    });
}

int vbe_pool_poll(vbe_pool* pool, int32_t ticket, int32_t* done) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool || !done) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        for (auto& lane : pool->lanes) {
            if (!lane.busy || lane.ticket != ticket) continue;
            const hipError_t q = hipEventQuery(lane.done);
            if (q == hipErrorNotReady) {
                *done = 0;
                return VBE_OK;
            }
            check(q, "hipEventQuery");
            *done = 1;
            return VBE_OK;
        }
        throw StatusError(VBE_ERR_ARGUMENT, "unknown or already completed ticket");
    });
}

int vbe_pool_wait(vbe_pool* pool, int32_t ticket, double* energies, double* grads) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool || !energies) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        for (auto& lane : pool->lanes) {
            if (!lane.busy || lane.ticket != ticket) continue;
            const size_t count = lane.slots.size();
            const bool g = lane.with_grad;
            finish(pool, lane);
            for (size_t i = 0; i < count; ++i) energies[i] = lane.h_energy[2 * i + 1];
            if (g && grads) std::memcpy(grads, lane.h_grad, lane.grad_offsets.back() * sizeof(double));
            return VBE_OK;
        }
        throw StatusError(VBE_ERR_ARGUMENT, "unknown or already completed ticket");
    });
}

int vbe_pool_autotune(vbe_pool* pool, int32_t count, const int32_t* slots, const double* C, int32_t repeats,
                      int32_t* fwd_target, int32_t* bwd_target, double* table) {
    if (!pool || !slots || !C || count <= 0 || repeats <= 0 || !fwd_target || !bwd_target) {
        if (pool) fail(pool->ctx, VBE_ERR_ARGUMENT, "vbe_pool_autotune: invalid argument");
        return VBE_ERR_ARGUMENT;
    }
    std::lock_guard<std::recursive_mutex> lk(pool->mu);
    size_t total = 0;
    for (int i = 0; i < count; ++i) {
        if (slots[i] < 0 || slots[i] >= pool->capacity || !pool->slots[slots[i]].loaded)
            return fail(pool->ctx, VBE_ERR_ARGUMENT, "slot not loaded");
        total += (size_t)pool->n_act * pool->slots[slots[i]].n_bf;
    }
    std::vector<double> e(count), g(total);
    auto run = [&](double& t_fwd, double& t_bwd) -> int {
        int32_t ticket = 0;
        int rc = vbe_pool_submit(pool, count, slots, C, 1, &ticket);
        if (rc != VBE_OK) return rc;
        rc = vbe_pool_wait(pool, ticket, e.data(), g.data());
        t_fwd = pool->times.pair_forward_ms;
        t_bwd = pool->times.pair_backward_ms;
        return rc;
    };
    const int saved_f = pool->fwd_target, saved_b = pool->bwd_target;
    double tf = 0.0, tb = 0.0;
    int rc = run(tf, tb);  // warm-up at the current shape
    double best_f = 1e300, best_b = 1e300;
    int pick_f = saved_f, pick_b = saved_b;
    for (int c = 0; c < VBE_AUTOTUNE_CANDIDATES && rc == VBE_OK; ++c) {
        pool->fwd_target = pool->bwd_target = kAutotuneTargets[c];
        double mf = 1e300, mb = 1e300;
        for (int r = 0; r < repeats && rc == VBE_OK; ++r) {
            rc = run(tf, tb);
            mf = std::min(mf, tf);
            mb = std::min(mb, tb);
        }
        if (table) {
            table[2 * c] = mf;
            table[2 * c + 1] = mb;
        }
        if (mf < best_f) best_f = mf, pick_f = kAutotuneTargets[c];
        if (mb < best_b) best_b = mb, pick_b = kAutotuneTargets[c];
    }
    if (rc != VBE_OK) {
        pool->fwd_target = saved_f;
        pool->bwd_target = saved_b;
        return rc;
    }
    pool->fwd_target = pick_f;
    pool->bwd_target = pick_b;
    *fwd_target = pick_f;
    *bwd_target = pick_b;
    return VBE_OK;
}

int vbe_pool_phase_times(const vbe_pool* pool, vbe_phase_times* times) {
    if (!pool || !times) return VBE_ERR_ARGUMENT;
    *times = pool->times;
    return VBE_OK;
}

// ---- single problems (a pool of one slot) -------------------------------------------------

int vbe_problem_upload(vbe_context* ctx, const vbe_dims* dims, const double* S_ao, const double* h_ao,
                       const double* B_packed, const vbe_topology_view* topo, double e_nuc, vbe_problem** out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !dims || !S_ao || !h_ao || !B_packed || !topo || !out)
            throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (dims->n_bf <= 0 || dims->n_aux <= 0) throw StatusError(VBE_ERR_ARGUMENT, "dimensions must be positive");
        auto prob = std::make_unique<vbe_problem>();
        prob->pool = make_pool(ctx, *topo, dims->n_act, dims->n_alpha, dims->n_beta, 1);
        load_slot(prob->pool.get(), 0, dims->n_bf, dims->n_aux, S_ao, h_ao, B_packed, e_nuc);
        *out = prob.release();
        return VBE_OK;
    });
}

void vbe_problem_release(vbe_problem* prob) { delete prob; }

int vbe_problem_device_bytes(const vbe_problem* prob, uint64_t* bytes) {
    if (!prob || !bytes) return VBE_ERR_ARGUMENT;
    return vbe_pool_device_bytes(prob->pool.get(), bytes);
}

int vbe_eval(vbe_context* ctx, vbe_problem* prob, const double* C_raw, const uint8_t* support, vbe_flags flags,
             double* E_total, double* E_states, double* grad_raw) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !C_raw || !E_total) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (prob->pool->ctx != ctx) throw StatusError(VBE_ERR_ARGUMENT, "problem does not belong to this context");
        check_flags_args(flags, prob->pool.get());
        prob->pool->eigensolver = flags.eigensolver;
        if (flags.want_gradient && !grad_raw) throw StatusError(VBE_ERR_ARGUMENT, "null gradient buffer");
        vbe_pool* p = prob->pool.get();
        Slot& s = p->slots[0];
        const size_t nc = (size_t)p->n_act * s.n_bf;
        hipStream_t st = p->lanes[0].stream;
        check(hipSetDevice(ctx->device), "hipSetDevice");
        upload(s.C_buf, C_raw, nc, st);
        if (support) upload(s.support_buf, support, nc, st);
        *E_total = eval_problem(prob, s.C_buf, support ? s.support_buf : nullptr, flags.want_gradient != 0,
                                s.grad_buf, true);
        if (E_states) E_states[0] = *E_total;
        if (flags.want_gradient) {
            check(hipMemcpyAsync(grad_raw, s.grad_buf, nc * sizeof(double), hipMemcpyDeviceToHost, st), "gradient D2H");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
        }
        return VBE_OK;
    });
}

int vbe_eval_device(vbe_context* ctx, vbe_problem* prob, const double* d_C_raw, const uint8_t* d_support,
                    vbe_flags flags, double* E_total, double* d_grad_raw) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !d_C_raw || !E_total) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (prob->pool->ctx != ctx) throw StatusError(VBE_ERR_ARGUMENT, "problem does not belong to this context");
        check_flags_args(flags, prob->pool.get());
        prob->pool->eigensolver = flags.eigensolver;
        if (flags.want_gradient && !d_grad_raw) throw StatusError(VBE_ERR_ARGUMENT, "null gradient buffer");
        *E_total = eval_problem(prob, d_C_raw, d_support, flags.want_gradient != 0, d_grad_raw, true);
        return VBE_OK;
    });
}

int vbe_eval_forward(vbe_context* ctx, vbe_problem* prob, const double* C_raw, const uint8_t* support,
                     double* E_total) {
    vbe_flags f{prob ? prob->pool->ri_precision : VBE_PRECISION_FP64,
                prob ? prob->pool->eigensolver : VBE_EIGENSOLVER_DENSE, 0};
    return vbe_eval(ctx, prob, C_raw, support, f, E_total, nullptr, nullptr);
}

int vbe_complete_gradient(vbe_context* ctx, vbe_problem* prob, double* grad_raw) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !grad_raw) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (!prob->forward_valid) throw StatusError(VBE_ERR_ARGUMENT, "no retained forward evaluation");
        vbe_pool* p = prob->pool.get();
        Slot& s = p->slots[0];
        // the retained descriptor table on the device still describes this slot's forward; the
        // backward writes the gradient through its grad pointer, which is the slot buffer for
        // host-buffer evaluations
        p->lanes[0].h_desc[0].grad = s.grad_buf;
        upload(p->lanes[0].d_desc, p->lanes[0].h_desc, 1, p->lanes[0].stream);
        eval_problem(prob, s.C_buf, p->lanes[0].h_desc[0].support, true, s.grad_buf, false);
        const size_t nc = (size_t)p->n_act * s.n_bf;
        hipStream_t st = p->lanes[0].stream;
        check(hipMemcpyAsync(grad_raw, s.grad_buf, nc * sizeof(double), hipMemcpyDeviceToHost, st), "gradient D2H");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        return VBE_OK;
    });
}

int vbe_last_intermediates(vbe_context* ctx, vbe_problem* prob, double* H, double* S, double* coef, double* S_act,
                           double* h_act, double* g_packed) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (!prob->forward_valid) throw StatusError(VBE_ERR_ARGUMENT, "no retained forward evaluation");
        vbe_pool* p = prob->pool.get();
        const SlotDesc& d = p->slots[0].desc;
        hipStream_t st = p->lanes[0].stream;
        const size_t ns = p->topo.n_str, n = p->n_act;
        auto get = [&](double* dst, const double* src, size_t count) {
            if (dst) check(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
        };
        get(H, d.Hs, ns * ns);
        get(S, d.Ss, ns * ns);
        get(coef, d.coef, ns);
        get(S_act, d.sact, n * n);
        get(h_act, d.hact, n * n);
        get(g_packed, d.G, p->topo.ng);
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        return VBE_OK;
    });
}

void vbe_optimizer_defaults(vbe_optimizer_options* options) {
    if (options) xmvb_amd::optimizer_defaults(options);
}

int vbe_optimize(vbe_context* ctx, vbe_problem* prob, double* C_raw, const uint8_t* support,
                 const vbe_optimizer_options* options, vbe_optimizer_result* result, double* grad_raw,
                 double* energy_trace, int32_t trace_capacity) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !C_raw || !options || !result) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        if (prob->pool->ctx != ctx) throw StatusError(VBE_ERR_ARGUMENT, "problem does not belong to this context");
        if (trace_capacity < 0) throw StatusError(VBE_ERR_ARGUMENT, "negative trace capacity");
        vbe_pool* p = prob->pool.get();
        xmvb_amd::OptProblem op;
        op.n_bf = p->slots[0].n_bf;
        op.S_ao = p->slots[0].S_host.data();
        op.support = support;
        op.C = C_raw;
        op.grad = grad_raw;
        op.trace = energy_trace;
        op.trace_capacity = trace_capacity;
        op.result = result;
        std::vector<xmvb_amd::OptProblem> batch{op};
        optimize_slots(p, prob, {0}, batch, *options);
        return VBE_OK;
    });
}

int vbe_pool_optimize(vbe_pool* pool, int32_t count, const int32_t* slots, double* C_raw, const uint8_t* support,
                      const vbe_optimizer_options* options, vbe_optimizer_result* results, double* grad_raw) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool || !slots || !C_raw || !options || !results || count <= 0)
            throw StatusError(VBE_ERR_ARGUMENT, "null argument or empty slot list");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        if (count > pool->capacity) throw StatusError(VBE_ERR_ARGUMENT, "more slots than the pool holds");
        for (const auto& l : pool->lanes)
            if (l.busy) throw StatusError(VBE_ERR_ARGUMENT, "the pool has submissions in flight; wait for them first");
        std::vector<int> list(slots, slots + count);
        std::vector<char> seen(pool->capacity, 0);
        std::vector<xmvb_amd::OptProblem> batch(count);
        size_t off = 0;
        for (int j = 0; j < count; ++j) {
            const int sl = list[j];
            if (sl < 0 || sl >= pool->capacity || !pool->slots[sl].loaded)
                throw StatusError(VBE_ERR_ARGUMENT, "slot not loaded");
            if (seen[sl]) throw StatusError(VBE_ERR_ARGUMENT, "a slot appears twice");
            seen[sl] = 1;
            const Slot& s = pool->slots[sl];
            auto& op = batch[j];
            op.n_bf = s.n_bf;
            op.S_ao = s.S_host.data();
            op.support = support ? support + off : nullptr;
            op.C = C_raw + off;
            op.grad = grad_raw ? grad_raw + off : nullptr;
            op.result = results + j;
            off += (size_t)pool->n_act * s.n_bf;
        }
        optimize_slots(pool, nullptr, list, batch, *options);
        return VBE_OK;
    });
}

namespace {

// total energies (E_nuc added) of the selected states of a slot's last evaluation
void state_energies(vbe_pool* p, int slot, double e_total, double* out) {
    const auto& T = p->topo;
    if (T.n_states == 1) {
        out[0] = e_total;
        return;
    }
    const Slot& s = p->slots[slot];
    check(hipSetDevice(p->ctx->device), "hipSetDevice");
    check(hipMemcpy(out, s.desc.state_energy, sizeof(double) * T.n_states, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (int k = 0; k < T.n_states; ++k) out[k] += s.e_nuc;
}

}  // namespace

int vbe_last_state_energies(vbe_context* ctx, vbe_problem* prob, double* E_states) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !prob || !E_states) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        vbe_pool* p = prob->pool.get();
        if (!prob->forward_valid) throw StatusError(VBE_ERR_ARGUMENT, "no evaluation of this problem to report");
        state_energies(p, 0, p->lanes[0].h_energy[1], E_states);
        return VBE_OK;
    });
}

int vbe_pool_state_energies(vbe_pool* pool, int32_t slot, double* E_states) {
    return guarded(pool ? pool->ctx : nullptr, [&]() -> int {
        if (!pool || !E_states) throw StatusError(VBE_ERR_ARGUMENT, "null argument");
        std::lock_guard<std::recursive_mutex> lk(pool->mu);
        if (slot < 0 || slot >= pool->capacity || !pool->slots[slot].loaded)
            throw StatusError(VBE_ERR_ARGUMENT, "slot not loaded");
        check_slots_free(pool, &slot, 1);
        if (pool->topo.n_states == 1)
            throw StatusError(VBE_ERR_ARGUMENT, "one selected state: its energy is the evaluation's energy");
        state_energies(pool, slot, 0.0, E_states);
        return VBE_OK;
    });
}

int vbe_last_phase_times(const vbe_problem* prob, vbe_phase_times* times) {
    if (!prob || !times) return VBE_ERR_ARGUMENT;
    *times = prob->pool->times;
    return VBE_OK;
}

int vbe_last_singular_pairs(const vbe_problem* prob, int64_t* count) {
    if (!prob || !count) return VBE_ERR_ARGUMENT;
    *count = prob->pool->last_singular;
    return VBE_OK;
}

}  // extern "C"
