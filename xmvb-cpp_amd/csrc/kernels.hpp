// Device-side descriptors and launch wrappers of the engine's HIP kernels (kernels.hip).
//
// Every kernel evaluates a set of resident problem slots at once: blockIdx.z selects an entry
// of a per-launch SlotDesc table (device memory), which carries that slot's static data,
// current orbitals, workspace and outputs. All slots of one launch share a TopoDesc (same
// structure set), while n_bf and n_aux may differ per slot.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace xmvb_amd {

// Largest active space handled by the LDS-resident determinant-pair kernels.
constexpr int MAX_ACT = 14;
// Largest number of selected structure states of a state-averaged evaluation.
constexpr int kMaxStates = 8;
// Largest occupied block (alpha or beta electrons): one element per lane of a 64-wide wavefront.
constexpr int MAX_SPIN_BLOCK = 8;

// device error flags (per slot)
enum DeviceFlag : int {
    FLAG_NORMALIZATION = 0,   // 1 + index of an orbital whose S-norm vanished
    FLAG_ACTIVE_OVERLAP = 1,  // S_act not positive definite
    FLAG_SINGULAR_PAIRS = 2,  // pairs with a rank-deficient overlap block (SlotDesc::singular)
    FLAG_METRIC = 3,          // structure metric not positive definite
    FLAG_DEFERRED = 4,        // pairs left by the lane kernel for the lane-group kernel (SlotDesc::defer)
    FLAG_EIGEN = 5,           // Davidson stopped without meeting its tolerances (1 + reason)
    FLAG_COUNT = 8
};

// HIP events recorded between phases (optional)
enum PhaseEvent : int {
    EV_START, EV_PREP, EV_ONE_INDEX, EV_TRANSFORM, EV_PAIR_FWD_MAIN, EV_PAIR_FWD, EV_STRUCT, EV_EIGEN, EV_BWD_BEGIN,
    EV_BWD_FIELD, EV_BWD_ADJOINT, EV_BWD_KERNEL, EV_PAIR_BWD, EV_TAIL, EV_COUNT
};

struct SlotDesc {
    // molecule-static data
    const double* S;
    const double* h;
    const double* B;         // RI factors, FP64 storage (or null)
    const float* B32;        // RI factors, FP32 storage or the FP32 part of compensated storage (or null)
    const signed char* Blo;  // compensated storage: 8-bit correction of each FP32 element (or null)
    int b_tiled;             // B, B32, Blo in the aux-tiled layout (ri_tiled.hip) instead of n_aux x npair
    int n_bf, n_aux;
    double e_nuc;
    // current evaluation: inputs and outputs
    const double* C;
    const uint8_t* support;
    double* grad;
    double* energy;  // [0] eigenvalue, [1] total energy
    int* flags;      // FLAG_COUNT ints
    // workspace
    double *Cn, *SC, *HC, *nu2, *sact, *hact, *T, *L, *G;
    double* CnT;  // Cn as [mu][kCnStride] for rows mu < n_bf rounded up to 32, zeros past n and n_bf
    double *Hdet, *Sdet, *Yh, *Ys, *Hs, *Ss, *gep, *coef, *dcoef;
    double *partials, *dsum, *dL, *gcn;
    int* defer;     // deferred unordered pair indices (capacity: all pairs)
    int* singular;  // pairs with a rank-deficient block (capacity: all pairs)
    uint8_t* defer_flag;     // per lane work item: deferred by the lane kernel (listed into defer in order)
    uint8_t* singular_flag;  // per lane-group entry: rank deficient (listed into singular in order)
    int* chunk_count;        // list compaction: entries listed per chunk of kCompactChunk, then offsets
    double* state_energy;    // eigenvalue of each selected state (kMaxStates; n_states > 1 only)
    int seeded;     // coef holds this slot's previous structure eigenvector (Davidson start)
};

struct TopoDesc {
    int n, np, ng, n_det, n_str;
    const uint64_t* amask;
    const uint64_t* bmask;
    const int* det_a;
    const int* det_b;
    const long long* off;   // determinant -> terms CSR
    const int* ts;
    const double* tc;
    const long long* soff;  // structure -> determinants CSR
    long long pair_mul;     // backward lane order: work item i is taken as (i * pair_mul) mod n_lane_pairs (coprime)
    const int* sdet;
    const double* scoef;
    // Spin-flip pairing (n_alpha == n_beta and a determinant set closed under exchanging the alpha
    // and beta strings): flip[d] is that partner of determinant d. Pairs {d, e} and {flip d, flip e}
    // have the same two blocks with the spins exchanged, hence the same H, S and the same adjoint
    // contributions; the lane kernels evaluate the one with the smaller packed index (listed in
    // lane_pairs) for both. flip == nullptr: every pair is evaluated (lane_pairs == nullptr).
    const int* flip;
    const int* lane_pairs;
    long long n_lane_pairs;  // work items of the lane kernels
    // Selected structure states (roots by index, ascending eigenvalue) and their normalised weights.
    // One state: dcoef holds its determinant coefficients and energy[0] its eigenvalue. Several:
    // dcoef holds n_states blocks of n_det, coef n_states blocks of n_str, state_energy the
    // eigenvalues and energy[0] their weighted sum.
    int n_states;
    int state_idx[kMaxStates];
    double state_w[kMaxStates];
};

struct LaunchShape {
    int max_n_bf, max_n_aux;
    int pair_blocks_fwd, pair_blocks_bwd;  // blocks per slot
    int mmax;                              // largest electron count of one spin
    int n_alpha, n_beta;
    bool lane_fwd, lane_bwd;               // one-pair-per-lane kernels for this block shape
    int lane_blocks_fwd, lane_blocks_bwd;  // their blocks per slot
    int list_blocks;                       // lane-group blocks per slot for the deferred pairs
    double lane_cond_limit;                // lane kernel keeps blocks with max|X| max|X^-1| below this
    bool b32;                              // RI factors stored in FP32
    bool b32c;                             // RI factors stored as compensated FP32 (FP32 + 8-bit correction)
    bool any_tiled, any_flat;              // slots of the launch with the aux-tiled / the flat layout
    int max_n_bf_tiled;                    // largest n_bf among the slots with the aux-tiled layout
    bool davidson;                         // structure eigenproblem by Davidson instead of dense
    // Pinned launch shape: every choice that changes a slot's summation order (blocks per slot of the
    // backward pair kernels, the few-problem or batched form of the Gram product, gradient tail and
    // large eigensolve) is made independently of how many slots the launch holds, so a slot's
    // energy and gradient are the same bits alone and in any batch.
    bool pinned;
    double dav_energy_tol, dav_residual_tol, dav_stationarity_tol;
};

// Aux-tiled RI factors (ri_tiled.hip): tiles of kRiTileAux aux indices, used for n_bf up to
// 32 * kRiTiledMaxBands (the one-index transform then reads every packed element once).
constexpr int kRiTileAux = 8;
constexpr int kCnStride = 16;  // row length of SlotDesc::CnT (>= MAX_ACT)
constexpr int kRiTiledMaxBands = 7;
constexpr int kRiTiledMaxBf = 32 * kRiTiledMaxBands;
hipError_t configure_ri_tiled();
void launch_ri_one_index_tiled(hipStream_t st, const SlotDesc* slots, int count, int n, int max_n_bf, int max_n_aux,
                               bool compensated);
// Writes flat FP64 factors (n_aux x npair) in a slot's storage: FP64 tiled (B), or compensated FP32
// (B32 + Blo) tiled or flat, whichever pointers are non-null.
void launch_ri_relayout(hipStream_t st, const double* flat, int n_aux, long long npair, bool tiled, double* B,
                        float* B32, signed char* Blo);

// largest condition bound under which a row-pivoted lane block is provably regular under the
// complete-pivoting rank rule (see pairs_lane.hip)
constexpr double kLaneConditionMax = 1.0e9;
// default bound: blocks conditioned worse than this follow the reference's pivot order exactly, so
// their rounding matches the complete-pivoting LU closely (typically well under 1% of all pairs)
constexpr double kLaneConditionDefault = 1.0e5;

// blocks per slot that evaluate the deferred-pair list (and own partial rows after the lane kernel's):
// at most kListBlocks; kListBlocksBatched in launches of many problems. A list block whose first
// pair lies beyond the list does no work and writes no partial row (the reduction skips it).
constexpr int kListBlocks = 64;
constexpr int kListBlocksBatched = 4;
// entries per chunk of the ordered pair-list compaction; chunks handled per launch per slot at most
constexpr int kCompactChunk = 256;
constexpr int kCompactBlocks = 64;

// blocks per slot of the singular-pair kernel (partial rows after the pair kernels' rows)
constexpr int kSingularBlocks = 4;

// Front half of the eigensolve in global memory (Cholesky of S, L^-1, A = L^-1 H L^-T; SlotDesc::gep
// holds L at 0, A at n^2, L^-1 column-major at 2 n^2); per_step spreads the Cholesky of large spaces
// over many workgroups.
void launch_gep_front(hipStream_t st, const SlotDesc* slots, unsigned z, const TopoDesc& topo, bool per_step);

// dense eigensolver workspace: 3 n^2 + 9 n doubles (in LDS for small n), plus two n x n scratch
// matrices for the multi-workgroup front half of the large-n path
inline std::size_t gep_core_doubles(int n) { return 3 * (std::size_t)n * n + 9 * (std::size_t)n; }
inline std::size_t gep_work_doubles(int n) { return gep_core_doubles(n) + 2 * (std::size_t)n * n; }

// Davidson subspace limit for one root: min(n, max(32, 8 n_roots)) (make_davidson_options,
// linalg/eigensolver.cpp:321-361)
constexpr int kDavidsonMaxSubspace = 32;
// lowest structure root by Davidson in the metric's factored coordinates (davidson.hip): the
// eigensolver's front half forms A = L^-1 H L^-T, the iteration works from 3 n^2 on in SlotDesc::gep
void launch_davidson(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape);

int pair_group(int mmax);
bool pairs_lane_supported(int n_alpha, int n_beta, bool backward);
std::size_t pairs_lane_lds_bytes(int n, int ng, bool backward, int m0);
hipError_t configure_pairs_lane();
std::size_t pairs_singular_lds_bytes(int n, int ng, bool backward);
hipError_t configure_pairs_singular();
void launch_pairs_singular(hipStream_t st, bool backward, int part_offset, const SlotDesc* slots, int count,
                           const TopoDesc& topo);
// which: 0 forward, 1 backward F blocks + S_act/h_act adjoints, 2 backward packed-kernel adjoint
void launch_pairs_lane(hipStream_t st, int which, int blocks, const SlotDesc* slots, int count, const TopoDesc& topo,
                       int n_alpha, int n_beta, double cond_limit);
bool pairs_full_dg(int n, int ng);
std::size_t gep_lds_bytes(int n);
bool gep_in_lds(int n);
// Lowest root of one selected state for n_str <= 64: one wavefront per problem, columns in
// registers (gep_wave.hip). Returns false, launching nothing, for larger structure spaces.
bool launch_gep_wave(hipStream_t st, const SlotDesc* slots, unsigned z, const TopoDesc& topo);
hipError_t configure_kernels(int n, int ng, int n_str, int mmax);

// A second stream for launches of few problems: kernels with disjoint outputs that each fill only
// a corner of the device run on it concurrently with the main stream (fork: an event recorded on
// the main stream that the side stream waits for; join: the reverse). Results do not depend on it.
struct SideStream {
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kForkMaxProblems = 8;  // launches of more problems fill the device kernel by kernel

// forward: orbital prep, RI transform, determinant pairs, structure H/S, eigensolve (+ a_d)
void launch_forward(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape,
                    hipEvent_t* ev, const SideStream* side = nullptr);
// backward: pair adjoints, partial reduction, RI/orbital pullback to the raw-slot gradient
void launch_backward(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape,
                     hipEvent_t* ev, const SideStream* side = nullptr);

}  // namespace xmvb_amd
