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
// Largest occupied block (alpha or beta electrons): one element per lane of a 64-wide wavefront.
constexpr int MAX_SPIN_BLOCK = 8;

// device error flags (per slot)
enum DeviceFlag : int {
    FLAG_NORMALIZATION = 0,   // 1 + index of an orbital whose S-norm vanished
    FLAG_ACTIVE_OVERLAP = 1,  // S_act not positive definite
    FLAG_SINGULAR_PAIRS = 2,  // rank-deficient determinant overlap blocks met
    FLAG_METRIC = 3,          // structure metric not positive definite
    FLAG_COUNT = 8
};

// HIP events recorded between phases (optional)
enum PhaseEvent : int {
    EV_START, EV_PREP, EV_ONE_INDEX, EV_TRANSFORM, EV_PAIR_FWD, EV_STRUCT, EV_EIGEN, EV_BWD_BEGIN, EV_BWD_KERNEL,
    EV_PAIR_BWD, EV_TAIL, EV_COUNT
};

struct SlotDesc {
    // molecule-static data
    const double* S;
    const double* h;
    const double* B;
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
    double *Hdet, *Sdet, *Yh, *Ys, *Hs, *Ss, *gep, *coef, *dcoef;
    double *partials, *dsum, *dL, *gcn;
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
    const int* sdet;
    const double* scoef;
};

struct LaunchShape {
    int max_n_bf, max_n_aux;
    int pair_blocks_fwd, pair_blocks_bwd;  // blocks per slot
    int mmax;                              // largest electron count of one spin
};

inline std::size_t gep_work_doubles(int n) { return 3 * (std::size_t)n * n + 9 * (std::size_t)n; }

int pair_group(int mmax);
bool pairs_full_dg(int n, int ng);
std::size_t gep_lds_bytes(int n);
bool gep_in_lds(int n);
hipError_t configure_kernels(int n, int ng, int n_str, int mmax);

// forward: orbital prep, RI transform, determinant pairs, structure H/S, eigensolve (+ a_d)
void launch_forward(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape,
                    hipEvent_t* ev);
// backward: pair adjoints, partial reduction, RI/orbital pullback to the raw-slot gradient
void launch_backward(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape,
                     hipEvent_t* ev);

}  // namespace xmvb_amd
