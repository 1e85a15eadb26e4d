// Launch wrappers of the engine's HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace xmvb_amd {

// Largest active space handled by the LDS-resident determinant-pair kernels.
constexpr int MAX_ACT = 14;
// Largest occupied block (alpha or beta electrons): one element per lane of a 64-wide wavefront.
constexpr int MAX_SPIN_BLOCK = 8;

// device error flags
enum DeviceFlag : int {
    FLAG_NORMALIZATION = 0,   // 1 + index of an orbital whose S-norm vanished
    FLAG_ACTIVE_OVERLAP = 1,  // S_act not positive definite
    FLAG_SINGULAR_PAIRS = 2,  // rank-deficient determinant overlap blocks met
    FLAG_METRIC = 3,          // structure metric not positive definite
    FLAG_COUNT = 8
};

inline std::size_t gep_work_doubles(int n) { return 3 * (std::size_t)n * n + 9 * (std::size_t)n; }

std::size_t pairs_lds_bytes(int n, int ng, bool backward);
hipError_t configure_kernels(int n, int ng, int n_str);

void launch_orbital_prep(hipStream_t st, const double* C, const uint8_t* support, const double* S, const double* h,
                         int n_bf, int n, double* Cn, double* SC, double* HC, double* nu2, int* flags);
void launch_active_matrices(hipStream_t st, const double* Cn, const double* SC, const double* HC, int n_bf, int n,
                            double* sact, double* hact, int* flags);
void launch_ri_one_index(hipStream_t st, const double* B, const double* Cn, int n_bf, int n, int n_aux, double* T);
void launch_ri_pair_stage(hipStream_t st, const double* T, const double* Cn, int n_bf, int n, int n_aux, double* L,
                          double* G);
void launch_pairs(hipStream_t st, bool backward, int n_blocks, const double* G, const double* sact, const double* hact,
                  int n, int ng, const uint64_t* amask, const uint64_t* bmask, const int* det_a, const int* det_b,
                  int n_det, double* Hdet, double* Sdet, const double* dcoef, const double* energy, double* partials,
                  int* flags);
void launch_reduce_partials(hipStream_t st, const double* partials, int n_blocks, int len, double* out);
void launch_structure(hipStream_t st, const double* Hdet, const double* Sdet, int n_det, int n_str,
                      const long long* soff, const int* sdet, const double* scoef, double* Yh, double* Ys, double* Hs,
                      double* Ss);
void launch_gep(hipStream_t st, const double* Hs, const double* Ss, int n, double e_nuc, double* work, double* coef,
                double* energy, int* flags, const long long* off, const int* ts, const double* tc, int n_det,
                double* dcoef);
void launch_det_coefficients(hipStream_t st, const long long* off, const int* ts, const double* tc, const double* coef,
                             int n_det, double* a);
void launch_gradient_tail(hipStream_t st, const double* T, const double* L, const double* dsum, const double* SC,
                          const double* HC, const double* Cn, const double* nu2, const uint8_t* support, int n_bf,
                          int n, int n_aux, double* dL, double* gcn, double* grad);

}  // namespace xmvb_amd
