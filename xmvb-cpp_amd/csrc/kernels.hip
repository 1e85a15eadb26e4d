// HIP/CDNA4 (gfx950) kernels of the VBSCF energy+gradient engine, FP64 path.
//
// Phases of one evaluation and the reference functions they replace:
//   k_orbital_prep        normalize_sparse_orbitals (orbitals/preparation/preparer.cpp:45-94)
//   k_active_matrices     S_act = C^T S C (preparer.cpp:371-375), raw h_act
//   k_ri_one_index        transform_ri_one_index (integrals/ao/ri_transform.cpp:10-37)
//   k_ri_pair_factors     transform_ri_active_integrals (integrals/active/ri.cpp:34-96), raw frame
//   k_ri_pair_gram        pack_ri_active_integrals (integrals/active/ri.cpp:14-31)
//   k_pairs<false>        DeterminantPairEvaluator (determinants/pairs/evaluator.cpp:328-384) +
//                         calc_same_spin_hamiltonian (determinants/algebra/hamiltonian.cpp:35-236)
//   k_structure_*         build_tiled_matrix_form_structure_matrices (structures/assembly/hamiltonian_overlap.cpp:889-1021)
//   k_gep_lowest          GeneralizedEigensolver::solve_dense (linalg/eigensolver.cpp:365-387)
//   k_pairs<true>         build_{same,opposite}_spin backward (derivatives/hessian/responses/*/backward.cpp)
//   k_ri_pair_adjoint,
//   k_orbital_gradient    backpropagate_ri_packed_active_two_electron_gradient (integrals/response/ri.cpp:687-694),
//                         pullback_active_matrices (integrals/active/matrix_adjoint.hpp:20-28)
//   k_normalization_pullback  backpropagate_normalization_to_raw_slots (orbitals/pullback/operator.cpp:139-215)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace xmvb_amd {

__device__ __forceinline__ int pair_index(int a, int b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

// wave-uniform broadcast of lane `src`
__device__ __forceinline__ double wave_bcast(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 64-bit DPP move (both halves with the same control)
template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), Ctrl, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// wave-uniform maximum: DPP within 16-lane rows (xor 1, xor 2, half-row mirror, row mirror),
// then the four row maxima through readlane
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fmax(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fmax(v, dpp_f64<0x141>(v));  // row_half_mirror
    v = fmax(v, dpp_f64<0x140>(v));  // row_mirror
    return fmax(fmax(wave_bcast(v, 0), wave_bcast(v, 16)), fmax(wave_bcast(v, 32), wave_bcast(v, 48)));
}

// wave-uniform sum: DPP within 16-lane rows, then the four row sums through readlane
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return (wave_bcast(v, 0) + wave_bcast(v, 16)) + (wave_bcast(v, 32) + wave_bcast(v, 48));
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    // wavefront sum, then across waves through LDS (red: >= 16 doubles); a single-wave block
    // needs no barrier
    v = wave_sum(v);
    if (blockDim.x <= 64) return v;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w];
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------------------------------------
// orbital preparation: one block per orbital
// ---------------------------------------------------------------------------------------------
__global__ void k_orbital_prep(const double* __restrict__ C, const uint8_t* __restrict__ support,
                               const double* __restrict__ S, const double* __restrict__ h, int n_bf, int n_act,
                               double* __restrict__ Cn, double* __restrict__ SC, double* __restrict__ HC,
                               double* __restrict__ nu2, int* __restrict__ flags) {
    extern __shared__ double lds[];
    double* c = lds;             // n_bf
    double* red = lds + n_bf;    // 16
    const int k = blockIdx.x;
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) {
        const size_t at = (size_t)k * n_bf + mu;
        c[mu] = (support && !support[at]) ? 0.0 : C[at];
    }
    __syncthreads();
    double q = 0.0;
    {
        // (S c)[mu] = sum_nu S[nu, mu] c[nu] (S and h are symmetric): one wavefront per row mu,
        // lanes over nu read contiguous columns
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int mu = wave; mu < n_bf; mu += nw) {
            double sv = 0.0, hv = 0.0;
            for (int nu = lane; nu < n_bf; nu += 64) {
                const double cv = c[nu];
                sv += S[nu + (size_t)n_bf * mu] * cv;
                hv += h[nu + (size_t)n_bf * mu] * cv;
            }
            sv = wave_sum(sv);
            hv = wave_sum(hv);
            if (lane == 0) {
                SC[mu + (size_t)n_bf * k] = sv;
                HC[mu + (size_t)n_bf * k] = hv;
                q += c[mu] * sv;
            }
        }
    }
    q = block_sum(q, red);
    const bool ok = isfinite(q) && q > 2.220446049250313e-16;
    const double f = ok ? 1.0 / sqrt(q) : 0.0;
    if (!ok && threadIdx.x == 0) atomicMax(&flags[FLAG_NORMALIZATION], k + 1);
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) {
        Cn[(size_t)k * n_bf + mu] = c[mu] * f;
        SC[mu + (size_t)n_bf * k] *= f;
        HC[mu + (size_t)n_bf * k] *= f;
    }
    if (threadIdx.x == 0) nu2[k] = q;
}

// S_act and raw h_act (n x n), symmetrised; thread 0 checks that S_act is positive definite.
__global__ void k_active_matrices(const double* __restrict__ Cn, const double* __restrict__ SC,
                                  const double* __restrict__ HC, int n_bf, int n, double* __restrict__ sact,
                                  double* __restrict__ hact, int* __restrict__ flags) {
    __shared__ double s_l[MAX_ACT * MAX_ACT], h_l[MAX_ACT * MAX_ACT];
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int i = 0; i < n; ++i)
            for (int j = wave; j < n; j += nw) {  // one wavefront per entry, lanes over mu
                double sv = 0.0, hv = 0.0;
                for (int mu = lane; mu < n_bf; mu += 64) {
                    const double ci = Cn[(size_t)i * n_bf + mu];
                    sv += ci * SC[mu + (size_t)n_bf * j];
                    hv += ci * HC[mu + (size_t)n_bf * j];
                }
                sv = wave_sum(sv);
                hv = wave_sum(hv);
                if (lane == 0) {
                    s_l[i + n * j] = sv;
                    h_l[i + n * j] = hv;
                }
            }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int i = t % n, j = t / n;
        sact[t] = 0.5 * (s_l[t] + s_l[j + n * i]);
        hact[t] = 0.5 * (h_l[t] + h_l[j + n * i]);
    }
    __syncthreads();
    // Cholesky of the symmetrised S_act (the reference's frame factorisation,
    // vbscf/integrals/active/space.cpp:99-113, fails the evaluation when it breaks down);
    // left-looking, wave 0, lane i owns row i.
    if (threadIdx.x < 64) {
        const int i = threadIdx.x;
        double* a = h_l;  // h_l is no longer needed: reuse it as the factor
        for (int t = i; t < n * n; t += 64) a[t] = 0.5 * (s_l[t] + s_l[(t / n) + n * (t % n)]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool ok = true;
        for (int j = 0; j < n; ++j) {
            double v = (i < n) ? a[i + n * j] : 0.0;
            for (int k = 0; k < j; ++k) v -= (i < n ? a[i + n * k] : 0.0) * a[j + n * k];
            const double djj = __shfl(v, j, 64);
            if (!(djj > 0.0)) {
                ok = false;
                break;
            }
            const double d = sqrt(djj);
            if (i == j) a[i + n * j] = d;
            else if (i > j && i < n) a[i + n * j] = v / d;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        if (!ok && i == 0) flags[FLAG_ACTIVE_OVERLAP] = 1;
    }
}

// T[A, q, mu] = sum_nu B[A, (mu nu)] Cn[nu, q]; T index A + n_aux*(q + n*mu).
// grid (ceil(n_aux/64), n_bf), four wavefronts per block splitting nu; lanes run over A so every
// packed B column is read as contiguous 512-byte segments (each column is read by its two AO rows).
__global__ void __launch_bounds__(256) k_ri_one_index(const double* __restrict__ B, const double* __restrict__ Cn,
                                                      int n_bf, int n, int n_aux, double* __restrict__ T) {
    __shared__ double part[4][MAX_ACT][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int A = blockIdx.x * 64 + lane;
    const int mu = blockIdx.y;
    const bool active = A < n_aux;
    double acc[MAX_ACT];
#pragma unroll
    for (int q = 0; q < MAX_ACT; ++q) acc[q] = 0.0;
    for (int nu = wave; nu < n_bf; nu += 4) {
        const double b = active ? B[(size_t)A + (size_t)n_aux * (size_t)pair_index(mu, nu)] : 0.0;
#pragma unroll
        for (int q = 0; q < MAX_ACT; ++q)
            if (q < n) acc[q] += b * Cn[(size_t)q * n_bf + nu];
    }
#pragma unroll
    for (int q = 0; q < MAX_ACT; ++q) part[wave][q][lane] = acc[q];
    __syncthreads();
    for (int q = wave; q < n; q += 4) {
        const double v = (part[0][q][lane] + part[1][q][lane]) + (part[2][q][lane] + part[3][q][lane]);
        if (active) T[(size_t)A + (size_t)n_aux * (q + (size_t)n * mu)] = v;
    }
}

// L[A, (p,q)] = sum_mu T[A,q,mu] Cn[mu,p], q <= p. One thread per (A, packed (p,q)):
// grid (ceil(n_aux/blockDim), n(n+1)/2).
__global__ void k_ri_pair_factors(const double* __restrict__ T, const double* __restrict__ Cn, int n_bf, int n,
                                  int n_aux, double* __restrict__ L) {
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = blockIdx.y;
    if (A >= n_aux) return;
    int p = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
    while (p * (p + 1) / 2 > P) --p;
    while ((p + 1) * (p + 2) / 2 <= P) ++p;
    const int q = P - p * (p + 1) / 2;
    const double* t = T + (size_t)A + (size_t)n_aux * q;
    const double* c = Cn + (size_t)p * n_bf;
    const size_t step = (size_t)n_aux * n;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int mu = 0;
    for (; mu + 3 < n_bf; mu += 4) {
        a0 += t[step * mu] * c[mu];
        a1 += t[step * (mu + 1)] * c[mu + 1];
        a2 += t[step * (mu + 2)] * c[mu + 2];
        a3 += t[step * (mu + 3)] * c[mu + 3];
    }
    for (; mu < n_bf; ++mu) a0 += t[step * mu] * c[mu];
    L[(size_t)A + (size_t)n_aux * P] = (a0 + a1) + (a2 + a3);
}

// packed G[(P,Q)] = sum_A L[A,P] L[A,Q]; one wavefront per packed entry.
__global__ void k_ri_pair_gram(const double* __restrict__ L, int n_aux, int np, double* __restrict__ G) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int ng = np * (np + 1) / 2;
    if (wave >= ng) return;
    int P = (int)((sqrt(8.0 * wave + 1.0) - 1.0) * 0.5);
    while (P * (P + 1) / 2 > wave) --P;
    while ((P + 1) * (P + 2) / 2 <= wave) ++P;
    const int Q = wave - P * (P + 1) / 2;
    double acc = 0.0;
    for (int A = lane; A < n_aux; A += 64) acc += L[(size_t)A + (size_t)n_aux * P] * L[(size_t)A + (size_t)n_aux * Q];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) G[wave] = acc;
}

// ---------------------------------------------------------------------------------------------
// determinant pairs: one wavefront per unordered determinant pair (d <= e)
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// k-th set bit of a mask (k < popcount)
__device__ __forceinline__ int nth_bit(uint64_t m, int k) {
    for (int i = 0; i < k; ++i) m &= m - 1;
    return __builtin_ctzll(m);
}

struct BlockInverse {
    double dens;    // this lane's D[r,l] = Y(l,r), Y = X^{-1}
    double det;     // wave-uniform determinant
    bool singular;  // wave-uniform: rank below the block size under the 1e-12 relative pivot rule
};

// Gauss-Jordan elimination with complete pivoting of the m x m block held one element per lane
// (lane = r + 8*l), augmented with the identity. Pivot choice and the rank rule follow Eigen's
// FullPivLU as used by DeterminantOverlapResolver (determinants/algebra/overlap.cpp:50-110):
// largest remaining |entry|; rank counts pivots above 1e-12 * max|pivot|.
__device__ BlockInverse invert_block(double x, int m, int lane) {
    const int r = lane & 7, l = lane >> 3;
    const bool valid = r < m && l < m;
    double aug = (valid && r == l) ? 1.0 : 0.0;
    uint32_t rows_used = 0, cols_used = 0;
    int inversions = 0;
    double det = 1.0, maxpiv = 0.0, minpiv = 1.0e308;
    int my_row = 0;
    double my_piv = 1.0;
    bool singular = false;
    for (int k = 0; k < m; ++k) {
        const double cand = (valid && !((rows_used >> r) & 1u) && !((cols_used >> l) & 1u)) ? fabs(x) : -1.0;
        const double best = wave_max(cand);
        if (!(best > 0.0)) {  // exactly zero corner (or non-finite): rank deficient
            singular = true;
            break;
        }
        // lowest lane holding the maximum (deterministic tie-break)
        const int idx = __builtin_ctzll(__ballot(cand == best));
        const int pr = idx & 7, pc = idx >> 3;
        const double p = wave_bcast(x, idx);
        const double rp = 1.0 / p;
        maxpiv = fmax(maxpiv, best);
        minpiv = fmin(minpiv, best);
        inversions += __builtin_popcount(rows_used >> (pr + 1)) + __builtin_popcount(cols_used >> (pc + 1));
        const double x_r_pc = __shfl(x, r + 8 * pc, 64);
        const double x_pr_l = __shfl(x, pr + 8 * l, 64);
        const double a_pr_l = __shfl(aug, pr + 8 * l, 64);
        if (valid && r != pr) {
            const double f = x_r_pc * rp;
            x -= f * x_pr_l;
            aug -= f * a_pr_l;
        }
        if (l == pc) {
            my_row = pr;
            my_piv = p;
        }
        det *= p;
        rows_used |= 1u << pr;
        cols_used |= 1u << pc;
    }
    if (!singular && minpiv <= 1.0e-12 * maxpiv) singular = true;
    BlockInverse out;
    out.singular = singular;
    out.det = (inversions & 1) ? -det : det;
    // Y(c, j) = M(pivot row of c, j) / pivot(c); lane (r,l) wants Y(l, r)
    const double a = __shfl(aug, (my_row & 7) + 8 * (r & 7), 64);
    out.dens = (valid && !singular) ? a / my_piv : 0.0;
    return out;
}

struct WaveSlab {
    double D[2][64];  // transition densities D_s[r + 8l] = Y_s(l, r)
    double F[2][64];  // dPhi/dD_s
    double Z[64];
    int PP[2][64];    // packed pair index (R_r, L_l) of each block entry
    int occL[2][8];
    int occR[2][8];
};

// One wavefront per unordered determinant pair (d <= e). The active two-electron kernel is held in
// LDS as the full symmetric np x np matrix G2[P + np*Q] so every (pq|rs) lookup is one index add.
// Forward: H_de = S_de * Phi_de, S_de = det X_alpha det X_beta (Lowdin rules in the regular
// inverse form, DeterminantPairEvaluator::combine_spin_pair_evaluations, pairs/evaluator.cpp:354-384).
// Backward: adjoints of S_act, h_act and the packed kernel for F = w (H_de - E S_de),
// accumulated per block in LDS (full np x np when FullDg, packed otherwise) and written as
// packed partials.
template <bool Backward, bool FullDg>
__global__ void __launch_bounds__(512) k_pairs(const double* __restrict__ G_glob, const double* __restrict__ sact_g,
                                               const double* __restrict__ hact_g, int n, int ng,
                                               const uint64_t* __restrict__ amask, const uint64_t* __restrict__ bmask,
                                               const int* __restrict__ det_a, const int* __restrict__ det_b, int n_det,
                                               long long n_pairs, double* __restrict__ Hdet, double* __restrict__ Sdet,
                                               const double* __restrict__ dcoef, const double* __restrict__ energy,
                                               double* __restrict__ partials, int* __restrict__ flags) {
    extern __shared__ double lds[];
    const int np = n * (n + 1) / 2;
    double* G2 = lds;
    double* sact = G2 + np * np;
    double* hact = sact + n * n;
    double* tail = hact + n * n;
    double *dg = nullptr, *ds = nullptr, *dh = nullptr;
    if (Backward) {
        dg = tail;
        ds = dg + (FullDg ? np * np : ng);
        dh = ds + n * n;
        tail = dh + n * n;
    }
    WaveSlab* slabs = reinterpret_cast<WaveSlab*>(tail);
    {
        const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, ln = threadIdx.x & 63;
        for (int Q = wave; Q < np; Q += nwaves)
            for (int P = ln; P < np; P += 64) G2[P + np * Q] = G_glob[pair_index(P, Q)];
    }
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        sact[t] = sact_g[t];
        hact[t] = hact_g[t];
    }
    if (Backward) {
        const int ndg = FullDg ? np * np : ng;
        for (int t = threadIdx.x; t < ndg; t += blockDim.x) dg[t] = 0.0;
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            ds[t] = 0.0;
            dh[t] = 0.0;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int waves_per_block = blockDim.x >> 6;
    WaveSlab& ws = slabs[threadIdx.x >> 6];
    const long long first = (long long)blockIdx.x * waves_per_block + (threadIdx.x >> 6);
    const long long stride = (long long)gridDim.x * waves_per_block;
    const double E = Backward ? energy[0] : 0.0;
    const int r = lane & 7, l = lane >> 3;

    for (long long t = first; t < n_pairs; t += stride) {
        // unordered pair t = e(e+1)/2 + d, d <= e; d is the left (bra) determinant
        long long e = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (e * (e + 1) / 2 > t) --e;
        while ((e + 1) * (e + 2) / 2 <= t) ++e;
        const int d = (int)(t - e * (e + 1) / 2);
        const int ee = (int)e;
        double weight = 0.0;
        if (Backward) {
            weight = dcoef[d] * dcoef[ee] * (d == ee ? 1.0 : 2.0);
            if (weight == 0.0) continue;
        }
        const uint64_t mL[2] = {amask[det_a[d]], bmask[det_b[d]]};
        const uint64_t mR[2] = {amask[det_a[ee]], bmask[det_b[ee]]};
        int m[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            m[s] = __builtin_popcountll(mL[s]);
            if (lane < m[s]) {
                ws.occL[s][lane] = nth_bit(mL[s], lane);
                ws.occR[s][lane] = nth_bit(mR[s], lane);
            }
        }
        wave_sync();
        // overlap blocks and their inverses
        double det[2];
        bool singular = false;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bool valid = r < m[s] && l < m[s];
            double x = 0.0;
            int ppv = 0;
            if (valid) {
                const int Rr = ws.occR[s][r], Ll = ws.occL[s][l];
                x = sact[Rr + n * Ll];
                ppv = pair_index(Rr, Ll);
            }
            ws.PP[s][lane] = ppv;
            if (m[s] == 0) {
                det[s] = 1.0;
                ws.D[s][lane] = 0.0;
                continue;
            }
            const BlockInverse inv = invert_block(x, m[s], lane);
            det[s] = inv.det;
            singular = singular || inv.singular;
            ws.D[s][lane] = inv.dens;
        }
        if (singular) {
            if (lane == 0) atomicAdd(&flags[FLAG_SINGULAR_PAIRS], 1);
            continue;
        }
        wave_sync();
        // F_s[r,l] = dPhi/dD_s[r,l] = h + sum (J - K) D_s + sum G D_other; Phi = 1/2 sum_s D_s.(h + F_s)
        double phi_part = 0.0;
        int cross_l[2][8], cross_r[2][8];  // pp(R_r, L_l2) and pp(R_r2, L_l) for this lane, per spin
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ms = m[s], mo = m[1 - s];
            double f = 0.0;
            const bool valid = r < ms && l < ms;
            const int Rr = valid ? ws.occR[s][r] : 0, Ll = valid ? ws.occL[s][l] : 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                cross_l[s][k] = (valid && k < ms) ? pair_index(Rr, ws.occL[s][k]) : 0;
                cross_r[s][k] = (valid && k < ms) ? pair_index(ws.occR[s][k], Ll) : 0;
            }
            if (valid) {
                const double hv = hact[Rr + n * Ll];
                const int P1 = ws.PP[s][lane];
                const double* g1 = G2 + (size_t)np * P1;
                f = hv;
#pragma unroll
                for (int l2 = 0; l2 < 8; ++l2) {
                    if (l2 < ms) {
                        const double* gx = G2 + (size_t)np * cross_l[s][l2];
#pragma unroll
                        for (int r2 = 0; r2 < 8; ++r2) {
                            if (r2 < ms) {
                                const int idx = r2 + 8 * l2;
                                const double J = g1[ws.PP[s][idx]];
                                const double K = gx[cross_r[s][r2]];
                                f += (J - K) * ws.D[s][idx];
                            }
                        }
                    }
                }
#pragma unroll
                for (int l2 = 0; l2 < 8; ++l2) {
                    if (l2 < mo) {
#pragma unroll
                        for (int r2 = 0; r2 < 8; ++r2) {
                            if (r2 < mo) {
                                const int idx = r2 + 8 * l2;
                                f += g1[ws.PP[1 - s][idx]] * ws.D[1 - s][idx];
                            }
                        }
                    }
                }
                phi_part += 0.5 * ws.D[s][lane] * (hv + f);
            }
            ws.F[s][lane] = f;
        }
        const double phi = wave_sum(phi_part);
        const double Dtot = det[0] * det[1];
        if (!Backward) {
            if (lane == 0) {
                const double H = Dtot * phi;
                Hdet[(size_t)d + (size_t)n_det * ee] = H;
                Hdet[(size_t)ee + (size_t)n_det * d] = H;
                Sdet[(size_t)d + (size_t)n_det * ee] = Dtot;
                Sdet[(size_t)ee + (size_t)n_det * d] = Dtot;
            }
            continue;
        }
        // ---- backward: F_pair = w (H - E S) ----
        const double wH = weight, wS = -E * weight;
        const double c = Dtot;
        const double cw = c * wH;
        wave_sync();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ms = m[s];
            if (ms == 0) continue;
            const bool valid = r < ms && l < ms;
            // Z[a,l] = sum_b F[b,a] D[b,l]
            double z = 0.0;
            if (valid) {
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    if (b < ms) z += ws.F[s][b + 8 * r] * ws.D[s][b + 8 * l];
            }
            ws.Z[lane] = z;
            wave_sync();
            if (valid) {
                double dfd = 0.0;  // (D F^T D)[r,l]
#pragma unroll
                for (int a = 0; a < 8; ++a)
                    if (a < ms) dfd += ws.D[s][r + 8 * a] * ws.Z[a + 8 * l];
                const double Drl = ws.D[s][lane];
                const int Rr = ws.occR[s][r], Ll = ws.occL[s][l];
                atomicAdd(&ds[Rr + n * Ll], c * ((wH * phi + wS) * Drl - wH * dfd));
                atomicAdd(&dh[Rr + n * Ll], cw * Drl);
                const int P1 = ws.PP[s][lane];
                // same-spin two-electron adjoint over r2 > r, l2 > l
#pragma unroll
                for (int l2 = 1; l2 < 8; ++l2) {
                    if (l2 > l && l2 < ms) {
#pragma unroll
                        for (int r2 = 1; r2 < 8; ++r2) {
                            if (r2 > r && r2 < ms) {
                                const int idx = r2 + 8 * l2;
                                const double mnr = Drl * ws.D[s][idx] - ws.D[s][r + 8 * l2] * ws.D[s][r2 + 8 * l];
                                if (FullDg) {
                                    atomicAdd(&dg[P1 + np * ws.PP[s][idx]], cw * mnr);
                                    atomicAdd(&dg[cross_l[s][l2] + np * cross_r[s][r2]], -cw * mnr);
                                } else {
                                    atomicAdd(&dg[pair_index(P1, ws.PP[s][idx])], cw * mnr);
                                    atomicAdd(&dg[pair_index(cross_l[s][l2], cross_r[s][r2])], -cw * mnr);
                                }
                            }
                        }
                    }
                }
                if (s == 0) {  // opposite spin, once per (alpha entry, beta entry)
                    const int mo = m[1];
#pragma unroll
                    for (int l2 = 0; l2 < 8; ++l2) {
                        if (l2 < mo) {
#pragma unroll
                            for (int r2 = 0; r2 < 8; ++r2) {
                                if (r2 < mo) {
                                    const int idx = r2 + 8 * l2;
                                    const double v = cw * Drl * ws.D[1][idx];
                                    if (FullDg) atomicAdd(&dg[P1 + np * ws.PP[1][idx]], v);
                                    else atomicAdd(&dg[pair_index(P1, ws.PP[1][idx])], v);
                                }
                            }
                        }
                    }
                }
            }
            wave_sync();
        }
    }
    if (Backward) {
        __syncthreads();
        const int stride_part = ng + 2 * n * n;
        double* out = partials + (size_t)blockIdx.x * stride_part;
        if (FullDg) {
            for (int t = threadIdx.x; t < ng; t += blockDim.x) {
                int P = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
                while (P * (P + 1) / 2 > t) --P;
                while ((P + 1) * (P + 2) / 2 <= t) ++P;
                const int Q = t - P * (P + 1) / 2;
                out[t] = (P == Q) ? dg[P + np * P] : dg[P + np * Q] + dg[Q + np * P];
            }
        } else {
            for (int t = threadIdx.x; t < ng; t += blockDim.x) out[t] = dg[t];
        }
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            out[ng + t] = ds[t];
            out[ng + n * n + t] = dh[t];
        }
    }
}

// sum block partials in a fixed order: out[i] = sum_b partials[b*len + i], one wavefront per i
__global__ void k_reduce_partials(const double* __restrict__ partials, int n_blocks, int len, double* __restrict__ out) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= len) return;
    double acc = 0.0;
    for (int b = lane; b < n_blocks; b += 64) acc += partials[(size_t)b * len + i];
    acc = wave_sum(acc);
    if (lane == 0) out[i] = acc;
}

// ---------------------------------------------------------------------------------------------
// structure matrices: Y = Hdet T (thread per (det, structure)), then T^T Y (thread per (I, J))
// ---------------------------------------------------------------------------------------------
__global__ void k_structure_right(const double* __restrict__ Hdet, const double* __restrict__ Sdet, int n_det,
                                  int n_str, const long long* __restrict__ soff, const int* __restrict__ sdet,
                                  const double* __restrict__ scoef, double* __restrict__ Yh, double* __restrict__ Ys) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)n_det * n_str) return;
    const int d = (int)(id % n_det), J = (int)(id / n_det);
    double h = 0.0, s = 0.0;
    for (long long t = soff[J]; t < soff[J + 1]; ++t) {
        const size_t at = (size_t)d + (size_t)n_det * sdet[t];
        h += scoef[t] * Hdet[at];
        s += scoef[t] * Sdet[at];
    }
    Yh[id] = h;
    Ys[id] = s;
}

__global__ void k_structure_left(const double* __restrict__ Yh, const double* __restrict__ Ys, int n_det, int n_str,
                                 const long long* __restrict__ soff, const int* __restrict__ sdet,
                                 const double* __restrict__ scoef, double* __restrict__ Hs, double* __restrict__ Ss) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)n_str * n_str) return;
    const int I = (int)(id % n_str), J = (int)(id / n_str);
    double h = 0.0, s = 0.0;
    for (long long t = soff[I]; t < soff[I + 1]; ++t) {
        const size_t at = (size_t)sdet[t] + (size_t)n_det * J;
        h += scoef[t] * Yh[at];
        s += scoef[t] * Ys[at];
    }
    Hs[id] = h;
    Ss[id] = s;
}

// a[d] = sum_terms coef * c[structure]
__global__ void k_det_coefficients(const long long* __restrict__ off, const int* __restrict__ ts,
                                   const double* __restrict__ tc, const double* __restrict__ coef, int n_det,
                                   double* __restrict__ a) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_det) return;
    double acc = 0.0;
    for (long long t = off[d]; t < off[d + 1]; ++t) acc += tc[t] * coef[ts[t]];
    a[d] = acc;
}

// ---------------------------------------------------------------------------------------------
// generalised eigenproblem H c = E S c, lowest root, one workgroup.
// Cholesky S = L L^T, M = L^{-1} H L^{-T}, Householder tridiagonalisation, multisection on
// Sturm counts for the lowest eigenvalue, inverse iteration on the tridiagonal,
// back-transformation, c = L^{-T} z so that c^T S c = 1.
// Storage: gep_work_doubles(n) = 3*n*n + 9*n doubles, in LDS when it fits (InLds) else in the
// global workspace. A single wavefront is used for small n so that the block barriers are free.
// ---------------------------------------------------------------------------------------------
template <bool InLds>
__global__ void __launch_bounds__(256) k_gep_lowest(const double* __restrict__ Hs, const double* __restrict__ Ss,
                                                    int n, double e_nuc, double* __restrict__ gwork,
                                                    double* __restrict__ coef, double* __restrict__ energy,
                                                    int* __restrict__ flags, const long long* __restrict__ off,
                                                    const int* __restrict__ ts, const double* __restrict__ tc,
                                                    int n_det, double* __restrict__ dcoef) {
    extern __shared__ double lds[];
    double* work = InLds ? lds : gwork;
    double* Lm = work;              // n x n, lower Cholesky factor
    double* M = work + n * n;       // n x n
    double* V = work + 2 * n * n;   // Householder vectors (column k holds v_k)
    double* dd = work + 3 * n * n;  // diagonal
    double* ee = dd + n;            // off-diagonal
    double* y = ee + n;             // tridiagonal eigenvector / scratch
    double* beta = y + n;           // Householder scalars
    double* tdl = beta + n;         // tridiagonal LU factors (inverse iteration)
    double* tdu = tdl + n;
    double* tdu2 = tdu + n;
    double* tdg = tdu2 + n;
    int* ipiv = reinterpret_cast<int*>(tdg + n);  // n ints in the last n doubles
    __shared__ double red[16];
    __shared__ double sh_scalar[4];
    __shared__ double lo_hi[2];
    __shared__ int sh_flag;
    __shared__ int first_true;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n; i += nt)
        for (int j = 0; j < n; ++j) {
            Lm[i + n * j] = 0.5 * (Ss[i + n * j] + Ss[j + n * i]);
            M[i + n * j] = 0.5 * (Hs[i + n * j] + Hs[j + n * i]);
        }
    if (tid == 0) sh_flag = 0;
    __syncthreads();
    // Cholesky (right-looking, lower); threads own rows, so no index division is needed
    for (int k = 0; k < n; ++k) {
        const double a = Lm[k + n * k];
        if (!(a > 0.0)) {
            if (tid == 0) flags[FLAG_METRIC] = 1;
            return;  // uniform: every thread read the same diagonal after the last barrier
        }
        const double dk = sqrt(a), rdk = 1.0 / dk;
        __syncthreads();
        for (int i = k + tid; i < n; i += nt) Lm[i + n * k] = (i == k) ? dk : Lm[i + n * k] * rdk;
        __syncthreads();
        for (int i = k + 1 + tid; i < n; i += nt) {
            const double lik = Lm[i + n * k];
            for (int j = k + 1; j <= i; ++j) Lm[i + n * j] -= lik * Lm[j + n * k];
        }
        __syncthreads();
    }
    // M <- L^{-1} H by forward substitution, one column per thread (no barriers inside)
    for (int j = tid; j < n; j += nt) {
        for (int i = 0; i < n; ++i) {
            double v0 = M[i + n * j], v1 = 0.0, v2 = 0.0, v3 = 0.0;
            int k = 0;
            for (; k + 3 < i; k += 4) {
                v0 -= Lm[i + n * k] * M[k + n * j];
                v1 -= Lm[i + n * (k + 1)] * M[k + 1 + n * j];
                v2 -= Lm[i + n * (k + 2)] * M[k + 2 + n * j];
                v3 -= Lm[i + n * (k + 3)] * M[k + 3 + n * j];
            }
            for (; k < i; ++k) v0 -= Lm[i + n * k] * M[k + n * j];
            M[i + n * j] = ((v0 + v1) + (v2 + v3)) / Lm[i + n * i];
        }
    }
    __syncthreads();
    // M <- L^{-1} M^T (transpose staged through V), then symmetrise
    for (int i = tid; i < n; i += nt)
        for (int j = 0; j < n; ++j) V[i + n * j] = M[j + n * i];
    __syncthreads();
    for (int j = tid; j < n; j += nt) {
        for (int i = 0; i < n; ++i) {
            double v0 = V[i + n * j], v1 = 0.0, v2 = 0.0, v3 = 0.0;
            int k = 0;
            for (; k + 3 < i; k += 4) {
                v0 -= Lm[i + n * k] * M[k + n * j];
                v1 -= Lm[i + n * (k + 1)] * M[k + 1 + n * j];
                v2 -= Lm[i + n * (k + 2)] * M[k + 2 + n * j];
                v3 -= Lm[i + n * (k + 3)] * M[k + 3 + n * j];
            }
            for (; k < i; ++k) v0 -= Lm[i + n * k] * M[k + n * j];
            M[i + n * j] = ((v0 + v1) + (v2 + v3)) / Lm[i + n * i];
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt)
        for (int j = 0; j < n; ++j) {
            V[i + n * j] = 0.0;
            if (i < j) {
                const double s = 0.5 * (M[i + n * j] + M[j + n * i]);
                M[i + n * j] = s;
                M[j + n * i] = s;
            }
        }
    __syncthreads();
    // Householder tridiagonalisation: reflect rows/cols k+1..n-1 of column k
    for (int k = 0; k < n - 2; ++k) {
        const int m0 = k + 1, len = n - m0;
        double part = 0.0;
        for (int i = m0 + tid; i < n; i += nt) part += M[i + n * k] * M[i + n * k];
        const double sigma = block_sum(part, red);
        const double x0 = M[m0 + n * k];
        const double alpha = sigma > 0.0 ? (x0 >= 0.0 ? -sqrt(sigma) : sqrt(sigma)) : 0.0;
        for (int i = m0 + tid; i < n; i += nt) V[i + n * k] = M[i + n * k] - (i == m0 ? alpha : 0.0);
        const double vtv = sigma - 2.0 * alpha * x0 + alpha * alpha;
        const double bk = (sigma > 0.0 && vtv > 0.0) ? 2.0 / vtv : 0.0;
        if (tid == 0) beta[k] = bk;
        __syncthreads();
        for (int i = m0 + tid; i < n; i += nt) {  // p = bk A v
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            int j = m0;
            for (; j + 3 < n; j += 4) {
                a0 += M[i + n * j] * V[j + n * k];
                a1 += M[i + n * (j + 1)] * V[j + 1 + n * k];
                a2 += M[i + n * (j + 2)] * V[j + 2 + n * k];
                a3 += M[i + n * (j + 3)] * V[j + 3 + n * k];
            }
            for (; j < n; ++j) a0 += M[i + n * j] * V[j + n * k];
            y[i] = bk * ((a0 + a1) + (a2 + a3));
        }
        __syncthreads();
        double kp = 0.0;
        for (int i = m0 + tid; i < n; i += nt) kp += V[i + n * k] * y[i];
        const double K = 0.5 * bk * block_sum(kp, red);
        for (int i = m0 + tid; i < n; i += nt) y[i] -= K * V[i + n * k];  // w = p - K v
        __syncthreads();
        for (int i = m0 + tid; i < n; i += nt) {  // A -= v w^T + w v^T
            const double vi = V[i + n * k], wi = y[i];
            for (int j = m0; j < n; ++j) M[i + n * j] -= vi * y[j] + wi * V[j + n * k];
        }
        if (bk != 0.0)
            for (int i = m0 + tid; i < n; i += nt) {
                const double v = (i == m0) ? alpha : 0.0;
                M[i + n * k] = v;
                M[k + n * i] = v;
            }
        __syncthreads();
    }
    for (int i = tid; i < n; i += nt) {
        dd[i] = M[i + n * i];
        ee[i] = (i + 1 < n) ? M[(i + 1) + n * i] : 0.0;
    }
    __syncthreads();
    // lowest eigenvalue by multisection on Sturm counts
    if (tid == 0) {
        double lo = 1e300, hi = -1e300;
        for (int i = 0; i < n; ++i) {
            const double rad = (i > 0 ? fabs(ee[i - 1]) : 0.0) + (i + 1 < n ? fabs(ee[i]) : 0.0);
            lo = fmin(lo, dd[i] - rad);
            hi = fmax(hi, dd[i] + rad);
        }
        const double pad = 1e-12 * fmax(fabs(lo), fabs(hi)) + 1e-300;
        lo_hi[0] = lo - pad;
        lo_hi[1] = hi + pad;
    }
    __syncthreads();
    for (int it = 0; it < 200; ++it) {
        const double lo = lo_hi[0], hi = lo_hi[1];
        if (!(hi - lo > 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)))) break;
        if (tid == 0) first_true = nt;
        __syncthreads();
        const double x = lo + (hi - lo) * (double)(tid + 1) / (double)nt;
        int count = 0;
        double q = 1.0;
        for (int i = 0; i < n; ++i) {  // number of eigenvalues below x (LDL^T pivots)
            const double e2 = i > 0 ? ee[i - 1] * ee[i - 1] : 0.0;
            q = dd[i] - x - (i > 0 ? e2 / q : 0.0);
            if (q == 0.0) q = -1e-300;
            if (q < 0.0) ++count;
        }
        if (count >= 1) atomicMin(&first_true, tid);
        __syncthreads();
        if (tid == 0) {
            const int f = first_true;
            lo_hi[0] = f == 0 ? lo : lo + (hi - lo) * (double)f / (double)nt;
            lo_hi[1] = f >= nt ? hi : lo + (hi - lo) * (double)(f + 1) / (double)nt;
        }
        __syncthreads();
    }
    const double lambda = 0.5 * (lo_hi[0] + lo_hi[1]);
    // inverse iteration on T - lambda I (LU with partial pivoting, LAPACK dgttrf/dgttrs scheme)
    if (tid == 0) {
        for (int i = 0; i < n; ++i) {
            tdg[i] = dd[i] - lambda;
            tdu2[i] = 0.0;
            if (i + 1 < n) {
                tdl[i] = ee[i];
                tdu[i] = ee[i];
            }
        }
        for (int i = 0; i + 1 < n; ++i) {
            if (fabs(tdg[i]) >= fabs(tdl[i])) {
                ipiv[i] = i;
                if (tdg[i] == 0.0) tdg[i] = 1e-300;
                const double f = tdl[i] / tdg[i];
                tdl[i] = f;
                tdg[i + 1] -= f * tdu[i];
            } else {
                ipiv[i] = i + 1;
                const double f = tdg[i] / tdl[i];
                tdg[i] = tdl[i];
                tdl[i] = f;
                const double t = tdu[i];
                tdu[i] = tdg[i + 1];
                tdg[i + 1] = t - f * tdg[i + 1];
                if (i + 2 < n) {
                    tdu2[i] = tdu[i + 1];
                    tdu[i + 1] = -f * tdu[i + 1];
                }
            }
        }
        if (n > 0) {
            ipiv[n - 1] = n - 1;
            if (tdg[n - 1] == 0.0) tdg[n - 1] = 1e-300;
        }
        for (int i = 0; i < n; ++i) y[i] = 1.0;
        for (int iter = 0; iter < 3; ++iter) {
            for (int i = 0; i + 1 < n; ++i) {
                if (ipiv[i] == i) {
                    y[i + 1] -= tdl[i] * y[i];
                } else {
                    const double t = y[i];
                    y[i] = y[i + 1];
                    y[i + 1] = t - tdl[i] * y[i];
                }
            }
            for (int i = n - 1; i >= 0; --i) {
                double v = y[i];
                if (i + 1 < n) v -= tdu[i] * y[i + 1];
                if (i + 2 < n) v -= tdu2[i] * y[i + 2];
                y[i] = v / tdg[i];
            }
            double nrm = 0.0;
            for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
            nrm = 1.0 / sqrt(nrm);
            for (int i = 0; i < n; ++i) y[i] *= nrm;
        }
        sh_scalar[1] = lambda;
    }
    __syncthreads();
    // back-transform z = H_0 H_1 ... H_{n-3} y
    for (int k = n - 3; k >= 0; --k) {
        const int m0 = k + 1;
        double part = 0.0;
        for (int i = m0 + tid; i < n; i += nt) part += V[i + n * k] * y[i];
        const double s = beta[k] * block_sum(part, red);
        for (int i = m0 + tid; i < n; i += nt) y[i] -= s * V[i + n * k];
        __syncthreads();
    }
    // c = L^{-T} z (back substitution over the columns of L)
    for (int k = n - 1; k >= 0; --k) {
        if (tid == 0) y[k] /= Lm[k + n * k];
        __syncthreads();
        const double ck = y[k];
        for (int i = tid; i < k; i += nt) y[i] -= Lm[k + n * i] * ck;
        __syncthreads();
    }
    for (int i = tid; i < n; i += nt) coef[i] = y[i];
    if (tid == 0) {
        energy[0] = sh_scalar[1];
        energy[1] = sh_scalar[1] + e_nuc;
    }
    // determinant coefficients a_d = sum_terms t_dI c_I for the backward weights
    for (int d = tid; d < n_det; d += nt) {
        double acc = 0.0;
        for (long long t = off[d]; t < off[d + 1]; ++t) acc += tc[t] * y[ts[t]];
        dcoef[d] = acc;
    }
}

// ---------------------------------------------------------------------------------------------
// gradient tail
// ---------------------------------------------------------------------------------------------

// dL[A,P] = sum_Q M[P,Q] L[A,Q], M[P,Q] = dg(P,Q) off the diagonal, 2 dg(P,P) on it.
__global__ void k_ri_pair_adjoint(const double* __restrict__ L, const double* __restrict__ dg, int n_aux, int np,
                                  double* __restrict__ dL) {
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = blockIdx.y;
    if (A >= n_aux) return;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int Q = 0; Q < np; ++Q) {
        const double m = (P == Q) ? 2.0 * dg[pair_index(P, P)] : dg[pair_index(P, Q)];
        acc[Q & 3] += m * L[(size_t)A + (size_t)n_aux * Q];
    }
    dL[(size_t)A + (size_t)n_aux * P] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// gcn[mu, p] = sum_i SC[mu,i] (ds+ds^T)[i,p] + HC[mu,i] (dh+dh^T)[i,p]
//            + sum_q f_pq sum_A dL[A,(p,q)] T[A,q,mu],  f_pq = 2 if p == q else 1.
// one block per AO index mu.
__global__ void __launch_bounds__(256) k_orbital_gradient(const double* __restrict__ T, const double* __restrict__ dL,
                                                          const double* __restrict__ SC, const double* __restrict__ HC,
                                                          const double* __restrict__ dsdh, int n_bf, int n, int n_aux,
                                                          int ng, double* __restrict__ gcn) {
    __shared__ double part[4][MAX_ACT];
    const int mu = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[MAX_ACT];
#pragma unroll
    for (int p = 0; p < MAX_ACT; ++p) acc[p] = 0.0;
    // contraction over the contiguous (A, q) plane of T for this mu
    const double* tmu = T + (size_t)n_aux * n * mu;
    for (int k = threadIdx.x; k < n_aux * n; k += blockDim.x) {
        const int q = k / n_aux, A = k - q * n_aux;
        const double t = tmu[k];
#pragma unroll
        for (int p = 0; p < MAX_ACT; ++p)
            if (p < n) acc[p] += (p == q ? 2.0 : 1.0) * dL[(size_t)A + (size_t)n_aux * pair_index(p, q)] * t;
    }
#pragma unroll
    for (int p = 0; p < MAX_ACT; ++p) {
        const double s = wave_sum(acc[p]);
        if (lane == 0 && p < n) part[wave][p] = s;
    }
    __syncthreads();
    const double* ds = dsdh + ng;
    const double* dh = dsdh + ng + n * n;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        double g = (part[0][p] + part[1][p]) + (part[2][p] + part[3][p]);
        for (int i = 0; i < n; ++i) {
            g += SC[mu + (size_t)n_bf * i] * (ds[i + n * p] + ds[p + n * i]);
            g += HC[mu + (size_t)n_bf * i] * (dh[i + n * p] + dh[p + n * i]);
        }
        gcn[mu + (size_t)n_bf * p] = g;
    }
}

// raw-slot gradient: grad[k, mu] = (g - (g.cn_k) S cn_k)[mu] / sqrt(nu2_k) on present slots.
__global__ void k_normalization_pullback(const double* __restrict__ gcn, const double* __restrict__ Cn,
                                         const double* __restrict__ SC, const double* __restrict__ nu2,
                                         const uint8_t* __restrict__ support, int n_bf, double* __restrict__ grad) {
    __shared__ double red[16];
    const int k = blockIdx.x;
    double part = 0.0;
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) part += gcn[mu + (size_t)n_bf * k] * Cn[(size_t)k * n_bf + mu];
    const double dot = block_sum(part, red);
    const double f = 1.0 / sqrt(nu2[k]);
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) {
        const size_t at = (size_t)k * n_bf + mu;
        grad[at] = (support && !support[at]) ? 0.0 : f * (gcn[mu + (size_t)n_bf * k] - dot * SC[mu + (size_t)n_bf * k]);
    }
}

// ---------------------------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------------------------
// LDS of the pair kernels: full np x np kernel, S_act, h_act, per-wave slabs and, for the
// backward, the adjoint accumulators (full np x np when that still fits, else packed).
constexpr size_t kPairLdsLimit = 160 * 1024;

bool pairs_full_dg(int n, int ng) {
    const size_t np = (size_t)n * (n + 1) / 2;
    const size_t bytes = (2 * np * np + 4 * (size_t)n * n) * sizeof(double) + 4 * sizeof(WaveSlab);
    (void)ng;
    return bytes <= 150 * 1024;
}

size_t pairs_lds_bytes(int n, int ng, bool backward, int waves) {
    const size_t np = (size_t)n * (n + 1) / 2;
    size_t doubles = np * np + 2 * (size_t)n * n;
    if (backward) doubles += (pairs_full_dg(n, ng) ? np * np : (size_t)ng) + 2 * (size_t)n * n;
    return doubles * sizeof(double) + (size_t)waves * sizeof(WaveSlab);
}

// eight wavefronts per block when the LDS allows it, else four
int pairs_waves(int n, int ng, bool backward) {
    return pairs_lds_bytes(n, ng, backward, 8) <= kPairLdsLimit ? 8 : 4;
}

size_t pairs_lds_bytes(int n, int ng, bool backward) {
    return pairs_lds_bytes(n, ng, backward, pairs_waves(n, ng, backward));
}

size_t gep_lds_bytes(int n);
bool gep_in_lds(int n);

hipError_t configure_kernels(int n, int ng, int n_str) {
    const size_t fwd = pairs_lds_bytes(n, ng, false), bwd = pairs_lds_bytes(n, ng, true);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs<false, false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)fwd);
    if (e != hipSuccess) return e;
    e = pairs_full_dg(n, ng)
            ? hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs<true, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bwd)
            : hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs<true, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bwd);
    if (e != hipSuccess) return e;
    if (gep_in_lds(n_str))
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gep_lowest<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)gep_lds_bytes(n_str));
    return e;
}

void launch_orbital_prep(hipStream_t st, const double* C, const uint8_t* support, const double* S, const double* h,
                         int n_bf, int n, double* Cn, double* SC, double* HC, double* nu2, int* flags) {
    const size_t lds = ((size_t)n_bf + 16) * sizeof(double);
    hipLaunchKernelGGL(k_orbital_prep, dim3(n), dim3(256), lds, st, C, support, S, h, n_bf, n, Cn, SC, HC, nu2, flags);
}

void launch_active_matrices(hipStream_t st, const double* Cn, const double* SC, const double* HC, int n_bf, int n,
                            double* sact, double* hact, int* flags) {
    hipLaunchKernelGGL(k_active_matrices, dim3(1), dim3(256), 0, st, Cn, SC, HC, n_bf, n, sact, hact, flags);
}

void launch_ri_one_index(hipStream_t st, const double* B, const double* Cn, int n_bf, int n, int n_aux, double* T) {
    hipLaunchKernelGGL(k_ri_one_index, dim3((n_aux + 63) / 64, n_bf), dim3(256), 0, st, B, Cn, n_bf, n, n_aux, T);
}

void launch_ri_pair_stage(hipStream_t st, const double* T, const double* Cn, int n_bf, int n, int n_aux, double* L,
                          double* G) {
    const int tb = 64;
    const int np = n * (n + 1) / 2, ng = np * (np + 1) / 2;
    hipLaunchKernelGGL(k_ri_pair_factors, dim3((n_aux + tb - 1) / tb, np), dim3(tb), 0, st, T, Cn, n_bf, n, n_aux, L);
    hipLaunchKernelGGL(k_ri_pair_gram, dim3((ng + 3) / 4), dim3(256), 0, st, L, n_aux, np, G);
}

void launch_pairs(hipStream_t st, bool backward, int n_blocks, const double* G, const double* sact, const double* hact,
                  int n, int ng, const uint64_t* amask, const uint64_t* bmask, const int* det_a, const int* det_b,
                  int n_det, double* Hdet, double* Sdet, const double* dcoef, const double* energy, double* partials,
                  int* flags) {
    const long long n_pairs = (long long)n_det * (n_det + 1) / 2;
    const size_t lds = pairs_lds_bytes(n, ng, backward);
    const int threads = 64 * pairs_waves(n, ng, backward);
    if (!backward)
        hipLaunchKernelGGL((k_pairs<false, false>), dim3(n_blocks), dim3(threads), lds, st, G, sact, hact, n, ng, amask,
                           bmask, det_a, det_b, n_det, n_pairs, Hdet, Sdet, dcoef, energy, partials, flags);
    else if (pairs_full_dg(n, ng))
        hipLaunchKernelGGL((k_pairs<true, true>), dim3(n_blocks), dim3(threads), lds, st, G, sact, hact, n, ng, amask,
                           bmask, det_a, det_b, n_det, n_pairs, Hdet, Sdet, dcoef, energy, partials, flags);
    else
        hipLaunchKernelGGL((k_pairs<true, false>), dim3(n_blocks), dim3(threads), lds, st, G, sact, hact, n, ng, amask,
                           bmask, det_a, det_b, n_det, n_pairs, Hdet, Sdet, dcoef, energy, partials, flags);
}

void launch_reduce_partials(hipStream_t st, const double* partials, int n_blocks, int len, double* out) {
    hipLaunchKernelGGL(k_reduce_partials, dim3((len + 3) / 4), dim3(256), 0, st, partials, n_blocks, len, out);
}

void launch_structure(hipStream_t st, const double* Hdet, const double* Sdet, int n_det, int n_str,
                      const long long* soff, const int* sdet, const double* scoef, double* Yh, double* Ys, double* Hs,
                      double* Ss) {
    const long long a = (long long)n_det * n_str, b = (long long)n_str * n_str;
    hipLaunchKernelGGL(k_structure_right, dim3((unsigned)((a + 255) / 256)), dim3(256), 0, st, Hdet, Sdet, n_det, n_str,
                       soff, sdet, scoef, Yh, Ys);
    hipLaunchKernelGGL(k_structure_left, dim3((unsigned)((b + 255) / 256)), dim3(256), 0, st, Yh, Ys, n_det, n_str,
                       soff, sdet, scoef, Hs, Ss);
}

size_t gep_lds_bytes(int n) { return gep_work_doubles(n) * sizeof(double); }

bool gep_in_lds(int n) { return gep_lds_bytes(n) <= 144 * 1024; }

void launch_gep(hipStream_t st, const double* Hs, const double* Ss, int n, double e_nuc, double* work, double* coef,
                double* energy, int* flags, const long long* off, const int* ts, const double* tc, int n_det,
                double* dcoef) {
    const int threads = n <= 64 ? 64 : 256;
    if (gep_in_lds(n))
        hipLaunchKernelGGL(k_gep_lowest<true>, dim3(1), dim3(threads), gep_lds_bytes(n), st, Hs, Ss, n, e_nuc, work,
                           coef, energy, flags, off, ts, tc, n_det, dcoef);
    else
        hipLaunchKernelGGL(k_gep_lowest<false>, dim3(1), dim3(threads), 0, st, Hs, Ss, n, e_nuc, work, coef, energy,
                           flags, off, ts, tc, n_det, dcoef);
}

void launch_det_coefficients(hipStream_t st, const long long* off, const int* ts, const double* tc, const double* coef,
                             int n_det, double* a) {
    hipLaunchKernelGGL(k_det_coefficients, dim3((n_det + 255) / 256), dim3(256), 0, st, off, ts, tc, coef, n_det, a);
}

void launch_gradient_tail(hipStream_t st, const double* T, const double* L, const double* dsum, const double* SC,
                          const double* HC, const double* Cn, const double* nu2, const uint8_t* support, int n_bf,
                          int n, int n_aux, double* dL, double* gcn, double* grad) {
    const int np = n * (n + 1) / 2, ng = np * (np + 1) / 2;
    const int tb = 64;
    hipLaunchKernelGGL(k_ri_pair_adjoint, dim3((n_aux + tb - 1) / tb, np), dim3(tb), 0, st, L, dsum, n_aux, np, dL);
    hipLaunchKernelGGL(k_orbital_gradient, dim3(n_bf), dim3(256), 0, st, T, dL, SC, HC, dsum, n_bf, n, n_aux, ng, gcn);
    hipLaunchKernelGGL(k_normalization_pullback, dim3(n), dim3(256), 0, st, gcn, Cn, SC, nu2, support, n_bf, grad);
}

}  // namespace xmvb_amd
