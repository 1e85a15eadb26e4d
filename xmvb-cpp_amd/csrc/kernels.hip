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

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "wave.hpp"

namespace xmvb_amd {


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
__global__ void k_orbital_prep(const SlotDesc* __restrict__ slots, int n_act) {
    int bx, bz;
    problem_block(bx, bz);  // the n_act workgroups of a problem on one XCD: S and h read once
    const SlotDesc& sd = slots[bz];
    const auto C = to_global(sd.C);
    const auto support = to_global(sd.support);
    const auto S = to_global(sd.S);
    const auto h = to_global(sd.h);
    const int n_bf = sd.n_bf;
    const auto Cn = to_global(sd.Cn);
    const auto SC = to_global(sd.SC);
    const auto HC = to_global(sd.HC);
    const auto nu2 = to_global(sd.nu2);
    const auto flags = to_global(sd.flags);
    extern __shared__ double lds[];
    double* c = lds;             // n_bf
    double* red = lds + n_bf;    // 16
    const int k = bx;
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) {
        const size_t at = (size_t)k * n_bf + mu;
        c[mu] = (support && !support[at]) ? 0.0 : C[at];
    }
    __syncthreads();
    double q = 0.0;
    // (S c)[mu] = sum_nu S[mu, nu] c[nu]: one thread per row mu, consecutive threads on consecutive
    // rows of each column (coalesced), c[nu] an LDS broadcast; no cross-lane reduction, and the
    // loads of successive nu are independent of the accumulation, so they stay in flight
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) {
        double sv = 0.0, hv = 0.0;
#pragma unroll 8
        for (int nu = 0; nu < n_bf; ++nu) {
            const double cv = c[nu];
            sv += S[mu + (size_t)n_bf * nu] * cv;
            hv += h[mu + (size_t)n_bf * nu] * cv;
        }
        SC[mu + (size_t)n_bf * k] = sv;
        HC[mu + (size_t)n_bf * k] = hv;
        q += c[mu] * sv;
    }
    q = block_sum(q, red);
    const bool ok = isfinite(q) && q > 2.220446049250313e-16;
    const double f = ok ? 1.0 / sqrt(q) : 0.0;
    if (!ok && threadIdx.x == 0) atomicMax((int*)&flags[FLAG_NORMALIZATION], k + 1);
    const auto CnT = to_global(sd.CnT);
    for (int mu = threadIdx.x; mu < n_bf; mu += blockDim.x) {
        Cn[(size_t)k * n_bf + mu] = c[mu] * f;
        CnT[(size_t)mu * kCnStride + k] = c[mu] * f;
        SC[mu + (size_t)n_bf * k] *= f;
        HC[mu + (size_t)n_bf * k] *= f;
    }
    if (k == 0) {  // the zeros of CnT: orbitals past n_act, rows past n_bf up to a multiple of 32
        const int rows = (n_bf + 31) / 32 * 32;
        for (int t = threadIdx.x; t < rows * kCnStride; t += blockDim.x) {
            const int mu = t / kCnStride, q = t - mu * kCnStride;
            if (mu >= n_bf || q >= n_act) CnT[t] = 0.0;
        }
    }
    if (threadIdx.x == 0) nu2[k] = q;
}

// S_act and raw h_act (n x n), symmetrised; thread 0 checks that S_act is positive definite.
__global__ void k_active_matrices(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto Cn = to_global(sd.Cn);
    const auto SC = to_global(sd.SC);
    const auto HC = to_global(sd.HC);
    const int n_bf = sd.n_bf;
    const auto sact = to_global(sd.sact);
    const auto hact = to_global(sd.hact);
    const auto flags = to_global(sd.flags);
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
// grid (ceil(n_aux/64), ceil(n_bf/16), slots), four wavefronts per block. A wavefront owns a 64-wide
// aux tile (lanes run over A, so every packed B column is read as one contiguous 512-byte segment)
// and four AO rows mu, so each Cn[nu, q] read from LDS feeds four accumulators. Every column is read
// by both of its AO rows; the second read usually comes from the last-level cache because the
// blocks of one slot are dispatched together. Cn is staged per chunk of nu as [nu][16] with zeros
// beyond n, so NQ (n rounded up to even) accumulators per row need no guards.
constexpr int kOneIndexChunk = 64;
constexpr int kOneIndexRows = 4;  // AO rows per wavefront
constexpr int kOneIndexNu = 4;    // nu per inner step (loads in flight per row); divides kOneIndexChunk

// storage of the flat factors read by k_ri_one_index: FP64, FP32, compensated FP32 (FP32 + an
// 8-bit correction in units of 2^(e(hi) - 158))
struct RiFp32c {};

template <typename BT>
__device__ __forceinline__ double flat_b(const SlotDesc& sd, size_t at) {
    if constexpr (std::is_same_v<BT, double>) {
        return to_global(sd.B)[at];
    } else if constexpr (std::is_same_v<BT, float>) {
        return (double)to_global(sd.B32)[at];
    } else {
        const float hi = to_global(sd.B32)[at];
        unsigned e = (__float_as_uint(hi) >> 23) & 0xffu;
        e = e ? e : 1u;
        const double scale = __hiloint2double((int)((e + 1023u - 158u) << 20), 0);
        return fma((double)to_global(sd.Blo)[at], scale, (double)hi);
    }
}

template <int NQ, typename BT>
__global__ void __launch_bounds__(256) k_ri_one_index(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_bf = sd.n_bf, n_aux = sd.n_aux;
    if (sd.b_tiled) return;  // read by k_ri_one_index_tiled
    if ((int)blockIdx.y * 4 * kOneIndexRows >= n_bf || (int)blockIdx.x * 64 >= n_aux) return;  // block-uniform
    const auto Cn = to_global(sd.Cn);
    __shared__ double cT[kOneIndexChunk][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int A = blockIdx.x * 64 + lane;
    const int Ac = A < n_aux ? A : n_aux - 1;  // clamped: loads stay in bounds, results are dropped
    const int mu0 = (blockIdx.y * 4 + wave) * kOneIndexRows;
    const bool rows = mu0 < n_bf;            // wave-uniform
    int mu[kOneIndexRows];
#pragma unroll
    for (int u = 0; u < kOneIndexRows; ++u) mu[u] = min(mu0 + u, n_bf - 1);
    double acc[kOneIndexRows][NQ];
#pragma unroll
    for (int u = 0; u < kOneIndexRows; ++u)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[u][q] = 0.0;
    for (int c0 = 0; c0 < n_bf; c0 += kOneIndexChunk) {
        const int len = min(kOneIndexChunk, n_bf - c0);
        __syncthreads();
        for (int t = threadIdx.x; t < kOneIndexChunk * 16; t += blockDim.x) {
            const int v = t >> 4, q = t & 15;
            cT[v][q] = (q < n && v < len) ? Cn[(size_t)q * n_bf + c0 + v] : 0.0;
        }
        __syncthreads();
        if (!rows) continue;
        for (int v = 0; v < len; v += kOneIndexNu) {
            // kOneIndexNu nu per step (past the chunk end: zero-weight repeats of its last nu; the
            // staged rows from len on are zeros, and v + kOneIndexNu <= kOneIndexChunk)
            double b[kOneIndexNu][kOneIndexRows];
#pragma unroll
            for (int t = 0; t < kOneIndexNu; ++t) {
                const int nu = c0 + min(v + t, len - 1);
#pragma unroll
                for (int u = 0; u < kOneIndexRows; ++u)
                    b[t][u] = flat_b<BT>(sd, (size_t)Ac + (size_t)n_aux * pair_index(mu[u], nu));
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double x[kOneIndexNu];
#pragma unroll
                for (int t = 0; t < kOneIndexNu; ++t) x[t] = cT[v + t][q];
#pragma unroll
                for (int u = 0; u < kOneIndexRows; ++u) {
                    double s = acc[u][q];
#pragma unroll
                    for (int t = 0; t < kOneIndexNu; ++t) s += b[t][u] * x[t];
                    acc[u][q] = s;
                }
            }
        }
    }
    if (!rows || A >= n_aux) return;
#pragma unroll
    for (int u = 0; u < kOneIndexRows; ++u) {
        if (mu0 + u >= n_bf) break;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (q < n) to_global(sd.T)[(size_t)A + (size_t)n_aux * (q + (size_t)n * (mu0 + u))] = acc[u][q];
    }
}

// L[A, (p,q)] = sum_mu T[A,q,mu] Cn[mu,p], q <= p. grid (ceil(n_aux/64), ceil(n/4), slots): a
// wavefront owns 64 aux rows and one q, so every T element is read once; Cn is staged in LDS as
// [mu][16] chunks (zeros beyond n) and read as broadcasts; NQ = n rounded up to even.
template <int NQ>
__global__ void __launch_bounds__(256) k_ri_pair_factors(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_bf = sd.n_bf, n_aux = sd.n_aux;
    if ((int)blockIdx.x * 64 >= n_aux) return;  // block-uniform
    const auto Cn = to_global(sd.Cn);
    __shared__ double cT[kOneIndexChunk][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int A = blockIdx.x * 64 + lane;
    const int Ac = A < n_aux ? A : n_aux - 1;
    const int q = blockIdx.y * 4 + wave;
    const int qc = q < n ? q : n - 1;
    const size_t step = (size_t)n_aux * n;
    const auto t = to_global(sd.T) + (size_t)Ac + (size_t)n_aux * qc;
    double acc[NQ];
#pragma unroll
    for (int p = 0; p < NQ; ++p) acc[p] = 0.0;
    for (int c0 = 0; c0 < n_bf; c0 += kOneIndexChunk) {
        const int len = min(kOneIndexChunk, n_bf - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < kOneIndexChunk * 16; i += blockDim.x) {
            const int v = i >> 4, p = i & 15;
            cT[v][p] = (p < n && v < len) ? Cn[(size_t)p * n_bf + c0 + v] : 0.0;
        }
        __syncthreads();
        int v = 0;
        for (; v + 4 <= len; v += 4) {
            double tv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) tv[u] = t[step * (c0 + v + u)];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int p = 0; p < NQ; ++p) acc[p] += tv[u] * cT[v + u][p];
        }
        for (; v < len; ++v) {
            const double tv = t[step * (c0 + v)];
#pragma unroll
            for (int p = 0; p < NQ; ++p) acc[p] += tv * cT[v][p];
        }
    }
    if (q >= n || A >= n_aux) return;
    const auto L = to_global(sd.L);
#pragma unroll
    for (int p = 0; p < NQ; ++p)
        if (p >= q && p < n) L[(size_t)A + (size_t)n_aux * (p * (p + 1) / 2 + q)] = acc[p];
}

// packed G[(P,Q)] = sum_A L[A,P] L[A,Q]; one wavefront per packed entry.
__global__ void k_ri_pair_gram(const SlotDesc* __restrict__ slots, int np) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_aux = sd.n_aux;
    const auto L = to_global(sd.L);
    const auto G = to_global(sd.G);
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

// The same for batched launches: one block per problem; L is staged in LDS by chunks of 64 aux
// rows as [a][P] (row length kGramLd, zeros past np), and each thread accumulates 2x2 tiles
// (P0, P0+1) x (Q0, Q0+1), P0 >= Q0, over the chunks, so each L element is read from memory once.
constexpr int kGramChunk = 64;
constexpr int kGramLd = MAX_ACT * (MAX_ACT + 1) / 2 + 3;  // np <= 105: room for the 2-wide tiles
constexpr int kGramTilesPerThread = 6;                    // 53 * 54 / 2 tile pairs at np = 105
constexpr unsigned kGramFewProblems = 16;                 // up to this many problems: k_ri_pair_gram

__global__ void __launch_bounds__(256) k_ri_pair_gram_tiled(const SlotDesc* __restrict__ slots, int np) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_aux = sd.n_aux;
    const auto L = to_global(sd.L);
    __shared__ double Ls[kGramChunk * kGramLd];
    const int nt = (np + 1) / 2, ntp = nt * (nt + 1) / 2;
    int p0[kGramTilesPerThread], q0[kGramTilesPerThread];
    double acc[kGramTilesPerThread][4];
#pragma unroll
    for (int k = 0; k < kGramTilesPerThread; ++k) {
        const int t = threadIdx.x + 256 * k;
        int tq = 0, tp = 0;
        if (t < ntp) unpack_pair(t, tq, tp);  // tq <= tp
        p0[k] = 2 * tp;
        q0[k] = 2 * tq;
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[k][j] = 0.0;
    }
    for (int A0 = 0; A0 < n_aux; A0 += kGramChunk) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < kGramChunk * kGramLd; idx += blockDim.x) {
            const int a = idx % kGramChunk, P = idx / kGramChunk, A = A0 + a;
            Ls[a * kGramLd + P] = (A < n_aux && P < np) ? L[(size_t)A + (size_t)n_aux * P] : 0.0;
        }
        __syncthreads();
        for (int a = 0; a < kGramChunk; ++a) {
            const double* row = Ls + a * kGramLd;
#pragma unroll
            for (int k = 0; k < kGramTilesPerThread; ++k) {
                if (threadIdx.x + 256 * k >= ntp) break;  // later k are past the tile count too
                const double x0 = row[p0[k]], x1 = row[p0[k] + 1], y0 = row[q0[k]], y1 = row[q0[k] + 1];
                acc[k][0] += x0 * y0;
                acc[k][1] += x0 * y1;
                acc[k][2] += x1 * y0;
                acc[k][3] += x1 * y1;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kGramTilesPerThread; ++k) {
        if (threadIdx.x + 256 * k >= ntp) break;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int P = p0[k] + (j >> 1), Q = q0[k] + (j & 1);
            if (P < np && Q <= P) to_global(sd.G)[pair_index(P, Q)] = acc[k][j];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// determinant pairs: one wavefront per unordered determinant pair (d <= e)
// ---------------------------------------------------------------------------------------------

// k-th set bit of a mask (k < popcount)
__device__ __forceinline__ int nth_bit(uint64_t m, int k) {
    for (int i = 0; i < k; ++i) m &= m - 1;
    return __builtin_ctzll(m);
}

// Lanes of a wavefront are split into groups of GROUP lanes and each group evaluates one
// unordered determinant pair. An m x m overlap block (m <= MB) sits one entry per lane at
// lane-in-group r + MB*l; blocks up to 4x4 use 16-lane groups (four pairs per wavefront), 5x5
// 32-lane groups (two pairs), 6x6..8x8 the whole wavefront.
template <int GROUP>
struct GroupShape;
template <>
struct GroupShape<16> {
    static constexpr int MB = 4;
};
template <>
struct GroupShape<32> {
    static constexpr int MB = 5;
};
template <>
struct GroupShape<64> {
    static constexpr int MB = 8;
};

// per-group maximum of an int: DPP within 16-lane rows, then shuffles across rows
template <int GROUP>
__device__ __forceinline__ int group_max(int v) {
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false));  // row_mirror
    if (GROUP >= 32) v = max(v, __shfl_xor(v, 16, 64));
    if (GROUP >= 64) v = max(v, __shfl_xor(v, 32, 64));
    return v;
}

// per-group sum; every lane of a group receives the bitwise-identical value
template <int GROUP>
__device__ __forceinline__ double group_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    if (GROUP >= 32) v += __shfl_xor(v, 16, 64);
    if (GROUP >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

struct BlockInverse {
    double dens;    // this lane's D[r,l] = Y(l,r), Y = X^{-1}
    double det;     // group-uniform determinant
    bool singular;  // group-uniform: rank below the block size under the 1e-12 relative pivot rule
};

// Gauss-Jordan elimination with complete pivoting of each group's m x m block, augmented with
// the identity. Pivot choice and the rank rule follow Eigen's FullPivLU as used by
// DeterminantOverlapResolver (determinants/algebra/overlap.cpp:50-110): the pivot is a largest
// remaining |entry| (compared on the high 32 bits of |x|, lowest lane on ties) and the rank counts
// pivots above 1e-12 * max|pivot|; an exactly zero remaining corner is rank deficient. The loop
// runs m steps for every group; a group found singular finishes with meaningless values that the
// caller discards. The inverse is read back per lane as D[r,l] = M(pivot row of l, r)/pivot(l).
template <int GROUP>
__device__ BlockInverse invert_group(double x, int m, int gbase, int r, int l) {
    constexpr int MB = GroupShape<GROUP>::MB;
    const uint64_t gmask = GROUP == 64 ? ~0ull : ((1ull << GROUP) - 1);
    const bool valid = r < m && l < m;
    double aug = (valid && r == l) ? 1.0 : 0.0;
    uint32_t rows_used = 0, cols_used = 0;
    int inversions = 0;
    double det = 1.0, maxpiv = 0.0, minpiv = 1.0e308;
    int my_row = 0;
    double my_piv = 1.0;
    bool singular = false;
    for (int k = 0; k < m; ++k) {
        const bool cand = valid && !((rows_used >> r) & 1u) && !((cols_used >> l) & 1u);
        const int key = cand ? (int)(__double_as_longlong(fabs(x)) >> 32) : -1;
        const int best = group_max<GROUP>(key);
        const uint64_t bal = __ballot(key == best);
        const int idx = __builtin_ctzll(((bal >> gbase) & gmask) | (1ull << (GROUP - 1)));
        const int pr = idx % MB, pc = idx / MB;
        const double p = __shfl(x, gbase + idx, 64);
        if (!(best > 0)) singular = true;  // exactly zero (or denormal) corner
        const double ap = fabs(p);
        maxpiv = fmax(maxpiv, ap);
        minpiv = fmin(minpiv, ap);
        inversions += __builtin_popcount(rows_used >> (pr + 1)) + __builtin_popcount(cols_used >> (pc + 1));
        const double rp = singular ? 0.0 : 1.0 / p;
        const double x_r_pc = __shfl(x, gbase + r + MB * pc, 64);
        const double x_pr_l = __shfl(x, gbase + pr + MB * l, 64);
        const double a_pr_l = __shfl(aug, gbase + pr + MB * l, 64);
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
    const double a = __shfl(aug, gbase + my_row + MB * r, 64);
    out.dens = (valid && !singular) ? a / my_piv : 0.0;
    return out;
}

template <int GROUP>
struct GroupSlab {
    double D[2][GROUP];  // transition densities D_s[r + MB l] = Y_s(l, r)
    double F[2][GROUP];  // dPhi/dD_s
    double Z[GROUP];
    int PP[2][GROUP];    // packed pair index (R_r, L_l) of each block entry
    int occL[2][8];
    int occR[2][8];
};

template <int GROUP>
constexpr size_t slab_bytes_per_wave() {
    return (64 / GROUP) * sizeof(GroupSlab<GROUP>);
}

// Determinant pairs. The active two-electron kernel is held in LDS as the full symmetric np x np
// matrix G2[P + np*Q] so every (pq|rs) lookup is one index add.
// Forward: H_de = S_de * Phi_de, S_de = det X_alpha det X_beta (Lowdin rules in the regular
// inverse form, DeterminantPairEvaluator::combine_spin_pair_evaluations, pairs/evaluator.cpp:354-384).
// Backward: adjoints of S_act, h_act and the packed kernel for F = w (H_de - E S_de),
// accumulated per block in LDS (full np x np when FullDg, packed otherwise) and written as
// packed partials.
// Listed: evaluate only the pairs of the slot's deferred list (left by the one-pair-per-lane kernel);
// backward partial rows then start at part_offset.
template <bool Backward, bool FullDg, int GROUP, bool Listed>
__global__ void __launch_bounds__(512) k_pairs(const SlotDesc* __restrict__ slots, TopoDesc topo, int part_offset) {
    constexpr int MB = GroupShape<GROUP>::MB;
    constexpr int GPW = 64 / GROUP;  // groups per wavefront
    const SlotDesc& sd = slots[blockIdx.z];
    const auto G_glob = to_global(sd.G);
    const int n = topo.n, ng = topo.ng, n_det = topo.n_det;
    const int np = n * (n + 1) / 2;
    const uint64_t* __restrict__ amask = topo.amask;
    const uint64_t* __restrict__ bmask = topo.bmask;
    const int* __restrict__ det_a = topo.det_a;
    const int* __restrict__ det_b = topo.det_b;
    const auto flags = to_global(sd.flags);
    const long long n_pairs = Listed ? (long long)flags[FLAG_DEFERRED] : (long long)n_det * (n_det + 1) / 2;
    // a list block with no pair stays idle and writes no partial row (k_reduce_partials skips it)
    if (Listed && (long long)blockIdx.x * (blockDim.x / GROUP) >= n_pairs) return;

    extern __shared__ double lds[];
    double* G2 = lds;
    double* sact = G2 + np * np;
    const auto hact = sact + n * n;
    double* tail = hact + n * n;
    double *dg = nullptr, *ds = nullptr, *dh = nullptr;
    if (Backward) {
        dg = tail;
        ds = dg + (FullDg ? np * np : ng);
        dh = ds + n * n;
        tail = dh + n * n;
    }
    GroupSlab<GROUP>* slabs = reinterpret_cast<GroupSlab<GROUP>*>(tail);
    if (n_pairs > 0) {
        const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, ln = threadIdx.x & 63;
        for (int Q = wave; Q < np; Q += nwaves)
            for (int P = ln; P < np; P += 64) G2[P + np * Q] = G_glob[pair_index(P, Q)];
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            sact[t] = to_global(sd.sact)[t];
            hact[t] = to_global(sd.hact)[t];
        }
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
    const int wave = threadIdx.x >> 6;
    const int waves_per_block = blockDim.x >> 6;
    const int grp = lane / GROUP, lig = lane % GROUP, gbase = grp * GROUP;
    const int r = lig % MB, l = lig / MB;
    GroupSlab<GROUP>& gs = slabs[wave * GPW + grp];
    const long long first = ((long long)blockIdx.x * waves_per_block + wave) * GPW;
    const long long stride = (long long)gridDim.x * waves_per_block * GPW;
    // every string of a topology has the same electron count per spin
    const int m[2] = {__builtin_popcountll(amask[0]), __builtin_popcountll(bmask[0])};

    for (long long tb = first; tb < n_pairs; tb += stride) {
        const long long ti = tb + grp;
        bool live = ti < n_pairs;
        int d = 0, ee = 0;
        // unordered pair t = e(e+1)/2 + d, d <= e; d is the left (bra) determinant
        const long long tp = live ? (Listed ? (long long)to_global(sd.defer)[ti] : ti) : 0;
        if (live) unpack_pair(tp, d, ee);
        double wH = 0.0, wS = 0.0;
        if (Backward) {
            if (live) pair_weights(sd, topo, d, ee, d, ee, false, d == ee ? 1.0 : 2.0, wH, wS);
            live = live && (wH != 0.0 || wS != 0.0);
        }
        const uint64_t mL[2] = {amask[det_a[d]], bmask[det_b[d]]};
        const uint64_t mR[2] = {amask[det_a[ee]], bmask[det_b[ee]]};
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (lig < m[s]) {
                gs.occL[s][lig] = nth_bit(mL[s], lig);
                gs.occR[s][lig] = nth_bit(mR[s], lig);
            }
        wave_sync();
        // overlap blocks and their inverses
        double det[2], dens[2];
        bool singular = false;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bool valid = r < m[s] && l < m[s];
            double x = 0.0;
            int ppv = 0;
            if (valid) {
                const int Rr = gs.occR[s][r], Ll = gs.occL[s][l];
                x = sact[Rr + n * Ll];
                ppv = pair_index(Rr, Ll);
            }
            if (lig < GROUP) gs.PP[s][lig] = ppv;
            if (m[s] == 0) {
                det[s] = 1.0;
                dens[s] = 0.0;
            } else {
                const BlockInverse inv = invert_group<GROUP>(x, m[s], gbase, r, l);
                det[s] = inv.det;
                dens[s] = inv.dens;
                singular = singular || inv.singular;
            }
            gs.D[s][lig] = dens[s];
        }
        if (!Backward && ti < n_pairs && lig == 0) {
            // rank-deficient under the complete-pivoting rule: evaluated by k_pairs_singular, listed in
            // entry order by k_compact_pairs
            to_global(sd.singular_flag)[ti] = live && singular ? 1 : 0;
        }
        live = live && !singular;
        wave_sync();
        // F_s[r,l] = dPhi/dD_s[r,l] = h + sum (J - K) D_s + sum G D_other; Phi = 1/2 sum_s D_s.(h + F_s)
        double phi_part = 0.0;
        int cross_l[2][MB], cross_r[2][MB];  // pp(R_r, L_l2) and pp(R_r2, L_l) for this lane, per spin
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ms = m[s], mo = m[1 - s];
            double f = 0.0;
            const bool valid = live && r < ms && l < ms;
            const int Rr = valid ? gs.occR[s][r] : 0, Ll = valid ? gs.occL[s][l] : 0;
#pragma unroll
            for (int k = 0; k < MB; ++k) {
                cross_l[s][k] = (valid && k < ms) ? pair_index(Rr, gs.occL[s][k]) : 0;
                cross_r[s][k] = (valid && k < ms) ? pair_index(gs.occR[s][k], Ll) : 0;
            }
            if (valid) {
                const double hv = hact[Rr + n * Ll];
                const double* g1 = G2 + (size_t)np * gs.PP[s][lig];
                f = hv;
#pragma unroll
                for (int l2 = 0; l2 < MB; ++l2) {
                    if (l2 < ms) {
                        const double* gx = G2 + (size_t)np * cross_l[s][l2];
#pragma unroll
                        for (int r2 = 0; r2 < MB; ++r2) {
                            if (r2 < ms) {
                                const int idx = r2 + MB * l2;
                                const double J = g1[gs.PP[s][idx]];
                                const double K = gx[cross_r[s][r2]];
                                f += (J - K) * gs.D[s][idx];
                            }
                        }
                    }
                }
#pragma unroll
                for (int l2 = 0; l2 < MB; ++l2) {
                    if (l2 < mo) {
#pragma unroll
                        for (int r2 = 0; r2 < MB; ++r2) {
                            if (r2 < mo) {
                                const int idx = r2 + MB * l2;
                                f += g1[gs.PP[1 - s][idx]] * gs.D[1 - s][idx];
                            }
                        }
                    }
                }
                phi_part += 0.5 * dens[s] * (hv + f);
            }
            gs.F[s][lig] = f;
        }
        const double phi = group_sum<GROUP>(phi_part);
        const double Dtot = det[0] * det[1];
        if (!Backward) {
            if (live && lig == 0) {
                const double H = Dtot * phi;
                to_global(sd.Hdet)[(size_t)d + (size_t)n_det * ee] = H;
                to_global(sd.Hdet)[(size_t)ee + (size_t)n_det * d] = H;
                to_global(sd.Sdet)[(size_t)d + (size_t)n_det * ee] = Dtot;
                to_global(sd.Sdet)[(size_t)ee + (size_t)n_det * d] = Dtot;
            }
            wave_sync();
            continue;
        }
        // ---- backward: F_pair = wH H + wS S (one state: w (H - E S)) ----
        const double c = Dtot;
        const double cw = c * wH;
        wave_sync();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ms = m[s];
            if (ms == 0) continue;
            const bool valid = live && r < ms && l < ms;
            // Z[a,l] = sum_b F[b,a] D[b,l]
            double z = 0.0;
            if (valid) {
#pragma unroll
                for (int b = 0; b < MB; ++b)
                    if (b < ms) z += gs.F[s][b + MB * r] * gs.D[s][b + MB * l];
            }
            gs.Z[lig] = z;
            wave_sync();
            if (valid) {
                double dfd = 0.0;  // (D F^T D)[r,l]
#pragma unroll
                for (int a2 = 0; a2 < MB; ++a2)
                    if (a2 < ms) dfd += gs.D[s][r + MB * a2] * gs.Z[a2 + MB * l];
                const double Drl = dens[s];
                const int Rr = gs.occR[s][r], Ll = gs.occL[s][l];
                atomicAdd(&ds[Rr + n * Ll], c * ((wH * phi + wS) * Drl - wH * dfd));
                atomicAdd(&dh[Rr + n * Ll], cw * Drl);
                const int P1 = gs.PP[s][lig];
                // same-spin two-electron adjoint over r2 > r, l2 > l
#pragma unroll
                for (int l2 = 1; l2 < MB; ++l2) {
                    if (l2 > l && l2 < ms) {
#pragma unroll
                        for (int r2 = 1; r2 < MB; ++r2) {
                            if (r2 > r && r2 < ms) {
                                const int idx = r2 + MB * l2;
                                const double mnr =
                                    Drl * gs.D[s][idx] - gs.D[s][r + MB * l2] * gs.D[s][r2 + MB * l];
                                const int pu = gs.PP[s][idx];
                                if (FullDg) {
                                    atomicAdd(&dg[P1 + np * pu], cw * mnr);
                                    atomicAdd(&dg[cross_l[s][l2] + np * cross_r[s][r2]], -cw * mnr);
                                } else {
                                    atomicAdd(&dg[pair_index(P1, pu)], cw * mnr);
                                    atomicAdd(&dg[pair_index(cross_l[s][l2], cross_r[s][r2])], -cw * mnr);
                                }
                            }
                        }
                    }
                }
                if (s == 0) {  // opposite spin, once per (alpha entry, beta entry)
                    const int mo = m[1];
#pragma unroll
                    for (int l2 = 0; l2 < MB; ++l2) {
                        if (l2 < mo) {
#pragma unroll
                            for (int r2 = 0; r2 < MB; ++r2) {
                                if (r2 < mo) {
                                    const int idx = r2 + MB * l2;
                                    const double v = cw * Drl * gs.D[1][idx];
                                    const int pu = gs.PP[1][idx];
                                    if (FullDg) atomicAdd(&dg[P1 + np * pu], v);
                                    else atomicAdd(&dg[pair_index(P1, pu)], v);
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
        double* out = sd.partials + (size_t)(blockIdx.x + part_offset) * stride_part;
        if (FullDg) {
            for (int t = threadIdx.x; t < ng; t += blockDim.x) {
                int Q, P;
                unpack_pair(t, Q, P);
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

// sum block partials in a fixed order: out[i] = sum_b partials[b*len + i], one wavefront per i.
// Rows [list_begin, list_begin + list_rows) belong to the deferred-list blocks, which take
// list_per_row pairs per sweep; a row whose block had no pair was not written and is skipped.
__global__ void k_reduce_partials(const SlotDesc* __restrict__ slots, int n_blocks, int len, int list_begin,
                                  int list_rows, int list_per_row) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto partials = to_global(sd.partials);
    const auto out = to_global(sd.dsum);
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= len) return;
    const long long deferred = list_rows > 0 ? (long long)to_global(sd.flags)[FLAG_DEFERRED] : 0;
    double acc = 0.0;
    for (int b = lane; b < n_blocks; b += 64) {
        const int j = b - list_begin;
        if (j >= 0 && j < list_rows && (long long)j * list_per_row >= deferred) continue;
        acc += partials[(size_t)b * len + i];
    }
    acc = wave_sum(acc);
    if (lane == 0) out[i] = acc;
}

// The same sum for launches with few partial rows (at most kReduceRowsFew): one thread per element,
// the rows added in order, consecutive threads on consecutive elements of a row (coalesced), the
// loads of several rows in flight.
constexpr int kReduceRowsFew = 64;
__global__ void __launch_bounds__(256) k_reduce_partials_few(const SlotDesc* __restrict__ slots, int n_blocks, int len,
                                                             int list_begin, int list_rows, int list_per_row) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto partials = to_global(sd.partials);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const long long deferred = list_rows > 0 ? (long long)to_global(sd.flags)[FLAG_DEFERRED] : 0;
    double acc = 0.0;
#pragma unroll 8
    for (int b = 0; b < n_blocks; ++b) {
        const int j = b - list_begin;
        const bool idle = j >= 0 && j < list_rows && (long long)j * list_per_row >= deferred;
        const double v = partials[(size_t)b * len + i];  // every row is allocated: an idle row is read, not added
        acc += idle ? 0.0 : v;
    }
    to_global(sd.dsum)[i] = acc;
}

// Ordered compaction of a pair list: the flagged entries of a per-entry byte array, listed in entry
// order. Deferred == true: entries are the lane kernels' work items and each flagged item lists its
// pair and, when it has one, the spin-flip partner pair (the lane kernels defer the two together);
// the count goes to FLAG_DEFERRED. Deferred == false: entries are the lane-group kernel's pairs (the
// deferred list, or every pair when the shape has no lane kernel) and each flagged entry lists its
// pair; the count goes to FLAG_SINGULAR_PAIRS. Three launches: per-chunk counts, their exclusive
// scan (one block per slot), and the scatter, each chunk of kCompactChunk entries by one block.
template <bool Deferred>
struct CompactView {
    const uint8_t* flag;
    int* out;
    long long entries;
    __device__ CompactView(const SlotDesc& sd, const TopoDesc& topo, int listed) {
        flag = Deferred ? sd.defer_flag : sd.singular_flag;
        out = Deferred ? sd.defer : sd.singular;
        entries = Deferred ? topo.n_lane_pairs
                  : listed ? (long long)to_global(sd.flags)[FLAG_DEFERRED]
                           : (long long)topo.n_det * (topo.n_det + 1) / 2;
    }
    // pairs listed for entry i (0, 1 or 2)
    __device__ int emit(const SlotDesc& sd, const TopoDesc& topo, int listed, long long i, int& first,
                        int& second) const {
        if (i >= entries || !flag[i]) return 0;
        if (Deferred) {
            const long long t = topo.lane_pairs ? (long long)topo.lane_pairs[i] : i;
            long long tf = t;
            if (topo.flip) {
                int d, e;
                unpack_pair(t, d, e);
                const int a = topo.flip[d], b = topo.flip[e];
                const long long lo = a < b ? a : b, hi = a < b ? b : a;
                tf = hi * (hi + 1) / 2 + lo;
            }
            first = (int)t;
            second = (int)tf;
            return tf != t ? 2 : 1;
        }
        first = listed ? to_global(sd.defer)[i] : (int)i;
        return 1;
    }
};

// wave-level exclusive prefix of per-lane counts (0..2) and the block's per-wave totals (LDS)
__device__ __forceinline__ int compact_prefix(int count, int* wave_total, int& block_before_wave, int& block_total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const unsigned long long one = __ballot(count >= 1), two = __ballot(count == 2);
    const int before = __popcll(one & below) + __popcll(two & below);
    if (lane == 0) wave_total[wave] = __popcll(one) + __popcll(two);
    __syncthreads();
    block_before_wave = 0;
    block_total = 0;
    for (int w = 0; w < nw; ++w) {
        if (w < wave) block_before_wave += wave_total[w];
        block_total += wave_total[w];
    }
    __syncthreads();
    return before;
}

template <bool Deferred>
__global__ void __launch_bounds__(kCompactChunk) k_compact_count(const SlotDesc* __restrict__ slots, TopoDesc topo,
                                                                 int listed) {
    const SlotDesc& sd = slots[blockIdx.z];
    const CompactView<Deferred> v(sd, topo, listed);
    __shared__ int wave_total[kCompactChunk / 64];
    const long long chunks = (v.entries + kCompactChunk - 1) / kCompactChunk;
    for (long long c = blockIdx.x; c < chunks; c += gridDim.x) {
        int first, second;
        const int count = v.emit(sd, topo, listed, c * kCompactChunk + threadIdx.x, first, second);
        int bw, total;
        (void)compact_prefix(count, wave_total, bw, total);
        if (threadIdx.x == 0) to_global(sd.chunk_count)[c] = total;
    }
}

// exclusive scan of the chunk counts in place (one block per slot), total to the slot's flag
template <bool Deferred>
__global__ void __launch_bounds__(1024) k_compact_scan(const SlotDesc* __restrict__ slots, TopoDesc topo, int listed) {
    const SlotDesc& sd = slots[blockIdx.z];
    const CompactView<Deferred> v(sd, topo, listed);
    const long long chunks = (v.entries + kCompactChunk - 1) / kCompactChunk;
    __shared__ int part[1024];
    long long carry = 0;
    for (long long base = 0; base < chunks; base += blockDim.x) {
        const long long c = base + threadIdx.x;
        const int x = c < chunks ? to_global(sd.chunk_count)[c] : 0;
        part[threadIdx.x] = x;
        __syncthreads();
        // Hillis-Steele inclusive scan over the block
        for (int off = 1; off < (int)blockDim.x; off <<= 1) {
            const int y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        if (c < chunks) to_global(sd.chunk_count)[c] = (int)(carry + part[threadIdx.x] - x);
        carry += part[blockDim.x - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) to_global(sd.flags)[Deferred ? FLAG_DEFERRED : FLAG_SINGULAR_PAIRS] = (int)carry;
}

template <bool Deferred>
__global__ void __launch_bounds__(kCompactChunk) k_compact_scatter(const SlotDesc* __restrict__ slots, TopoDesc topo,
                                                                   int listed) {
    const SlotDesc& sd = slots[blockIdx.z];
    const CompactView<Deferred> v(sd, topo, listed);
    __shared__ int wave_total[kCompactChunk / 64];
    const long long chunks = (v.entries + kCompactChunk - 1) / kCompactChunk;
    for (long long c = blockIdx.x; c < chunks; c += gridDim.x) {
        int first = 0, second = 0;
        const int count = v.emit(sd, topo, listed, c * kCompactChunk + threadIdx.x, first, second);
        int bw, total;
        const int before = compact_prefix(count, wave_total, bw, total);
        const long long at = (long long)to_global(sd.chunk_count)[c] + bw + before;
        if (count >= 1) v.out[at] = first;
        if (count == 2) v.out[at + 1] = second;
    }
}

// The deferred (Deferred) or rank-deficient list of every slot of a launch, in entry order. The
// singular list's entry count is read on the device (the deferred count), so it runs after the
// deferred list's kernels on the same stream.
template <bool Deferred>
void compact_pairs(hipStream_t st, const SlotDesc* slots, unsigned z, const TopoDesc& topo, int listed,
                   long long max_entries) {
    const unsigned blocks = (unsigned)std::max<long long>(
        1, std::min<long long>(kCompactBlocks, (max_entries + kCompactChunk - 1) / kCompactChunk));
    hipLaunchKernelGGL(k_compact_count<Deferred>, dim3(blocks, 1, z), dim3(kCompactChunk), 0, st, slots, topo, listed);
    hipLaunchKernelGGL(k_compact_scan<Deferred>, dim3(1, 1, z), dim3(1024), 0, st, slots, topo, listed);
    hipLaunchKernelGGL(k_compact_scatter<Deferred>, dim3(blocks, 1, z), dim3(kCompactChunk), 0, st, slots, topo,
                       listed);
}

// ---------------------------------------------------------------------------------------------
// structure matrices: Y = Hdet T (thread per (det, structure)), then T^T Y (thread per (I, J))
// ---------------------------------------------------------------------------------------------
__global__ void k_structure_right(const SlotDesc* __restrict__ slots, TopoDesc topo) {
    int bx, bz;
    problem_block(bx, bz);  // a problem's workgroups on one XCD: each Hdet/Sdet column comes from HBM once
    const SlotDesc& sd = slots[bz];
    const auto Hdet = to_global(sd.Hdet);
    const auto Sdet = to_global(sd.Sdet);
    const int n_det = topo.n_det, n_str = topo.n_str;
    const long long* __restrict__ soff = topo.soff;
    const int* __restrict__ sdet = topo.sdet;
    const double* __restrict__ scoef = topo.scoef;
    const auto Yh = to_global(sd.Yh);
    const auto Ys = to_global(sd.Ys);
    const long long id = (long long)bx * blockDim.x + threadIdx.x;
    if (id >= (long long)n_det * n_str) return;
    const int d = (int)(id % n_det), J = (int)(id / n_det);
    double h = 0.0, s = 0.0;
    // unrolled: the column loads of several determinants in flight at once
#pragma unroll 8
    for (long long t = soff[J]; t < soff[J + 1]; ++t) {
        const size_t at = (size_t)d + (size_t)n_det * sdet[t];
        h += scoef[t] * Hdet[at];
        s += scoef[t] * Sdet[at];
    }
    Yh[id] = h;
    Ys[id] = s;
}

__global__ void k_structure_left(const SlotDesc* __restrict__ slots, TopoDesc topo) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto Yh = to_global(sd.Yh);
    const auto Ys = to_global(sd.Ys);
    const int n_det = topo.n_det, n_str = topo.n_str;
    const long long* __restrict__ soff = topo.soff;
    const int* __restrict__ sdet = topo.sdet;
    const double* __restrict__ scoef = topo.scoef;
    const auto Hs = to_global(sd.Hs);
    const auto Ss = to_global(sd.Ss);
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)n_str * n_str) return;
    const int I = (int)(id % n_str), J = (int)(id / n_str);
    double h = 0.0, s = 0.0;
#pragma unroll 8
    for (long long t = soff[I]; t < soff[I + 1]; ++t) {
        const size_t at = (size_t)sdet[t] + (size_t)n_det * J;
        h += scoef[t] * Yh[at];
        s += scoef[t] * Ys[at];
    }
    Hs[id] = h;
    Ss[id] = s;
}


// ---------------------------------------------------------------------------------------------
// generalised eigenproblem H c = E S c, lowest root, one workgroup.
// Cholesky S = L L^T, explicit L^{-1}, M = L^{-1} H L^{-T} by two triangular products,
// Householder tridiagonalisation, multisection on Sturm counts for the lowest eigenvalue, inverse
// iteration on the tridiagonal, back-transformation, c = L^{-T} z so that c^T S c = 1.
// Storage: gep_work_doubles(n) = 3*n*n + 9*n doubles, in LDS when it fits (InLds) else in the
// global workspace. A single wavefront is used for small n so that the block barriers are free.
// ---------------------------------------------------------------------------------------------
// Front half of the eigensolver for n_str whose matrices live in global memory, over many
// workgroups: k_gep_chol (one workgroup per problem: symmetrised S and H, Cholesky S = L L^T in
// place), k_gep_linv (L^-1 by independent columns, one wavefront per column: X1 holds it row-major and
// V column-major), k_gep_hlt (U = H L^-T into X2) and k_gep_lul (A = L^-1 U into M), after which
// k_gep_lowest<false, 3> continues from the symmetrisation of A (launches of few problems replace
// the Cholesky and the Householder stage by the per-step kernels further down). Workspace layout
// as in k_gep_lowest (Lm, M, V, 9 n vectors) followed by X1, X2.
__device__ __forceinline__ size_t gep_core_dev(int n) { return 3 * (size_t)n * n + 9 * (size_t)n; }

__global__ void __launch_bounds__(1024) k_gep_chol(const SlotDesc* __restrict__ slots, TopoDesc topo) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n_str;
    const auto Lm = to_global(sd.gep);
    const auto M = Lm + (size_t)n * n;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int TX = 16, TY = nt / TX, tx = tid % TX, ty = tid / TX;
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        Lm[idx] = 0.5 * (sd.Ss[i + n * j] + to_global(sd.Ss)[j + n * i]);
        M[idx] = 0.5 * (sd.Hs[i + n * j] + to_global(sd.Hs)[j + n * i]);
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double a = Lm[k + n * k];
        if (!(a > 0.0)) {
            if (tid == 0) to_global(sd.flags)[FLAG_METRIC] = 1;
            return;  // uniform; the later kernels run on meaningless values the host discards
        }
        const double dk = sqrt(a), rdk = 1.0 / dk;
        __syncthreads();
        for (int i = k + tid; i < n; i += nt) Lm[i + n * k] = (i == k) ? dk : Lm[i + n * k] * rdk;
        __syncthreads();
        for (int j = k + 1 + ty; j < n; j += TY) {
            const double ljk = Lm[j + n * k];
            for (int i = j + tx; i < n; i += TX) Lm[i + n * j] -= Lm[i + n * k] * ljk;
        }
        __syncthreads();
    }
}

// X = L^-1 by columns, one wavefront per column j: X[i][j] = (delta_ij - sum_{k=j}^{i-1} L[i][k]
// X[k][j]) / L[i][i] for i = j..n-1 in order, each sum split over the 64 lanes, the column kept in
// LDS as it grows. grid (ceil(n/4), 1, slots), 256 threads, dynamic LDS 4 n doubles
// (n <= kLinvWaveMaxN; larger n take k_gep_linv_lane).
constexpr int kLinvWaveMaxN = 2048;

__global__ void __launch_bounds__(256) k_gep_linv(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto L = to_global(sd.gep);
    const auto V = to_global(sd.gep) + 2 * (size_t)n * n;
    const auto X1 = to_global(sd.gep) + gep_core_dev(n);
    extern __shared__ double xs[];  // four columns of n
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = blockIdx.x * 4 + wave;
    if (j >= n) return;  // wave-uniform; the kernel has no block barrier
    double* x = xs + (size_t)wave * n;
    for (int i = lane; i < j; i += 64) {
        X1[j + (size_t)n * i] = 0.0;
        V[i + (size_t)n * j] = 0.0;
    }
    for (int i = j; i < n; ++i) {
        double s = 0.0;
        for (int k = j + lane; k < i; k += 64) s += L[i + (size_t)n * k] * x[k];
        s = wave_sum(s);
        const double xi = ((i == j) ? 1.0 - s : -s) / L[i + (size_t)n * i];
        if (lane == 0) {
            x[i] = xi;
            X1[j + (size_t)n * i] = xi;  // row-major: X1[j + n i] = X[i][j]
            V[i + (size_t)n * j] = xi;   // column-major
        }
        wave_sync();  // x[i] is visible to every lane before row i + 1
    }
}

// The same with one lane per column (no LDS; for n above kLinvWaveMaxN).
__global__ void __launch_bounds__(64) k_gep_linv_lane(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto L = to_global(sd.gep);
    const auto V = to_global(sd.gep) + 2 * (size_t)n * n;
    const auto X1 = to_global(sd.gep) + gep_core_dev(n);
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= n) return;
    for (int i = 0; i < j; ++i) {
        X1[j + (size_t)n * i] = 0.0;
        V[i + (size_t)n * j] = 0.0;
    }
    for (int i = j; i < n; ++i) {
        double s0 = (i == j) ? 1.0 : 0.0, s1 = 0.0;
        int k = j;
        for (; k + 1 < i; k += 2) {
            s0 -= L[i + (size_t)n * k] * X1[j + (size_t)n * k];
            s1 -= L[i + (size_t)n * (k + 1)] * X1[j + (size_t)n * (k + 1)];
        }
        for (; k < i; ++k) s0 -= L[i + (size_t)n * k] * X1[j + (size_t)n * k];
        const double x = (s0 + s1) / L[i + (size_t)n * i];
        X1[j + (size_t)n * i] = x;  // row-major: X1[j + n i] = X[i][j]
        V[i + (size_t)n * j] = x;   // column-major
    }
}

// U = H L^-T: U[i][j] = sum_{k<=j} H[i][k] X[j][k] (thread = row i, grid y = column j)
__global__ void __launch_bounds__(256) k_gep_hlt(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto H = to_global(sd.gep) + (size_t)n * n;  // M holds the symmetrised H
    const auto X1 = to_global(sd.gep) + gep_core_dev(n);
    const auto X2 = to_global(sd.gep) + gep_core_dev(n) + (size_t)n * n;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    double s0 = 0.0, s1 = 0.0;
    int k = 0;
    for (; k + 1 <= j; k += 2) {
        s0 += H[i + (size_t)n * k] * X1[k + (size_t)n * j];
        s1 += H[i + (size_t)n * (k + 1)] * X1[k + 1 + (size_t)n * j];
    }
    for (; k <= j; ++k) s0 += H[i + (size_t)n * k] * X1[k + (size_t)n * j];
    X2[i + (size_t)n * j] = s0 + s1;
}

// A = L^-1 U: A[i][j] = sum_{k<=i} X[i][k] U[k][j] into M (thread = row i, grid y = column j)
__global__ void __launch_bounds__(256) k_gep_lul(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto A = to_global(sd.gep) + (size_t)n * n;
    const auto V = to_global(sd.gep) + 2 * (size_t)n * n;
    const auto X2 = to_global(sd.gep) + gep_core_dev(n) + (size_t)n * n;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    double s0 = 0.0, s1 = 0.0;
    int k = 0;
    for (; k + 1 <= i; k += 2) {
        s0 += V[i + (size_t)n * k] * X2[k + (size_t)n * j];
        s1 += V[i + (size_t)n * (k + 1)] * X2[k + 1 + (size_t)n * j];
    }
    for (; k <= i; ++k) s0 += V[i + (size_t)n * k] * X2[k + (size_t)n * j];
    A[i + (size_t)n * j] = s0 + s1;
}

// Blocked Cholesky of the large-n path for launches of few problems: k_gep_load symmetrises S
// and H into the workspace (many blocks); then per panel of kCholPanel columns j0.. j0+jb-1:
// k_chol_diag factors the diagonal tile in LDS (one block; a pivot a <= 0 sets the metric flag
// and the solve goes on with values the host discards), k_chol_trsm solves the rows below it
// against the tile (one thread per row), k_chol_syrk subtracts the panel's product from the
// trailing lower triangle (16 x 16 tiles). About 3 n / kCholPanel launches instead of 2 n.
__global__ void __launch_bounds__(256) k_gep_load(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto Lm = to_global(sd.gep);
    const auto M = Lm + (size_t)n * n;
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    Lm[idx] = 0.5 * (sd.Ss[i + (size_t)n * j] + to_global(sd.Ss)[j + (size_t)n * i]);
    M[idx] = 0.5 * (sd.Hs[i + (size_t)n * j] + to_global(sd.Hs)[j + (size_t)n * i]);
}

constexpr int kCholPanel = 32;

__global__ void __launch_bounds__(256) k_chol_diag(const SlotDesc* __restrict__ slots, int n, int j0, int jb) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto L = to_global(sd.gep);
    __shared__ double t[kCholPanel][kCholPanel + 1];  // t[i][j], lower triangle
    for (int idx = threadIdx.x; idx < jb * jb; idx += blockDim.x) {
        const int i = idx % jb, j = idx / jb;
        t[i][j] = L[(size_t)(j0 + i) + (size_t)n * (j0 + j)];
    }
    __syncthreads();
    for (int k = 0; k < jb; ++k) {
        const double a = t[k][k];
        if (!(a > 0.0) && threadIdx.x == 0) to_global(sd.flags)[FLAG_METRIC] = 1;
        const double dk = sqrt(a), rdk = 1.0 / dk;
        __syncthreads();  // every thread has read the pivot before it is replaced
        for (int i = k + threadIdx.x; i < jb; i += blockDim.x) t[i][k] = (i == k) ? dk : t[i][k] * rdk;
        __syncthreads();
        const int m = jb - k - 1;
        for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
            const int i = k + 1 + idx % m, j = k + 1 + idx / m;
            if (j <= i) t[i][j] -= t[i][k] * t[j][k];
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < jb * jb; idx += blockDim.x) {
        const int i = idx % jb, j = idx / jb;
        if (j <= i) L[(size_t)(j0 + i) + (size_t)n * (j0 + j)] = t[i][j];
    }
}

// rows i >= j0 + jb: L[i][j0..j0+jb) = A[i][j0..j0+jb) L_tile^-T (forward substitution per row)
__global__ void __launch_bounds__(256) k_chol_trsm(const SlotDesc* __restrict__ slots, int n, int j0, int jb) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto L = to_global(sd.gep);
    __shared__ double t[kCholPanel][kCholPanel + 1];
    for (int idx = threadIdx.x; idx < jb * jb; idx += blockDim.x) {
        const int i = idx % jb, j = idx / jb;
        t[i][j] = j <= i ? L[(size_t)(j0 + i) + (size_t)n * (j0 + j)] : 0.0;
    }
    __syncthreads();
    const int i = j0 + jb + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[kCholPanel];
#pragma unroll
    for (int j = 0; j < kCholPanel; ++j) {
        if (j < jb) {
            double v = L[(size_t)i + (size_t)n * (j0 + j)];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= x[k] * t[j][k];
            x[j] = v / t[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < kCholPanel; ++j)
        if (j < jb) L[(size_t)i + (size_t)n * (j0 + j)] = x[j];
}

// trailing update: A[i][j] -= sum_k L[i][j0+k] L[j][j0+k] for j0 + jb <= j <= i < n, 16 x 16 tiles
__global__ void __launch_bounds__(256) k_chol_syrk(const SlotDesc* __restrict__ slots, int n, int j0, int jb) {
    const int ti = blockIdx.x, tj = blockIdx.y;
    if (tj > ti) return;  // block-uniform: tiles above the diagonal
    const SlotDesc& sd = slots[blockIdx.z];
    const auto L = to_global(sd.gep);
    __shared__ double li[16][kCholPanel + 1], lj[16][kCholPanel + 1];
    const int r0 = j0 + jb + 16 * ti, c0 = j0 + jb + 16 * tj;
    for (int idx = threadIdx.x; idx < 16 * jb; idx += blockDim.x) {
        const int r = idx % 16, k = idx / 16;
        li[r][k] = r0 + r < n ? L[(size_t)(r0 + r) + (size_t)n * (j0 + k)] : 0.0;
        lj[r][k] = c0 + r < n ? L[(size_t)(c0 + r) + (size_t)n * (j0 + k)] : 0.0;
    }
    __syncthreads();
    const int r = threadIdx.x % 16, c = threadIdx.x / 16;
    const int i = r0 + r, j = c0 + c;
    if (i >= n || j > i) return;
    double acc = 0.0;
    for (int k = 0; k < jb; ++k) acc += li[r][k] * lj[c][k];
    L[(size_t)i + (size_t)n * j] -= acc;
}

// Householder step k with the previous step's rank-2 update applied lazily, one launch per step.
// Step k - 1 left p_{k-1} = beta_{k-1} A_{k-1} v_{k-1} (in pbuf[(k-1) & 1]); every block forms
// w_{k-1} = p_{k-1} - K v_{k-1} (K = beta_{k-1} v_{k-1}^T p_{k-1} / 2) and v_k from column k of
// A_k = A_{k-1} - v w^T - w v^T (corrected on the fly; column k itself is never written back).
// Wavefront i (i > k) then reads column i once: applies the step k-1 update to rows >= k+1, writes
// it back and accumulates p_i = beta_k (A_k v_k)_i. Block 0 stores v_k, beta_k, alpha_k, the
// corrected diagonal A[k][k], and the reduced column and row k - 1 (alpha_{k-1} on the
// subdiagonal, zeros below), which no block reads during this step. Each column is read and
// written by its own wavefront only. k_hh_finish applies the last step to the trailing 2 x 2
// block and writes the last reduced column. Workspace: M, V as in k_gep_lowest; p in y / tdl by
// parity of k, alpha_k in tdu.
__device__ __forceinline__ void hh_pending(gptr<const double> V, gptr<const double> pbuf, gptr<const double> beta,
                                           int n, int k, double* vprev, double* wprev, double* red) {
    // v_{k-1} and w_{k-1} into LDS (zeros above row k and for k == 0)
    const int tid = threadIdx.x, nt = blockDim.x;
    double part = 0.0;
    for (int i = tid; i < n; i += nt) {
        const double v = (k > 0 && i >= k) ? V[i + (size_t)n * (k - 1)] : 0.0;
        vprev[i] = v;
        if (k > 0 && i >= k) part += v * pbuf[i];
    }
    const double K = k > 0 ? 0.5 * beta[k - 1] * block_sum(part, red) : 0.0;
    for (int i = tid; i < n; i += nt) wprev[i] = (k > 0 && i >= k) ? pbuf[i] - K * vprev[i] : 0.0;
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_hh_step(const SlotDesc* __restrict__ slots, int n, int k) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto M = to_global(sd.gep) + (size_t)n * n;
    const auto V = to_global(sd.gep) + 2 * (size_t)n * n;
    const auto vec = to_global(sd.gep) + 3 * (size_t)n * n;  // dd, ee, y, beta, tdl, tdu, ...
    const auto beta = vec + 3 * (size_t)n;
    const auto alphas = vec + 5 * (size_t)n;
    const auto pprev = vec + ((k - 1) & 1 ? 4 : 2) * (size_t)n;
    const auto pnow = vec + (k & 1 ? 4 : 2) * (size_t)n;
    extern __shared__ double loc[];  // v_k, v_{k-1}, w_{k-1}: 3 n
    double* vloc = loc;
    double* vprev = loc + n;
    double* wprev = loc + 2 * n;
    __shared__ double red[16];
    const int m0 = k + 1, tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
    hh_pending(V, pprev, beta, n, k, vprev, wprev, red);
    const double vk = vprev[k], wk = wprev[k];
    double part = 0.0;
    for (int i = m0 + tid; i < n; i += nt) {
        const double x = M[i + (size_t)n * k] - vprev[i] * wk - wprev[i] * vk;
        vloc[i] = x;
        part += x * x;
    }
    const double sigma = block_sum(part, red);  // ends with a barrier: vloc is complete
    const double x0 = vloc[m0];
    const double alpha = sigma > 0.0 ? (x0 >= 0.0 ? -sqrt(sigma) : sqrt(sigma)) : 0.0;
    const double vtv = sigma - 2.0 * alpha * x0 + alpha * alpha;
    const double bk = (sigma > 0.0 && vtv > 0.0) ? 2.0 / vtv : 0.0;
    __syncthreads();  // every thread has read x0 before it becomes v_m0
    if (tid == 0) vloc[m0] = x0 - alpha;
    for (int i = tid; i < m0; i += nt) vloc[i] = 0.0;
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int i = m0 + tid; i < n; i += nt) V[i + (size_t)n * k] = vloc[i];
        if (tid == 0) {
            beta[k] = bk;
            alphas[k] = alpha;
            M[k + (size_t)n * k] -= 2.0 * vk * wk;
        }
        if (k > 0)
            for (int i = k + tid; i < n; i += nt) {
                const double e = (i == k) ? alphas[k - 1] : 0.0;
                M[i + (size_t)n * (k - 1)] = e;
                M[(k - 1) + (size_t)n * i] = e;
            }
    }
    const int i = m0 + (int)blockIdx.x * 4 + wave;
    if (i < n) {  // wave-uniform: column i, rows j >= m0 on consecutive lanes
        const double vi = vprev[i], wi = wprev[i];
        double a = 0.0;
        for (int j = m0 + lane; j < n; j += 64) {
            const double mji = M[j + (size_t)n * i] - vprev[j] * wi - wprev[j] * vi;
            M[j + (size_t)n * i] = mji;
            a += mji * vloc[j];
        }
        a = wave_sum(a);
        if (lane == 0) pnow[i] = bk * a;
    }
}

__global__ void __launch_bounds__(256) k_hh_finish(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto M = to_global(sd.gep) + (size_t)n * n;
    const auto V = to_global(sd.gep) + 2 * (size_t)n * n;
    const auto vec = to_global(sd.gep) + 3 * (size_t)n * n;
    const auto beta = vec + 3 * (size_t)n;
    const auto alphas = vec + 5 * (size_t)n;
    const int k = n - 2;  // the step whose update is pending is k - 1 = n - 3
    const auto pprev = vec + ((k - 1) & 1 ? 4 : 2) * (size_t)n;
    extern __shared__ double loc[];  // v_{k-1}, w_{k-1}: 2 n
    double* vprev = loc;
    double* wprev = loc + n;
    __shared__ double red[16];
    hh_pending(V, pprev, beta, n, k, vprev, wprev, red);
    const int tid = threadIdx.x;
    if (tid < 4) {  // the trailing 2 x 2 block (rows and columns n-2, n-1)
        const int j = k + (tid & 1), c = k + (tid >> 1);
        M[j + (size_t)n * c] -= vprev[j] * wprev[c] + wprev[j] * vprev[c];
    }
    for (int i = k + tid; i < n; i += blockDim.x) {
        const double e = (i == k) ? alphas[k - 1] : 0.0;
        if (i > k) {
            M[i + (size_t)n * (k - 1)] = e;
            M[(k - 1) + (size_t)n * i] = e;
        }
    }
    if (tid == 0) {
        M[k + (size_t)n * (k - 1)] = alphas[k - 1];
        M[(k - 1) + (size_t)n * k] = alphas[k - 1];
    }
}

// Part 0: the whole solve (matrices in LDS). Large-n path (matrices in global memory, after the
// multi-workgroup front half): Part 3 is the rest on one workgroup per problem (batched launches);
// for launches of few problems, Part 1 is the symmetrisation and Part 2 the rest after the
// per-step Householder kernels (k_hh_step, k_hh_finish).
template <bool InLds, int Part = 0>
__global__ void __launch_bounds__(InLds ? 256 : 1024) k_gep_lowest(const SlotDesc* __restrict__ slots, TopoDesc topo) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto Hs = to_global(sd.Hs);
    const auto Ss = to_global(sd.Ss);
    const int n = topo.n_str;
    const double e_nuc = sd.e_nuc;
    double* __restrict__ gwork = sd.gep;
    const auto coef = to_global(sd.coef);
    const auto energy = to_global(sd.energy);
    const auto flags = to_global(sd.flags);
    const long long* __restrict__ off = topo.off;
    const int* __restrict__ ts = topo.ts;
    const double* __restrict__ tc = topo.tc;
    const int n_det = topo.n_det;
    const auto dcoef = to_global(sd.dcoef);
    extern __shared__ double lds[];
    // LDS or device memory, typed as such so that neither instantiation uses flat accesses
    const auto work = [&] {
        if constexpr (InLds) return static_cast<double*>(lds);
        else return to_global(gwork);
    }();
    const auto Lm = work;              // n x n, lower Cholesky factor
    const auto M = work + n * n;       // n x n
    const auto V = work + 2 * n * n;   // Householder vectors (column k holds v_k)
    const auto dd = work + 3 * n * n;  // diagonal
    const auto ee = dd + n;            // off-diagonal
    const auto y = ee + n;             // tridiagonal eigenvector / scratch
    const auto beta = y + n;           // Householder scalars
    const auto tdl = beta + n;         // tridiagonal LU factors (inverse iteration)
    const auto tdu = tdl + n;
    const auto tdu2 = tdu + n;
    const auto tdg = tdu2 + n;
    const auto ipiv = [&] {  // n ints in the last n doubles
        if constexpr (InLds) return reinterpret_cast<int*>(tdg + n);
        else return reinterpret_cast<gptr<int>>(tdg + n);
    }();
    __shared__ double red[16];
    __shared__ double sh_scalar[4];
    __shared__ double lo_hi[2];
    __shared__ int first_true;
    const int tid = threadIdx.x, nt = blockDim.x;
    // 2D thread layout for element-parallel updates: TX consecutive lanes share a row
    const int TX = nt == 64 ? 8 : 16, TY = nt / TX, tx = tid % TX, ty = tid / TX;
    if constexpr (Part == 0) {
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        Lm[idx] = 0.5 * (Ss[i + n * j] + Ss[j + n * i]);
        M[idx] = 0.5 * (Hs[i + n * j] + Hs[j + n * i]);
        V[idx] = 0.0;
    }
    __syncthreads();
    // Cholesky S = L L^T (right-looking; the lower triangle of Lm becomes L), trailing update
    // spread over all threads element by element
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
        for (int j = k + 1 + ty; j < n; j += TY) {  // column j, rows i >= j: consecutive lanes, consecutive i
            const double ljk = Lm[j + n * k];
            for (int i = j + tx; i < n; i += TX) Lm[i + n * j] -= Lm[i + n * k] * ljk;
        }
        __syncthreads();
    }
    // Linv = L^{-1} into V (lower triangle), one row per sweep: V[i,j] = -(sum_{k=j}^{i-1} L[i,k] V[k,j]) / L[i,i]
    for (int i = 0; i < n; ++i) {
        const double rl = 1.0 / Lm[i + n * i];
        for (int j = tid; j <= i; j += nt) {
            if (j == i) {
                V[i + n * i] = rl;
            } else {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                int k = j;
                for (; k + 3 < i; k += 4) {
                    s0 += Lm[i + n * k] * V[k + n * j];
                    s1 += Lm[i + n * (k + 1)] * V[k + 1 + n * j];
                    s2 += Lm[i + n * (k + 2)] * V[k + 2 + n * j];
                    s3 += Lm[i + n * (k + 3)] * V[k + 3 + n * j];
                }
                for (; k < i; ++k) s0 += Lm[i + n * k] * V[k + n * j];
                V[i + n * j] = -((s0 + s1) + (s2 + s3)) * rl;
            }
        }
        __syncthreads();
    }
    // standard form A = Linv H Linv^T: T1 = Linv H into Lm (L is not needed any more), A into M.
    // Consecutive lanes take consecutive columns of one T1 row so that the trip count (i + 1) is
    // uniform across them; H is symmetric, so H[k,j] is read as H[j,k] (unit stride across lanes).
    for (int idx = tid; idx < n * n; idx += nt) {
        const int j = idx % n, i = idx / n;
        double s0 = 0.0, s1 = 0.0;
        int k = 0;
        for (; k + 1 <= i; k += 2) {
            s0 += V[i + n * k] * M[j + n * k];
            s1 += V[i + n * (k + 1)] * M[j + n * (k + 1)];
        }
        for (; k <= i; ++k) s0 += V[i + n * k] * M[j + n * k];
        Lm[i + n * j] = s0 + s1;
    }
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        double s0 = 0.0, s1 = 0.0;
        int k = 0;
        for (; k + 1 <= j; k += 2) {
            s0 += Lm[i + n * k] * V[j + n * k];
            s1 += Lm[i + n * (k + 1)] * V[j + n * (k + 1)];
        }
        for (; k <= j; ++k) s0 += Lm[i + n * k] * V[j + n * k];
        M[idx] = s0 + s1;
    }
    __syncthreads();
    }
    // symmetrise A; keep Linv in Lm (for c = Linv^T z at the end) and clear V for the reflectors
    if constexpr (Part != 2) {
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        Lm[idx] = (i >= j) ? V[idx] : 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        V[idx] = 0.0;
        if (i < j) {
            const double s2 = 0.5 * (M[i + n * j] + M[j + n * i]);
            M[i + n * j] = s2;
            M[j + n * i] = s2;
        }
    }
    __syncthreads();
    }
    if constexpr (Part == 1) return;
    // Householder tridiagonalisation: reflect rows/cols k+1..n-1 of column k
    if constexpr (Part == 0 || Part == 3) {
    for (int k = 0; k < n - 2; ++k) {
        const int m0 = k + 1;
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
        // p = bk A v: each row split over the TX lanes of a row group, summed across those lanes
        for (int i0 = m0; i0 < n; i0 += TY) {
            const int i = i0 + ty;
            double a0 = 0.0;
            if (i < n)  // A is symmetric: row i read as column i, consecutive across the lanes
                for (int j = m0 + tx; j < n; j += TX) a0 += M[j + n * i] * V[j + n * k];
            a0 += dpp_f64<0xB1>(a0);
            a0 += dpp_f64<0x4E>(a0);
            a0 += dpp_f64<0x141>(a0);
            if (TX == 16) a0 += dpp_f64<0x140>(a0);
            if (i < n && tx == 0) y[i] = bk * a0;
        }
        __syncthreads();
        double kp = 0.0;
        for (int i = m0 + tid; i < n; i += nt) kp += V[i + n * k] * y[i];
        const double K = 0.5 * bk * block_sum(kp, red);
        for (int i = m0 + tid; i < n; i += nt) y[i] -= K * V[i + n * k];  // w = p - K v
        __syncthreads();
        for (int i = m0 + ty; i < n; i += TY) {  // A -= v w^T + w v^T, column i, consecutive rows j
            const double vi = V[i + n * k], wi = y[i];
            for (int j = m0 + tx; j < n; j += TX) M[j + n * i] -= V[j + n * k] * wi + y[j] * vi;
        }
        if (bk != 0.0)
            for (int i = m0 + tid; i < n; i += nt) {
                const double v = (i == m0) ? alpha : 0.0;
                M[i + n * k] = v;
                M[k + n * i] = v;
            }
        __syncthreads();
    }
    }
    for (int i = tid; i < n; i += nt) {
        dd[i] = M[i + n * i];
        ee[i] = (i + 1 < n) ? M[(i + 1) + n * i] : 0.0;
    }
    __syncthreads();
    // Lowest eigenvalue by multisection on Sturm counts, then its eigenvector by inverse
    // iteration on T - lambda I (LU with partial pivoting, LAPACK dgttrf/dgttrs scheme).
    // For n <= 64 the tridiagonal lives in registers, lane i holding row i, and the serial
    // recurrences read and write single lanes (readlane/writelane) instead of going through LDS.
    if (n <= 64) {
        // the first wavefront alone (tridiag_lowest_wave), then the back-transform
        __shared__ __align__(16) double t_scratch[288];
        if (tid < 64) {
            const int ln = tid;
            const double d_r = ln < n ? dd[ln] : 0.0;
            const double e_r = ln + 1 < n ? ee[ln] : 0.0;  // T[ln+1, ln]
            double yv;
            const double lambda = tridiag_lowest_wave(ln, n, d_r, e_r, t_scratch, yv);
            for (int k = n - 3; k >= 0; --k) {  // back-transform z = H_0 H_1 ... H_{n-3} y
                const double vk = ln > k && ln < n ? V[ln + n * k] : 0.0;
                yv -= beta[k] * wave_sum(vk * yv) * vk;
            }
            if (ln < n) y[ln] = yv;
            if (tid == 0) sh_scalar[1] = lambda;
        }
        __syncthreads();
    } else {
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
    }
    // c = Linv^T z (Lm holds Linv, lower triangle)
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        double c0 = 0.0, c1 = 0.0;
        int k = i;
        for (; k + 1 < n; k += 2) {
            c0 += Lm[k + n * i] * y[k];
            c1 += Lm[k + 1 + n * i] * y[k + 1];
        }
        for (; k < n; ++k) c0 += Lm[k + n * i] * y[k];
        tdg[i] = c0 + c1;  // tdg is free after the inverse iteration
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        y[i] = tdg[i];
        coef[i] = tdg[i];
    }
    __syncthreads();
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

// Selected states (state-averaged evaluations, workflow/evaluator.cpp:22-32,65-111): after the
// global-memory dense solve (k_gep_lowest<false, 2 or 3>), whose workspace still holds L^-1 (Lm),
// the Householder vectors and scalars (V, beta) and the tridiagonal (dd, ee), one workgroup per
// problem finds eigenvalue state_idx[k] of the tridiagonal for each selected state by multisection
// on Sturm counts (the smallest x with at least state_idx[k] + 1 eigenvalues below it), its vector
// by inverse iteration (LAPACK dgttrf/dgttrs scheme; three iterations from a fixed +-1 start, each
// orthogonalised against the vectors of the states already found, as dstein does within a
// cluster), back-transforms it, maps it to c = L^-T z (c^T S c = 1), expands determinant
// coefficients and forms E = sum_k w_k lambda_k. The tridiagonal vectors live in the workspace
// past the core arrays (free after the front half). State indices are distinct, so
// n_states <= n_str.
__global__ void __launch_bounds__(256) k_gep_states(const SlotDesc* __restrict__ slots, TopoDesc topo) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n_str, S = topo.n_states, n_det = topo.n_det;
    if (sd.flags[FLAG_METRIC]) return;
    const auto work = to_global(sd.gep);
    const auto Lm = work;            // L^-1, lower triangle
    const auto V = work + 2 * (size_t)n * n;
    const auto dd = work + 3 * (size_t)n * n;
    const auto ee = dd + n;
    const auto beta = ee + 2 * (size_t)n;
    const auto tdl = beta + n;
    const auto tdu = tdl + n;
    const auto tdu2 = tdu + n;
    const auto tdg = tdu2 + n;
    const auto ipiv = reinterpret_cast<gptr<int>>(tdg + n);
    const auto Z = work + gep_core_dev(n);  // n_states tridiagonal vectors of n
    __shared__ double red[16];
    __shared__ double lo_hi[2];
    __shared__ double lam_sh[kMaxStates];
    __shared__ int first_true;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 0; k < S; ++k) {
        const int target = topo.state_idx[k] + 1;
        if (tid == 0) {  // Gershgorin interval
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
            for (int i = 0; i < n; ++i) {  // eigenvalues below x (LDL^T pivots)
                const double e2 = i > 0 ? ee[i - 1] * ee[i - 1] : 0.0;
                q = dd[i] - x - (i > 0 ? e2 / q : 0.0);
                if (q == 0.0) q = -1e-300;
                if (q < 0.0) ++count;
            }
            if (count >= target) atomicMin(&first_true, tid);
            __syncthreads();
            if (tid == 0) {
                const int f = first_true;
                lo_hi[0] = f == 0 ? lo : lo + (hi - lo) * (double)f / (double)nt;
                lo_hi[1] = f >= nt ? hi : lo + (hi - lo) * (double)(f + 1) / (double)nt;
            }
            __syncthreads();
        }
        if (tid == 0) {
            const double lambda = 0.5 * (lo_hi[0] + lo_hi[1]);
            lam_sh[k] = lambda;
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
            ipiv[n - 1] = n - 1;
            if (tdg[n - 1] == 0.0) tdg[n - 1] = 1e-300;
            const auto y = Z + (size_t)k * n;
            unsigned long long st = 0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL * (unsigned long long)k;
            for (int i = 0; i < n; ++i) {
                st ^= st << 13;
                st ^= st >> 7;
                st ^= st << 17;
                y[i] = (st & 1ULL) ? 1.0 : -1.0;
            }
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
                for (int j = 0; j < k; ++j) {  // against the states already found
                    const auto zj = Z + (size_t)j * n;
                    double dot = 0.0;
                    for (int i = 0; i < n; ++i) dot += zj[i] * y[i];
                    for (int i = 0; i < n; ++i) y[i] -= dot * zj[i];
                }
                double nrm = 0.0;
                for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
                nrm = 1.0 / sqrt(nrm);
                for (int i = 0; i < n; ++i) y[i] *= nrm;
            }
        }
        __syncthreads();
    }
    double eavg = 0.0;
    for (int k = 0; k < S; ++k) {
        const auto y = Z + (size_t)k * n;
        for (int j = n - 3; j >= 0; --j) {  // z = H_0 ... H_{n-3} y
            const int m0 = j + 1;
            double part = 0.0;
            for (int i = m0 + tid; i < n; i += nt) part += V[i + (size_t)n * j] * y[i];
            const double sv = beta[j] * block_sum(part, red);
            for (int i = m0 + tid; i < n; i += nt) y[i] -= sv * V[i + (size_t)n * j];
            __syncthreads();
        }
        double* c = sd.coef + (size_t)k * n;
        for (int i = tid; i < n; i += nt) {  // c = L^-T z
            double c0 = 0.0, c1 = 0.0;
            int m = i;
            for (; m + 1 < n; m += 2) {
                c0 += Lm[m + (size_t)n * i] * y[m];
                c1 += Lm[m + 1 + (size_t)n * i] * y[m + 1];
            }
            for (; m < n; ++m) c0 += Lm[m + (size_t)n * i] * y[m];
            c[i] = c0 + c1;
        }
        __syncthreads();
        double* a = sd.dcoef + (size_t)k * n_det;
        for (int d = tid; d < n_det; d += nt) {
            double acc = 0.0;
            for (long long t = topo.off[d]; t < topo.off[d + 1]; ++t) acc += topo.tc[t] * c[topo.ts[t]];
            a[d] = acc;
        }
        eavg += topo.state_w[k] * lam_sh[k];
    }
    if (tid < S) to_global(sd.state_energy)[tid] = lam_sh[tid];
    if (tid == 0) {
        to_global(sd.energy)[0] = eavg;
        to_global(sd.energy)[1] = eavg + sd.e_nuc;
    }
}

// ---------------------------------------------------------------------------------------------
// gradient tail
// ---------------------------------------------------------------------------------------------

// dL[A,P] = sum_Q M[P,Q] L[A,Q], M[P,Q] = dg(P,Q) off the diagonal, 2 dg(P,P) on it.
// grid (ceil(n_aux/256), ceil(np/8), slots): a thread owns one aux row and eight P, so each L
// element is read np/8 times instead of np times; the block's eight rows of M are staged in LDS
// once and read as broadcasts.
__global__ void __launch_bounds__(256) k_ri_pair_adjoint(const SlotDesc* __restrict__ slots, int np) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_aux = sd.n_aux;
    if ((int)blockIdx.x * 256 >= n_aux) return;  // block-uniform
    const auto L = to_global(sd.L);
    const auto dg = to_global(sd.dsum);
    __shared__ double m[8][kGramLd];
    const int P0 = blockIdx.y * 8;
    for (int idx = threadIdx.x; idx < 8 * np; idx += blockDim.x) {
        const int j = idx / np, Q = idx - j * np, P = P0 + j;
        m[j][Q] = P < np ? (P == Q ? 2.0 : 1.0) * dg[pair_index(P, Q)] : 0.0;
    }
    __syncthreads();
    const int A = blockIdx.x * 256 + threadIdx.x;
    const int Ac = A < n_aux ? A : n_aux - 1;
    double acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = 0.0;
#pragma unroll 8
    for (int Q = 0; Q < np; ++Q) {
        const double l = L[(size_t)Ac + (size_t)n_aux * Q];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += m[j][Q] * l;
    }
    if (A >= n_aux) return;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (P0 + j < np) to_global(sd.dL)[(size_t)A + (size_t)n_aux * (P0 + j)] = acc[j];
}

// gcn[mu, p] = sum_i SC[mu,i] (ds+ds^T)[i,p] + HC[mu,i] (dh+dh^T)[i,p]
//            + sum_q f_pq sum_A dL[A,(p,q)] T[A,q,mu],  f_pq = 2 if p == q else 1.
// grid (ceil(n_bf / (4 R)), 1, slots), R = grad_wave_rows<NQ>() (batched launches): wavefront w of
// a block owns the R AO indices mu0 + R w .. and its lanes take 64 consecutive aux indices A, so every T element
// is read once, with unit stride across the lanes. The weights f_pq dL[A,(p,q)] of a 64-aux chunk
// and QC values of q are staged in LDS as [q][p][a] (a = lane: no bank conflicts; at most 64 KB)
// and serve the block's 4 R rows. NQ = n rounded up to even.
template <int NQ>
constexpr int grad_wave_rows() { return NQ <= 10 ? 8 : 4; }
template <int NQ>
constexpr int grad_q_chunk() { return NQ <= 10 ? NQ : NQ / 2; }
constexpr unsigned kGradFewProblems = 16;  // up to this many problems per launch: k_orbital_gradient_few

template <int NQ>
__global__ void __launch_bounds__(256) k_orbital_gradient(const SlotDesc* __restrict__ slots, int n, int ng) {
    constexpr int R = grad_wave_rows<NQ>(), QC = grad_q_chunk<NQ>();
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_bf = sd.n_bf, n_aux = sd.n_aux;
    const int mu_blk = blockIdx.x * 4 * R;
    if (mu_blk >= n_bf) return;  // block-uniform
    const auto T = to_global(sd.T);
    const auto dL = to_global(sd.dL);
    __shared__ double wq[QC * NQ * 64];  // [q - q0][p][a]
    __shared__ double part[4 * R][NQ];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mu0 = mu_blk + wave * R;
    gptr<const double> tmu[R];
#pragma unroll
    for (int u = 0; u < R; ++u) tmu[u] = T + (size_t)n_aux * n * min(mu0 + u, n_bf - 1);
    double acc[R][NQ];
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
        for (int p = 0; p < NQ; ++p) acc[u][p] = 0.0;
    for (int A0 = 0; A0 < n_aux; A0 += 64)
    for (int q0 = 0; q0 < n; q0 += QC) {
        const int nq = min(QC, n - q0);
        __syncthreads();
        for (int idx = threadIdx.x; idx < nq * n * 64; idx += blockDim.x) {
            const int a = idx & 63, qp = idx >> 6, qq = qp / n, p = qp - qq * n, q = q0 + qq;
            const int A = A0 + a;
            wq[(qq * NQ + p) * 64 + a] =
                A < n_aux ? (p == q ? 2.0 : 1.0) * dL[(size_t)A + (size_t)n_aux * pair_index(p, q)] : 0.0;
        }
        __syncthreads();
        const int A = min(A0 + lane, n_aux - 1);  // past n_aux the weights are zero
        for (int qq = 0; qq < nq; ++qq) {
            const int q = q0 + qq;
            double t[R];
#pragma unroll
            for (int u = 0; u < R; ++u) t[u] = tmu[u][(size_t)n_aux * q + A];
            const double* w = wq + qq * NQ * 64 + lane;
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                const double wp = p < n ? w[p * 64] : 0.0;  // select, not branch: the LDS reads stay batched
#pragma unroll
                for (int u = 0; u < R; ++u) acc[u][p] += wp * t[u];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            const double s = wave_sum(acc[u][p]);
            if (lane == 0) part[wave * R + u][p] = s;
        }
    __syncthreads();
    const double* ds = sd.dsum + ng;
    const double* dh = ds + n * n;
    for (int k = threadIdx.x; k < 4 * R * n; k += blockDim.x) {
        const int u = k / n, p = k - u * n, mu = mu_blk + u;
        if (mu >= n_bf) continue;
        double g = part[u][p];
        for (int i = 0; i < n; ++i) {
            g += to_global(sd.SC)[mu + (size_t)n_bf * i] * (ds[i + n * p] + ds[p + n * i]);
            g += to_global(sd.HC)[mu + (size_t)n_bf * i] * (dh[i + n * p] + dh[p + n * i]);
        }
        to_global(sd.gcn)[mu + (size_t)n_bf * p] = g;
    }
}

// The same for launches of few problems, where blocks are scarce: grid (ceil(n_bf/4), 1, slots), a
// block owns four AO indices mu and its threads run over A with no staging (each dL element is
// read once per four mu).
constexpr int kGradRows = 4;

template <int NQ>
__global__ void __launch_bounds__(256) k_orbital_gradient_few(const SlotDesc* __restrict__ slots, int n, int ng) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_bf = sd.n_bf, n_aux = sd.n_aux;
    const int mu0 = blockIdx.x * kGradRows;
    if (mu0 >= n_bf) return;  // block-uniform
    const auto T = to_global(sd.T);
    const auto dL = to_global(sd.dL);
    __shared__ double part[4][kGradRows][NQ];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    gptr<const double> tmu[kGradRows];
#pragma unroll
    for (int u = 0; u < kGradRows; ++u) tmu[u] = T + (size_t)n_aux * n * min(mu0 + u, n_bf - 1);
    double acc[kGradRows][NQ];
#pragma unroll
    for (int u = 0; u < kGradRows; ++u)
#pragma unroll
        for (int p = 0; p < NQ; ++p) acc[u][p] = 0.0;
    for (int q = 0; q < n; ++q)
        for (int A = threadIdx.x; A < n_aux; A += blockDim.x) {
            double t[kGradRows];
#pragma unroll
            for (int u = 0; u < kGradRows; ++u) t[u] = tmu[u][(size_t)n_aux * q + A];
            // every weight is loaded unconditionally (p past n clamped, its weight zeroed by a select):
            // loads under a branch would each wait for all loads before them
            double wv[NQ];
#pragma unroll
            for (int p = 0; p < NQ; ++p) wv[p] = dL[(size_t)A + (size_t)n_aux * pair_index(p < n ? p : n - 1, q)];
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                const double w = p < n ? (p == q ? 2.0 : 1.0) * wv[p] : 0.0;
#pragma unroll
                for (int u = 0; u < kGradRows; ++u) acc[u][p] += w * t[u];
            }
        }
#pragma unroll
    for (int u = 0; u < kGradRows; ++u)
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            const double s = wave_sum(acc[u][p]);
            if (lane == 0) part[wave][u][p] = s;
        }
    __syncthreads();
    const double* ds = sd.dsum + ng;
    const double* dh = ds + n * n;
    for (int k = threadIdx.x; k < kGradRows * n; k += blockDim.x) {
        const int u = k / n, p = k - u * n, mu = mu0 + u;
        if (mu >= n_bf) continue;
        double g = (part[0][u][p] + part[1][u][p]) + (part[2][u][p] + part[3][u][p]);
        for (int i = 0; i < n; ++i) {
            g += to_global(sd.SC)[mu + (size_t)n_bf * i] * (ds[i + n * p] + ds[p + n * i]);
            g += to_global(sd.HC)[mu + (size_t)n_bf * i] * (dh[i + n * p] + dh[p + n * i]);
        }
        to_global(sd.gcn)[mu + (size_t)n_bf * p] = g;
    }
}

// raw-slot gradient: grad[k, mu] = (g - (g.cn_k) S cn_k)[mu] / sqrt(nu2_k) on present slots.
__global__ void k_normalization_pullback(const SlotDesc* __restrict__ slots) {
    const SlotDesc& sd = slots[blockIdx.z];
    const auto gcn = to_global(sd.gcn);
    const auto Cn = to_global(sd.Cn);
    const auto SC = to_global(sd.SC);
    const auto nu2 = to_global(sd.nu2);
    const auto support = to_global(sd.support);
    const int n_bf = sd.n_bf;
    const auto grad = to_global(sd.grad);
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
// LDS of the pair kernels: full np x np kernel, S_act, h_act, per-wave group slabs and, for the
// backward, the adjoint accumulators (full np x np when that still fits, else packed).
constexpr size_t kPairLdsLimit = 160 * 1024;
constexpr size_t kGepLdsLimit = 144 * 1024;
constexpr unsigned kGepPerStepMax = 8;  // problems per launch up to which the large-n solve runs per step
constexpr int kCholPerStepMinN = 256;   // below it one workgroup factors S faster than n launch pairs

// lanes per determinant pair for blocks of at most `mmax` electrons of one spin
int pair_group(int mmax) { return mmax <= 4 ? 16 : (mmax == 5 ? 32 : 64); }

size_t slab_bytes(int group) {
    return group == 16 ? slab_bytes_per_wave<16>() : group == 32 ? slab_bytes_per_wave<32>() : slab_bytes_per_wave<64>();
}

bool pairs_full_dg(int n, int ng) {
    const size_t np = (size_t)n * (n + 1) / 2;
    const size_t bytes = (2 * np * np + 4 * (size_t)n * n) * sizeof(double) + 4 * slab_bytes(64);
    (void)ng;
    return bytes <= 150 * 1024;
}

size_t pairs_lds_bytes(int n, int ng, bool backward, int waves, int group) {
    const size_t np = (size_t)n * (n + 1) / 2;
    size_t doubles = np * np + 2 * (size_t)n * n;
    if (backward) doubles += (pairs_full_dg(n, ng) ? np * np : (size_t)ng) + 2 * (size_t)n * n;
    return doubles * sizeof(double) + (size_t)waves * slab_bytes(group);
}

// eight wavefronts per block when the LDS allows it, else four
int pairs_waves(int n, int ng, bool backward, int group) {
    return pairs_lds_bytes(n, ng, backward, 8, group) <= kPairLdsLimit ? 8 : 4;
}

template <bool B, bool F, bool L>
hipError_t set_pairs_lds(int group, size_t bytes) {
    const void* fn = group == 16   ? reinterpret_cast<const void*>(&k_pairs<B, F, 16, L>)
                     : group == 32 ? reinterpret_cast<const void*>(&k_pairs<B, F, 32, L>)
                                   : reinterpret_cast<const void*>(&k_pairs<B, F, 64, L>);
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <bool B, bool F, bool L>
void launch_pairs_variant(hipStream_t st, int group, dim3 grid, int threads, size_t lds, const SlotDesc* slots,
                          const TopoDesc& topo, int part_offset) {
    if (group == 16)
        hipLaunchKernelGGL((k_pairs<B, F, 16, L>), grid, dim3(threads), lds, st, slots, topo, part_offset);
    else if (group == 32)
        hipLaunchKernelGGL((k_pairs<B, F, 32, L>), grid, dim3(threads), lds, st, slots, topo, part_offset);
    else
        hipLaunchKernelGGL((k_pairs<B, F, 64, L>), grid, dim3(threads), lds, st, slots, topo, part_offset);
}

// group kernel over all pairs, or over the deferred list after the lane kernel
int launch_pairs_group(hipStream_t st, bool backward, bool listed, int blocks, int part_offset, const SlotDesc* slots,
                       int count, const TopoDesc& topo, int mmax) {
    const int n = topo.n, ng = topo.ng;
    const int group = pair_group(mmax);
    // the backward accumulates into the block's LDS with atomics: one wavefront per block keeps their
    // order that of the wavefront's own instructions, hence the same bits on every run
    const int waves = backward ? 1 : pairs_waves(n, ng, backward, group);
    const size_t lds = pairs_lds_bytes(n, ng, backward, waves, group);
    const dim3 grid(blocks, 1, count);
    const int thr = 64 * waves;
    if (!backward) {
        if (listed) launch_pairs_variant<false, false, true>(st, group, grid, thr, lds, slots, topo, part_offset);
        else launch_pairs_variant<false, false, false>(st, group, grid, thr, lds, slots, topo, part_offset);
    } else if (pairs_full_dg(n, ng)) {
        if (listed) launch_pairs_variant<true, true, true>(st, group, grid, thr, lds, slots, topo, part_offset);
        else launch_pairs_variant<true, true, false>(st, group, grid, thr, lds, slots, topo, part_offset);
    } else {
        if (listed) launch_pairs_variant<true, false, true>(st, group, grid, thr, lds, slots, topo, part_offset);
        else launch_pairs_variant<true, false, false>(st, group, grid, thr, lds, slots, topo, part_offset);
    }
    return thr / group;  // pairs per block sweep
}

// Raise the dynamic-LDS ceiling of every LDS-heavy kernel variant to its limit. The attribute is
// per function and process-wide, so pools of different active spaces never lower it for each other.
hipError_t configure_kernels(int n, int ng, int n_str, int mmax) {
    (void)n;
    (void)ng;
    (void)n_str;
    (void)mmax;
    hipError_t e = hipSuccess;
    for (int group : {16, 32, 64}) {
        if ((e = set_pairs_lds<false, false, false>(group, kPairLdsLimit)) != hipSuccess) return e;
        if ((e = set_pairs_lds<true, true, false>(group, kPairLdsLimit)) != hipSuccess) return e;
        if ((e = set_pairs_lds<true, false, false>(group, kPairLdsLimit)) != hipSuccess) return e;
        if ((e = set_pairs_lds<false, false, true>(group, kPairLdsLimit)) != hipSuccess) return e;
        if ((e = set_pairs_lds<true, true, true>(group, kPairLdsLimit)) != hipSuccess) return e;
        if ((e = set_pairs_lds<true, false, true>(group, kPairLdsLimit)) != hipSuccess) return e;
    }
    if ((e = configure_pairs_lane()) != hipSuccess) return e;
    if ((e = configure_ri_tiled()) != hipSuccess) return e;
    if ((e = configure_pairs_singular()) != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gep_lowest<true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGepLdsLimit);
}


size_t gep_lds_bytes(int n) { return gep_core_doubles(n) * sizeof(double); }

bool gep_in_lds(int n) { return gep_lds_bytes(n) <= kGepLdsLimit; }

// Front half of the eigensolve in global memory: Cholesky S = L L^T (blocked over many workgroups
// when per_step and the space is large, else one workgroup per problem), L^-1 and the standard-form
// matrix A = L^-1 H L^-T. Workspace (SlotDesc::gep): L at 0, A at n^2, L^-1 column-major at 2 n^2.
void launch_gep_front(hipStream_t st, const SlotDesc* slots, unsigned z, const TopoDesc& topo, bool per_step) {
    const int ns = topo.n_str;
    if (per_step && ns >= kCholPerStepMinN) {
        hipLaunchKernelGGL(k_gep_load, dim3((unsigned)(((long long)ns * ns + 255) / 256), 1, z), dim3(256), 0, st,
                           slots, ns);
        for (int j0 = 0; j0 < ns; j0 += kCholPanel) {
            const int jb = std::min(kCholPanel, ns - j0), rest = ns - j0 - jb;
            hipLaunchKernelGGL(k_chol_diag, dim3(1, 1, z), dim3(256), 0, st, slots, ns, j0, jb);
            if (rest > 0) {
                hipLaunchKernelGGL(k_chol_trsm, dim3((unsigned)((rest + 255) / 256), 1, z), dim3(256), 0, st, slots, ns,
                                   j0, jb);
                const unsigned tiles = (unsigned)((rest + 15) / 16);
                hipLaunchKernelGGL(k_chol_syrk, dim3(tiles, tiles, z), dim3(256), 0, st, slots, ns, j0, jb);
            }
        }
    } else {
        hipLaunchKernelGGL(k_gep_chol, dim3(1, 1, z), dim3(1024), 0, st, slots, topo);
    }
    if (ns <= kLinvWaveMaxN)
        hipLaunchKernelGGL(k_gep_linv, dim3((unsigned)((ns + 3) / 4), 1, z), dim3(256), 4 * (size_t)ns * sizeof(double),
                           st, slots, ns);
    else
        hipLaunchKernelGGL(k_gep_linv_lane, dim3((unsigned)((ns + 63) / 64), 1, z), dim3(64), 0, st, slots, ns);
    hipLaunchKernelGGL(k_gep_hlt, dim3((unsigned)((ns + 255) / 256), (unsigned)ns, z), dim3(256), 0, st, slots, ns);
    hipLaunchKernelGGL(k_gep_lul, dim3((unsigned)((ns + 255) / 256), (unsigned)ns, z), dim3(256), 0, st, slots, ns);
}

void launch_forward(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape,
                    hipEvent_t* ev, const SideStream* side) {
    const int n = topo.n, np = topo.np, ng = topo.ng;
    const unsigned z = (unsigned)count;
    auto mark = [&](int k) {
        if (ev) (void)hipEventRecord(ev[k], st);
    };
    // few problems: the active-space matrices (read first by the pair kernels) beside the RI transform
    const bool fork = side && count <= kForkMaxProblems;
    mark(EV_START);
    hipLaunchKernelGGL(k_orbital_prep, dim3(n, 1, z), dim3(256), ((size_t)shape.max_n_bf + 16) * sizeof(double), st,
                       slots, n);
    if (fork) {
        (void)hipEventRecord(side->fork, st);
        (void)hipStreamWaitEvent(side->st, side->fork, 0);
        hipLaunchKernelGGL(k_active_matrices, dim3(1, 1, z), dim3(256), 0, side->st, slots, n);
        (void)hipEventRecord(side->join, side->st);
    } else {
        hipLaunchKernelGGL(k_active_matrices, dim3(1, 1, z), dim3(256), 0, st, slots, n);
    }
    mark(EV_PREP);
    if (shape.any_tiled)
        launch_ri_one_index_tiled(st, slots, count, n, shape.max_n_bf_tiled, shape.max_n_aux, shape.b32c);
    if (shape.any_flat) {
        const dim3 grid((shape.max_n_aux + 63) / 64, (shape.max_n_bf + 4 * kOneIndexRows - 1) / (4 * kOneIndexRows), z);
        auto go = [&](auto nq) {
            constexpr int NQ = decltype(nq)::value;
            if (shape.b32) hipLaunchKernelGGL((k_ri_one_index<NQ, float>), grid, dim3(256), 0, st, slots, n);
            else if (shape.b32c) hipLaunchKernelGGL((k_ri_one_index<NQ, RiFp32c>), grid, dim3(256), 0, st, slots, n);
            else hipLaunchKernelGGL((k_ri_one_index<NQ, double>), grid, dim3(256), 0, st, slots, n);
        };
        switch ((n + 1) / 2) {
            case 1: go(std::integral_constant<int, 2>{}); break;
            case 2: go(std::integral_constant<int, 4>{}); break;
            case 3: go(std::integral_constant<int, 6>{}); break;
            case 4: go(std::integral_constant<int, 8>{}); break;
            case 5: go(std::integral_constant<int, 10>{}); break;
            case 6: go(std::integral_constant<int, 12>{}); break;
            default: go(std::integral_constant<int, 14>{}); break;
        }
    }
    mark(EV_ONE_INDEX);
    {
        const dim3 grid((shape.max_n_aux + 63) / 64, (n + 3) / 4, z);
        switch ((n + 1) / 2) {
            case 1: hipLaunchKernelGGL(k_ri_pair_factors<2>, grid, dim3(256), 0, st, slots, n); break;
            case 2: hipLaunchKernelGGL(k_ri_pair_factors<4>, grid, dim3(256), 0, st, slots, n); break;
            case 3: hipLaunchKernelGGL(k_ri_pair_factors<6>, grid, dim3(256), 0, st, slots, n); break;
            case 4: hipLaunchKernelGGL(k_ri_pair_factors<8>, grid, dim3(256), 0, st, slots, n); break;
            case 5: hipLaunchKernelGGL(k_ri_pair_factors<10>, grid, dim3(256), 0, st, slots, n); break;
            case 6: hipLaunchKernelGGL(k_ri_pair_factors<12>, grid, dim3(256), 0, st, slots, n); break;
            default: hipLaunchKernelGGL(k_ri_pair_factors<14>, grid, dim3(256), 0, st, slots, n); break;
        }
    }
    if (shape.pinned || z <= kGramFewProblems)
        hipLaunchKernelGGL(k_ri_pair_gram, dim3((ng + 3) / 4, 1, z), dim3(256), 0, st, slots, np);
    else
        hipLaunchKernelGGL(k_ri_pair_gram_tiled, dim3(1, 1, z), dim3(256), 0, st, slots, np);
    if (fork) (void)hipStreamWaitEvent(st, side->join, 0);
    mark(EV_TRANSFORM);
    if (shape.lane_fwd) {
        launch_pairs_lane(st, 0, shape.lane_blocks_fwd, slots, count, topo, shape.n_alpha, shape.n_beta,
                          shape.lane_cond_limit);
        mark(EV_PAIR_FWD_MAIN);
        const long long all_pairs = (long long)topo.n_det * (topo.n_det + 1) / 2;
        compact_pairs<true>(st, slots, z, topo, 1, topo.n_lane_pairs);
        launch_pairs_group(st, false, true, shape.list_blocks, 0, slots, count, topo, shape.mmax);
        compact_pairs<false>(st, slots, z, topo, 1, all_pairs);
    } else {
        launch_pairs_group(st, false, false, shape.pair_blocks_fwd, 0, slots, count, topo, shape.mmax);
        mark(EV_PAIR_FWD_MAIN);
        compact_pairs<false>(st, slots, z, topo, 0, (long long)topo.n_det * (topo.n_det + 1) / 2);
    }
    launch_pairs_singular(st, false, 0, slots, count, topo);
    mark(EV_PAIR_FWD);
    const long long a = (long long)topo.n_det * topo.n_str, b = (long long)topo.n_str * topo.n_str;
    hipLaunchKernelGGL(k_structure_right, dim3((unsigned)((a + 255) / 256), 1, z), dim3(256), 0, st, slots, topo);
    hipLaunchKernelGGL(k_structure_left, dim3((unsigned)((b + 255) / 256), 1, z), dim3(256), 0, st, slots, topo);
    mark(EV_STRUCT);
    const int threads = 256;
    // one selected state that is the lowest root: the lowest-root solvers; otherwise the
    // global-memory dense solve followed by the selected-states kernel
    const bool states = topo.n_states > 1 || topo.state_idx[0] != 0;
    if (shape.davidson)
        launch_davidson(st, slots, z, topo, shape);
    else if (!states && launch_gep_wave(st, slots, z, topo)) {
    } else if (gep_in_lds(topo.n_str) && !states)
        hipLaunchKernelGGL(k_gep_lowest<true>, dim3(1, 1, z), dim3(threads), gep_lds_bytes(topo.n_str), st, slots,
                           topo);
    else {
        // matrices in global memory: few problems per launch take the per-step Cholesky and
        // Householder kernels (many blocks each); batched launches keep one workgroup per problem
        // for those two stages, which the number of problems already parallelises
        const int ns = topo.n_str;
        const bool per_step = shape.pinned || z <= kGepPerStepMax;
        launch_gep_front(st, slots, z, topo, per_step);
        if (per_step) {
            hipLaunchKernelGGL((k_gep_lowest<false, 1>), dim3(1, 1, z), dim3(1024), 0, st, slots, topo);
            for (int k = 0; k + 2 < ns; ++k) {
                const unsigned rows = (unsigned)((ns - (k + 1) + 3) / 4);
                hipLaunchKernelGGL(k_hh_step, dim3(rows, 1, z), dim3(256), 3 * (size_t)ns * sizeof(double), st, slots,
                                   ns, k);
            }
            hipLaunchKernelGGL(k_hh_finish, dim3(1, 1, z), dim3(256), 2 * (size_t)ns * sizeof(double), st, slots, ns);
            hipLaunchKernelGGL((k_gep_lowest<false, 2>), dim3(1, 1, z), dim3(1024), 0, st, slots, topo);
        } else {
            hipLaunchKernelGGL((k_gep_lowest<false, 3>), dim3(1, 1, z), dim3(1024), 0, st, slots, topo);
        }
        if (states) hipLaunchKernelGGL(k_gep_states, dim3(1, 1, z), dim3(256), 0, st, slots, topo);
    }
    mark(EV_EIGEN);
}

void launch_backward(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo, const LaunchShape& shape,
                     hipEvent_t* ev, const SideStream* side) {
    const int n = topo.n, np = topo.np, ng = topo.ng;
    const unsigned z = (unsigned)count;
    auto mark = [&](int k) {
        if (ev) (void)hipEventRecord(ev[k], st);
    };
    mark(EV_BWD_BEGIN);
    // The pair-adjoint kernels write disjoint columns (field: ds/dh; adjoint: dg) or disjoint rows
    // (deferred list, singular pairs) of the partial rows, so for few problems the adjoint kernel and
    // the singular-pair kernel run on the side stream beside the field and deferred-list kernels.
    const bool fork = side && count <= kForkMaxProblems && shape.lane_bwd;
    hipStream_t st2 = st;
    if (fork) {
        (void)hipEventRecord(side->fork, st);
        (void)hipStreamWaitEvent(side->st, side->fork, 0);
        st2 = side->st;
    }
    int part_rows = shape.pair_blocks_bwd, list_begin = 0, list_rows = 0, list_per_row = 1;
    if (shape.lane_bwd) {
        launch_pairs_lane(st, 1, shape.lane_blocks_bwd, slots, count, topo, shape.n_alpha, shape.n_beta,
                          shape.lane_cond_limit);
        mark(EV_BWD_FIELD);
        launch_pairs_lane(st2, 2, shape.lane_blocks_bwd, slots, count, topo, shape.n_alpha, shape.n_beta,
                          shape.lane_cond_limit);
        if (!fork) mark(EV_BWD_ADJOINT);
        list_per_row = launch_pairs_group(st, true, true, shape.list_blocks, shape.lane_blocks_bwd, slots, count, topo,
                                          shape.mmax);
        list_begin = shape.lane_blocks_bwd;
        list_rows = shape.list_blocks;
        part_rows = shape.lane_blocks_bwd + shape.list_blocks;
    } else {
        launch_pairs_group(st, true, false, shape.pair_blocks_bwd, 0, slots, count, topo, shape.mmax);
        mark(EV_BWD_FIELD);
        mark(EV_BWD_ADJOINT);
    }
    launch_pairs_singular(st2, true, part_rows, slots, count, topo);
    part_rows += kSingularBlocks;
    if (fork) {
        (void)hipEventRecord(side->join, side->st);
        (void)hipStreamWaitEvent(st, side->join, 0);
        mark(EV_BWD_ADJOINT);  // field and deferred list done on the main stream, adjoint joined
    }
    mark(EV_BWD_KERNEL);
    const int len = ng + 2 * n * n;
    if (part_rows <= kReduceRowsFew)
        hipLaunchKernelGGL(k_reduce_partials_few, dim3((len + 255) / 256, 1, z), dim3(256), 0, st, slots, part_rows,
                           len, list_begin, list_rows, list_per_row);
    else
        hipLaunchKernelGGL(k_reduce_partials, dim3((len + 3) / 4, 1, z), dim3(256), 0, st, slots, part_rows, len,
                           list_begin, list_rows, list_per_row);
    mark(EV_PAIR_BWD);
    hipLaunchKernelGGL(k_ri_pair_adjoint, dim3((shape.max_n_aux + 255) / 256, (np + 7) / 8, z), dim3(256), 0, st, slots,
                       np);
    {
        auto grad = [&]<int NQ>() {
            if (shape.pinned || z <= kGradFewProblems) {
                hipLaunchKernelGGL(k_orbital_gradient_few<NQ>,
                                   dim3((unsigned)((shape.max_n_bf + kGradRows - 1) / kGradRows), 1, z), dim3(256), 0,
                                   st, slots, n, ng);
            } else {
                const int rows = 4 * grad_wave_rows<NQ>();
                hipLaunchKernelGGL(k_orbital_gradient<NQ>,
                                   dim3((unsigned)((shape.max_n_bf + rows - 1) / rows), 1, z), dim3(256), 0, st, slots,
                                   n, ng);
            }
        };
        switch ((n + 1) / 2) {
            case 1: grad.template operator()<2>(); break;
            case 2: grad.template operator()<4>(); break;
            case 3: grad.template operator()<6>(); break;
            case 4: grad.template operator()<8>(); break;
            case 5: grad.template operator()<10>(); break;
            case 6: grad.template operator()<12>(); break;
            default: grad.template operator()<14>(); break;
        }
    }
    hipLaunchKernelGGL(k_normalization_pullback, dim3(n, 1, z), dim3(256), 0, st, slots);
    mark(EV_TAIL);
}

}  // namespace xmvb_amd
