// Determinant pairs, one pair per lane.
//
// Each lane evaluates one unordered determinant pair entirely in registers: the two spin blocks
// X_s(r, l) = S_act(R_r, L_l) are gathered from LDS, inverted by Gauss-Jordan elimination with row
// pivoting, and the matrix element is assembled from the LDS-resident active kernel
//   H_de = det X_a det X_b * Phi,  Phi = sum_s D_s.h + sum_s sum_{r<r2,l<l2} A_s minor_s + D_a.J[D_b],
// the regular-inverse form of DeterminantPairEvaluator::combine_spin_pair_evaluations
// (determinants/pairs/evaluator.cpp:354-384) with the same-spin Lowdin terms of
// calc_same_spin_hamiltonian (determinants/algebra/hamiltonian.cpp:35-236) written over 2x2 minors.
//
// The reference decides regularity with a complete-pivoting LU (DeterminantOverlapResolver,
// determinants/algebra/overlap.cpp:50-110: rank counts pivots above 1e-12 max|pivot|). A lane only
// keeps a pair whose blocks satisfy max|X| max|X^-1| < cond_limit (at most 1e9): every
// complete-pivoting pivot is at least 1/(m max|X^-1|) and none exceeds 16 max|X| for m <= 8
// (Wilkinson's growth bound), so such a block has full rank under that rule. Any other pair is
// appended to the slot's deferred list and evaluated by the lane-group kernel, which applies the
// complete-pivoting rule itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device.hpp"
#include "kernels.hpp"

namespace xmvb_amd {

// Wavefronts per SIMD the compiler must fit (register budget 512 / w per lane). Two for the
// forward up to 6-electron blocks and for the field kernel up to 5: measured faster than one
// wavefront with more registers (medium batch: forward 2.66 -> 2.22 ms, field 4.92 -> 3.82 ms);
// the 6-electron field kernel and every adjoint kernel are slower when held to two.
constexpr int lane_fwd_waves(int m0) { return m0 <= 6 ? 2 : 1; }
constexpr int lane_field_waves(int m0) { return m0 <= 5 ? 2 : 1; }

template <int M>
struct LaneBlock {
    static constexpr int K = M > 0 ? M : 1;
    int R[K], L[K];  // occupied orbitals of the right (ket) and left (bra) strings
    double a[K][K];  // X on entry; X^{-1} after inversion, so D[r][l] = a[l][r]
    double det;
};

template <int M>
__device__ __forceinline__ void occupied(uint32_t mask, int* out) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
        out[k] = __builtin_ctz(mask);
        mask &= mask - 1;
    }
}

// In-place Gauss-Jordan inversion with row pivoting. A row interchange of X is an interchange of
// two ket orbitals, so it is applied to R as well and never undone: the result is the inverse of
// the row-permuted block, paired with the permuted R, which every consumer sums over in any order
// (same-spin terms are symmetric under r <-> r2; the compound W index sorts the pair itself).
// det keeps the sign of the unpermuted block. Returns whether the block stays on the lane path.
template <int M>
__device__ __forceinline__ bool lane_invert(const double* __restrict__ sact, int n, double cond_limit, LaneBlock<M>& b) {
    if constexpr (M == 0) {
        b.det = 1.0;
        return true;
    } else {
        auto& a = b.a;
        double amax = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) {
                a[i][j] = sact[b.R[i] + n * b.L[j]];
                amax = fmax(amax, fabs(a[i][j]));
            }
        double det = 1.0;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            int p = k;
            double best = fabs(a[k][k]);
#pragma unroll
            for (int i = k + 1; i < M; ++i) {
                const double v = fabs(a[i][k]);
                const bool g = v > best;
                best = g ? v : best;
                p = g ? i : p;
            }
#pragma unroll
            for (int i = k + 1; i < M; ++i) {
                const bool s = p == i;
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    const double t = a[k][j];
                    a[k][j] = s ? a[i][j] : t;
                    a[i][j] = s ? t : a[i][j];
                }
                const int tr = b.R[k];
                b.R[k] = s ? b.R[i] : tr;
                b.R[i] = s ? tr : b.R[i];
            }
            const double pv = a[k][k];
            ok = ok && best > 0.0;
            det = (p != k) ? -det * pv : det * pv;
            const double inv = best > 0.0 ? 1.0 / pv : 0.0;
            a[k][k] = 1.0;
#pragma unroll
            for (int j = 0; j < M; ++j) a[k][j] *= inv;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                if (i == k) continue;
                const double f = a[i][k];
                a[i][k] = 0.0;
#pragma unroll
                for (int j = 0; j < M; ++j) a[i][j] -= f * a[k][j];
            }
        }
        double ymax = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) ymax = fmax(ymax, fabs(a[i][j]));
        b.det = det;
        return ok && amax * ymax < cond_limit;
    }
}

// Same-spin two-electron energy over 2x2 minors: sum_{r<r2,l<l2} [(rl|r2l2) - (rl2|r2l)] minor.
// pidx(r, l) is the packed pair index of block entry (R_r, L_l).
template <int M, typename Idx>
__device__ __forceinline__ double same_spin_energy(const double* __restrict__ G2, int np, const LaneBlock<M>& b,
                                                   Idx pidx) {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int r2 = r + 1; r2 < M; ++r2)
#pragma unroll
            for (int l = 0; l < M; ++l)
#pragma unroll
                for (int l2 = l + 1; l2 < M; ++l2) {
                    const double J = G2[pidx(r, l) * np + pidx(r2, l2)];
                    const double K = G2[pidx(r, l2) * np + pidx(r2, l)];
                    // D[r][l] = a[l][r]
                    const double minor = b.a[l][r] * b.a[l2][r2] - b.a[l2][r] * b.a[l][r2];
                    acc += (J - K) * minor;
                }
    return acc;
}

// Same-spin part of F = dPhi/dD: F[rl] += A(rl, r2l2) D[r2l2], A the antisymmetrised kernel.
template <int M, typename Idx>
__device__ __forceinline__ void same_spin_field(const double* __restrict__ G2, int np, const LaneBlock<M>& b, Idx pidx,
                                                double (&F)[LaneBlock<M>::K][LaneBlock<M>::K]) {
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int r2 = r + 1; r2 < M; ++r2)
#pragma unroll
            for (int l = 0; l < M; ++l)
#pragma unroll
                for (int l2 = l + 1; l2 < M; ++l2) {
                    const double A = G2[pidx(r, l) * np + pidx(r2, l2)] - G2[pidx(r, l2) * np + pidx(r2, l)];
                    F[r][l] += A * b.a[l2][r2];
                    F[r2][l2] += A * b.a[l][r];
                    F[r][l2] -= A * b.a[l][r2];
                    F[r2][l] -= A * b.a[l2][r];
                }
}

// Same-spin adjoint of the packed kernel: cw * minor at (rl, r2l2), minus it at (rl2, r2l).
template <bool FullDg, int M, typename Idx>
__device__ __forceinline__ void same_spin_dg(double* dg, int np, const LaneBlock<M>& b, Idx pidx, double cw) {
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int r2 = r + 1; r2 < M; ++r2)
#pragma unroll
            for (int l = 0; l < M; ++l)
#pragma unroll
                for (int l2 = l + 1; l2 < M; ++l2) {
                    const double v = cw * (b.a[l][r] * b.a[l2][r2] - b.a[l2][r] * b.a[l][r2]);
                    if (FullDg) {
                        atomicAdd(&dg[pidx(r, l) * np + pidx(r2, l2)], v);
                        atomicAdd(&dg[pidx(r, l2) * np + pidx(r2, l)], -v);
                    } else {
                        atomicAdd(&dg[pair_index(pidx(r, l), pidx(r2, l2))], v);
                        atomicAdd(&dg[pair_index(pidx(r, l2), pidx(r2, l))], -v);
                    }
                }
}

// ds += c [ (wH Phi + wS) D - wH D F^T D ], dh += c wH D over one spin block
template <int M>
__device__ __forceinline__ void overlap_adjoint(double* ds, double* dh, int n, const LaneBlock<M>& b,
                                                const double (&F)[LaneBlock<M>::K][LaneBlock<M>::K], double c, double sc,
                                                double wH) {
#pragma unroll
    for (int l = 0; l < M; ++l) {
        double z[LaneBlock<M>::K];
#pragma unroll
        for (int a2 = 0; a2 < M; ++a2) {
            double s = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < M; ++b2) s += F[b2][a2] * b.a[l][b2];
            z[a2] = s;
        }
#pragma unroll
        for (int r = 0; r < M; ++r) {
            double dfd = 0.0;
#pragma unroll
            for (int a2 = 0; a2 < M; ++a2) dfd += b.a[a2][r] * z[a2];
            const double Drl = b.a[l][r];
            const int at = b.R[r] + n * b.L[l];
            atomicAdd(&ds[at], c * (sc * Drl - wH * dfd));
            atomicAdd(&dh[at], c * wH * Drl);
        }
    }
}

// Launch-uniform view of the two spins: spin 0 holds M0 >= M1 electrons.
struct SpinTables {
    const uint64_t* mask0;
    const uint64_t* mask1;
    const int* det0;
    const int* det1;
};

__device__ __forceinline__ SpinTables spin_tables(const TopoDesc& topo, int swap_spins) {
    SpinTables t;
    t.mask0 = swap_spins ? topo.bmask : topo.amask;
    t.mask1 = swap_spins ? topo.amask : topo.bmask;
    t.det0 = swap_spins ? topo.det_b : topo.det_a;
    t.det1 = swap_spins ? topo.det_a : topo.det_b;
    return t;
}

// Packed index of the spin-flip partner pair of {d, e} (t itself when the topology has no
// flip pairing or the pair is its own partner); fd, fe: the partner determinants.
__device__ __forceinline__ long long flip_partner(const TopoDesc& topo, long long t, int d, int e, int& fd, int& fe) {
    if (!topo.flip) {
        fd = d;
        fe = e;
        return t;
    }
    const int a = topo.flip[d], b = topo.flip[e];
    fd = a < b ? a : b;
    fe = a < b ? b : a;
    return (long long)fe * (fe + 1) / 2 + fd;
}

// Work item i of the lane kernels -> packed pair index.
__device__ __forceinline__ long long lane_pair(const TopoDesc& topo, long long i) {
    return topo.lane_pairs ? (long long)topo.lane_pairs[i] : i;
}

// Same-spin adjoint of the kernel in compound space: every same-spin kernel term of a pair is
// [(R_r L_l|R_r2 L_l2) - (R_r L_l2|R_r2 L_l)] times a 2x2 minor of D with r < r2, l < l2, so the
// adjoint is carried by W[(p,s),(q,t)] (p = R_r < s = R_r2, q = L_l < t = L_l2), one atomic per
// minor, and expanded into dg once per block (see w_to_dg).
__device__ __forceinline__ int compound_index(int p, int s) { return s * (s - 1) / 2 + p; }  // p < s

template <int M>
__device__ __forceinline__ void same_spin_w(double* W, int nc, const LaneBlock<M>& b, double cw) {
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int r2 = r + 1; r2 < M; ++r2) {
            // R is in pivot order: sort the ket pair, the minor changes sign with it
            const bool up = b.R[r] < b.R[r2];
            const int row = compound_index(up ? b.R[r] : b.R[r2], up ? b.R[r2] : b.R[r]) * nc;
            const double cws = up ? cw : -cw;
#pragma unroll
            for (int l = 0; l < M; ++l)
#pragma unroll
                for (int l2 = l + 1; l2 < M; ++l2) {
                    const double v = cws * (b.a[l][r] * b.a[l2][r2] - b.a[l2][r] * b.a[l][r2]);
                    atomicAdd(&W[row + compound_index(b.L[l], b.L[l2])], v);
                }
        }
}

template <bool FullDg>
__device__ __forceinline__ void dg_add(double* dg, int np, int P, int Q, double v) {
    if (FullDg) atomicAdd(&dg[P * np + Q], v);
    else atomicAdd(&dg[pair_index(P, Q)], v);
}

__device__ __forceinline__ void compound_pair(int i, int& p, int& s) {
    s = (int)((1.0 + sqrt(1.0 + 8.0 * i)) * 0.5);
    while (s * (s - 1) / 2 > i) --s;
    while ((s + 1) * s / 2 <= i) ++s;
    p = i - s * (s - 1) / 2;
}

// Adjoint of the ordered kernel entry (X, Y), X and Y packed orbital pairs, from the same-spin
// compound accumulator: W[(p,s),(q,t)] (p < s, q < t, summed over `copies` copies of nc x nc)
// enters with + at (X, Y) = ((p,q),(s,t)) and with - at ((p,t),(s,q)). Terms are taken in a fixed
// order.
__device__ __forceinline__ double w_terms(const double* W, int copies, int nc, int X, int Y) {
    int x0, x1, y0, y1;
    unpack_pair(X, x0, x1);
    unpack_pair(Y, y0, y1);
    const int xa[2] = {x0, x1}, ya[2] = {y0, y1};
    const int nx = x0 == x1 ? 1 : 2, ny = y0 == y1 ? 1 : 2;
    double acc = 0.0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const int u = xa[i], u2 = xa[1 - i], v = ya[j], v2 = ya[1 - j];
            // + term: p = u, q = u2, s = v, t = v2
            if (u < v && u2 < v2) {
                const int at = compound_index(u, v) * nc + compound_index(u2, v2);
                for (int c = 0; c < copies; ++c) acc += W[c * nc * nc + at];
            }
            // - term: p = u, t = u2, s = v, q = v2
            if (u < v && v2 < u2) {
                const int at = compound_index(u, v) * nc + compound_index(v2, u2);
                for (int c = 0; c < copies; ++c) acc -= W[c * nc * nc + at];
            }
        }
    return acc;
}

// dg += the expansion of W, every adjoint element computed by one thread (no atomics);
// block-cooperative. FullDg: dg[X * np + Y] ordered; packed: dg[pair_index(X, Y)] takes both orders.
template <bool FullDg>
__device__ __forceinline__ void w_gather(const double* W, int copies, int nc, double* dg, int np) {
    const int targets = FullDg ? np * np : np * (np + 1) / 2;
    for (int idx = threadIdx.x; idx < targets; idx += blockDim.x) {
        double acc;
        if (FullDg) {
            acc = w_terms(W, copies, nc, idx / np, idx % np);
        } else {
            int Y, X;
            unpack_pair(idx, Y, X);
            acc = w_terms(W, copies, nc, X, Y);
            if (X != Y) acc += w_terms(W, copies, nc, Y, X);
        }
        dg[idx] += acc;
    }
}

// Runs body() for each wavefront of the block in turn, wavefront 0 first, with a block barrier
// between turns, so that accumulations into shared LDS arrive in a fixed wavefront order. Every
// thread of the block must reach the call (block-uniform control flow).
template <typename F>
__device__ __forceinline__ void wave_turns(F&& body) {
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
    for (int t = 0; t < nw; ++t) {
        __syncthreads();
        if (w == t) body();
    }
    __syncthreads();
}

typedef double lane_double4 __attribute__((ext_vector_type(4)));

// Opposite-spin adjoint of the kernel as outer products on the matrix cores:
//   dg[P, Q] += sum_pairs cw * ga[P] * gb[Q],
// ga / gb the alpha / beta transition densities scattered to packed pair indices (np <= 64, so at
// most 4 x 4 tiles of 16). The wavefront's 64 pairs go through in four chunks of 16: the chunk's
// lanes scatter their rows into zeroed LDS rows (accumulating where two block entries share a
// packed index), every 16x16 tile is updated by four v_mfma_f64_16x16x4 steps (A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i%4), element i/4) into accumulators
// that stay in registers for the whole kernel, and the rows are zeroed again. Called by all lanes.
constexpr int kMfmaTiles = 4;  // tiles per side kept for the whole kernel (np16 <= 64)
constexpr int kMfmaTilesWide = 5;  // 6-electron blocks: np16 <= 80, tiles flushed every iteration

// Only tiles (tp, tq) with tp <= tq are used: the packed adjoint needs M + M^T of the outer-product
// sum M = sum_pairs u0 u1^T, so a diagonal tile accumulates A_p B_p^T and an off-diagonal one
// A_p B_q^T + B_p A_q^T (its transpose partner), TN(TN+1)/2 accumulators instead of TN^2.
template <int TN>
using MfmaTilesN = lane_double4[TN][TN];

template <int M0, int M1, int TN>
__device__ __forceinline__ void opposite_dg_mfma(double* bufA, double* bufB, int np16, bool live, const LaneBlock<M0>& b0,
                                 const LaneBlock<M1>& b1, const int (&P1)[LaneBlock<M1>::K][LaneBlock<M1>::K],
                                 double cw, MfmaTilesN<TN>& acc) {
    const int lane = threadIdx.x & 63;
    const int k4 = lane >> 4, i16 = lane & 15;
    const int tiles = np16 >> 4;
#pragma unroll 1
    for (int chunk = 0; chunk < 4; ++chunk) {
        const bool mine = live && k4 == chunk;
        double* rowA = bufA + i16 * np16;
        double* rowB = bufB + i16 * np16;
        if (mine) {
#pragma unroll
            for (int r = 0; r < M0; ++r)
#pragma unroll
                for (int l = 0; l < M0; ++l) rowA[pair_index(b0.R[r], b0.L[l])] += b0.a[l][r];
#pragma unroll
            for (int r = 0; r < M1; ++r)
#pragma unroll
                for (int l = 0; l < M1; ++l) rowB[P1[r][l]] += cw * b1.a[l][r];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const double* ra = bufA + (4 * kk + k4) * np16 + i16;
            const double* rb = bufB + (4 * kk + k4) * np16 + i16;
            double av[TN], bv[TN];
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                av[t] = t < tiles ? ra[16 * t] : 0.0;
                bv[t] = t < tiles ? rb[16 * t] : 0.0;
            }
#pragma unroll
            for (int tp = 0; tp < TN; ++tp)
#pragma unroll
                for (int tq = tp; tq < TN; ++tq)
                    if (tq < tiles) {
                        acc[tp][tq] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[tp], bv[tq], acc[tp][tq], 0, 0, 0);
                        if (tq != tp)
                            acc[tp][tq] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[tp], av[tq], acc[tp][tq], 0, 0, 0);
                    }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (mine) {
#pragma unroll
            for (int r = 0; r < M0; ++r)
#pragma unroll
                for (int l = 0; l < M0; ++l) rowA[pair_index(b0.R[r], b0.L[l])] = 0.0;
#pragma unroll
            for (int r = 0; r < M1; ++r)
#pragma unroll
                for (int l = 0; l < M1; ++l) rowB[P1[r][l]] = 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// adds a wavefront's MFMA tiles (tp <= tq) into the block's dg accumulator and, if asked, clears them
template <bool FullDg, int TN>
__device__ __forceinline__ void flush_mfma_tiles(MfmaTilesN<TN>& acc, int np16, int np, double* dg, bool clear) {
    const int lane = threadIdx.x & 63;
    const int k4 = lane >> 4, i16 = lane & 15;
    const int tiles = np16 >> 4;
#pragma unroll
    for (int tp = 0; tp < TN; ++tp)
#pragma unroll
        for (int tq = tp; tq < TN; ++tq) {
            if (!(tq < tiles)) continue;
            const int Q = 16 * tq + i16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int P = 16 * tp + 4 * v + k4;
                if (P < np && Q < np && acc[tp][tq][v] != 0.0) dg_add<FullDg>(dg, np, P, Q, acc[tp][tq][v]);
            }
            if (clear) acc[tp][tq] = lane_double4{0.0, 0.0, 0.0, 0.0};
        }
}

// Optional LDS regions of the backward lane kernel (0 = not used).
struct LaneBwdLayout {
    int nc;       // n(n-1)/2: compound dimension of the same-spin W accumulator (0: no W, direct atomics)
    int np16;     // np rounded up to 16: row length of the MFMA staging rows (16 rows x 2 per wavefront)
    int full_dg;  // 1: the kernel adjoint is held as the full np x np matrix in LDS, 0: packed
    int wcopies;  // copies of W: one per wavefront (no ordering needed) or 1 (wavefronts take turns)
    int dcopies;  // copies of the kernel adjoint: one per wavefront or 1 (MFMA flushes take turns)
    int waves;    // wavefronts per block
};

__device__ __forceinline__ void load_pair_tables(const SlotDesc& sd, int n, int np, double* G2, double* sact,
                                                 double* hact) {
    const auto G_glob = to_global(static_cast<const double*>(sd.G));
    const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, ln = threadIdx.x & 63;
    for (int Q = wave; Q < np; Q += nwaves)
        for (int P = ln; P < np; P += 64) G2[P + np * Q] = G_glob[pair_index(P, Q)];
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        sact[t] = to_global(sd.sact)[t];
        hact[t] = to_global(sd.hact)[t];
    }
}

// Forward. LDS: full np x np kernel G2, S_act, h_act.
template <int M0, int M1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lane_fwd_waves(M0))))
    k_pairs_lane_fwd(const SlotDesc* __restrict__ slots, TopoDesc topo, int swap_spins, double cond_limit) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n, n_det = topo.n_det;
    const int np = n * (n + 1) / 2;
    extern __shared__ double lds[];
    double* G2 = lds;
    double* sact = G2 + np * np;
    double* hact = sact + n * n;
    load_pair_tables(sd, n, np, G2, sact, hact);
    __syncthreads();
    const SpinTables tab = spin_tables(topo, swap_spins);
    const long long n_work = topo.n_lane_pairs;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_work;
         i += (long long)gridDim.x * blockDim.x) {
        const long long t = lane_pair(topo, i);
        int d, e, fd, fe;
        unpack_pair(t, d, e);  // d: left (bra) determinant, e: right (ket)
        const long long tf = flip_partner(topo, t, d, e, fd, fe);
        LaneBlock<M0> b0;
        LaneBlock<M1> b1;
        occupied<M0>((uint32_t)tab.mask0[tab.det0[e]], b0.R);
        occupied<M0>((uint32_t)tab.mask0[tab.det0[d]], b0.L);
        occupied<M1>((uint32_t)tab.mask1[tab.det1[e]], b1.R);
        occupied<M1>((uint32_t)tab.mask1[tab.det1[d]], b1.L);
        const bool ok0 = lane_invert<M0>(sact, n, cond_limit, b0);
        const bool ok1 = lane_invert<M1>(sact, n, cond_limit, b1);
        // deferred work items (the partner pair has the same blocks and goes with it) are flagged
        // and listed in work-item order by k_compact_pairs, so the list does not depend on timing
        to_global(sd.defer_flag)[i] = ok0 && ok1 ? 0 : 1;
        if (!(ok0 && ok1)) continue;
        // spin-1 packed indices are kept (the opposite-spin loop reuses each one M0^2 times);
        // spin-0 indices are recomputed where needed, which costs VALU only
        int P1[LaneBlock<M1>::K][LaneBlock<M1>::K];
#pragma unroll
        for (int r = 0; r < M1; ++r)
#pragma unroll
            for (int l = 0; l < M1; ++l) P1[r][l] = pair_index(b1.R[r], b1.L[l]);
        auto pidx0 = [&](int r, int l) { return pair_index(b0.R[r], b0.L[l]); };
        auto pidx1 = [&](int r, int l) { return P1[r][l]; };
        const double c = b0.det * b1.det;
        double phi = 0.0;
#pragma unroll
        for (int r = 0; r < M0; ++r)
#pragma unroll
            for (int l = 0; l < M0; ++l) phi += b0.a[l][r] * hact[b0.R[r] + n * b0.L[l]];
#pragma unroll
        for (int r = 0; r < M1; ++r)
#pragma unroll
            for (int l = 0; l < M1; ++l) phi += b1.a[l][r] * hact[b1.R[r] + n * b1.L[l]];
        phi += same_spin_energy<M0>(G2, np, b0, pidx0);
        phi += same_spin_energy<M1>(G2, np, b1, pidx1);
#pragma unroll
        for (int r = 0; r < M0; ++r)
#pragma unroll
            for (int l = 0; l < M0; ++l) {
                const double* g = G2 + pidx0(r, l) * np;
                double inner = 0.0;
#pragma unroll
                for (int r2 = 0; r2 < M1; ++r2)
#pragma unroll
                    for (int l2 = 0; l2 < M1; ++l2) inner += g[P1[r2][l2]] * b1.a[l2][r2];
                phi += b0.a[l][r] * inner;
            }
        const double H = c * phi;
        const auto Hd = to_global(sd.Hdet), Sd = to_global(sd.Sdet);
        Hd[(size_t)d + (size_t)n_det * e] = H;
        Hd[(size_t)e + (size_t)n_det * d] = H;
        Sd[(size_t)d + (size_t)n_det * e] = c;
        Sd[(size_t)e + (size_t)n_det * d] = c;
        if (tf != t) {
            Hd[(size_t)fd + (size_t)n_det * fe] = H;
            Hd[(size_t)fe + (size_t)n_det * fd] = H;
            Sd[(size_t)fd + (size_t)n_det * fe] = c;
            Sd[(size_t)fe + (size_t)n_det * fd] = c;
        }
    }
}

// The backward of a pair, F_pair = w (H - E S), runs as two kernels over the same blocks and pairs,
// each recomputing the (cheap) inversions, so that neither holds everything in registers:
//   k_pairs_lane_bwd_f: the F blocks (reads of the kernel), then the S_act / h_act adjoints
//                       (ds, dh: the ds/dh columns of each block's partial row);
//   k_pairs_lane_bwd_g: the adjoint of the packed kernel (the dg columns of the partial row),
//                       same-spin through the compound W accumulator, opposite-spin on the matrix
//                       cores; it needs no kernel values at all.
// Both use the same wave-uniform pair loop, so a pair is skipped (zero weight, deferred) by both.

// LDS: G2, S_act, h_act, then ds, dh per wavefront: each wavefront accumulates its pairs into its
// own copy, in its own instruction order, and the copies are summed in wavefront order at the end,
// so the block's partial row does not depend on how the wavefronts interleave.
template <int M0, int M1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lane_field_waves(M0))))
    k_pairs_lane_bwd_f(const SlotDesc* __restrict__ slots, TopoDesc topo, int swap_spins, double cond_limit) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n, ng = topo.ng;
    const int np = n * (n + 1) / 2;
    extern __shared__ double lds[];
    double* G2 = lds;
    double* sact = G2 + np * np;
    double* hact = sact + n * n;
    double* ds_all = hact + n * n;  // one (ds, dh) copy per wavefront
    double* ds = ds_all + (threadIdx.x >> 6) * 2 * n * n;
    double* dh = ds + n * n;
    load_pair_tables(sd, n, np, G2, sact, hact);
    for (int t = threadIdx.x; t < 2 * n * n * (int)(blockDim.x >> 6); t += blockDim.x) ds_all[t] = 0.0;
    __syncthreads();
    const SpinTables tab = spin_tables(topo, swap_spins);
    const long long n_work = topo.n_lane_pairs;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_work;
         i += (long long)gridDim.x * blockDim.x) {
        const long long t =
            lane_pair(topo, (long long)(((unsigned long long)i * (unsigned long long)topo.pair_mul) % (unsigned long long)n_work));
        int d, e, fd, fe;
        unpack_pair(t, d, e);  // d: left (bra) determinant, e: right (ket)
        const long long tf = flip_partner(topo, t, d, e, fd, fe);
        double wH, wS;
        pair_weights(sd, topo, d, e, fd, fe, tf != t, d == e ? 1.0 : 2.0, wH, wS);
        if (wH == 0.0 && wS == 0.0) continue;
        LaneBlock<M0> b0;
        LaneBlock<M1> b1;
        occupied<M0>((uint32_t)tab.mask0[tab.det0[e]], b0.R);
        occupied<M0>((uint32_t)tab.mask0[tab.det0[d]], b0.L);
        occupied<M1>((uint32_t)tab.mask1[tab.det1[e]], b1.R);
        occupied<M1>((uint32_t)tab.mask1[tab.det1[d]], b1.L);
        const bool ok0 = lane_invert<M0>(sact, n, cond_limit, b0);
        const bool ok1 = lane_invert<M1>(sact, n, cond_limit, b1);
        if (!(ok0 && ok1)) continue;  // deferred: done by the lane-group kernel
        int P1[LaneBlock<M1>::K][LaneBlock<M1>::K];
#pragma unroll
        for (int r = 0; r < M1; ++r)
#pragma unroll
            for (int l = 0; l < M1; ++l) P1[r][l] = pair_index(b1.R[r], b1.L[l]);
        auto pidx0 = [&](int r, int l) { return pair_index(b0.R[r], b0.L[l]); };
        auto pidx1 = [&](int r, int l) { return P1[r][l]; };
        const double c = b0.det * b1.det;
        double F0[LaneBlock<M0>::K][LaneBlock<M0>::K];
        double F1[LaneBlock<M1>::K][LaneBlock<M1>::K];
        double phi_h = 0.0;
#pragma unroll
        for (int r = 0; r < M0; ++r)
#pragma unroll
            for (int l = 0; l < M0; ++l) {
                F0[r][l] = hact[b0.R[r] + n * b0.L[l]];
                phi_h += b0.a[l][r] * F0[r][l];
            }
#pragma unroll
        for (int r = 0; r < M1; ++r)
#pragma unroll
            for (int l = 0; l < M1; ++l) {
                F1[r][l] = hact[b1.R[r] + n * b1.L[l]];
                phi_h += b1.a[l][r] * F1[r][l];
            }
        same_spin_field<M0>(G2, np, b0, pidx0, F0);
        same_spin_field<M1>(G2, np, b1, pidx1, F1);
        // opposite spin: each kernel entry feeds both F blocks
#pragma unroll
        for (int r = 0; r < M0; ++r)
#pragma unroll
            for (int l = 0; l < M0; ++l) {
                const double* g = G2 + pidx0(r, l) * np;
                const double D0 = b0.a[l][r];
                double acc = 0.0;
#pragma unroll
                for (int r2 = 0; r2 < M1; ++r2)
#pragma unroll
                    for (int l2 = 0; l2 < M1; ++l2) {
                        const double gv = g[P1[r2][l2]];
                        acc += gv * b1.a[l2][r2];
                        F1[r2][l2] += gv * D0;
                    }
                F0[r][l] += acc;
            }
        double phi_f = 0.0;
#pragma unroll
        for (int r = 0; r < M0; ++r)
#pragma unroll
            for (int l = 0; l < M0; ++l) phi_f += b0.a[l][r] * F0[r][l];
#pragma unroll
        for (int r = 0; r < M1; ++r)
#pragma unroll
            for (int l = 0; l < M1; ++l) phi_f += b1.a[l][r] * F1[r][l];
        const double sc = wH * 0.5 * (phi_h + phi_f) + wS;
        overlap_adjoint<M0>(ds, dh, n, b0, F0, c, sc, wH);
        overlap_adjoint<M1>(ds, dh, n, b1, F1, c, sc, wH);
    }
    __syncthreads();
    const auto out = to_global(sd.partials) + (size_t)blockIdx.x * (ng + 2 * n * n) + ng;
    const int nw = blockDim.x >> 6;
    for (int t = threadIdx.x; t < 2 * n * n; t += blockDim.x) {
        double v = 0.0;
        for (int w = 0; w < nw; ++w) v += ds_all[w * 2 * n * n + t];
        out[t] = v;
    }
}

// LDS: S_act, dg (full np x np or packed), then the optional W (wcopies x nc x nc) and MFMA staging
// rows (per wavefront 2 x 16 x np16).
// Fast: W and the MFMA update (when both fit in LDS); otherwise direct LDS atomics into dg.
// Fixed summation order: every wavefront runs the same number of iterations (block-uniform loop);
// W is either private per wavefront (summed in wavefront order when it is expanded) or shared and
// updated by the wavefronts in turn; the MFMA tiles and the direct atomics enter dg in turn; the
// expansion of W into dg gathers per element. With a given launch shape the partial rows are
// therefore the same bits on every run.
template <bool FullDg, bool Fast, int M0, int M1>
__global__ void __launch_bounds__(256)
    k_pairs_lane_bwd_g(const SlotDesc* __restrict__ slots, TopoDesc topo, int swap_spins, double cond_limit,
                       LaneBwdLayout lay) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n, ng = topo.ng;
    const int np = n * (n + 1) / 2;
    extern __shared__ double lds[];
    double* sact = lds;
    const int dglen = FullDg ? np * np : ng;
    double* dg_all = sact + n * n;
    double* W_all = dg_all + lay.dcopies * dglen;
    double* bufs = W_all + lay.wcopies * lay.nc * lay.nc;
    const int wave = threadIdx.x >> 6;
    const bool private_w = lay.wcopies > 1, private_dg = lay.dcopies > 1;
    double* dg = dg_all + (private_dg ? wave * dglen : 0);
    double* W = W_all + (private_w ? wave * lay.nc * lay.nc : 0);
    double* bufA = bufs + (size_t)wave * 32 * lay.np16;
    double* bufB = bufA + 16 * lay.np16;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) sact[t] = to_global(sd.sact)[t];
    {
        const int nzero = lay.dcopies * dglen + lay.wcopies * lay.nc * lay.nc + (blockDim.x >> 6) * 32 * lay.np16;
        for (int t = threadIdx.x; t < nzero; t += blockDim.x) dg_all[t] = 0.0;
    }
    __syncthreads();
    const SpinTables tab = spin_tables(topo, swap_spins);
    const long long n_work = topo.n_lane_pairs;
    const long long stride = (long long)gridDim.x * blockDim.x;
    // 6-electron blocks need np16 = 80 (5 x 5 tiles); their tiles live only for one iteration and
    // are flushed into dg after it, the rest keep theirs in registers for the whole kernel
    constexpr int TN = M0 >= 6 ? kMfmaTilesWide : kMfmaTiles;
    constexpr bool FlushEach = M0 >= 6;
    MfmaTilesN<TN> tiles;
#pragma unroll
    for (int tp = 0; tp < TN; ++tp)
#pragma unroll
        for (int tq = 0; tq < TN; ++tq) tiles[tp][tq] = lane_double4{0.0, 0.0, 0.0, 0.0};
    // block-uniform trip count: every wavefront takes part in every turn, and the MFMA update needs
    // every lane of its wavefront
    for (long long base = (long long)blockIdx.x * blockDim.x; base < n_work; base += stride) {
        const long long i = base + threadIdx.x;
        int d = 0, e = 0;
        bool live = i < n_work;
        double w = 0.0;
        if (live) {
            const long long t = lane_pair(
                topo, (long long)(((unsigned long long)i * (unsigned long long)topo.pair_mul) % (unsigned long long)n_work));
            int fd, fe;
            unpack_pair(t, d, e);  // d: left (bra) determinant, e: right (ket)
            const long long tf = flip_partner(topo, t, d, e, fd, fe);
            double wS;
            pair_weights(sd, topo, d, e, fd, fe, tf != t, d == e ? 1.0 : 2.0, w, wS);  // the adjoint of G needs wH only
        }
        live = live && w != 0.0;
        LaneBlock<M0> b0;
        LaneBlock<M1> b1;
        occupied<M0>((uint32_t)tab.mask0[tab.det0[e]], b0.R);
        occupied<M0>((uint32_t)tab.mask0[tab.det0[d]], b0.L);
        occupied<M1>((uint32_t)tab.mask1[tab.det1[e]], b1.R);
        occupied<M1>((uint32_t)tab.mask1[tab.det1[d]], b1.L);
        const bool ok0 = lane_invert<M0>(sact, n, cond_limit, b0);
        const bool ok1 = lane_invert<M1>(sact, n, cond_limit, b1);
        live = live && ok0 && ok1;  // deferred pairs are done by the lane-group kernel
        int P1[LaneBlock<M1>::K][LaneBlock<M1>::K];
#pragma unroll
        for (int r = 0; r < M1; ++r)
#pragma unroll
            for (int l = 0; l < M1; ++l) P1[r][l] = pair_index(b1.R[r], b1.L[l]);
        auto pidx0 = [&](int r, int l) { return pair_index(b0.R[r], b0.L[l]); };
        auto pidx1 = [&](int r, int l) { return P1[r][l]; };
        const double cw = live ? b0.det * b1.det * w : 0.0;
        if constexpr (Fast) {
            // inlined at both call sites: an outlined lambda sees W only as a generic pointer, and
            // its LDS accesses become flat instructions
            auto same_spin = [&]() __attribute__((always_inline)) {
                if (live) {
                    same_spin_w<M0>(W, lay.nc, b0, cw);
                    same_spin_w<M1>(W, lay.nc, b1, cw);
                }
            };
            if (FlushEach) {  // pair state dies before the wide tiles come alive
                if (private_w) same_spin();
                else wave_turns(same_spin);
            }
            opposite_dg_mfma<M0, M1, TN>(bufA, bufB, lay.np16, live, b0, b1, P1, cw, tiles);
            if (!FlushEach) {
                if (private_w) same_spin();
                else wave_turns(same_spin);
            }
            if (FlushEach) {
                if (private_dg) flush_mfma_tiles<FullDg, TN>(tiles, lay.np16, np, dg, true);
                else wave_turns([&]() { flush_mfma_tiles<FullDg, TN>(tiles, lay.np16, np, dg, true); });
            }
        } else if (live) {
            // no room for W or a copy of dg per wavefront (n_act 13-14): shared LDS atomics, whose
            // order follows the wavefronts' interleaving (results reproducible to rounding only);
            // taking turns here would serialise the ~4,000 atomics per 7x7 pair
            same_spin_dg<FullDg, M0>(dg, np, b0, pidx0, cw);
            same_spin_dg<FullDg, M1>(dg, np, b1, pidx1, cw);
#pragma unroll
            for (int r = 0; r < M0; ++r)
#pragma unroll
                for (int l = 0; l < M0; ++l) {
                    const int P = pidx0(r, l);
                    const double cwD0 = cw * b0.a[l][r];
#pragma unroll
                    for (int r2 = 0; r2 < M1; ++r2)
#pragma unroll
                        for (int l2 = 0; l2 < M1; ++l2) dg_add<FullDg>(dg, np, P, P1[r2][l2], cwD0 * b1.a[l2][r2]);
                }
        }
    }
    if (Fast && !FlushEach) {
        if (private_dg) flush_mfma_tiles<FullDg, TN>(tiles, lay.np16, np, dg, false);
        else wave_turns([&]() { flush_mfma_tiles<FullDg, TN>(tiles, lay.np16, np, dg, false); });
    }
    __syncthreads();
    dg = dg_all;
    if (private_dg) {  // the wavefronts' copies, summed in wavefront order
        for (int t = threadIdx.x; t < dglen; t += blockDim.x) {
            double v = dg_all[t];
            for (int w = 1; w < lay.dcopies; ++w) v += dg_all[(size_t)w * dglen + t];
            dg_all[t] = v;
        }
        __syncthreads();
    }
    if (Fast) {
        w_gather<FullDg>(W_all, lay.wcopies, lay.nc, dg, np);
        __syncthreads();
    }
    const auto out = to_global(sd.partials) + (size_t)blockIdx.x * (ng + 2 * n * n);
    if (FullDg) {
        for (int t = threadIdx.x; t < ng; t += blockDim.x) {
            int Q, P;
            unpack_pair(t, Q, P);
            out[t] = (P == Q) ? dg[P + np * P] : dg[P + np * Q] + dg[Q + np * P];
        }
    } else {
        for (int t = threadIdx.x; t < ng; t += blockDim.x) out[t] = dg[t];
    }
}

// spin-block shapes with a lane kernel: M0 >= M1, M0 <= 7, M0 - M1 <= 2
template <int M0, int M1>
struct LaneShape {
    static constexpr bool forward = M0 >= M1 && M0 <= 7 && M0 - M1 <= 2;
    static constexpr bool backward = forward;
};

namespace {

constexpr size_t kLaneLdsLimit = 160 * 1024;
constexpr int kLaneWaves = 4;  // wavefronts per block (256 threads)

size_t lane_fwd_doubles(int n) {
    const size_t np = (size_t)n * (n + 1) / 2;
    return np * np + 2 * (size_t)n * n;
}

size_t lane_bwd_f_doubles(int n) { return lane_fwd_doubles(n) + (size_t)kLaneWaves * 2 * n * n; }

size_t lane_bwd_g_doubles(int n, int ng, const LaneBwdLayout& lay) {
    const size_t np = (size_t)n * (n + 1) / 2;
    return (size_t)n * n + (size_t)lay.dcopies * (lay.full_dg ? np * np : (size_t)ng) +
           (size_t)lay.wcopies * lay.nc * lay.nc + (size_t)lay.waves * 32 * lay.np16;
}

// W and the MFMA staging rows when both fit in LDS, else neither (direct atomics). The adjoint is
// held as the full np x np matrix when that fits too, else packed (n = 11, 12 with 6-electron
// blocks). m0: the larger spin block, which sets how many MFMA tiles the kernel can carry.
// Preference: four wavefronts with a private W each (the MFMA tiles enter the shared adjoint once,
// in turns, at the end); else two wavefronts with a private W and a private adjoint each (6-electron
// blocks, whose tiles are flushed after every iteration: no turns at all); else four wavefronts
// taking turns on one W and one adjoint.
LaneBwdLayout lane_bwd_layout(int n, int ng, int m0) {
    const int np = n * (n + 1) / 2;
    const int tn = m0 >= 6 ? kMfmaTilesWide : kMfmaTiles;
    const int full = pairs_full_dg(n, ng) ? 1 : 0;
    const int nc = n * (n - 1) / 2, np16 = (np + 15) / 16 * 16;
    if (np16 <= 16 * tn) {
        const int shapes[3][3] = {{kLaneWaves, kLaneWaves, 1}, {2, 2, 2}, {kLaneWaves, 1, 1}};  // waves, W, dg copies
        for (const auto& sh : shapes)
            for (const int f : {full, 0}) {
                const LaneBwdLayout fast{nc, np16, f, sh[1], sh[2], sh[0]};
                if (lane_bwd_g_doubles(n, ng, fast) * sizeof(double) <= kLaneLdsLimit) return fast;
            }
    }
    return {0, 0, full, 1, 1, kLaneWaves};
}

template <int M0, int M1>
hipError_t lane_fwd(bool run, hipStream_t st, dim3 grid, size_t lds, const SlotDesc* slots, const TopoDesc& topo,
                    int swap, double cond) {
    if (!run)
        return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_lane_fwd<M0, M1>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLaneLdsLimit);
    hipLaunchKernelGGL((k_pairs_lane_fwd<M0, M1>), grid, dim3(64 * kLaneWaves), lds, st, slots, topo, swap, cond);
    return hipSuccess;
}

template <bool F, bool Fast, int M0, int M1>
hipError_t lane_bwd(bool run, int which, hipStream_t st, dim3 grid, const SlotDesc* slots, const TopoDesc& topo,
                    int swap, double cond, LaneBwdLayout lay) {
    if (!run) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_lane_bwd_f<M0, M1>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLaneLdsLimit);
        if (e != hipSuccess) return e;
        return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_lane_bwd_g<F, Fast, M0, M1>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLaneLdsLimit);
    }
    if (which == 1) {
        const size_t lds_f = lane_bwd_f_doubles(topo.n) * sizeof(double);
        hipLaunchKernelGGL((k_pairs_lane_bwd_f<M0, M1>), grid, dim3(64 * kLaneWaves), lds_f, st, slots, topo, swap,
                           cond);
    } else {
        const size_t lds_g = lane_bwd_g_doubles(topo.n, topo.ng, lay) * sizeof(double);
        hipLaunchKernelGGL((k_pairs_lane_bwd_g<F, Fast, M0, M1>), grid, dim3(64 * lay.waves), lds_g, st, slots, topo,
                           swap, cond, lay);
    }
    return hipSuccess;
}

template <int M0, int M1>
hipError_t lane_dispatch_shape(bool run, int which, bool full_dg, hipStream_t st, dim3 grid, size_t lds,
                               const SlotDesc* slots, const TopoDesc& topo, int swap, double cond,
                               LaneBwdLayout lay) {
    if (which == 0) return lane_fwd<M0, M1>(run, st, grid, lds, slots, topo, swap, cond);
    if constexpr (LaneShape<M0, M1>::backward) {
        // configuration (run == false) prepares every variant
        const bool fast = lay.nc != 0;
        hipError_t e = hipSuccess;
        if (!run || (full_dg && fast))
            e = lane_bwd<true, true, M0, M1>(run, which, st, grid, slots, topo, swap, cond, lay);
        if (e == hipSuccess && (!run || (full_dg && !fast)))
            e = lane_bwd<true, false, M0, M1>(run, which, st, grid, slots, topo, swap, cond, lay);
        if (e == hipSuccess && (!run || (!full_dg && fast)))
            e = lane_bwd<false, true, M0, M1>(run, which, st, grid, slots, topo, swap, cond, lay);
        if (e == hipSuccess && (!run || (!full_dg && !fast)))
            e = lane_bwd<false, false, M0, M1>(run, which, st, grid, slots, topo, swap, cond, lay);
        return e;
    }
    return hipErrorInvalidValue;
}

// calls f.template operator()<M0, M1>() for the lane shape (m0, m1), or returns false
template <typename Fn>
bool for_lane_shape(int m0, int m1, Fn&& f) {
    bool hit = false;
    auto one = [&](auto a, auto b) {
        constexpr int A = decltype(a)::value, Bm = decltype(b)::value;
        if (!hit && A == m0 && Bm == m1) {
            hit = true;
            f.template operator()<A, Bm>();
        }
    };
    using std::integral_constant;
#define XMVB_LANE(A, B) one(integral_constant<int, A>{}, integral_constant<int, B>{})
    XMVB_LANE(1, 0); XMVB_LANE(1, 1); XMVB_LANE(2, 0); XMVB_LANE(2, 1); XMVB_LANE(2, 2);
    XMVB_LANE(3, 1); XMVB_LANE(3, 2); XMVB_LANE(3, 3); XMVB_LANE(4, 2); XMVB_LANE(4, 3);
    XMVB_LANE(4, 4); XMVB_LANE(5, 3); XMVB_LANE(5, 4); XMVB_LANE(5, 5); XMVB_LANE(6, 4);
    XMVB_LANE(6, 5); XMVB_LANE(6, 6); XMVB_LANE(7, 5); XMVB_LANE(7, 6); XMVB_LANE(7, 7);
#undef XMVB_LANE
    return hit;
}

}  // namespace

size_t pairs_lane_lds_bytes(int n, int ng, bool backward, int m0) {
    if (!backward) return lane_fwd_doubles(n) * sizeof(double);
    return std::max(lane_bwd_f_doubles(n), lane_bwd_g_doubles(n, ng, lane_bwd_layout(n, ng, m0))) * sizeof(double);
}

bool pairs_lane_supported(int n_alpha, int n_beta, bool backward) {
    const int m0 = n_alpha >= n_beta ? n_alpha : n_beta, m1 = n_alpha >= n_beta ? n_beta : n_alpha;
    bool ok = false;
    for_lane_shape(m0, m1, [&]<int A, int B>() { ok = backward ? LaneShape<A, B>::backward : LaneShape<A, B>::forward; });
    return ok;
}

hipError_t configure_pairs_lane() {
    hipError_t err = hipSuccess;
    for (int m0 = 1; m0 <= 7; ++m0)
        for (int m1 = 0; m1 <= m0; ++m1)
            for_lane_shape(m0, m1, [&]<int A, int B>() {
                for (int which = 0; which < 2; ++which) {
                    if (which && !LaneShape<A, B>::backward) continue;
                    const hipError_t e = lane_dispatch_shape<A, B>(false, which, true, nullptr, dim3(), 0, nullptr,
                                                                   TopoDesc{}, 0, 0.0, LaneBwdLayout{});
                    if (e != hipSuccess) err = e;
                }
            });
    return err;
}

void launch_pairs_lane(hipStream_t st, int which, int blocks, const SlotDesc* slots, int count, const TopoDesc& topo,
                       int n_alpha, int n_beta, double cond_limit) {
    const int swap = n_beta > n_alpha ? 1 : 0;
    const int m0 = swap ? n_beta : n_alpha, m1 = swap ? n_alpha : n_beta;
    const LaneBwdLayout lay = which ? lane_bwd_layout(topo.n, topo.ng, m0) : LaneBwdLayout{};
    const bool full = lay.full_dg != 0;
    const size_t lds = pairs_lane_lds_bytes(topo.n, topo.ng, which != 0, m0);
    const dim3 grid(blocks, 1, count);
    for_lane_shape(m0, m1, [&]<int A, int B>() {
        (void)lane_dispatch_shape<A, B>(true, which, full, st, grid, lds, slots, topo, swap, cond_limit, lay);
    });
}

}  // namespace xmvb_amd
