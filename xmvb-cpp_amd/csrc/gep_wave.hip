// One-wavefront dense eigensolve for small structure spaces (n_str <= 64): the lowest root of
// H c = lambda S c with c^T S c = 1, GeneralizedEigensolver::solve_dense semantics
// (linalg/eigensolver.cpp:365-387: Cholesky of the metric, standard form L^-1 H L^-T, lowest
// eigenpair, c = L^-T z), and the determinant coefficients a_d = sum_I t_dI c_I.
//
// Layout: lane j owns column j of every n x n matrix and keeps it in registers (a[0..N-1], N the
// compile-time bucket >= n, rows beyond n zero). Every loop over rows is unrolled at compile time,
// so register indices are constants; the values a step needs from all lanes (a column of L, a
// Householder vector) are written to LDS by their owning lanes and read back as wave-uniform
// broadcasts. One wavefront per problem and no block barrier anywhere: the serial chains of the
// factorisations cost a few LDS round trips per step instead of several block barriers.
//   Cholesky:        lane j holds S[:, j]; step k broadcasts column k of L (kept in LDS, Lc)
//   X = L^-1 H:      lane j runs the forward substitution on H[:, j]
//   A = L^-1 X^T:    rows of X through an LDS transpose, a second forward substitution, then
//                    A <- (A + A^T) / 2 through another transpose
//   tridiagonal:     Householder reflections, p = beta A v from lane j's own column (no cross-lane
//                    sum), vectors kept in LDS (Wb) for the back-transform
//   T y = lambda y:  tridiag_lowest_wave (multisection, inverse iteration, Rayleigh quotient)
//   c = L^-T z:      backward substitution, c_k broadcast by readlane
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"
#include "wave.hpp"

namespace xmvb_amd {

namespace {

// Rows are processed in chunks of kChunk with the wave-uniform LDS operands of the next chunk
// issued before the FMAs of the current one (scheduling barriers keep that order), so a chunk's
// loads are in flight while the previous chunk computes.
constexpr int kChunk = 16;

template <int N, int K0>
__device__ __forceinline__ void load_chunk(double (&lv)[kChunk], const double* col, int c0) {
#pragma unroll
    for (int t = 0; t < kChunk; ++t)
        if (c0 + t >= K0 && c0 + t < N) lv[t] = col[c0 + t];
}

// a[i] -= col[i] * s for the rows i > k (col: wave-uniform column in LDS); k is a compile-time
// constant after unrolling, which the first-chunk test relies on
template <int N>
__device__ __forceinline__ void axpy_below(double (&a)[N], const double* col, double s, int k) {
    constexpr int NC = (N + kChunk - 1) / kChunk;
    double lv[2][kChunk];
    const int first = (k + 1) / kChunk;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (c < first) continue;
        if (c == first) {
#pragma unroll
            for (int t = 0; t < kChunk; ++t)
                if (c * kChunk + t > k && c * kChunk + t < N) lv[c & 1][t] = col[c * kChunk + t];
        }
        if (c + 1 < NC) {
#pragma unroll
            for (int t = 0; t < kChunk; ++t)
                if ((c + 1) * kChunk + t < N) lv[(c + 1) & 1][t] = col[(c + 1) * kChunk + t];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < kChunk; ++t)
            if (c * kChunk + t > k && c * kChunk + t < N) a[c * kChunk + t] -= lv[c & 1][t] * s;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// 1 / sqrt(x) for x > 0: v_rsq_f64 and two Newton steps (quadratic: each squares the relative error)
__device__ __forceinline__ double rsqrt_f64(double x) {
    double r = __builtin_amdgcn_rsq(x);
    r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
    r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
    return r;
}

// lane j: a[i] = (M[i, j] + M[j, i]) / 2 for an n x n column-major M, read by rows (coalesced)
// and transposed through LDS (W: N x P doubles)
template <int N, int P>
__device__ __forceinline__ void load_symmetric(double (&a)[N], gptr<const double> M, int n, int j, double* W) {
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = (j < n && i < n) ? M[j + n * i] : 0.0;  // M[j, i]
    if (j < N) {
#pragma unroll
        for (int r = 0; r < N; ++r) W[r * P + j] = a[r];
    }
    wave_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = 0.5 * (a[i] + (j < N ? W[j * P + i] : 0.0));  // W[j P + i] = M[i, j]
    wave_sync();
}

template <int N>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) k_gep_wave(const SlotDesc* __restrict__ slots, TopoDesc topo) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n_str;
    const int j = threadIdx.x;
    constexpr int P = N + 1;               // row stride of the transposes
    __shared__ double Lc[N * N];           // Lc[k N + i] = L[i, k]
    __shared__ double Wb[N * P];           // transposes, then Householder vector k at Wb[k N + i]
    __shared__ double wv[N];               // Householder w
    __shared__ double rl[N];               // 1 / L[k, k]
    __shared__ double hb[N];               // Householder beta
    __shared__ __align__(16) double tri[288];  // tridiagonal solver scratch
    const auto Ss = to_global(sd.Ss);
    const auto Hs = to_global(sd.Hs);
    double a[N];
    load_symmetric<N, P>(a, Ss, n, j, Wb);
    // Cholesky S = L L^T, right-looking: lane j's a[k] is S[k, j] = S[j, k], so L[j, k] is local.
    // Inner loops run over all N rows with the row condition constant after unrolling. The next
    // pivot, S[k+1, k+1] - L[k+1, k]^2, is lane k+1's own update, so it is formed (and its reciprocal
    // square root started) before the trailing update instead of after it.
    bool definite = true;  // wave-uniform
    double akk = wave_bcast(a[0], 0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k < n && definite) {
            definite = akk > 0.0;
            const double piv = definite ? akk : 1.0;
            const double rdk = rsqrt_f64(piv), dk = piv * rdk;
            const double l = j == k ? dk : (j > k ? a[k] * rdk : 0.0);
            if (j < N) Lc[k * N + j] = l;
            if (j == 0) rl[k] = rdk;
            if (k + 1 < N) akk = wave_bcast(a[k + 1] - l * l, (k + 1) & 63);
            wave_sync();
            axpy_below<N>(a, Lc + k * N, l, k);
        }
    }
    if (!definite) {
        if (j == 0) to_global(sd.flags)[FLAG_METRIC] = 1;
        return;
    }
    // X = L^-1 H, column j
    load_symmetric<N, P>(a, Hs, n, j, Wb);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k < n) {
            const double x = a[k] * rl[k];
            a[k] = x;
            axpy_below<N>(a, Lc + k * N, x, k);
        }
    }
    // row j of X: Wb[r P + c] = X[r, c]
    if (j < N) {
#pragma unroll
        for (int r = 0; r < N; ++r) Wb[r * P + j] = a[r];
    }
    wave_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = j < N ? Wb[j * P + i] : 0.0;
    // A[:, j] = L^-1 (row j of X)^T
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k < n) {
            const double x = a[k] * rl[k];
            a[k] = x;
            axpy_below<N>(a, Lc + k * N, x, k);
        }
    }
    wave_sync();
    if (j < N) {
#pragma unroll
        for (int r = 0; r < N; ++r) Wb[r * P + j] = a[r];
    }
    wave_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = 0.5 * (a[i] + (j < N ? Wb[j * P + i] : 0.0));
    wave_sync();
    // Householder tridiagonalisation; lane j collects d_j and e_j = T[j+1, j]
    double d_r = 0.0, e_r = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k < n) {
            d_r = wave_put(wave_bcast(a[k], k), k, d_r);
            if (k + 2 < n) {
                const int m0 = k + 1;
                const double x = j > k ? a[k] : 0.0;  // A[j, k]
                const double sigma = wave_sum(x * x);
                const double x0 = wave_bcast(a[k], m0);
                const double rs = sigma > 0.0 ? sigma * rsqrt_f64(sigma) : 0.0;  // sqrt(sigma)
                const double alpha = x0 >= 0.0 ? -rs : rs;
                const double v = x - (j == m0 ? alpha : 0.0);
                const double vtv = sigma - 2.0 * alpha * x0 + alpha * alpha;
                const double bk = (sigma > 0.0 && vtv > 0.0) ? fast_div(2.0, vtv) : 0.0;
                if (j < N) Wb[k * N + j] = v;
                if (j == 0) hb[k] = bk;
                wave_sync();
                double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;  // (A v)_j from column j
                {
                    constexpr int NC = (N + kChunk - 1) / kChunk;
                    double lv[2][kChunk];
                    const int first = m0 / kChunk;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        if (c < first) continue;
                        if (c == first) load_chunk<N, 0>(lv[c & 1], Wb + k * N, c * kChunk);
                        if (c + 1 < NC) load_chunk<N, 0>(lv[(c + 1) & 1], Wb + k * N, (c + 1) * kChunk);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int t = 0; t < kChunk; ++t) {
                            const int i = c * kChunk + t;
                            if (i < m0 || i >= N) continue;
                            const double q = a[i] * lv[c & 1][t];
                            if ((i & 3) == 0) p0 += q;
                            else if ((i & 3) == 1) p1 += q;
                            else if ((i & 3) == 2) p2 += q;
                            else p3 += q;
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                const double p = bk * ((p0 + p1) + (p2 + p3));
                const double K = 0.5 * bk * wave_sum(v * p);
                const double w = p - K * v;
                if (j < N) wv[j] = w;
                wave_sync();
                {
                    constexpr int NC = (N + kChunk - 1) / kChunk;
                    double lv[2][kChunk], lw[2][kChunk];
                    const int first = m0 / kChunk;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        if (c < first) continue;
                        if (c == first) {
                            load_chunk<N, 0>(lv[c & 1], Wb + k * N, c * kChunk);
                            load_chunk<N, 0>(lw[c & 1], wv, c * kChunk);
                        }
                        if (c + 1 < NC) {
                            load_chunk<N, 0>(lv[(c + 1) & 1], Wb + k * N, (c + 1) * kChunk);
                            load_chunk<N, 0>(lw[(c + 1) & 1], wv, (c + 1) * kChunk);
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int t = 0; t < kChunk; ++t) {
                            const int i = c * kChunk + t;
                            if (i >= m0 && i < N) a[i] -= lv[c & 1][t] * w + lw[c & 1][t] * v;
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                e_r = wave_put(bk != 0.0 ? alpha : x0, k, e_r);
                wave_sync();
            } else if (k + 1 < n) {
                e_r = wave_put(wave_bcast(a[k], (k + 1) & 63), k, e_r);
            }
        }
    }
    double yv;
    const double lambda = tridiag_lowest_wave(j, n, d_r, e_r, tri, yv);
    for (int k = n - 3; k >= 0; --k) {  // back-transform z = H_0 H_1 ... H_{n-3} y
        const double vk = j < N ? Wb[k * N + j] : 0.0;
        yv -= hb[k] * wave_sum(vk * yv) * vk;
    }
    // c = L^-T z: c_k = (z_k - sum_{i > k} L[i, k] c_i) / L[k, k], z updated column by column
    double c = 0.0, z = yv;
    for (int k = n - 1; k >= 0; --k) {
        const double ck = wave_bcast(z, k) * rl[k];
        if (j == k) c = ck;
        if (j < k) z -= Lc[j * N + k] * ck;
    }
    wave_sync();
    if (j < n) {
        to_global(sd.coef)[j] = c;
        wv[j] = c;
    }
    if (j == 0) {
        to_global(sd.energy)[0] = lambda;
        to_global(sd.energy)[1] = lambda + sd.e_nuc;
    }
    wave_sync();
    const auto off = to_global(topo.off);
    const auto ts = to_global(topo.ts);
    const auto tc = to_global(topo.tc);
    for (int d = j; d < topo.n_det; d += 64) {
        double acc = 0.0;
        for (long long t = off[d]; t < off[d + 1]; ++t) acc += tc[t] * wv[ts[t]];
        to_global(sd.dcoef)[d] = acc;
    }
}

}  // namespace

bool launch_gep_wave(hipStream_t st, const SlotDesc* slots, unsigned z, const TopoDesc& topo) {
    const int n = topo.n_str;
    const dim3 grid(1, 1, z), block(64);
    if (n <= 16)
        hipLaunchKernelGGL(k_gep_wave<16>, grid, block, 0, st, slots, topo);
    else if (n <= 32)
        hipLaunchKernelGGL(k_gep_wave<32>, grid, block, 0, st, slots, topo);
    else if (n <= 48)
        hipLaunchKernelGGL(k_gep_wave<48>, grid, block, 0, st, slots, topo);
    else if (n <= 64)
        hipLaunchKernelGGL(k_gep_wave<64>, grid, block, 0, st, slots, topo);
    else
        return false;
    return true;
}

}  // namespace xmvb_amd
