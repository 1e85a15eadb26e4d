// Lowest root of the structure eigenproblem H c = E S c by Davidson iteration in the factored
// coordinates of the structure metric, the reference's solve_factored_structure_roots
// (derivatives/gradient/active_space/evaluator.cpp:339-391) with n_roots = 1 and the Davidson core
// of linalg/davidson_core.hpp:271-560 in its identity-overlap form.
//
// With S = L L^T (the metric's Cholesky factor; the reference keeps its upper factor U = L^T) and
// y = L^T c the problem becomes A y = E y, A = L^-1 H L^-T, with the identity metric. The reference
// applies A matrix-free as U^-T H U^-1 y; here H and S are resident, so the front half of the
// eigensolve (launch_gep_front: Cholesky, L^-1, A) forms A once and every Davidson action is one
// dense product A y. Options as structure_solve_options with the physical stopping of
// StructureSolveAccuracy::configure_physical_stopping (core/contracts/eigensolver.hpp:51-56): energy
// tolerance min(energy_tolerance, 1e-8), stationarity tolerance its square root; at most 2 n
// iterations and min(n, 32) basis vectors (make_davidson_options, linalg/eigensolver.cpp:321-361).
//
// One 256-thread workgroup per problem slot runs the iteration: orthonormal basis extension with two
// classical Gram-Schmidt passes and the 64-epsilon dependence floor (append_s_orthonormal_block,
// davidson_core.hpp:101-180, identity metric), the projected matrix (at most 32 x 32) solved by the
// first wavefront (Householder in LDS, multisection on Sturm counts for the two lowest eigenvalues,
// inverse iteration, back-transform), the EnergyAndResidual stopping test (davidson_core.hpp:
// 452-487: stationarity residual |r| / |y| at most the tolerance and the Ritz energy stable to the
// energy tolerance, or resolved at the first iteration, or the basis complete), the identity
// preconditioner of the factored solve (the correction is the residual itself), projected against
// the basis with the rounding bound of project_davidson_corrections (davidson_core.hpp:186-213).
//
// Start: the slot's previous eigenvector when it has one (root recycling,
// gradient/active_space/evaluator.cpp:233-246), mapped as y = L^T c, else the reference's initial
// block (make_davidson_initial_vectors: the unit vector at the smallest H_ii/S_ii and a
// deterministic xorshift +-1 vector) mapped the same way. Restart (basis full): the basis shrinks to
// the current Ritz vector and the previous one (the reference keeps the 16 lowest Ritz vectors);
// this changes iteration counts, not the stopping test. The result c = L^-T y satisfies c^T S c = 1.

#include "gmem.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace xmvb_amd {
namespace {

constexpr int kDavThreads = 256;
constexpr int MS = kDavidsonMaxSubspace;  // LDS leading dimension of the projected matrices

template <int Ctrl>
__device__ __forceinline__ double dav_dpp(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), Ctrl, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double dav_lane(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double dav_put(double v, int dst, double old) {
    return (int)(threadIdx.x & 63) == dst ? v : old;
}

__device__ __forceinline__ double dav_wave_sum(double v) {
    v += dav_dpp<0xB1>(v);
    v += dav_dpp<0x4E>(v);
    v += dav_dpp<0x141>(v);
    v += dav_dpp<0x140>(v);
    return (dav_lane(v, 0) + dav_lane(v, 16)) + (dav_lane(v, 32) + dav_lane(v, 48));
}

__device__ __forceinline__ double dav_wave_max(double v) {
    v = fmax(v, dav_dpp<0xB1>(v));
    v = fmax(v, dav_dpp<0x4E>(v));
    v = fmax(v, dav_dpp<0x141>(v));
    v = fmax(v, dav_dpp<0x140>(v));
    return fmax(fmax(dav_lane(v, 0), dav_lane(v, 16)), fmax(dav_lane(v, 32), dav_lane(v, 48)));
}

// block sum of up to 4 values at once (red: >= 4 * 4 doubles); every thread gets the totals
__device__ void dav_block_sums(double* v, int k, double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int t = 0; t < k; ++t) v[t] = dav_wave_sum(v[t]);
    __syncthreads();
    if (lane == 0)
        for (int t = 0; t < k; ++t) red[t * 4 + wave] = v[t];
    __syncthreads();
    for (int t = 0; t < k; ++t) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red[t * 4 + w];
        v[t] = s;
    }
    __syncthreads();
}

// c[j] = sum_rows X[:, j] . y for j < cols (X column-major n x cols), block-wide; c in LDS
__device__ void dav_project(const double* X, int n, int cols, const double* y, double* c, double* red32) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    double acc[MS];
#pragma unroll
    for (int j = 0; j < MS; ++j) acc[j] = 0.0;
    for (int i = tid; i < n; i += blockDim.x) {
        const double yi = y[i];
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < cols) acc[j] += X[i + (size_t)n * j] * yi;
    }
#pragma unroll
    for (int j = 0; j < MS; ++j)
        if (j < cols) acc[j] = dav_wave_sum(acc[j]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < cols) red32[wave * MS + j] = acc[j];
    __syncthreads();
    if (tid < cols) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red32[w * MS + tid];
        c[tid] = s;
    }
    __syncthreads();
}

// y[i] -= sum_j X[i, j] c[j] over the rows this thread owns
__device__ __forceinline__ void dav_subtract(double* y, const double* X, int n, int cols, const double* c) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double a = 0.0;
        for (int j = 0; j < cols; ++j) a += X[i + (size_t)n * j] * c[j];
        y[i] -= a;
    }
}

// Lowest two eigenvalues lam[0..1] (lam[1] only when m > 1) and the lowest eigenvector y[0..m) of
// the symmetric m x m matrix A (LDS, column-major, leading dimension MS; destroyed), m <= 32.
// Called by all 64 lanes of one wavefront; lane l owns row l of A.
__device__ void dav_small_lowest(double* A, int m, double* Vh, double* hb, double* y, double* lam) {
    const int l = threadIdx.x & 63;
    const bool row = l < m;
    // Householder tridiagonalisation: reflect rows/columns k+1..m-1 of column k
    for (int k = 0; k + 2 < m; ++k) {
        const int m0 = k + 1;
        const bool in = l >= m0 && l < m;
        const double x = in ? A[l + MS * k] : 0.0;
        const double sigma = dav_wave_sum(x * x);
        const double x0 = dav_lane(x, m0);
        const double alpha = sigma > 0.0 ? (x0 >= 0.0 ? -sqrt(sigma) : sqrt(sigma)) : 0.0;
        const double v = in ? x - (l == m0 ? alpha : 0.0) : 0.0;
        const double vtv = sigma - 2.0 * alpha * x0 + alpha * alpha;
        const double bk = (sigma > 0.0 && vtv > 0.0) ? 2.0 / vtv : 0.0;
        if (l < MS) Vh[l + MS * k] = v;
        if (l == 0) hb[k] = bk;
        double p = 0.0;
        for (int j = m0; j < m; ++j) {
            const double vj = dav_lane(v, j);
            if (in) p += A[l + MS * j] * vj;
        }
        p *= bk;
        const double K = 0.5 * bk * dav_wave_sum(v * p);
        const double w = p - K * v;
        for (int j = m0; j < m; ++j) {
            const double vj = dav_lane(v, j), wj = dav_lane(w, j);
            if (in) A[l + MS * j] -= v * wj + w * vj;
        }
        if (l == m0) A[m0 + MS * k] = alpha;
    }
    __builtin_amdgcn_wave_barrier();  // subdiagonal entries were written by the lanes below
    const double d_r = row ? A[l + MS * l] : 0.0;
    const double e_r = l + 1 < m ? A[(l + 1) + MS * l] : 0.0;     // T[l+1, l]
    const double em1 = l > 0 && row ? A[l + MS * (l - 1)] : 0.0;  // T[l, l-1]
    const double e2_r = em1 * em1;
    const double rad = fabs(em1) + fabs(e_r);
    const double glo = -dav_wave_max(row ? rad - d_r : -1e300);
    const double ghi = dav_wave_max(row ? d_r + rad : -1e300);
    const double pad = 1e-12 * fmax(fabs(glo), fabs(ghi)) + 1e-300;
    // multisection on Sturm counts: the k-th lowest eigenvalue (k = 0, 1)
    double found[2] = {0.0, 0.0};
    for (int kth = 0; kth < 2 && kth < m; ++kth) {
        double lo = glo - pad, hi = ghi + pad;
        const double rel = kth == 0 ? 4.0 * 2.220446049250313e-16 : 1e-12;
        for (int it = 0; it < 200; ++it) {
            if (!(hi - lo > rel * fmax(fabs(lo), fabs(hi)))) break;
            const double xs = lo + (hi - lo) * (double)(l + 1) / 64.0;
            int count = 0;
            double q = dav_lane(d_r, 0) - xs;
            if (q == 0.0) q = -1e-300;
            if (q < 0.0) ++count;
            for (int i = 1; i < m; ++i) {
                q = dav_lane(d_r, i) - xs - dav_lane(e2_r, i) / q;
                if (q == 0.0) q = -1e-300;
                if (q < 0.0) ++count;
            }
            const unsigned long long mask = __ballot(count >= kth + 1);
            const int f = mask ? __ffsll((long long)mask) - 1 : 64;
            const double nlo = f == 0 ? lo : lo + (hi - lo) * (double)f / 64.0;
            const double nhi = f >= 64 ? hi : lo + (hi - lo) * (double)(f + 1) / 64.0;
            lo = nlo;
            hi = nhi;
        }
        found[kth] = 0.5 * (lo + hi);  // uniform across the wavefront
    }
    const double lambda = found[0];
    if (l == 0) {
        lam[0] = found[0];
        lam[1] = found[1];
    }
    // inverse iteration on T - lambda I (LU with partial pivoting, dgttrf/dgttrs order)
    double g = row ? d_r - lambda : 1.0;
    double lo_ = e_r, up = e_r, up2 = 0.0;
    int swp = 0;
    for (int i = 0; i + 1 < m; ++i) {
        double gi = dav_lane(g, i);
        const double li = dav_lane(lo_, i), ui = dav_lane(up, i), gi1 = dav_lane(g, i + 1);
        if (fabs(gi) >= fabs(li)) {
            if (gi == 0.0) gi = 1e-300;
            const double f = li / gi;
            g = dav_put(gi, i, g);
            lo_ = dav_put(f, i, lo_);
            g = dav_put(gi1 - f * ui, i + 1, g);
        } else {
            const double f = gi / li;
            g = dav_put(li, i, g);
            lo_ = dav_put(f, i, lo_);
            up = dav_put(gi1, i, up);
            g = dav_put(ui - f * gi1, i + 1, g);
            if (i + 2 < m) {
                const double ui1 = dav_lane(up, i + 1);
                up2 = dav_put(ui1, i, up2);
                up = dav_put(-f * ui1, i + 1, up);
            }
            if (l == i) swp = 1;
        }
    }
    if (dav_lane(g, m - 1) == 0.0) g = dav_put(1e-300, m - 1, g);
    double z = row ? 1.0 : 0.0;
    for (int iter = 0; iter < 3; ++iter) {
        for (int i = 0; i + 1 < m; ++i) {
            const double zi = dav_lane(z, i), zi1 = dav_lane(z, i + 1), li = dav_lane(lo_, i);
            if (__builtin_amdgcn_readlane(swp, i) == 0) {
                z = dav_put(zi1 - li * zi, i + 1, z);
            } else {
                z = dav_put(zi1, i, z);
                z = dav_put(zi - li * zi1, i + 1, z);
            }
        }
        for (int i = m - 1; i >= 0; --i) {
            double v = dav_lane(z, i);
            if (i + 1 < m) v -= dav_lane(up, i) * dav_lane(z, i + 1);
            if (i + 2 < m) v -= dav_lane(up2, i) * dav_lane(z, i + 2);
            z = dav_put(v / dav_lane(g, i), i, z);
        }
        z *= 1.0 / sqrt(dav_wave_sum(z * z));
    }
    __builtin_amdgcn_wave_barrier();  // reflector scalars were written by lane 0
    for (int k = m - 3; k >= 0; --k) {  // back-transform z = H_0 ... H_{m-3} z
        const double vk = l > k && row ? Vh[l + MS * k] : 0.0;
        z -= hb[k] * dav_wave_sum(vk * z) * vk;
    }
    if (l < MS) y[l] = row ? z : 0.0;
}

struct DavidsonParams {
    double energy_tol, stationarity_tol;
};

// gep layout after launch_gep_front: L at 0, A at n^2, L^-1 (column-major) at 2 n^2; the basis and
// its vectors from 3 n^2 on (2 n ms + 6 n <= 2 n^2 + 9 n doubles, inside gep_work_doubles(n))
__global__ void __launch_bounds__(kDavThreads)
k_davidson_factored(const SlotDesc* __restrict__ slots, TopoDesc topo, DavidsonParams prm) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n_str;
    if (sd.flags[FLAG_METRIC]) return;  // the metric's Cholesky failed: the host reports it
    const int ms = n < MS ? n : MS;
    const double* __restrict__ Lc = sd.gep;                       // L, lower triangle
    const double* __restrict__ A = sd.gep + (size_t)n * n;        // L^-1 H L^-T
    const double* __restrict__ Li = sd.gep + 2 * (size_t)n * n;   // L^-1, column-major
    double* V = sd.gep + 3 * (size_t)n * n;  // n x ms orthonormal basis
    double* AV = V + (size_t)n * ms;         // A V
    double* root = AV + (size_t)n * ms;      // Ritz vector and its image
    double* aroot = root + n;
    double* pv = aroot + n;                  // candidate / correction and its image
    double* ap = pv + n;
    double* prv = ap + n;                    // previous Ritz vector and its image (restart)
    double* aprv = prv + n;
    const double* __restrict__ H = sd.Hs;
    const double* __restrict__ S = sd.Ss;
    __shared__ double P[MS * MS];        // projected matrix V^T A V
    __shared__ double Aw[MS * MS];       // its working copy for the small eigensolver
    __shared__ double Vh[MS * MS];
    __shared__ double hb[MS], ys[MS], yprev[MS], coef[MS], bnorm[MS], lam[2];
    __shared__ double red[16], red32[4 * MS];
    __shared__ int state[2];             // [0] active dimension, [1] scratch index
    const int tid = threadIdx.x, nt = blockDim.x;
    constexpr double eps = 2.220446049250313e-16;

    // ap = A pv (rows owned by each thread; A is symmetric, read by columns for unit stride)
    auto action = [&]() {
        __syncthreads();
        for (int i = tid; i < n; i += nt) {
            double a0 = 0.0, a1 = 0.0;
            int j = 0;
            for (; j + 1 < n; j += 2) {
                a0 += A[i + (size_t)n * j] * pv[j];
                a1 += A[i + (size_t)n * (j + 1)] * pv[j + 1];
            }
            if (j < n) a0 += A[i + (size_t)n * j] * pv[j];
            ap[i] = a0 + a1;
        }
        __syncthreads();
    };
    // append_s_orthonormal_block for the single candidate (pv, ap), identity metric
    auto append = [&]() -> bool {
        const int active = state[0];
        double t[1] = {0.0};
        for (int i = tid; i < n; i += nt) t[0] += pv[i] * pv[i];
        dav_block_sums(t, 1, red);
        const double initial_scale = t[0];
        for (int pass = 0; pass < 2 && active > 0; ++pass) {
            dav_project(V, n, active, pv, coef, red32);
            dav_subtract(pv, V, n, active, coef);
            dav_subtract(ap, AV, n, active, coef);
            __syncthreads();
        }
        double u[1] = {0.0};
        for (int i = tid; i < n; i += nt) u[0] += pv[i] * pv[i];
        dav_block_sums(u, 1, red);
        const double metric2 = u[0];
        const double threshold = 64.0 * eps * fmax(initial_scale, 2.2250738585072014e-308);
        if (!isfinite(metric2) || metric2 <= threshold || metric2 <= 0.0) return false;
        const double inv = 1.0 / sqrt(metric2);
        for (int i = tid; i < n; i += nt) {
            V[i + (size_t)n * active] = pv[i] * inv;
            AV[i + (size_t)n * active] = ap[i] * inv;
        }
        __syncthreads();
        if (tid == 0) {
            bnorm[active] = 1.0;
            state[0] = active + 1;
        }
        __syncthreads();
        return true;
    };
    // projected matrix columns [first, active): V^T (A V[:, c]), symmetrised block
    auto update_projected = [&](int first) {
        const int active = state[0];
        for (int c = first; c < active; ++c) {
            dav_project(V, n, active, AV + (size_t)n * c, coef, red32);
            if (tid < active) {
                P[tid + MS * c] = coef[tid];
                if (tid < first) P[c + MS * tid] = coef[tid];
            }
            __syncthreads();
        }
        if (tid == 0)
            for (int a = first; a < active; ++a)
                for (int b = first; b < a; ++b) {
                    const double s = 0.5 * (P[a + MS * b] + P[b + MS * a]);
                    P[a + MS * b] = s;
                    P[b + MS * a] = s;
                }
        __syncthreads();
    };
    // pv = L^T x (x in prv): seeds and recycled roots enter the factored coordinates
    auto to_factored = [&]() {
        __syncthreads();
        for (int i = tid; i < n; i += nt) {
            double a = 0.0;
            for (int k = i; k < n; ++k) a += Lc[k + (size_t)n * i] * prv[k];
            pv[i] = a;
        }
        __syncthreads();
    };

    if (tid == 0) {
        state[0] = 0;
        state[1] = 0;
    }
    __syncthreads();
    const int n_initial = sd.seeded ? 1 : (ms < 2 ? ms : 2);
    for (int c = 0; c < n_initial; ++c) {
        if (sd.seeded) {
            for (int i = tid; i < n; i += nt) prv[i] = to_global(sd.coef)[i];
        } else if (c == 0) {
            // argmin_j H_jj / S_jj (ties: lowest index), by the first wavefront
            if (tid < 64) {
                double best = 1e300;
                int at = 0;
                for (int j = tid; j < n; j += 64) {
                    const double r = H[j + (size_t)n * j] / S[j + (size_t)n * j];
                    if (r < best) {
                        best = r;
                        at = j;
                    }
                }
                const double lowest = -dav_wave_max(-best);
                int cand = best == lowest ? at : 0x7fffffff;
                for (int o = 32; o >= 1; o >>= 1) cand = min(cand, __shfl_xor(cand, o));
                if (tid == 0) state[1] = cand;
            }
            __syncthreads();
            const int jmin = state[1];
            for (int i = tid; i < n; i += nt) prv[i] = i == jmin ? 1.0 : 0.0;
        } else {
            if (tid == 0) {
                unsigned long long st = 0x9e3779b97f4a7c15ULL;
                for (int i = 0; i < n; ++i) {
                    st ^= st << 13;
                    st ^= st >> 7;
                    st ^= st << 17;
                    prv[i] = (st & 1ULL) == 0ULL ? -1.0 : 1.0;
                }
            }
        }
        to_factored();
        action();
        append();
    }
    if (state[0] < 1) {
        if (tid == 0) to_global(sd.flags)[FLAG_EIGEN] = 1;  // initial vectors dependent
        return;
    }
    update_projected(0);

    double lam_prev = 0.0;
    bool have_prev = false;
    const int max_iterations = 2 * n;
    for (int iteration = 1; iteration <= max_iterations; ++iteration) {
        const int active = state[0];
        if (tid < 64) {
            for (int e = tid; e < MS * MS; e += 64) Aw[e] = P[e];
            __builtin_amdgcn_wave_barrier();
            dav_small_lowest(Aw, active, Vh, hb, ys, lam);
        }
        __syncthreads();
        const double lambda = lam[0];
        for (int i = tid; i < n; i += nt) {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < active; ++j) {
                a += V[i + (size_t)n * j] * ys[j];
                b += AV[i + (size_t)n * j] * ys[j];
            }
            root[i] = a;
            aroot[i] = b;
        }
        double t[2] = {0.0, 0.0};
        for (int i = tid; i < n; i += nt) {
            const double r = aroot[i] - lambda * root[i];
            t[0] += r * r;
            t[1] += root[i] * root[i];
        }
        dav_block_sums(t, 2, red);
        const double stationarity = sqrt(t[0]) / sqrt(t[1]);
        const bool stable = have_prev && fabs(lambda - lam_prev) <= prm.energy_tol;
        const bool initial_resolved = iteration == 1 && stationarity <= prm.energy_tol;
        if (stationarity <= prm.stationarity_tol && (stable || initial_resolved || active == n)) {
            // c = L^-T y / |y| (c^T S c = 1), the energy, the determinant coefficients
            const double inv = 1.0 / sqrt(t[1]);
            for (int i = tid; i < n; i += nt) {
                double a = 0.0;
                for (int k = i; k < n; ++k) a += Li[k + (size_t)n * i] * root[k];
                to_global(sd.coef)[i] = a * inv;
            }
            if (tid == 0) {
                to_global(sd.energy)[0] = lambda;
                to_global(sd.energy)[1] = lambda + sd.e_nuc;
            }
            __syncthreads();
            for (int d = tid; d < topo.n_det; d += nt) {
                double acc = 0.0;
                for (long long q = topo.off[d]; q < topo.off[d + 1]; ++q) acc += topo.tc[q] * to_global(sd.coef)[topo.ts[q]];
                to_global(sd.dcoef)[d] = acc;
            }
            return;
        }
        if (active == n) {
            if (tid == 0) to_global(sd.flags)[FLAG_EIGEN] = 2;  // complete basis cannot meet the residual
            return;
        }
        lam_prev = lambda;
        have_prev = true;
        // identity preconditioner: the correction is the residual
        for (int i = tid; i < n; i += nt) pv[i] = aroot[i] - lambda * root[i];
        __syncthreads();
        if (active + 1 > ms) {
            // restart: the basis becomes the current Ritz vector and the previous one
            for (int i = tid; i < n; i += nt) {
                double a = 0.0, b = 0.0;
                for (int j = 0; j < active; ++j) {
                    a += V[i + (size_t)n * j] * yprev[j];
                    b += AV[i + (size_t)n * j] * yprev[j];
                }
                prv[i] = a;
                aprv[i] = b;
            }
            __syncthreads();
            const double rn = 1.0 / sqrt(t[1]);
            for (int i = tid; i < n; i += nt) {
                V[i] = root[i] * rn;
                AV[i] = aroot[i] * rn;
            }
            if (tid == 0) {
                state[0] = 1;
                bnorm[0] = 1.0;
            }
            __syncthreads();
            // the previous Ritz vector, orthogonalised against the current one (the correction is
            // kept aside in prv meanwhile)
            for (int i = tid; i < n; i += nt) {
                const double c = pv[i];
                pv[i] = prv[i];
                prv[i] = c;
                ap[i] = aprv[i];
            }
            __syncthreads();
            append();
            for (int i = tid; i < n; i += nt) pv[i] = prv[i];
            __syncthreads();
            update_projected(0);
        }
        // project the correction against the basis (project_davidson_corrections)
        {
            const int act = state[0];
            double t0[1] = {0.0};
            for (int i = tid; i < n; i += nt) t0[0] += pv[i] * pv[i];
            dav_block_sums(t0, 1, red);
            const double initial_norm = sqrt(t0[0]);
            dav_project(V, n, act, pv, coef, red32);
            double terms = 0.0;
            for (int j = 0; j < act; ++j) terms += fabs(coef[j]) * bnorm[j];
            const double ops = (double)(2 * act + 1) * eps;
            const double bound = (ops < 1.0 ? ops / (1.0 - ops) : 1e300) * (initial_norm + terms);
            dav_subtract(pv, V, n, act, coef);
            __syncthreads();
            double t1[1] = {0.0};
            for (int i = tid; i < n; i += nt) t1[0] += pv[i] * pv[i];
            dav_block_sums(t1, 1, red);
            const double pn = sqrt(t1[0]);
            const double sc = pn <= bound ? 0.0 : 1.0 / pn;
            for (int i = tid; i < n; i += nt) pv[i] *= sc;
        }
        if (tid < MS) yprev[tid] = tid < active ? ys[tid] : 0.0;
        action();
        const int before = state[0];
        if (!append()) {
            if (tid == 0) to_global(sd.flags)[FLAG_EIGEN] = 3;  // correction space became linearly dependent
            return;
        }
        update_projected(before);
        if (before < active) {
            // after a restart the current Ritz vector is the first basis vector
            if (tid < MS) yprev[tid] = tid == 0 ? 1.0 : 0.0;
            __syncthreads();
        }
    }
    if (tid == 0) to_global(sd.flags)[FLAG_EIGEN] = 4;  // iteration limit
}

}  // namespace

void launch_davidson(hipStream_t st, const SlotDesc* slots, int count, const TopoDesc& topo,
                     const LaunchShape& shape) {
    // StructureSolveAccuracy::configure_physical_stopping
    // (the caller's stationarity tolerance can only tighten the square-root bound)
    const double etol = std::min(shape.dav_energy_tol, 1.0e-8);
    const DavidsonParams prm{etol, std::min(std::sqrt(etol), shape.dav_stationarity_tol)};
    launch_gep_front(st, slots, (unsigned)count, topo, shape.pinned || count <= 8);
    hipLaunchKernelGGL(k_davidson_factored, dim3(1, 1, count), dim3(kDavThreads), 0, st, slots, topo, prm);
}

}  // namespace xmvb_amd
