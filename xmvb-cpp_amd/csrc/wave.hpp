// Wavefront-level helpers (64-lane broadcasts, DPP reductions, a fast reciprocal divide) and the
// one-wavefront lowest eigenpair of a symmetric tridiagonal matrix, shared by the eigensolver
// kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xmvb_amd {

// wave-uniform broadcast of lane `src`
__device__ __forceinline__ double wave_bcast(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// returns `old` with lane `dst` replaced by the wave-uniform value v (one compare and select)
__device__ __forceinline__ double wave_put(double v, int dst, double old) {
    return (int)(threadIdx.x & 63) == dst ? v : old;
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

// a / b to a few ulp: v_rcp_f64 refined by two Newton steps. Used where only a small relative
// error matters (Sturm counts, the pivots of a tridiagonal LU), not for results.
__device__ __forceinline__ double fast_div(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return a * r;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Sturm sequence of T - x I in product form, p_0 = 1, p_1 = d_0 - x,
// p_{i+1} = (d_i - x) p_i - e_{i-1}^2 p_{i-1}: its sign changes count the eigenvalues below x, and
// q_i = p_{i+1} / p_i are the pivots of the LDL^T factorisation of T - x I. The chain costs one
// multiply and one FMA per row (the ratio form has a division on it); the pair (p_i, p_{i-1}) is
// rescaled by a power of two every eight rows, which leaves ratios and signs unchanged. An exact zero
// is replaced by a tiny value of the sign opposite to its predecessor, i.e. counted as a negative
// pivot, as the ratio form does (q = 0 -> -tiny). de: wave-uniform LDS pairs (d_i, e_{i-1}^2) for
// rows 0..n+2, rows from n on zero; the rows are read four at a time, and rows beyond n - 1 enter
// the chain as (d - x, e^2) = (1, 0), which changes nothing, so the loop has no per-row branch.
// Returns the count; with WithPivots lane i receives q_i.
template <bool WithPivots>
__device__ __forceinline__ int sturm_count(int ln, int n, const double2* de, double x, double& q) {
    double pm = 1.0, pc = de[0].x - x;
    if (pc == 0.0) pc = -1e-300;
    int count = pc < 0.0;
    double cap_n = pc, cap_c = 1.0;  // (p_{i+1}, p_i) of the lane's own row
    for (int i = 1; i < n; i += 4) {
        double2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = de[i + t];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double dv = i + t < n ? v[t].x - x : 1.0;
            double pn = dv * pc - v[t].y * pm;
            if (pn == 0.0) pn = pc < 0.0 ? 1e-300 : -1e-300;
            count += (pn < 0.0) != (pc < 0.0);
            if (WithPivots && ln == i + t) {
                cap_n = pn;
                cap_c = pc;
            }
            pm = pc;
            pc = pn;
        }
        if ((i & 7) == 5) {  // every eight rows
            const int ex = ilogb(fmax(fabs(pc), fabs(pm)));
            pc = ldexp(pc, -ex);
            pm = ldexp(pm, -ex);
        }
    }
    if (WithPivots) q = cap_n / cap_c;
    return count;
}

// Lowest eigenpair of the symmetric tridiagonal T of order n <= 64 (the last stage of
// GeneralizedEigensolver::solve_dense, linalg/eigensolver.cpp:365-387, after the Householder
// reduction), on one wavefront with no block barrier: lane i holds row i (d_r = T[i,i], e_r =
// T[i+1,i], zero beyond n). Multisection on Sturm counts (64 points per step) brackets the lowest
// eigenvalue in [lo, hi] to a relative width of 1e-10 with no eigenvalue below lo; T - lo I is then
// positive definite, so its LDL^T factorisation needs no pivoting (its pivots are the Sturm ratios
// at lo, every one positive), and two steps of inverse iteration with it give the eigenvector
// (each step: forward and backward substitution, one FMA chain each, on wave-uniform row pairs
// read from LDS four rows at a time). The eigenvalue is the Rayleigh quotient of that vector.
// t: 288 doubles of LDS, 16-byte aligned. Returns the eigenvalue (wave-uniform); yv receives the
// lane's component of the unit eigenvector.
__device__ inline double tridiag_lowest_wave(int ln, int n, double d_r, double e_r, double* t, double& yv) {
    double2* de = reinterpret_cast<double2*>(t);        // rows 0..71: (d_i, e_{i-1}^2)
    double2* sv = reinterpret_cast<double2*>(t) + 72;   // rows -4..67 at 4..71: solve operands
    const bool row = ln < n;
    const double e_dn = __shfl_up(e_r, 1);  // T[ln, ln-1]
    const double e2_r = ln > 0 && row ? e_dn * e_dn : 0.0;
    de[ln] = make_double2(row ? d_r : 0.0, e2_r);
    if (ln < 8) {
        de[64 + ln] = make_double2(0.0, 0.0);
        sv[ln < 4 ? ln : 64 + ln] = make_double2(0.0, 0.0);
    }
    const double e_prev = ln > 0 && row ? fabs(e_dn) : 0.0;
    const double rad = e_prev + fabs(e_r);
    double lo = -wave_max(row ? rad - d_r : -1e300);
    double hi = wave_max(row ? d_r + rad : -1e300);
    const double pad = 1e-12 * fmax(fabs(lo), fabs(hi)) + 1e-300;
    lo -= pad;
    hi += pad;
    wave_sync();
    double unused = 0.0;
    for (int it = 0; it < 64; ++it) {
        if (!(hi - lo > 1e-10 * fmax(fabs(lo), fabs(hi)))) break;
        const double x = lo + (hi - lo) * (double)(ln + 1) / 64.0;
        const int count = sturm_count<false>(ln, n, de, x, unused);
        const uint64_t hit = __ballot(count >= 1);
        const int f = hit ? __builtin_ctzll(hit) : 64;
        const double nlo = f == 0 ? lo : lo + (hi - lo) * (double)f / 64.0;
        const double nhi = f >= 64 ? hi : lo + (hi - lo) * (double)(f + 1) / 64.0;
        lo = nlo;
        hi = nhi;
    }
    // T - lo I = L D L^T: D = q (lane i: q_i > 0), L[i+1, i] = e_i / q_i
    double q = 1.0;
    (void)sturm_count<true>(ln, n, de, lo, q);
    if (!(q > 0.0)) q = 1e-300;  // rounding at a bracket end that touches an eigenvalue
    const double rq = row ? 1.0 / q : 0.0;
    const double l_r = ln + 1 < n ? e_r * rq : 0.0;
    const double l_up = __shfl_up(l_r, 1);
    const double l_in = ln > 0 ? l_up : 0.0;  // l_{i-1}
    double y = row ? 1.0 : 0.0;
    for (int iter = 0; iter < 2; ++iter) {
        // forward: z_i = b_i - l_{i-1} z_{i-1}; rows beyond n are (0, 0)
        sv[4 + ln] = make_double2(y, l_in);
        wave_sync();
        double z = 0.0, cur = 0.0;
        for (int i = 0; i < n; i += 4) {
            double2 v[4];
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) v[t4] = sv[4 + i + t4];
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                cur = v[t4].x - v[t4].y * cur;
                if (ln == i + t4) z = cur;
            }
        }
        wave_sync();
        // backward: y_i = z_i / q_i - l_i y_{i+1}; rows below 0 are (0, 0)
        sv[4 + ln] = make_double2(z * rq, l_r);
        wave_sync();
        cur = 0.0;
        for (int i = n - 1; i >= 0; i -= 4) {
            double2 v[4];
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) v[t4] = sv[4 + i - t4];
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                cur = v[t4].x - v[t4].y * cur;
                if (ln == i - t4) y = cur;
            }
        }
        wave_sync();
        y *= 1.0 / sqrt(wave_sum(y * y));
    }
    // Rayleigh quotient y^T T y: (T y)_i = e_{i-1} y_{i-1} + d_i y_i + e_i y_{i+1}
    const double y_dn = __shfl_up(y, 1), y_up = __shfl_down(y, 1);
    const double ty = row ? d_r * y + (ln > 0 ? e_dn * y_dn : 0.0) + (ln + 1 < n ? e_r * y_up : 0.0) : 0.0;
    yv = y;
    return wave_sum(y * ty);
}

}  // namespace xmvb_amd
