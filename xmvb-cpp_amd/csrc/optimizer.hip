// Projected L-BFGS orbital optimisation (run_plbfgs_backend, src/vbscf/optimization/backends/
// plbfgs.cpp:20-226) with device-resident orbitals, gradients and secant history.
//
// Chart. The energy is invariant under rescaling each orbital, so the optimisation variables of
// orbital i are the tangent directions t_i on its support with c_i^T S t_i = 0 (S restricted to
// the support, c_i normalised to c_i^T S c_i = 1), measured in the S metric: reduced coordinates
// are the coefficients of t_i in any S-orthonormal basis of that tangent space. This is the
// whitened product-sphere chart of OrbitalChart (orbitals/charts/chart.cpp:560-660 with
// whitening.cpp); L-BFGS with a scaled-identity initial inverse is invariant under orthogonal
// changes of reduced basis, so the iterates do not depend on which S-orthonormal basis is used
// and none is formed. In raw coefficients the reduced-space algebra becomes:
//   reduced gradient      <->  covector r_i = g_i - (S c_i)(c_i^T g_i)      (projected gradient)
//   reduced step          <->  vector   s_i - c_i ((S c_i)^T s_i)           (S-orthogonal projection)
//   reduced inner product <->  vector . covector (plain pairing); |v|^2 = v^T S v; |w|^2 = w^T S^-1 w
//   identity inverse      <->  t_i = S_i^-1 r_i (support-restricted inverse overlap)
// Secant pairs are stored as raw packed differences (steps and gradient changes of normalised
// orbitals) and projected into the current chart when used, as transport_secant_pairs_to_chart
// does (preconditioners/transported_lbfgs.cpp:107-131).
//
// Device work per iteration: one block per orbital for the chart (k_opt_point), the two-loop
// recursion in compact form (k_dir_*: dot products of the stored pairs with the gradient and the
// preconditioned vector over many workgroups, the two scalar loops on one wavefront, two streaming
// combinations), a streaming kernel per Armijo trial (k_opt_trial) and, per accepted point, the
// secant test and history append (k_opt_append, k_opt_gram: one new row and column of the ring's
// Gram entries). The host reads scalars only: energies, the directional derivative, the
// projected-gradient norm and the step size.

#include "gmem.hpp"
#include "optimizer.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace xmvb_amd {
namespace {

constexpr int kPointThreads = 256;
constexpr int kOrbStats = 4;  // per orbital: dual norm, max |c|, norm before normalisation, c^T g

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <int Ctrl>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), Ctrl, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double lane_read(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// wavefront reductions: DPP inside 16-lane rows, then the four row results through readlane
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}

__device__ __forceinline__ double wave_largest(double v) {
    v = fmax(v, dpp_move<0xB1>(v));
    v = fmax(v, dpp_move<0x4E>(v));
    v = fmax(v, dpp_move<0x141>(v));
    v = fmax(v, dpp_move<0x140>(v));
    return fmax(fmax(lane_read(v, 0), lane_read(v, 16)), fmax(lane_read(v, 32), lane_read(v, 48)));
}

// block reductions (blockDim a multiple of 64, at most 1024); red holds >= 16 doubles
__device__ double block_total(double v, double* red) {
    v = wave_total(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w];
    return t;
}

__device__ double block_largest(double v, double* red) {
    v = wave_largest(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = red[0];
    for (int w = 1; w < nw; ++w) t = fmax(t, red[w]);
    return t;
}

__device__ __forceinline__ bool on_support(const uint8_t* support, size_t at) { return !support || support[at]; }

// Per-problem view of the optimiser's device state for one batched launch (blockIdx.z indexes the
// descriptor array). cur / prev / trial rotate on the host, so descriptors are rebuilt per launch.
struct OptDesc {
    int n_bf;
    int N;                      // n_act * n_bf
    const double* S;            // AO overlap, n_bf x n_bf
    const double* Sinv;         // support-restricted inverse overlaps, one n_bf x n_bf per pattern
    const int* pat;             // pattern of each orbital
    const uint8_t* support;     // N mask or null
    double* C;                  // current point (normalised in place by the chart)
    double* G;
    const double* Cp;           // previous point (secant append)
    const double* Gp;
    double* Ct;                 // trial point (trial kernel output)
    double* Sc;                 // S c_i on the support
    double* R;                  // projected gradient
    double* T;                  // search direction
    double* q;                  // two-loop work vectors
    double* z;
    double* apx;                // 6 per orbital: secant sums and max |s_i| (append kernels)
    double* gsy;                // Gram entries s_a,i . y_b,i of the ring: cap x cap x n_act (slots)
    double* sy;                 // SY of the pairs in use: cap x cap (ages)
    double* dws;                // direction workspace: 5 cap + 4 + 2 x combine workgroups
    double* stats;              // kOrbStats * n_act chart statistics
    double* out;                // kOutSlots direction / append scalars
    double* hs;                 // secant ring: steps, gradient changes, per-orbital statistics
    double* hy;
    double* hstat;
    double* proj;               // chart projections of the stored pairs: sig then bet, cap x n_act each
    int cap;                    // ring capacity (>= 1)
    int count;                  // pairs used by this direction (newest first)
    int newest;                 // ring slot of the newest pair
    int slot;                   // ring slot the append writes
    double step;                // trial step
};

constexpr int kOutSlots = 8;  // [0] g.t [1] max|t| [2] pairs used [3] gamma [4] bad | [5] accepted [6] max|s|

// Chart at an accepted point, one block per orbital i (grid x = orbital, z = problem): normalise
// c_i in the S metric (the gradient of the scale-invariant energy scales inversely), form S c_i on
// the support, the projected gradient r_i = g_i - (S c_i)(c_i^T g_i) and its dual norm
// sqrt(r_i^T S_i^-1 r_i). stats[4 i + 0..3] = dual norm, max |c_i|, norm before normalisation
// (<= 0: failure), c_i^T g_i.
__global__ void __launch_bounds__(kPointThreads) k_opt_point(const OptDesc* __restrict__ descs) {
    const OptDesc& o = descs[blockIdx.z];
    const int n_bf = o.n_bf;
    extern __shared__ double lds[];
    double* cs = lds;          // c_i, n_bf
    double* rs = lds + n_bf;   // r_i, n_bf
    __shared__ double red[16];
    const int i = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const size_t row = (size_t)i * n_bf;
    const uint8_t* support = o.support;
    double* __restrict__ C = o.C;
    double* __restrict__ G = o.G;
    double* __restrict__ Sc = o.Sc;
    for (int k = tid; k < n_bf; k += nt) cs[k] = on_support(support, row + k) ? C[row + k] : 0.0;
    __syncthreads();
    double part = 0.0;
    const auto Sg = to_global(o.S);
    for (int k = tid; k < n_bf; k += nt) {
        double a = 0.0;
        // unconditional, unrolled: the S column loads of several j in flight (off the support the
        // result is discarded by the select below)
#pragma unroll 8
        for (int j = 0; j < n_bf; ++j) a += Sg[k + (size_t)n_bf * j] * cs[j];
        a = on_support(support, row + k) ? a : 0.0;
        Sc[row + k] = a;
        part += cs[k] * a;
    }
    const double rho2 = block_total(part, red);
    if (!(rho2 > 0.0) || !isfinite(rho2)) {
        if (tid == 0) o.stats[kOrbStats * i + 2] = -1.0;
        return;  // uniform
    }
    const double rho = sqrt(rho2), inv = 1.0 / rho;
    double bpart = 0.0, cmax = 0.0;
    for (int k = tid; k < n_bf; k += nt) {
        const double c = cs[k] * inv;
        const double g = on_support(support, row + k) ? G[row + k] * rho : 0.0;
        C[row + k] = c;
        G[row + k] = g;
        Sc[row + k] *= inv;
        bpart += c * g;
        cmax = fmax(cmax, fabs(c));
    }
    const double beta = block_total(bpart, red);
    const double cmax_all = block_largest(cmax, red);
    for (int k = tid; k < n_bf; k += nt) {
        const double r = G[row + k] - Sc[row + k] * beta;
        o.R[row + k] = r;
        rs[k] = r;
    }
    __syncthreads();
    const auto P = to_global(o.Sinv) + (size_t)to_global(o.pat)[i] * n_bf * n_bf;
    double npart = 0.0;
    for (int k = tid; k < n_bf; k += nt) {
        double a = 0.0;
#pragma unroll 8
        for (int j = 0; j < n_bf; ++j) a += P[k + (size_t)n_bf * j] * rs[j];
        npart += rs[k] * a;
    }
    const double nu2 = block_total(npart, red);
    if (tid == 0) {
        o.stats[kOrbStats * i + 0] = sqrt(fmax(nu2, 0.0));
        o.stats[kOrbStats * i + 1] = cmax_all;
        o.stats[kOrbStats * i + 2] = rho;
        o.stats[kOrbStats * i + 3] = beta;
    }
}

// Chart projections of the stored pairs for the two-loop: sig[a,i] = (S c_i)^T s_a,i and
// bet[a,i] = c_i^T y_a,i (a = age, newest first), one wavefront per dot product over the orbital's
// n_bf coefficients, spread over many blocks (grid x over tasks, y = which, z = problem).
__global__ void __launch_bounds__(256) k_opt_project(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    const int task = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (task >= o.count * n_act) return;  // wave-uniform
    const int which = blockIdx.y, a = task / n_act, i = task % n_act, n_bf = o.n_bf;
    const int slot = (o.newest - a + o.cap) % o.cap;
    const double* v = (which == 0 ? o.hs : o.hy) + (size_t)slot * o.N + (size_t)i * n_bf;
    const double* w = (which == 0 ? o.Sc : o.C) + (size_t)i * n_bf;
    double part = 0.0;
    for (int k = lane; k < n_bf; k += 64) part += v[k] * w[k];
    part = wave_total(part);
    if (lane == 0) o.proj[(size_t)which * o.cap * n_act + task] = part;
}

// Search direction t = -H r of the two-loop recursion (TransportedReducedLbfgsPreconditioner::
// apply, transported_lbfgs.cpp:41-86) over the secant pairs projected into the current chart,
// with the scaled-identity initial inverse gamma S^-1 (gamma = s'y'/y'y' of the newest kept pair)
// and the try_add_pair curvature filter (transported_lbfgs.cpp:14-32); pairs are addressed by
// age (0 = newest) in the ring. The recursion is evaluated in compact form, which gives the same
// coefficients without a reduction over the full vectors per pair: with s^_a, y^_a the projected
// pairs (s^ = s - c (S c)^T s, y^ = y - (S c) c^T y per orbital) and SY[a][b] = s^_a . y^_b,
//   first loop   alpha_a = rho_a (s^_a . r - sum_{b newer} alpha_b SY[a][b]),  q = r - sum_b alpha_b y^_b
//   z0           = gamma S_i^-1 q_i
//   second loop  coef_a = alpha_a - rho_a (y^_a . z0 + sum_{c older} coef_c SY[c][a])
//   t            = -(z0 + sum_c coef_c s^_c)
// so the vector work is four streaming passes over the stored pairs (s^ . r, q, y^ . z0, t) and
// the two loops run on scalars. SY comes from chart-independent per-orbital Gram entries
// Gamma[a][b][i] = s_a,i . y_b,i kept with the ring (k_opt_gram, one new row and column per
// appended pair): s^_a,i . y^_b,i = Gamma[a][b][i] - sig_a,i bet_b,i, because (S c_i)^T c_i = 1.
// One block per problem.
// The direction runs as ten small launches over many workgroups (a single-problem direction
// otherwise occupies one CU and waits on memory latency for every pair).
constexpr int kPairsMax = 256;   // secant pairs the recursions hold (four per lane)
constexpr int kCombineElems = 64;  // elements per combine workgroup (x 4 pair chunks = 256 threads)

// per-problem direction workspace (o.dws): rho, gam, alp, cof, dts by age (cap each), then the
// scalars [gamma, used, bad] and the final kernel's per-workgroup partials (dd, tinf)
__device__ __forceinline__ size_t dws_scalars(int cap) { return (size_t)5 * cap; }
__device__ __forceinline__ size_t dws_partials(int cap) { return (size_t)5 * cap + 4; }

__device__ __forceinline__ double sig_of(const OptDesc& o, int n_act, int a, int i) {
    return to_global(static_cast<const double*>(o.proj))[(size_t)a * n_act + i];
}
__device__ __forceinline__ double bet_of(const OptDesc& o, int n_act, int a, int i) {
    return to_global(static_cast<const double*>(o.proj))[(size_t)o.cap * n_act + (size_t)a * n_act + i];
}

// rho_a = 1 / (s^_a . y^_a) and gam_a = s^'y^ / y^'y^ of the pairs that pass the curvature test
// (0 otherwise), gamma of the newest one kept, the count of pairs kept. One workgroup per problem.
__global__ void __launch_bounds__(256) k_dir_prep(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    const int cap = o.cap, count = o.count, newest = o.newest;
    const auto hstat = to_global(static_cast<const double*>(o.hstat));
    const auto w = to_global(o.dws);
    __shared__ double rho_s[kPairsMax], gam_s[kPairsMax];
    for (int a = threadIdx.x; a < count; a += blockDim.x) {
        const int slot = (newest - a + cap) % cap;
        double ss = 0.0, yy = 0.0, sy = 0.0;
        for (int i = 0; i < n_act; ++i) {
            const auto st = hstat + ((size_t)slot * n_act + i) * 3;
            const double s_ = sig_of(o, n_act, a, i), b_ = bet_of(o, n_act, a, i);
            ss += st[0] - s_ * s_;
            yy += st[1] - b_ * b_;
            sy += st[2] - s_ * b_;
        }
        const double sn = sqrt(fmax(ss, 0.0)), yn = sqrt(fmax(yy, 0.0));
        const bool ok = sn > 0.0 && yn > 0.0 && isfinite(sn) && isfinite(yn) && isfinite(sy) &&
                        sy > 1.4901161193847656e-08 * sn * yn;  // sqrt(DBL_EPSILON)
        rho_s[a] = ok ? 1.0 / sy : 0.0;
        gam_s[a] = ok ? sy / yy : 0.0;
        w[a] = rho_s[a];
        w[cap + a] = gam_s[a];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int used = 0;
        double gamma = 1.0;
        for (int a = 0; a < count; ++a)
            if (rho_s[a] != 0.0) {
                if (used == 0) gamma = gam_s[a];
                ++used;
            }
        w[dws_scalars(cap) + 0] = gamma;
        w[dws_scalars(cap) + 1] = used;
        w[dws_scalars(cap) + 2] = 0.0;  // bad
    }
}

// dts[a] = s^_a . r (Mode 0) or y^_a . z0 (Mode 1); one workgroup per pair (grid x = pair)
template <int Mode>
__global__ void __launch_bounds__(256) k_dir_dots(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    const int a = blockIdx.x, cap = o.cap, N = o.N, n_bf = o.n_bf;
    if (a >= o.count) return;  // block-uniform
    const auto w = to_global(o.dws);
    __shared__ double red[16];
    double part = 0.0;
    if (w[a] != 0.0) {  // rho: block-uniform
        const int slot = (o.newest - a + cap) % cap;
        if (Mode == 0) {
            const auto s = to_global(static_cast<const double*>(o.hs)) + (size_t)slot * N;
            const auto C = to_global(static_cast<const double*>(o.C));
            const auto R = to_global(static_cast<const double*>(o.R));
            for (int e = threadIdx.x; e < N; e += blockDim.x)
                part += (s[e] - C[e] * sig_of(o, n_act, a, e / n_bf)) * R[e];
        } else {
            const auto y = to_global(static_cast<const double*>(o.hy)) + (size_t)slot * N;
            const auto Sc = to_global(static_cast<const double*>(o.Sc));
            const auto z = to_global(static_cast<const double*>(o.z));
            for (int e = threadIdx.x; e < N; e += blockDim.x)
                part += (y[e] - Sc[e] * bet_of(o, n_act, a, e / n_bf)) * z[e];
        }
    }
    part = block_total(part, red);
    if (threadIdx.x == 0) w[4 * (size_t)cap + a] = part;
}

// SY[a][b] = s^_a . y^_b = sum_i Gamma[a][b][i] - sig_a,i bet_b,i for the pairs kept (0 otherwise)
__global__ void __launch_bounds__(256) k_dir_sy(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    const int cap = o.cap, count = o.count, newest = o.newest;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * count) return;
    const int a = t / count, b = t - a * count;
    const auto w = to_global(static_cast<const double*>(o.dws));
    double v = 0.0;
    if (w[a] != 0.0 && w[b] != 0.0) {
        const auto g = to_global(static_cast<const double*>(o.gsy)) +
                       ((size_t)((newest - a + cap) % cap) * cap + (newest - b + cap) % cap) * n_act;
        for (int i = 0; i < n_act; ++i) v += g[i] - sig_of(o, n_act, a, i) * bet_of(o, n_act, b, i);
    }
    to_global(o.sy)[(size_t)a * cap + b] = v;
}

// The two loops on scalars, one wavefront per problem: lane l owns pairs a = l + 64 r (registers),
// and a pair's accumulator is broadcast by readlane when its turn comes (column by column).
//   Mode 0 (newest to oldest): alpha_b = rho_b acc_b, acc_a -= SY[a][b] alpha_b for a older
//   Mode 1 (oldest to newest): coef_d = alpha_d - rho_d acc_d, acc_a += SY[d][a] coef_d for a newer
// Staged: SY copied to LDS first (count <= kSyLdsMax), row a at a * count.
constexpr int kSyLdsMax = 128;

template <int Mode, bool Staged>
__global__ void __launch_bounds__(64) k_dir_rec(const OptDesc* __restrict__ descs) {
    const OptDesc& o = descs[blockIdx.z];
    const int cap = o.cap, count = o.count, lane = threadIdx.x;
    const auto w = to_global(o.dws);
    const auto syg = to_global(static_cast<const double*>(o.sy));
    extern __shared__ double syl[];
    __shared__ double rho_l[kPairsMax], alp_l[kPairsMax];  // per-step scalars from LDS, not global loads
    for (int t = lane; t < count; t += 64) {
        rho_l[t] = w[t];
        if (Mode == 1) alp_l[t] = w[2 * (size_t)cap + t];
    }
    if (Staged) {  // eight loads per lane in flight (indices past the end clamped), then the stores
        const int total = count * count;
        for (int t0 = 0; t0 < total; t0 += 64 * 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = min(t0 + u * 64 + lane, total - 1), a = t / count, b = t - a * count;
                v[u] = syg[(size_t)a * cap + b];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (t0 + u * 64 + lane < total) syl[t0 + u * 64 + lane] = v[u];
        }
    }
    __syncthreads();
    auto sy_at = [&](int a, int b) { return Staged ? syl[a * count + b] : syg[(size_t)a * cap + b]; };
    const auto alp = w + 2 * (size_t)cap, cof = w + 3 * (size_t)cap, dts = w + 4 * (size_t)cap;
    double acc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = lane + 64 * r < count ? dts[lane + 64 * r] : 0.0;
    bool bad = false;
    auto take = [&](int b) {  // acc of pair b, wave-uniform
        const int r = b >> 6, l = b & 63;
        const double v = r == 0 ? acc[0] : r == 1 ? acc[1] : r == 2 ? acc[2] : acc[3];
        return lane_read(v, l);
    };
    if (Mode == 0) {
        for (int b = 0; b < count; ++b) {
            const double ab = rho_l[b] * take(b);
            bad |= !isfinite(ab);
            if (lane == 0) alp[b] = ab;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = lane + 64 * r;
                if (a > b && a < count) acc[r] -= sy_at(a, b) * ab;
            }
        }
    } else {
        for (int d = count - 1; d >= 0; --d) {
            const double bd = rho_l[d] * take(d);
            const double c = alp_l[d] - bd;
            bad |= !isfinite(bd);
            if (lane == 0) cof[d] = c;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = lane + 64 * r;
                if (a < d) acc[r] += sy_at(d, a) * c;
            }
        }
    }
    if (__ballot(bad) != 0 && lane == 0) w[dws_scalars(cap) + 2] = 1.0;
}

// Mode 0: q = r - sum_b alpha_b y^_b; Mode 1: t = -(z0 + sum_c coef_c s^_c) with per-workgroup
// partials of g.t and max |t|. A workgroup owns kCombineElems elements; its four wavefronts take
// interleaved quarters of the pairs, whose sums are added in wavefront order.
template <int Mode>
__global__ void __launch_bounds__(256) k_dir_combine(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    const int cap = o.cap, count = o.count, newest = o.newest, N = o.N, n_bf = o.n_bf;
    if (blockIdx.x * kCombineElems >= N) return;  // block-uniform: past this problem's elements
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = blockIdx.x * kCombineElems + lane;
    const int e_c = e < N ? e : N - 1;
    const int i = e_c / n_bf;
    const auto w = to_global(o.dws);
    const auto coefs = Mode == 0 ? w + 2 * (size_t)cap : w + 3 * (size_t)cap;
    const auto vec = to_global(static_cast<const double*>(Mode == 0 ? o.hy : o.hs));
    const double base = Mode == 0 ? to_global(static_cast<const double*>(o.Sc))[e_c]
                                  : to_global(static_cast<const double*>(o.C))[e_c];
    __shared__ double part[4][kCombineElems];
    double v = 0.0;
    for (int b = wave; b < count; b += 4) {
        const double cb = coefs[b];
        const double pr = Mode == 0 ? bet_of(o, n_act, b, i) : sig_of(o, n_act, b, i);
        v += cb * (vec[(size_t)((newest - b + cap) % cap) * N + e_c] - base * pr);
    }
    part[wave][lane] = v;
    __syncthreads();
    if (wave != 0) return;
    const double sum = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    double dd = 0.0, tinf = 0.0;
    if (Mode == 0) {
        if (e < N) to_global(o.q)[e] = to_global(static_cast<const double*>(o.R))[e] - sum;
    } else if (e < N) {
        const double t = -(to_global(static_cast<const double*>(o.z))[e] + sum);
        to_global(o.T)[e] = t;
        dd = to_global(static_cast<const double*>(o.G))[e] * t;
        tinf = fabs(t);
    }
    if (Mode == 1) {
        dd = wave_total(dd);
        tinf = wave_largest(tinf);
        if (lane == 0) {
            w[dws_partials(cap) + 2 * (size_t)blockIdx.x] = dd;
            w[dws_partials(cap) + 2 * (size_t)blockIdx.x + 1] = tinf;
        }
    }
}

// z0 = gamma S_i^-1 q_i: one workgroup per orbital (grid x), q_i staged in LDS, one thread per
// coefficient with the S_i^-1 column loads unrolled so that several are in flight
__global__ void __launch_bounds__(256) k_dir_precond(const OptDesc* __restrict__ descs) {
    const OptDesc& o = descs[blockIdx.z];
    const int n_bf = o.n_bf, i = blockIdx.x;
    if (i * n_bf >= o.N) return;  // block-uniform: past this problem's orbitals
    extern __shared__ double ql[];
    const auto qi = to_global(static_cast<const double*>(o.q)) + (size_t)i * n_bf;
    for (int j = threadIdx.x; j < n_bf; j += blockDim.x) ql[j] = qi[j];
    __syncthreads();
    const auto P = to_global(o.Sinv) + (size_t)to_global(o.pat)[i] * n_bf * n_bf;
    const double gamma = to_global(static_cast<const double*>(o.dws))[dws_scalars(o.cap)];
    for (int k = threadIdx.x; k < n_bf; k += blockDim.x) {
        double acc = 0.0;
#pragma unroll 8
        for (int j = 0; j < n_bf; ++j) acc += P[k + (size_t)n_bf * j] * ql[j];
        to_global(o.z)[(size_t)i * n_bf + k] = gamma * acc;
    }
}

// out[0..4] = g.t, max |t|, pairs used, gamma, non-finite flag
__global__ void __launch_bounds__(64) k_dir_out(const OptDesc* __restrict__ descs) {
    const OptDesc& o = descs[blockIdx.z];
    if (threadIdx.x != 0) return;
    const auto w = to_global(static_cast<const double*>(o.dws));
    const int blocks = (o.N + kCombineElems - 1) / kCombineElems;
    double dd = 0.0, tinf = 0.0;
    for (int b = 0; b < blocks; ++b) {
        dd += w[dws_partials(o.cap) + 2 * (size_t)b];
        tinf = fmax(tinf, w[dws_partials(o.cap) + 2 * (size_t)b + 1]);
    }
    const auto out = to_global(o.out);
    out[0] = dd;
    out[1] = tinf;
    out[2] = w[dws_scalars(o.cap) + 1];
    out[3] = w[dws_scalars(o.cap)];
    out[4] = w[dws_scalars(o.cap) + 2];
}

size_t dws_doubles(int cap, int N) { return (size_t)5 * cap + 4 + 2 * (size_t)((N + kCombineElems - 1) / kCombineElems); }

// Gram entries of a new pair (append time): Gamma[n][b][i] = s_n,i . y_b,i and
// Gamma[b][n][i] = s_b,i . y_n,i for the pairs b still kept after the append, and Gamma[n][n][i].
// Grid x = kept pairs + 1 (the last block: the pair with itself), one wavefront per orbital.
__global__ void __launch_bounds__(256) k_opt_gram(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    const int cap = o.cap, N = o.N, n_bf = o.n_bf;
    const int kept = min(o.count, cap - 1), x = blockIdx.x;
    if (x > kept) return;  // block-uniform
    const int sn = o.slot, sb = x == kept ? sn : (o.newest - x + cap) % cap;
    const auto hs = to_global(static_cast<const double*>(o.hs));
    const auto hy = to_global(static_cast<const double*>(o.hy));
    const auto gsy = to_global(o.gsy);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = wave; i < n_act; i += nw) {
        const auto s_n = hs + (size_t)sn * N + (size_t)i * n_bf, y_n = hy + (size_t)sn * N + (size_t)i * n_bf;
        const auto s_b = hs + (size_t)sb * N + (size_t)i * n_bf, y_b = hy + (size_t)sb * N + (size_t)i * n_bf;
        double d1 = 0.0, d2 = 0.0;
        for (int k = lane; k < n_bf; k += 64) {
            d1 += s_n[k] * y_b[k];
            d2 += s_b[k] * y_n[k];
        }
        d1 = wave_total(d1);
        d2 = wave_total(d2);
        if (lane == 0) {
            gsy[((size_t)sn * cap + sb) * n_act + i] = d1;
            gsy[((size_t)sb * cap + sn) * n_act + i] = d2;
        }
    }
}

// Armijo trial points C + step T (grid x over elements, z = problem).
__global__ void k_opt_trial(const OptDesc* __restrict__ descs) {
    const OptDesc& o = descs[blockIdx.z];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < o.N) o.Ct[e] = o.C[e] + o.step * o.T[e];
}

// Candidate secant pair (append_projected_secant_pair, transported_lbfgs.cpp:147-176) between the
// previous and the new accepted point (both normalised): s = C - Cp, y = G - Gp are written to
// the ring slot with their per-orbital chart-independent statistics s^T S s, y^T S^-1 y, s^T y;
// the curvature test runs in the new chart. One wavefront per orbital (grid x = orbital, z =
// problem); the per-orbital sums and max |s_i| go to apx (6 per orbital) for k_opt_append_finish,
// which forms out[5] = 1 if accepted and out[6] = max |s|.
__global__ void __launch_bounds__(64) k_opt_append(const OptDesc* __restrict__ descs) {
    const OptDesc& o = descs[blockIdx.z];
    const int n_bf = o.n_bf, N = o.N, i = blockIdx.x, lane = threadIdx.x;
    const size_t row = (size_t)i * n_bf;
    double* __restrict__ s = o.hs + (size_t)o.slot * N + row;
    double* __restrict__ y = o.hy + (size_t)o.slot * N + row;
    double* __restrict__ hstat = o.hstat + (size_t)o.slot * (N / n_bf) * 3;
    extern __shared__ double lds[];
    double* sl = lds;         // s_i, n_bf
    double* yl = lds + n_bf;  // y_i, n_bf
    double sinf = 0.0;
    for (int k = lane; k < n_bf; k += 64) {
        const double sk = o.C[row + k] - o.Cp[row + k], yk = o.G[row + k] - o.Gp[row + k];
        s[k] = sk;
        y[k] = yk;
        sl[k] = sk;
        yl[k] = yk;
        sinf = fmax(sinf, fabs(sk));
    }
    sinf = wave_largest(sinf);
    __syncthreads();  // publishes sl and yl to the wavefront
    const double* __restrict__ P = o.Sinv + (size_t)o.pat[i] * n_bf * n_bf;
    const double* __restrict__ S = o.S;
    double sss = 0.0, ysy = 0.0, sy = 0.0, sg = 0.0, bt = 0.0;
    for (int k = lane; k < n_bf; k += 64) {
        double a = 0.0, b = 0.0;
        // the operands from LDS: the S and S^-1 loads of several j are in flight at once
#pragma unroll 8
        for (int j = 0; j < n_bf; ++j) {
            a += S[k + (size_t)n_bf * j] * sl[j];
            b += P[k + (size_t)n_bf * j] * yl[j];
        }
        sss += sl[k] * a;
        ysy += yl[k] * b;
        sy += sl[k] * yl[k];
        sg += o.Sc[row + k] * sl[k];
        bt += o.C[row + k] * yl[k];
    }
    sss = wave_total(sss);
    ysy = wave_total(ysy);
    sy = wave_total(sy);
    sg = wave_total(sg);
    bt = wave_total(bt);
    if (lane == 0) {
        hstat[i * 3 + 0] = sss;
        hstat[i * 3 + 1] = ysy;
        hstat[i * 3 + 2] = sy;
        double* w = o.apx + (size_t)i * 6;
        w[0] = sss;
        w[1] = ysy;
        w[2] = sy;
        w[3] = sg;
        w[4] = bt;
        w[5] = sinf;
    }
}

__global__ void __launch_bounds__(64) k_opt_append_finish(const OptDesc* __restrict__ descs, int n_act) {
    const OptDesc& o = descs[blockIdx.z];
    if (threadIdx.x != 0) return;
    double ss = 0.0, yy = 0.0, sy = 0.0, sinf = 0.0;
    for (int i = 0; i < n_act; ++i) {
        const double* w = o.apx + (size_t)i * 6;
        ss += w[0] - w[3] * w[3];
        yy += w[1] - w[4] * w[4];
        sy += w[2] - w[3] * w[4];
        sinf = fmax(sinf, w[5]);
    }
    const double sn = sqrt(fmax(ss, 0.0)), yn = sqrt(fmax(yy, 0.0));
    const bool ok = sn > 0.0 && yn > 0.0 && isfinite(sn) && isfinite(yn) && isfinite(sy) &&
                    sy > 1.4901161193847656e-08 * sn * yn;
    o.out[5] = ok ? 1.0 : 0.0;
    o.out[6] = sinf;
}

// Support-restricted inverse overlap of one support pattern, zero-padded to n_bf x n_bf
// (column-major): Cholesky of S[idx, idx], then the inverse column by column.
std::vector<double> restricted_inverse(const double* S, int n_bf, const std::vector<int>& idx) {
    const int m = (int)idx.size();
    std::vector<double> L((size_t)m * m, 0.0);  // row-major lower factor
    for (int a = 0; a < m; ++a)
        for (int b = 0; b <= a; ++b) {
            double v = S[idx[a] + (size_t)n_bf * idx[b]];
            for (int k = 0; k < b; ++k) v -= L[(size_t)a * m + k] * L[(size_t)b * m + k];
            if (a == b) {
                if (!(v > 0.0))
                    throw std::runtime_error("AO overlap restricted to an orbital support is not positive definite");
                L[(size_t)a * m + a] = std::sqrt(v);
            } else {
                L[(size_t)a * m + b] = v / L[(size_t)b * m + b];
            }
        }
    std::vector<double> out((size_t)n_bf * n_bf, 0.0), col(m);
    for (int c = 0; c < m; ++c) {
        for (int a = 0; a < m; ++a) {  // L y = e_c
            double v = a == c ? 1.0 : 0.0;
            for (int k = 0; k < a; ++k) v -= L[(size_t)a * m + k] * col[k];
            col[a] = v / L[(size_t)a * m + a];
        }
        for (int a = m - 1; a >= 0; --a) {  // L^T x = y
            double v = col[a];
            for (int k = a + 1; k < m; ++k) v -= L[(size_t)k * m + a] * col[k];
            col[a] = v / L[(size_t)a * m + a];
        }
        for (int a = 0; a < m; ++a) out[idx[a] + (size_t)n_bf * idx[c]] = col[a];
    }
    return out;
}

// One block for the device state of every problem of a batch: the caller's (hooks.device_block)
// when it provides one, else allocated here and freed on scope exit.
struct DeviceBlock {
    void* base = nullptr;
    size_t used = 0, size = 0;
    bool owned = false;
    DeviceBlock(size_t bytes, const std::function<void*(size_t)>& provider) : size(std::max<size_t>(bytes, 1)) {
        if (provider) {
            base = provider(size);
            if (!base) throw std::runtime_error("optimizer device block unavailable");
        } else {
            hip_check(hipMalloc(&base, size), "hipMalloc");
            owned = true;
        }
    }
    DeviceBlock(const DeviceBlock&) = delete;
    DeviceBlock& operator=(const DeviceBlock&) = delete;
    ~DeviceBlock() {
        if (owned && base) (void)hipFree(base);
    }
    template <typename T>
    T* take(size_t count) {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        if (used + bytes > size) throw std::logic_error("optimizer device block overflow");
        T* p = reinterpret_cast<T*>(static_cast<char*>(base) + used);
        used += bytes;
        return p;
    }
    static size_t bytes_for(size_t count, size_t elem) { return (std::max<size_t>(count, 1) * elem + 255) & ~size_t(255); }
};

// Host side of one problem: the control flow of run_plbfgs_backend (plbfgs.cpp:20-226) unrolled
// into phases, so that a batch advances all its problems with one launch per phase.
struct Problem {
    enum Phase { Direction, Trial, Rechart, Done };
    OptProblem* p = nullptr;
    int n_bf = 0, N = 0;
    long long reduced_size = 0;
    std::vector<int> pat;
    std::vector<std::vector<int>> patterns;
    // device state
    double *S = nullptr, *Sinv = nullptr, *Sc = nullptr, *R = nullptr, *T = nullptr, *q = nullptr, *z = nullptr;
    double* apx = nullptr;  // per-orbital secant sums between the two append kernels
    double *gsy = nullptr, *sy = nullptr, *dws = nullptr;  // ring Gram entries, SY, direction workspace
    double *stats = nullptr, *out = nullptr, *hs = nullptr, *hy = nullptr, *hstat = nullptr, *proj = nullptr;
    int* d_pat = nullptr;
    uint8_t* d_support = nullptr;
    double* pts[3] = {};
    double* grs[3] = {};
    int cur = 0, prev = 1, trial = 2;
    // host copies of the scalars read back after a phase
    const double* h_stats = nullptr;
    const double* h_out = nullptr;
    // control state
    Phase phase = Direction;
    int count = 0, newest = 0, slot = 0, use = 0;
    bool stall_mode = false, first_trial = false, speculate = false, with_grad = false;
    double E = 0.0, Et = 0.0, previous_energy = 0.0, reference_energy = 0.0;
    double gnorm = 0.0, cmax = 0.0, gnorm_before = 0.0, cmax_before = 0.0;
    double dd = 0.0, tinf = 0.0, step = 0.0;
    int n_trace = 0;
    vbe_optimizer_result res{};
    void record(double e) {
        if (p->trace && n_trace < p->trace_capacity) p->trace[n_trace] = e;
        ++n_trace;
    }
    void finish(int termination, bool converged) {
        res.termination = termination;
        res.converged = converged ? 1 : 0;
        phase = Done;
    }
};

}  // namespace

void optimizer_defaults(vbe_optimizer_options* o) {
    o->max_iterations = 2000;
    o->history_size = 100;
    o->gradient_tolerance = 1.0e-3;
    o->energy_tolerance = 1.0e-7;
    o->initial_step_size = 1.0;
    o->minimum_step_size = 1.0e-7;
    o->armijo_constant = 1.0e-4;
}

void run_plbfgs(const OptimizerHooks& h, std::vector<OptProblem>& problems, const vbe_optimizer_options& opt) {
    if (opt.max_iterations < 0 || opt.history_size < 0) throw std::invalid_argument("negative iteration or history count");
    if (opt.history_size > kPairsMax)
        throw std::invalid_argument("at most " + std::to_string(kPairsMax) + " secant pairs (history_size)");
    if (!(opt.gradient_tolerance >= 0.0) || !(opt.energy_tolerance >= 0.0) || !(opt.minimum_step_size > 0.0) ||
        !(opt.initial_step_size > 0.0) || !(opt.armijo_constant > 0.0 && opt.armijo_constant < 1.0))
        throw std::invalid_argument("optimizer tolerances and step sizes must be positive (Armijo constant in (0,1))");
    const int n_act = h.n_act;
    if (n_act > 16) throw std::invalid_argument("at most 16 orbitals per optimisation");
    const int np = (int)problems.size();
    if (np == 0) return;
    const bool single = np == 1 && h.energy_only_trials;
    hipStream_t st = h.stream;
    constexpr double kEps = std::numeric_limits<double>::epsilon();
    const double zero_step = 128.0 * kEps;
    const int cap = std::max(opt.history_size, 1);
    for (const void* f : {reinterpret_cast<const void*>(&k_dir_rec<0, true>), reinterpret_cast<const void*>(&k_dir_rec<1, true>)})
        hip_check(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(sizeof(double) * kSyLdsMax * kSyLdsMax)),
                  "hipFuncSetAttribute");

    // support patterns, restricted inverse overlaps (host, once per distinct S and pattern)
    std::vector<Problem> P(np);
    std::vector<std::vector<double>> sinv_host(np);
    size_t bytes = 0;
    for (int k = 0; k < np; ++k) {
        Problem& s = P[k];
        s.p = &problems[k];
        s.n_bf = s.p->n_bf;
        s.N = n_act * s.n_bf;
        s.pat.resize(n_act);
        std::map<std::vector<uint8_t>, int> pattern_of;
        for (int i = 0; i < n_act; ++i) {
            std::vector<uint8_t> mask(s.n_bf, 1);
            if (s.p->support)
                for (int b = 0; b < s.n_bf; ++b) mask[b] = s.p->support[(size_t)i * s.n_bf + b] ? 1 : 0;
            auto it = pattern_of.find(mask);
            if (it == pattern_of.end()) {
                std::vector<int> idx;
                for (int b = 0; b < s.n_bf; ++b)
                    if (mask[b]) idx.push_back(b);
                if (idx.empty()) throw std::invalid_argument("an orbital has an empty support");
                it = pattern_of.emplace(mask, (int)s.patterns.size()).first;
                s.patterns.push_back(idx);
            }
            s.pat[i] = it->second;
            s.reduced_size += (long long)s.patterns[s.pat[i]].size() - 1;
        }
        const size_t nn = (size_t)s.n_bf * s.n_bf;
        sinv_host[k].resize(s.patterns.size() * nn);
        bytes += DeviceBlock::bytes_for(nn, 8) + DeviceBlock::bytes_for(sinv_host[k].size(), 8) +
                 DeviceBlock::bytes_for(n_act, 4) + (s.p->support ? DeviceBlock::bytes_for(s.N, 1) : 0) +
                 11 * DeviceBlock::bytes_for(s.N, 8) + DeviceBlock::bytes_for((size_t)kOrbStats * n_act, 8) +
                 DeviceBlock::bytes_for(kOutSlots, 8) + 2 * DeviceBlock::bytes_for((size_t)cap * s.N, 8) +
                 DeviceBlock::bytes_for((size_t)cap * n_act * 3, 8) + DeviceBlock::bytes_for((size_t)cap * n_act * 2, 8) +
                 DeviceBlock::bytes_for((size_t)6 * n_act, 8) +
                 DeviceBlock::bytes_for((size_t)cap * cap * n_act, 8) + DeviceBlock::bytes_for((size_t)cap * cap, 8) +
                 DeviceBlock::bytes_for(dws_doubles(cap, s.N), 8);
    }
    {  // the inverses, spread over host threads
        std::vector<std::pair<int, int>> jobs;
        for (int k = 0; k < np; ++k)
            for (int t = 0; t < (int)P[k].patterns.size(); ++t) jobs.emplace_back(k, t);
        std::atomic<size_t> next{0};
        std::exception_ptr failure;
        std::mutex failure_lock;
        auto work = [&]() {
            for (size_t j; (j = next.fetch_add(1)) < jobs.size();) {
                const auto [k, t] = jobs[j];
                try {
                    const std::vector<double> inv = restricted_inverse(problems[k].S_ao, P[k].n_bf, P[k].patterns[t]);
                    std::copy(inv.begin(), inv.end(), sinv_host[k].begin() + (size_t)t * P[k].n_bf * P[k].n_bf);
                } catch (...) {
                    std::lock_guard<std::mutex> g(failure_lock);
                    if (!failure) failure = std::current_exception();
                }
            }
        };
        const int n_threads = (int)std::min<size_t>(jobs.size(), std::max(1u, std::min(8u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (int t = 1; t < n_threads; ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
        if (failure) std::rethrow_exception(failure);
    }

    const size_t scal_per = (size_t)kOrbStats * n_act + kOutSlots;
    const size_t scal_bytes = DeviceBlock::bytes_for(scal_per * np, 8), desc_bytes = DeviceBlock::bytes_for((size_t)np, sizeof(OptDesc));
    DeviceBlock dev(bytes + scal_bytes + desc_bytes, h.device_block);
    // pinned read-back of every problem's stats and out, and the descriptors' host copy: the
    // caller's block (hooks.host_block) when it provides one, else allocated here
    void* pinned = nullptr;
    bool pinned_owned = false;
    if (h.host_block) {
        pinned = h.host_block(scal_bytes + desc_bytes);
        if (!pinned) throw std::runtime_error("optimizer host block unavailable");
    } else {
        hip_check(hipHostMalloc(&pinned, scal_bytes + desc_bytes, hipHostMallocDefault), "hipHostMalloc");
        pinned_owned = true;
    }
    struct PinnedFree {
        void* a;
        bool owned;
        ~PinnedFree() {
            if (owned) (void)hipHostFree(a);
        }
    } pinned_free{pinned, pinned_owned};
    double* h_scal = static_cast<double*>(pinned);
    OptDesc* h_desc = reinterpret_cast<OptDesc*>(static_cast<char*>(pinned) + scal_bytes);
    double* d_scal = dev.take<double>(scal_per * np);
    OptDesc* d_desc = dev.take<OptDesc>(np);
    std::vector<double> c0;
    for (int k = 0; k < np; ++k) {
        Problem& s = P[k];
        const size_t nn = (size_t)s.n_bf * s.n_bf;
        s.S = dev.take<double>(nn);
        s.Sinv = dev.take<double>(sinv_host[k].size());
        s.d_pat = dev.take<int>(n_act);
        s.d_support = s.p->support ? dev.take<uint8_t>(s.N) : nullptr;
        for (int b = 0; b < 3; ++b) {
            s.pts[b] = dev.take<double>(s.N);
            s.grs[b] = dev.take<double>(s.N);
        }
        s.Sc = dev.take<double>(s.N);
        s.R = dev.take<double>(s.N);
        s.T = dev.take<double>(s.N);
        s.q = dev.take<double>(s.N);
        s.z = dev.take<double>(s.N);
        s.apx = dev.take<double>((size_t)6 * n_act);
        s.gsy = dev.take<double>((size_t)cap * cap * n_act);
        s.sy = dev.take<double>((size_t)cap * cap);
        s.dws = dev.take<double>(dws_doubles(cap, s.N));
        s.stats = d_scal + scal_per * k;
        s.out = s.stats + (size_t)kOrbStats * n_act;
        s.h_stats = h_scal + scal_per * k;
        s.h_out = s.h_stats + (size_t)kOrbStats * n_act;
        s.hs = dev.take<double>((size_t)cap * s.N);
        s.hy = dev.take<double>((size_t)cap * s.N);
        s.hstat = dev.take<double>((size_t)cap * n_act * 3);
        s.proj = dev.take<double>((size_t)cap * n_act * 2);
        s.newest = cap - 1;
        c0.assign(s.p->C, s.p->C + s.N);
        if (s.p->support)
            for (int e = 0; e < s.N; ++e)
                if (!s.p->support[e]) c0[e] = 0.0;
        hip_check(hipMemcpyAsync(s.S, s.p->S_ao, sizeof(double) * nn, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(s.Sinv, sinv_host[k].data(), sizeof(double) * sinv_host[k].size(),
                                 hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(s.d_pat, s.pat.data(), sizeof(int) * n_act, hipMemcpyHostToDevice, st), "H2D");
        if (s.d_support) hip_check(hipMemcpyAsync(s.d_support, s.p->support, s.N, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(s.pts[0], c0.data(), sizeof(double) * s.N, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");  // c0 is reused
    }

    int max_n_bf = 0, max_N = 0;
    for (auto& s : P) {
        max_n_bf = std::max(max_n_bf, s.n_bf);
        max_N = std::max(max_N, s.N);
    }
    auto describe = [&](const std::vector<int>& list) {
        for (size_t j = 0; j < list.size(); ++j) {
            Problem& s = P[list[j]];
            OptDesc d{};
            d.n_bf = s.n_bf;
            d.N = s.N;
            d.S = s.S;
            d.Sinv = s.Sinv;
            d.pat = s.d_pat;
            d.support = s.d_support;
            d.C = s.pts[s.cur];
            d.G = s.grs[s.cur];
            d.Cp = s.pts[s.prev];
            d.Gp = s.grs[s.prev];
            d.Ct = s.pts[s.trial];
            d.Sc = s.Sc;
            d.R = s.R;
            d.T = s.T;
            d.q = s.q;
            d.apx = s.apx;
            d.gsy = s.gsy;
            d.sy = s.sy;
            d.dws = s.dws;
            d.z = s.z;
            d.stats = s.stats;
            d.out = s.out;
            d.hs = s.hs;
            d.hy = s.hy;
            d.hstat = s.hstat;
            d.proj = s.proj;
            d.cap = cap;
            d.count = s.use;
            d.newest = s.newest;
            d.slot = s.slot;
            d.step = s.step;
            h_desc[j] = d;
        }
        hip_check(hipMemcpyAsync(d_desc, h_desc, sizeof(OptDesc) * list.size(), hipMemcpyHostToDevice, st), "H2D");
    };
    auto read_back = [&]() {
        hip_check(hipMemcpyAsync(h_scal, d_scal, sizeof(double) * scal_per * np, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    };
    // chart (+ optional secant append) of the listed problems at their current point
    auto chart = [&](const std::vector<int>& list, bool append) {
        if (list.empty()) return;
        describe(list);
        hipLaunchKernelGGL(k_opt_point, dim3(n_act, 1, (unsigned)list.size()), dim3(kPointThreads),
                           sizeof(double) * 2 * max_n_bf, st, d_desc);
        hip_check(hipGetLastError(), "k_opt_point");
        if (append) {
            hipLaunchKernelGGL(k_opt_append, dim3(n_act, 1, (unsigned)list.size()), dim3(64),
                               sizeof(double) * 2 * max_n_bf, st, d_desc);
            hipLaunchKernelGGL(k_opt_append_finish, dim3(1, 1, (unsigned)list.size()), dim3(64), 0, st, d_desc,
                               n_act);
            int kept = 0;
            for (int k : list) kept = std::max(kept, std::min(P[k].use, cap - 1));
            hipLaunchKernelGGL(k_opt_gram, dim3((unsigned)kept + 1, 1, (unsigned)list.size()), dim3(256), 0, st,
                               d_desc, n_act);
            hip_check(hipGetLastError(), "k_opt_append");
        }
        read_back();
        for (int k : list) {
            Problem& s = P[k];
            s.gnorm = 0.0;
            s.cmax = 0.0;
            for (int i = 0; i < n_act; ++i) {
                if (!(s.h_stats[kOrbStats * i + 2] > 0.0))
                    throw std::runtime_error("orbital " + std::to_string(i) + " has no positive norm on its support");
                s.gnorm = std::max(s.gnorm, s.h_stats[kOrbStats * i + 0]);
                s.cmax = std::max(s.cmax, s.h_stats[kOrbStats * i + 1]);
            }
        }
    };
    auto direction = [&](const std::vector<int>& list) {  // uses each problem's `use` newest pairs
        if (list.empty()) return;
        describe(list);
        int max_use = 0;
        for (int k : list) max_use = std::max(max_use, P[k].use);
        const unsigned nz = (unsigned)list.size();
        const bool staged = max_use <= kSyLdsMax;
        const size_t sy_lds = staged ? sizeof(double) * (size_t)max_use * max_use : 0;
        if (max_use > 0) {
            hipLaunchKernelGGL(k_opt_project, dim3((unsigned)((max_use * n_act + 3) / 4), 2, nz), dim3(256), 0, st,
                               d_desc, n_act);
        }
        hipLaunchKernelGGL(k_dir_prep, dim3(1, 1, nz), dim3(256), 0, st, d_desc, n_act);
        if (max_use > 0) {
            hipLaunchKernelGGL(k_dir_dots<0>, dim3((unsigned)max_use, 1, nz), dim3(256), 0, st, d_desc, n_act);
            hipLaunchKernelGGL(k_dir_sy, dim3((unsigned)((max_use * max_use + 255) / 256), 1, nz), dim3(256), 0, st,
                               d_desc, n_act);
            if (staged) hipLaunchKernelGGL((k_dir_rec<0, true>), dim3(1, 1, nz), dim3(64), sy_lds, st, d_desc);
            else hipLaunchKernelGGL((k_dir_rec<0, false>), dim3(1, 1, nz), dim3(64), 0, st, d_desc);
        }
        const unsigned cblocks = (unsigned)((max_N + kCombineElems - 1) / kCombineElems);
        hipLaunchKernelGGL(k_dir_combine<0>, dim3(cblocks, 1, nz), dim3(256), 0, st, d_desc, n_act);
        hipLaunchKernelGGL(k_dir_precond, dim3((unsigned)n_act, 1, nz), dim3(128), sizeof(double) * max_n_bf, st,
                           d_desc);
        if (max_use > 0) {
            hipLaunchKernelGGL(k_dir_dots<1>, dim3((unsigned)max_use, 1, nz), dim3(256), 0, st, d_desc, n_act);
            if (staged) hipLaunchKernelGGL((k_dir_rec<1, true>), dim3(1, 1, nz), dim3(64), sy_lds, st, d_desc);
            else hipLaunchKernelGGL((k_dir_rec<1, false>), dim3(1, 1, nz), dim3(64), 0, st, d_desc);
        }
        hipLaunchKernelGGL(k_dir_combine<1>, dim3(cblocks, 1, nz), dim3(256), 0, st, d_desc, n_act);
        hipLaunchKernelGGL(k_dir_out, dim3(1, 1, nz), dim3(64), 0, st, d_desc);
        hip_check(hipGetLastError(), "direction kernels");
        read_back();
        for (int k : list) {
            Problem& s = P[k];
            if (s.h_out[4] != 0.0)
                throw std::runtime_error("transported L-BFGS recursion produced a non-finite coefficient");
            s.dd = s.h_out[0];
            s.tinf = s.h_out[1];
        }
    };
    auto evaluate = [&](const std::vector<int>& list, bool trial_point, bool with_grad) {
        std::vector<const double*> Cs;
        std::vector<const uint8_t*> sup;
        std::vector<double*> Gs;
        for (int k : list) {
            Problem& s = P[k];
            const int b = trial_point ? s.trial : s.cur;
            Cs.push_back(s.pts[b]);
            sup.push_back(s.d_support);
            Gs.push_back(s.grs[b]);
        }
        std::vector<double> E(list.size());
        h.evaluate(list, Cs, sup, Gs, with_grad, E);
        return E;
    };
    auto descent = [](const Problem& s) { return std::isfinite(s.dd) && s.dd < 0.0; };
    auto zero_direction = [&](const Problem& s) { return s.tinf <= zero_step * std::max(1.0, s.cmax); };

    // starting point: energy, gradient and chart
    std::vector<int> all(np);
    for (int k = 0; k < np; ++k) all[k] = k;
    {
        const std::vector<double> E = evaluate(all, false, true);
        for (int k = 0; k < np; ++k) {
            if (!std::isfinite(E[k])) throw std::runtime_error("the starting orbitals cannot be evaluated");
            P[k].E = E[k];
            P[k].res.energy_evaluations = 1;
            P[k].res.gradient_evaluations = 1;
        }
        chart(all, false);
        for (auto& s : P) {
            s.record(s.E);
            s.previous_energy = s.E;
            s.res.termination = VBE_TERM_MAX_ITERATIONS;
        }
    }

    for (;;) {
        // -- search directions: the iteration's opening checks, then t = -H r with fallbacks
        std::vector<int> dir;
        for (int k = 0; k < np; ++k) {
            Problem& s = P[k];
            if (s.phase != Problem::Direction) continue;
            if (!s.stall_mode) {
                if (s.res.iterations >= opt.max_iterations) {
                    s.finish(VBE_TERM_MAX_ITERATIONS, false);
                    continue;
                }
                if (s.reduced_size == 0) {
                    s.finish(VBE_TERM_EMPTY_SPACE, false);
                    continue;
                }
                s.gnorm_before = s.gnorm;
                if (s.res.iterations == 0 && s.gnorm < opt.gradient_tolerance) {
                    s.finish(VBE_TERM_INITIAL_TOLERANCE, true);
                    continue;
                }
            }
            s.use = s.stall_mode || opt.history_size == 0 ? 0 : s.count;
            dir.push_back(k);
        }
        direction(dir);
        std::vector<int> redo;
        for (int k : dir) {
            Problem& s = P[k];
            if (s.stall_mode) continue;
            if (!descent(s) || zero_direction(s)) {  // drop the history and take steepest descent
                s.count = 0;
                if (s.use > 0) {
                    s.use = 0;
                    redo.push_back(k);
                }
            }
        }
        direction(redo);
        for (int k : dir) {
            Problem& s = P[k];
            if (s.stall_mode) {
                // Armijo steepest descent from the previous point (plbfgs.cpp:166-191)
                s.step = std::max(opt.minimum_step_size,
                                  std::min(std::min(1.0, opt.initial_step_size), 1.0 / std::max(1.0, s.gnorm_before)));
                if (!descent(s)) {
                    s.E = s.reference_energy;
                    s.finish(VBE_TERM_LINE_SEARCH_STALLED, false);
                    continue;
                }
            } else {
                if (!descent(s)) {
                    s.finish(VBE_TERM_NON_DESCENT, false);
                    continue;
                }
                s.reference_energy = s.E;
                s.cmax_before = s.cmax;
                s.step = std::max(opt.minimum_step_size, std::min(1.0, opt.initial_step_size));
            }
            s.first_trial = true;
            s.phase = Problem::Trial;
        }

        // -- one Armijo trial per problem in a line search (line_search.cpp:33-95)
        std::vector<int> tri;
        for (int k = 0; k < np; ++k) {
            Problem& s = P[k];
            if (s.phase != Problem::Trial) continue;
            while (s.step >= opt.minimum_step_size && s.step * s.tinf <= zero_step * std::max(1.0, s.cmax))
                s.step *= 0.5;
            if (s.step < opt.minimum_step_size) {
                s.speculate = false;
                if (s.stall_mode) {
                    s.E = s.reference_energy;
                    s.finish(VBE_TERM_LINE_SEARCH_STALLED, false);
                } else {
                    s.finish(VBE_TERM_LINE_SEARCH_FAILED, false);
                }
                continue;
            }
            tri.push_back(k);
        }
        bool any_active = false;
        for (auto& s : P) any_active |= s.phase != Problem::Done;
        if (!any_active) break;
        if (tri.empty()) continue;
        describe(tri);
        hipLaunchKernelGGL(k_opt_trial, dim3((max_N + 255) / 256, 1, (unsigned)tri.size()), dim3(256), 0, st, d_desc);
        hip_check(hipGetLastError(), "k_opt_trial");
        // a batch always evaluates trials with their gradient; one problem alone runs the first
        // trial of a line search forward + backward when the previous search accepted its first
        // trial, otherwise energy only, with the backward after acceptance
        const bool with_grad = !single || (P[tri[0]].first_trial && P[tri[0]].speculate);
        const std::vector<double> E = evaluate(tri, true, with_grad);
        std::vector<int> post;
        for (size_t j = 0; j < tri.size(); ++j) {
            Problem& s = P[tri[j]];
            ++s.res.energy_evaluations;
            if (with_grad) ++s.res.gradient_evaluations;
            s.Et = E[j];
            const double E0 = s.stall_mode ? s.reference_energy : s.E;
            if (std::isfinite(s.Et) && s.Et <= E0 + opt.armijo_constant * s.step * s.dd) {
                if (!with_grad) {
                    h.complete(tri[j], s.pts[s.trial], s.grs[s.trial]);
                    ++s.res.gradient_evaluations;
                }
                s.speculate = s.first_trial;
                post.push_back(tri[j]);
            } else {
                s.first_trial = false;
                s.step *= 0.5;
            }
        }

        // -- accepted points: chart, candidate secant pair, stall test, convergence
        for (int k : post) {
            Problem& s = P[k];
            std::swap(s.prev, s.cur);   // prev = the point stepped from
            std::swap(s.cur, s.trial);  // cur = the accepted trial
            s.E = s.Et;
            s.slot = (s.newest + 1) % cap;
        }
        chart(post, true);
        std::vector<int> rechart;
        for (int k : post) {
            Problem& s = P[k];
            if (!s.stall_mode) {
                const double scale = std::max(1.0, std::fabs(s.reference_energy));
                const bool no_progress = std::isfinite(s.reference_energy) && std::isfinite(s.E) &&
                                         std::fabs(s.E - s.reference_energy) <= 4096.0 * kEps * scale &&
                                         !(s.gnorm_before < opt.gradient_tolerance) && s.gnorm >= 0.9 * s.gnorm_before;
                const bool stalled = s.h_out[6] <= zero_step * std::max(1.0, s.cmax_before) || no_progress;
                if (stalled && s.gnorm_before >= opt.gradient_tolerance) {
                    std::swap(s.cur, s.trial);  // discard the stalled point
                    std::swap(s.cur, s.prev);   // back at the point stepped from
                    s.stall_mode = true;
                    s.count = 0;
                    s.phase = Problem::Rechart;
                    rechart.push_back(k);
                    continue;
                }
            }
            s.stall_mode = false;
            ++s.res.iterations;
            s.record(s.E);
            const double energy_change = s.E - s.previous_energy;
            s.previous_energy = s.E;
            if (std::fabs(energy_change) < opt.energy_tolerance && s.gnorm < opt.gradient_tolerance) {
                s.finish(VBE_TERM_CONVERGED, true);
                continue;
            }
            if (opt.history_size > 0 && s.h_out[5] != 0.0) {
                s.newest = s.slot;
                s.count = std::min(s.count + 1, cap);
            }
            s.phase = Problem::Direction;
        }
        chart(rechart, false);
        for (int k : rechart) P[k].phase = Problem::Direction;
    }

    // results: orbitals, gradients, energies
    for (int k = 0; k < np; ++k) {
        Problem& s = P[k];
        s.res.energy = s.E;
        s.res.gradient_norm = s.gnorm;
        *s.p->result = s.res;
        hip_check(hipMemcpyAsync(s.p->C, s.pts[s.cur], sizeof(double) * s.N, hipMemcpyDeviceToHost, st), "D2H");
        if (s.p->grad)
            hip_check(hipMemcpyAsync(s.p->grad, s.grs[s.cur], sizeof(double) * s.N, hipMemcpyDeviceToHost, st), "D2H");
    }
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

}  // namespace xmvb_amd
