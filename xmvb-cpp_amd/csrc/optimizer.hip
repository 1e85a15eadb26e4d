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
// Device work per iteration: one block per orbital for the chart (k_opt_point), one block for the
// two-loop recursion with the chart projections folded into its dot products (k_opt_direction),
// a streaming kernel per Armijo trial (k_opt_trial) and one block for the secant test and history
// append (k_opt_append). The host reads scalars only: energies, the directional derivative, the
// projected-gradient norm and the step size.

#include "optimizer.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace xmvb_amd {
namespace {

constexpr int kPointThreads = 256;
constexpr int kBlockThreads = 1024;
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

// Chart at an accepted point, one block per orbital i: normalise c_i in the S metric (the
// gradient of the scale-invariant energy scales inversely), form S c_i on the support, the
// projected gradient r_i = g_i - (S c_i)(c_i^T g_i) and its dual norm sqrt(r_i^T S_i^-1 r_i).
// stats[4 i + 0..3] = dual norm, max |c_i|, norm before normalisation (<= 0: failure), c_i^T g_i.
__global__ void __launch_bounds__(kPointThreads)
k_opt_point(int n_bf, const double* __restrict__ S, const double* __restrict__ Sinv, const int* __restrict__ pat,
            const uint8_t* __restrict__ support, double* __restrict__ C, double* __restrict__ G,
            double* __restrict__ Sc, double* __restrict__ R, double* __restrict__ stats) {
    extern __shared__ double lds[];
    double* cs = lds;          // c_i, n_bf
    double* rs = lds + n_bf;   // r_i, n_bf
    __shared__ double red[16];
    const int i = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const size_t row = (size_t)i * n_bf;
    for (int k = tid; k < n_bf; k += nt) cs[k] = on_support(support, row + k) ? C[row + k] : 0.0;
    __syncthreads();
    double part = 0.0;
    for (int k = tid; k < n_bf; k += nt) {
        double a = 0.0;
        if (on_support(support, row + k))
            for (int j = 0; j < n_bf; ++j) a += S[k + (size_t)n_bf * j] * cs[j];
        Sc[row + k] = a;
        part += cs[k] * a;
    }
    const double rho2 = block_total(part, red);
    if (!(rho2 > 0.0) || !isfinite(rho2)) {
        if (tid == 0) stats[kOrbStats * i + 2] = -1.0;
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
        R[row + k] = r;
        rs[k] = r;
    }
    __syncthreads();
    const double* P = Sinv + (size_t)pat[i] * n_bf * n_bf;
    double npart = 0.0;
    for (int k = tid; k < n_bf; k += nt) {
        double a = 0.0;
        for (int j = 0; j < n_bf; ++j) a += P[k + (size_t)n_bf * j] * rs[j];
        npart += rs[k] * a;
    }
    const double nu2 = block_total(npart, red);
    if (tid == 0) {
        stats[kOrbStats * i + 0] = sqrt(fmax(nu2, 0.0));
        stats[kOrbStats * i + 1] = cmax_all;
        stats[kOrbStats * i + 2] = rho;
        stats[kOrbStats * i + 3] = beta;
    }
}

// Search direction t = -H r by the two-loop recursion (TransportedReducedLbfgsPreconditioner::
// apply, transported_lbfgs.cpp:41-86) over the secant pairs projected into the current chart,
// with the scaled-identity initial inverse gamma S^-1 (gamma = s'y'/y'y' of the newest kept pair)
// and the try_add_pair curvature filter (transported_lbfgs.cpp:14-32). Pairs are addressed by age
// (0 = newest) in a ring of `cap` slots whose newest slot is `newest`.
// out: [0] g . t, [1] max |t|, [2] pairs used, [3] gamma, [4] 1 if a recursion coefficient was
// not finite.
__global__ void __launch_bounds__(kBlockThreads)
k_opt_direction(int n_act, int n_bf, const double* __restrict__ C, const double* __restrict__ Sc,
                const double* __restrict__ G, const double* __restrict__ R, const double* __restrict__ Sinv,
                const int* __restrict__ pat, const double* __restrict__ hs, const double* __restrict__ hy,
                const double* __restrict__ hstat, int cap, int count, int newest, double* __restrict__ q,
                double* __restrict__ z, double* __restrict__ T, double* __restrict__ out) {
    extern __shared__ double lds[];
    const int N = n_act * n_bf;
    double* sig = lds;                          // (S c_i)^T s_a,i     count x n_act
    double* bet = sig + (size_t)count * n_act;  // c_i^T y_a,i         count x n_act
    double* alp = bet + (size_t)count * n_act;  // first-loop coefficients
    double* rho = alp + count;                  // 1 / s'^T y' (0: pair rejected)
    double* gam = rho + count;                  // s'^T y' / y'^T y'
    __shared__ double red[16];
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    if (tid == 0) bad = 0;
    // projection coefficients of every stored pair in this chart, one wavefront per dot product
    const int tasks = 2 * count * n_act;
    for (int task = wave; task < tasks; task += nw) {
        const int which = task / (count * n_act), rem = task % (count * n_act), a = rem / n_act, i = rem % n_act;
        const int slot = (newest - a + cap) % cap;
        const double* v = (which == 0 ? hs : hy) + (size_t)slot * N + (size_t)i * n_bf;
        const double* w = (which == 0 ? Sc : C) + (size_t)i * n_bf;
        double part = 0.0;
        for (int k = lane; k < n_bf; k += 64) part += v[k] * w[k];
        part = wave_total(part);
        if (lane == 0) (which == 0 ? sig : bet)[a * n_act + i] = part;
    }
    __syncthreads();
    for (int a = tid; a < count; a += nt) {
        const int slot = (newest - a + cap) % cap;
        double ss = 0.0, yy = 0.0, sy = 0.0;
        for (int i = 0; i < n_act; ++i) {
            const double* st = hstat + ((size_t)slot * n_act + i) * 3;
            const double s_ = sig[a * n_act + i], b_ = bet[a * n_act + i];
            ss += st[0] - s_ * s_;
            yy += st[1] - b_ * b_;
            sy += st[2] - s_ * b_;
        }
        const double sn = sqrt(fmax(ss, 0.0)), yn = sqrt(fmax(yy, 0.0));
        const bool ok = sn > 0.0 && yn > 0.0 && isfinite(sn) && isfinite(yn) && isfinite(sy) &&
                        sy > 1.4901161193847656e-08 * sn * yn;  // sqrt(DBL_EPSILON)
        rho[a] = ok ? 1.0 / sy : 0.0;
        gam[a] = ok ? sy / yy : 0.0;
    }
    for (int e = tid; e < N; e += nt) q[e] = R[e];
    __syncthreads();
    int used = 0;
    double gamma = 1.0;
    bool have_newest = false;
    for (int a = 0; a < count; ++a) {  // newest to oldest
        if (rho[a] == 0.0) continue;
        if (!have_newest) {
            gamma = gam[a];
            have_newest = true;
        }
        ++used;
        const int slot = (newest - a + cap) % cap;
        const double* s = hs + (size_t)slot * N;
        const double* y = hy + (size_t)slot * N;
        double part = 0.0;
        for (int e = tid; e < N; e += nt) part += (s[e] - C[e] * sig[a * n_act + e / n_bf]) * q[e];
        const double alpha = rho[a] * block_total(part, red);
        if (!isfinite(alpha) && tid == 0) bad = 1;
        if (tid == 0) alp[a] = alpha;
        for (int e = tid; e < N; e += nt) q[e] -= alpha * (y[e] - Sc[e] * bet[a * n_act + e / n_bf]);
    }
    __syncthreads();
    for (int e = tid; e < N; e += nt) {  // z = gamma S_i^-1 q_i
        const int i = e / n_bf, k = e - i * n_bf;
        const double* P = Sinv + (size_t)pat[i] * n_bf * n_bf;
        const double* qi = q + (size_t)i * n_bf;
        double acc = 0.0;
        for (int j = 0; j < n_bf; ++j) acc += P[k + (size_t)n_bf * j] * qi[j];
        z[e] = gamma * acc;
    }
    __syncthreads();
    for (int a = count - 1; a >= 0; --a) {  // oldest to newest
        if (rho[a] == 0.0) continue;
        const int slot = (newest - a + cap) % cap;
        const double* s = hs + (size_t)slot * N;
        const double* y = hy + (size_t)slot * N;
        double part = 0.0;
        for (int e = tid; e < N; e += nt) part += (y[e] - Sc[e] * bet[a * n_act + e / n_bf]) * z[e];
        const double b = rho[a] * block_total(part, red);
        if (!isfinite(b) && tid == 0) bad = 1;
        const double coef = alp[a] - b;
        for (int e = tid; e < N; e += nt) z[e] += (s[e] - C[e] * sig[a * n_act + e / n_bf]) * coef;
    }
    double dd = 0.0, tinf = 0.0;
    for (int e = tid; e < N; e += nt) {
        const double t = -z[e];
        T[e] = t;
        dd += G[e] * t;
        tinf = fmax(tinf, fabs(t));
    }
    dd = block_total(dd, red);
    tinf = block_largest(tinf, red);
    if (tid == 0) {
        out[0] = dd;
        out[1] = tinf;
        out[2] = used;
        out[3] = gamma;
        out[4] = bad;
    }
}

// Armijo trial point C + step * T.
__global__ void k_opt_trial(int N, const double* __restrict__ C, const double* __restrict__ T, double step,
                            double* __restrict__ Ct) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) Ct[e] = C[e] + step * T[e];
}

// Candidate secant pair (append_projected_secant_pair, transported_lbfgs.cpp:147-176) between the
// previous and the new accepted point (both normalised): s = C - Cp, y = G - Gp are written to
// the ring slot with their per-orbital chart-independent statistics s^T S s, y^T S^-1 y, s^T y;
// the curvature test runs in the new chart. out: [0] 1 if accepted, [1] max |s|.
__global__ void __launch_bounds__(kBlockThreads)
k_opt_append(int n_act, int n_bf, const double* __restrict__ S, const double* __restrict__ Sinv,
             const int* __restrict__ pat, const double* __restrict__ Cp, const double* __restrict__ Gp,
             const double* __restrict__ C, const double* __restrict__ G, const double* __restrict__ Sc,
             double* __restrict__ hs, double* __restrict__ hy, double* __restrict__ hstat, double* __restrict__ out) {
    __shared__ double red[16];
    __shared__ double orb[16 * 5];
    const int N = n_act * n_bf;
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    double sinf = 0.0;
    for (int e = tid; e < N; e += nt) {
        const double s = C[e] - Cp[e];
        hs[e] = s;
        hy[e] = G[e] - Gp[e];
        sinf = fmax(sinf, fabs(s));
    }
    sinf = block_largest(sinf, red);  // its barriers also publish hs / hy to the block
    for (int i = wave; i < n_act; i += nw) {
        const double* s = hs + (size_t)i * n_bf;
        const double* y = hy + (size_t)i * n_bf;
        const double* P = Sinv + (size_t)pat[i] * n_bf * n_bf;
        double sss = 0.0, ysy = 0.0, sy = 0.0, sg = 0.0, bt = 0.0;
        for (int k = lane; k < n_bf; k += 64) {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < n_bf; ++j) {
                a += S[k + (size_t)n_bf * j] * s[j];
                b += P[k + (size_t)n_bf * j] * y[j];
            }
            sss += s[k] * a;
            ysy += y[k] * b;
            sy += s[k] * y[k];
            sg += Sc[(size_t)i * n_bf + k] * s[k];
            bt += C[(size_t)i * n_bf + k] * y[k];
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
            orb[i * 5 + 0] = sss;
            orb[i * 5 + 1] = ysy;
            orb[i * 5 + 2] = sy;
            orb[i * 5 + 3] = sg;
            orb[i * 5 + 4] = bt;
        }
    }
    __syncthreads();
    if (tid == 0) {
        double ss = 0.0, yy = 0.0, sy = 0.0;
        for (int i = 0; i < n_act; ++i) {
            const double* o = orb + i * 5;
            ss += o[0] - o[3] * o[3];
            yy += o[1] - o[4] * o[4];
            sy += o[2] - o[3] * o[4];
        }
        const double sn = sqrt(fmax(ss, 0.0)), yn = sqrt(fmax(yy, 0.0));
        const bool ok = sn > 0.0 && yn > 0.0 && isfinite(sn) && isfinite(yn) && isfinite(sy) &&
                        sy > 1.4901161193847656e-08 * sn * yn;
        out[0] = ok ? 1.0 : 0.0;
        out[1] = sinf;
    }
}

// Device buffers of one optimisation, freed on scope exit.
struct DeviceBuffers {
    std::vector<void*> ptrs;
    template <typename T>
    T* alloc(size_t count) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~DeviceBuffers() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

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
                if (!(v > 0.0)) throw std::runtime_error("AO overlap restricted to an orbital support is not positive definite");
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

void run_plbfgs(const OptimizerHooks& h, const uint8_t* support, double* C_host, const vbe_optimizer_options& opt,
                vbe_optimizer_result& res, double* grad_host, double* trace, int trace_capacity) {
    if (opt.max_iterations < 0 || opt.history_size < 0) throw std::invalid_argument("negative iteration or history count");
    if (!(opt.gradient_tolerance >= 0.0) || !(opt.energy_tolerance >= 0.0) || !(opt.minimum_step_size > 0.0) ||
        !(opt.initial_step_size > 0.0) || !(opt.armijo_constant > 0.0 && opt.armijo_constant < 1.0))
        throw std::invalid_argument("optimizer tolerances and step sizes must be positive (Armijo constant in (0,1))");
    const int n_act = h.n_act, n_bf = h.n_bf, N = n_act * n_bf;
    if (n_act > 16) throw std::invalid_argument("at most 16 orbitals per optimisation");
    hipStream_t st = h.stream;
    constexpr double kEps = std::numeric_limits<double>::epsilon();

    // support patterns and their restricted inverse overlaps
    std::map<std::vector<uint8_t>, int> pattern_of;
    std::vector<std::vector<int>> patterns;
    std::vector<int> pat(n_act);
    long long reduced_size = 0;
    for (int i = 0; i < n_act; ++i) {
        std::vector<uint8_t> mask(n_bf, 1);
        if (support)
            for (int k = 0; k < n_bf; ++k) mask[k] = support[(size_t)i * n_bf + k] ? 1 : 0;
        auto it = pattern_of.find(mask);
        if (it == pattern_of.end()) {
            std::vector<int> idx;
            for (int k = 0; k < n_bf; ++k)
                if (mask[k]) idx.push_back(k);
            if (idx.empty()) throw std::invalid_argument("an orbital has an empty support");
            it = pattern_of.emplace(mask, (int)patterns.size()).first;
            patterns.push_back(idx);
        }
        pat[i] = it->second;
        reduced_size += (long long)patterns[pat[i]].size() - 1;
    }
    std::vector<double> sinv_host((size_t)patterns.size() * n_bf * n_bf);
    for (size_t p = 0; p < patterns.size(); ++p) {
        const std::vector<double> inv = restricted_inverse(h.S_ao, n_bf, patterns[p]);
        std::copy(inv.begin(), inv.end(), sinv_host.begin() + p * (size_t)n_bf * n_bf);
    }

    const int cap = opt.history_size;
    DeviceBuffers buf;
    double* d_S = buf.alloc<double>((size_t)n_bf * n_bf);
    double* d_Sinv = buf.alloc<double>(sinv_host.size());
    int* d_pat = buf.alloc<int>(n_act);
    uint8_t* d_support = support ? buf.alloc<uint8_t>(N) : nullptr;
    double* pts[3] = {buf.alloc<double>(N), buf.alloc<double>(N), buf.alloc<double>(N)};  // orbitals
    double* grs[3] = {buf.alloc<double>(N), buf.alloc<double>(N), buf.alloc<double>(N)};  // gradients
    double* d_Sc = buf.alloc<double>(N);
    double* d_R = buf.alloc<double>(N);
    double* d_T = buf.alloc<double>(N);
    double* d_q = buf.alloc<double>(N);
    double* d_z = buf.alloc<double>(N);
    double* d_stats = buf.alloc<double>((size_t)kOrbStats * n_act);
    double* d_out = buf.alloc<double>(8);
    double* d_hs = buf.alloc<double>((size_t)std::max(cap, 1) * N);
    double* d_hy = buf.alloc<double>((size_t)std::max(cap, 1) * N);
    double* d_hstat = buf.alloc<double>((size_t)std::max(cap, 1) * n_act * 3);

    std::vector<double> c0(C_host, C_host + N);
    if (support)
        for (int e = 0; e < N; ++e)
            if (!support[e]) c0[e] = 0.0;
    hip_check(hipMemcpyAsync(d_S, h.S_ao, sizeof(double) * n_bf * n_bf, hipMemcpyHostToDevice, st), "H2D");
    hip_check(hipMemcpyAsync(d_Sinv, sinv_host.data(), sizeof(double) * sinv_host.size(), hipMemcpyHostToDevice, st),
              "H2D");
    hip_check(hipMemcpyAsync(d_pat, pat.data(), sizeof(int) * n_act, hipMemcpyHostToDevice, st), "H2D");
    if (support) hip_check(hipMemcpyAsync(d_support, support, N, hipMemcpyHostToDevice, st), "H2D");
    hip_check(hipMemcpyAsync(pts[0], c0.data(), sizeof(double) * N, hipMemcpyHostToDevice, st), "H2D");

    int cur = 0, prev = 1, trial = 2;  // rotating indices into pts / grs
    std::vector<double> stats((size_t)kOrbStats * n_act), out(8);
    auto read = [&](std::vector<double>& dst, const double* src, size_t n) {
        hip_check(hipMemcpyAsync(dst.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    };
    double gnorm = 0.0, cmax = 0.0;
    auto launch_chart = [&](int at) {
        const size_t lds = sizeof(double) * 2 * n_bf;
        hipLaunchKernelGGL(k_opt_point, dim3(n_act), dim3(kPointThreads), lds, st, n_bf, d_S, d_Sinv, d_pat,
                           d_support, pts[at], grs[at], d_Sc, d_R, d_stats);
        hip_check(hipGetLastError(), "k_opt_point");
        hip_check(hipMemcpyAsync(stats.data(), d_stats, sizeof(double) * stats.size(), hipMemcpyDeviceToHost, st),
                  "D2H");
    };
    // normalise and project at pts[at] (sets gnorm / cmax); with `then`, enqueue more work before
    // the one synchronisation
    auto chart = [&](int at, const std::function<void()>& then = {}) {
        launch_chart(at);
        if (then) then();
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        gnorm = 0.0;
        cmax = 0.0;
        for (int i = 0; i < n_act; ++i) {
            if (!(stats[kOrbStats * i + 2] > 0.0))
                throw std::runtime_error("orbital " + std::to_string(i) + " has no positive norm on its support");
            gnorm = std::max(gnorm, stats[kOrbStats * i + 0]);
            cmax = std::max(cmax, stats[kOrbStats * i + 1]);
        }
    };
    int count = 0, newest = cap - 1;
    double dd = 0.0, tinf = 0.0;
    auto direction = [&](int at, int use) {  // search direction at pts[at] with the newest `use` pairs
        const size_t lds = sizeof(double) * ((size_t)2 * use * n_act + 3 * (size_t)use);
        hipLaunchKernelGGL(k_opt_direction, dim3(1), dim3(kBlockThreads), lds, st, n_act, n_bf, pts[at], d_Sc,
                           grs[at], d_R, d_Sinv, d_pat, d_hs, d_hy, d_hstat, std::max(cap, 1), use,
                           std::max(newest, 0), d_q, d_z, d_T, d_out);
        hip_check(hipGetLastError(), "k_opt_direction");
        read(out, d_out, 5);
        if (out[4] != 0.0) throw std::runtime_error("transported L-BFGS recursion produced a non-finite coefficient");
        dd = out[0];
        tinf = out[1];
    };
    const double zero_step = 128.0 * kEps;
    // The first trial of a line search runs forward and backward in one launch sequence when the
    // previous search accepted its first trial (the common case once the secant model is built),
    // which saves a synchronisation and a launch gap; a rejected speculative trial costs one
    // backward. Which trials are accepted does not depend on it.
    bool speculate = false;
    auto armijo = [&](int at, double E0, double step0, double& Et) -> bool {
        // try_armijo_backtracking_nonredundant_direction (globalization/line_search.cpp:33-95)
        if (!std::isfinite(dd) || dd >= 0.0) return false;
        double step = std::max(opt.minimum_step_size, step0);
        bool first = true;
        while (step >= opt.minimum_step_size) {
            if (step * tinf <= zero_step * std::max(1.0, cmax)) {
                step *= 0.5;
                continue;
            }
            hipLaunchKernelGGL(k_opt_trial, dim3((N + 255) / 256), dim3(256), 0, st, N, pts[at], d_T, step,
                               pts[trial]);
            hip_check(hipGetLastError(), "k_opt_trial");
            const bool with_grad = first && speculate;
            Et = with_grad ? h.energy_gradient(pts[trial], d_support, grs[trial])
                           : h.energy(pts[trial], d_support, grs[trial]);
            ++res.energy_evaluations;
            if (with_grad) ++res.gradient_evaluations;
            if (std::isfinite(Et) && Et <= E0 + opt.armijo_constant * step * dd) {
                if (!with_grad) {
                    h.gradient(pts[trial], d_support, grs[trial]);
                    ++res.gradient_evaluations;
                }
                speculate = first;
                return true;
            }
            first = false;
            step *= 0.5;
        }
        speculate = false;
        return false;
    };

    res = vbe_optimizer_result{};
    res.termination = VBE_TERM_MAX_ITERATIONS;
    int n_trace = 0;
    auto record = [&](double e) {
        if (trace && n_trace < trace_capacity) trace[n_trace] = e;
        ++n_trace;
    };
    double E = h.energy(pts[cur], d_support, grs[cur]);
    ++res.energy_evaluations;
    if (!std::isfinite(E)) throw std::runtime_error("the starting orbitals cannot be evaluated");
    h.gradient(pts[cur], d_support, grs[cur]);
    ++res.gradient_evaluations;
    chart(cur);
    record(E);
    double previous_energy = E;

    for (int iteration = 0; iteration < opt.max_iterations; ++iteration) {
        if (reduced_size == 0) {
            res.termination = VBE_TERM_EMPTY_SPACE;
            break;
        }
        const double gnorm_before = gnorm;
        if (iteration == 0 && gnorm < opt.gradient_tolerance) {
            res.converged = 1;
            res.termination = VBE_TERM_INITIAL_TOLERANCE;
            break;
        }
        direction(cur, count);
        if (!(std::isfinite(dd) && dd < 0.0)) {  // not a descent direction: drop the history
            count = 0;
            direction(cur, 0);
        }
        if (!(std::isfinite(dd) && dd < 0.0) || tinf <= zero_step * std::max(1.0, cmax)) {
            count = 0;
            direction(cur, 0);
        }
        if (!(std::isfinite(dd) && dd < 0.0)) {
            res.termination = VBE_TERM_NON_DESCENT;
            break;
        }
        const double reference_energy = E;
        const double cmax_before = cmax;
        double Et = 0.0;
        if (!armijo(cur, E, std::min(1.0, opt.initial_step_size), Et)) {
            res.termination = VBE_TERM_LINE_SEARCH_FAILED;
            break;
        }
        std::swap(prev, cur);   // prev = the point we stepped from
        std::swap(cur, trial);  // cur = the accepted trial
        E = Et;
        const int slot = cap > 0 ? (newest + 1) % cap : 0;
        auto append = [&]() {  // enqueue only; the chart's synchronisation completes it
            hipLaunchKernelGGL(k_opt_append, dim3(1), dim3(kBlockThreads), 0, st, n_act, n_bf, d_S, d_Sinv, d_pat,
                               pts[prev], grs[prev], pts[cur], grs[cur], d_Sc, d_hs + (size_t)slot * N,
                               d_hy + (size_t)slot * N, d_hstat + (size_t)slot * n_act * 3, d_out);
            hip_check(hipGetLastError(), "k_opt_append");
            hip_check(hipMemcpyAsync(out.data(), d_out, sizeof(double) * 2, hipMemcpyDeviceToHost, st), "D2H");
        };
        chart(cur, append);  // the candidate pair also measures the accepted step
        // a step that changed nothing or made no progress: Armijo steepest descent from the
        // previous point with a fresh history (plbfgs.cpp:146-191)
        const double energy_scale = std::max(1.0, std::fabs(reference_energy));
        const bool no_progress = std::isfinite(reference_energy) && std::isfinite(E) &&
                                 std::fabs(E - reference_energy) <= 4096.0 * kEps * energy_scale &&
                                 !(gnorm_before < opt.gradient_tolerance) && gnorm >= 0.9 * gnorm_before;
        const bool stalled = out[1] <= zero_step * std::max(1.0, cmax_before) || no_progress;
        if (stalled && gnorm_before >= opt.gradient_tolerance) {
            std::swap(cur, trial);  // discard the stalled point; trial slot is free again
            std::swap(cur, prev);   // back at the previous point
            chart(cur);
            count = 0;
            direction(cur, 0);
            const double sd_step = std::max(opt.minimum_step_size,
                                            std::min(std::min(1.0, opt.initial_step_size), 1.0 / std::max(1.0, gnorm)));
            if (!armijo(cur, reference_energy, sd_step, Et)) {
                E = reference_energy;
                res.termination = VBE_TERM_LINE_SEARCH_STALLED;
                break;
            }
            std::swap(prev, cur);
            std::swap(cur, trial);
            E = Et;
            chart(cur, append);
        }
        ++res.iterations;
        record(E);
        const double energy_change = E - previous_energy;
        previous_energy = E;
        if (std::fabs(energy_change) < opt.energy_tolerance && gnorm < opt.gradient_tolerance) {
            res.converged = 1;
            res.termination = VBE_TERM_CONVERGED;
            break;
        }
        if (cap > 0 && out[0] != 0.0) {
            newest = slot;
            count = std::min(count + 1, cap);
        }
    }
    res.energy = E;
    res.gradient_norm = gnorm;
    hip_check(hipMemcpyAsync(C_host, pts[cur], sizeof(double) * N, hipMemcpyDeviceToHost, st), "D2H");
    if (grad_host) hip_check(hipMemcpyAsync(grad_host, grs[cur], sizeof(double) * N, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

}  // namespace xmvb_amd
