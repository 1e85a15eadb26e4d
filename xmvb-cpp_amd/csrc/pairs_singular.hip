// Determinant pairs with a rank-deficient overlap block.
//
// The lane-group pair kernel lists every pair one of whose spin blocks fails the complete-pivoting
// rank rule of DeterminantOverlapResolver::resolve_matrix (determinants/algebra/overlap.cpp:50-110:
// rank = #pivots above 1e-12 max|pivot|). As there, such a block is taken in its SVD chart
// X = U diag(s) V^T and singular values below 1e-12 count as null (nullity k); with k = 0 the SVD
// inverse is used. The reference evaluates the remaining pairs with first- and second-order
// cofactors (pairs/contractions.cpp:172-236, 901-983). The pair function is multilinear in the
// singular values of each block in that chart, so its value and first derivatives at the null
// point equal the mean over the 2^k chart nodes where each null singular value is +tau or -tau
// (tau: the largest singular value, or 1 when all vanish); every node is regular and is evaluated
// in the regular inverse form. The value vanishes for k_alpha + k_beta > 2 and every first
// derivative for k_alpha + k_beta > 3 (each term carries at most a second-order cofactor per
// block), so those pairs are skipped. One thread per listed pair: a rare path, written for clarity.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"
#include "kernels.hpp"

namespace xmvb_amd {

namespace {

constexpr int SB = MAX_SPIN_BLOCK;
constexpr double kRankThreshold = 1.0e-12;

struct Chart {
    int m;
    int R[SB], L[SB];  // occupied orbitals of the right (ket) and left (bra) strings
    bool lu_regular;   // full rank under the complete-pivoting rule
    int nullity;       // singular values below the threshold (SVD chart only)
    double det;        // LU path: det X
    double y[SB][SB];  // LU path: Y = X^{-1}, y[l][r]
    double u[SB][SB], v[SB][SB], s[SB];  // SVD chart, columns u[.][q], v[.][q]
    double parity;     // sign(det U det V)
};

__device__ void occupied_list(uint64_t mask, int* out) {
    int k = 0;
    for (int b = 0; b < 64 && mask; ++b, mask >>= 1)
        if (mask & 1) out[k++] = b;
}

// Gauss-Jordan with complete pivoting on [X | I]; rank rule as FullPivLU with threshold 1e-12.
__device__ bool fullpiv_inverse(int m, double (&a)[SB][SB], double (&y)[SB][SB], double& det) {
    double aug[SB][SB];
    int col_of[SB];
    for (int i = 0; i < m; ++i) {
        col_of[i] = i;
        for (int j = 0; j < m; ++j) aug[i][j] = i == j ? 1.0 : 0.0;
    }
    double maxpiv = 0.0, minpiv = 1.0e308, sign = 1.0, prod = 1.0;
    for (int k = 0; k < m; ++k) {
        int p = k, q = k;
        double best = -1.0;
        for (int j = k; j < m; ++j)
            for (int i = k; i < m; ++i)
                if (fabs(a[i][j]) > best) {
                    best = fabs(a[i][j]);
                    p = i;
                    q = j;
                }
        if (!(best > 0.0)) return false;
        if (p != k) {
            sign = -sign;
            for (int j = 0; j < m; ++j) {
                double t = a[k][j];
                a[k][j] = a[p][j];
                a[p][j] = t;
                t = aug[k][j];
                aug[k][j] = aug[p][j];
                aug[p][j] = t;
            }
        }
        if (q != k) {
            sign = -sign;
            for (int i = 0; i < m; ++i) {
                const double t = a[i][k];
                a[i][k] = a[i][q];
                a[i][q] = t;
            }
            const int t = col_of[k];
            col_of[k] = col_of[q];
            col_of[q] = t;
        }
        const double piv = a[k][k];
        maxpiv = fmax(maxpiv, fabs(piv));
        minpiv = fmin(minpiv, fabs(piv));
        prod *= piv;
        for (int i = 0; i < m; ++i) {
            if (i == k) continue;
            const double f = a[i][k] / piv;
            for (int j = 0; j < m; ++j) {
                a[i][j] -= f * a[k][j];
                aug[i][j] -= f * aug[k][j];
            }
        }
    }
    if (minpiv <= kRankThreshold * maxpiv) return false;
    // (X Q) rows reduced to diag: X^{-1} row col_of[k] = aug row k / pivot k
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j) y[col_of[k]][j] = aug[k][j] / a[k][k];
    det = sign * prod;
    return true;
}

// One-sided Jacobi SVD X = U diag(s) V^T, s descending, null left vectors completed to an
// orthonormal basis.
__device__ void jacobi_svd(int m, double (&a)[SB][SB], Chart& c) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) c.v[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < m - 1; ++i)
            for (int j = i + 1; j < m; ++j) {
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int k = 0; k < m; ++k) {
                    al += a[k][i] * a[k][i];
                    be += a[k][j] * a[k][j];
                    ga += a[k][i] * a[k][j];
                }
                if (ga == 0.0 || fabs(ga) <= 1e-17 * sqrt(al * be)) continue;
                off = fmax(off, fabs(ga) / sqrt(al * be));
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                for (int k = 0; k < m; ++k) {
                    const double ai = a[k][i], aj = a[k][j];
                    a[k][i] = cs * ai - sn * aj;
                    a[k][j] = sn * ai + cs * aj;
                    const double vi = c.v[k][i], vj = c.v[k][j];
                    c.v[k][i] = cs * vi - sn * vj;
                    c.v[k][j] = sn * vi + cs * vj;
                }
            }
        if (off < 1e-16) break;
    }
    double sv[SB];
    int order[SB];
    for (int j = 0; j < m; ++j) {
        double t = 0.0;
        for (int k = 0; k < m; ++k) t += a[k][j] * a[k][j];
        sv[j] = sqrt(t);
        order[j] = j;
    }
    for (int i = 1; i < m; ++i)  // stable insertion sort, descending
        for (int j = i; j > 0 && sv[order[j]] > sv[order[j - 1]]; --j) {
            const int t = order[j];
            order[j] = order[j - 1];
            order[j - 1] = t;
        }
    double vv[SB][SB];
    bool have[SB];
    const double smax = m ? sv[order[0]] : 0.0;
    for (int q = 0; q < m; ++q) {
        const int j = order[q];
        c.s[q] = sv[j];
        for (int k = 0; k < m; ++k) vv[k][q] = c.v[k][j];
        have[q] = sv[j] > 1e-300 && sv[j] > 1e-15 * smax;
        for (int k = 0; k < m; ++k) c.u[k][q] = have[q] ? a[k][j] / sv[j] : 0.0;
    }
    for (int k = 0; k < m; ++k)
        for (int q = 0; q < m; ++q) c.v[k][q] = vv[k][q];
    // complete missing left vectors: the unit vector with the largest residual after two
    // Gram-Schmidt passes against the present ones
    for (int q = 0; q < m; ++q) {
        if (have[q]) continue;
        double best[SB], best_norm = -1.0;
        for (int e = 0; e < m; ++e) {
            double w[SB];
            for (int k = 0; k < m; ++k) w[k] = k == e ? 1.0 : 0.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int d = 0; d < m; ++d) {
                    if (!have[d]) continue;
                    double dot = 0.0;
                    for (int k = 0; k < m; ++k) dot += c.u[k][d] * w[k];
                    for (int k = 0; k < m; ++k) w[k] -= dot * c.u[k][d];
                }
            double nn = 0.0;
            for (int k = 0; k < m; ++k) nn += w[k] * w[k];
            if (nn > best_norm) {
                best_norm = nn;
                for (int k = 0; k < m; ++k) best[k] = w[k];
            }
        }
        const double inv = 1.0 / sqrt(best_norm);
        for (int k = 0; k < m; ++k) c.u[k][q] = best[k] * inv;
        have[q] = true;
    }
    // parity = sign(det U det V) from row-pivoted eliminations of copies
    auto det_sign = [&](const double (&mat)[SB][SB]) {
        double w[SB][SB];
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) w[i][j] = mat[i][j];
        double sgn = 1.0;
        for (int k = 0; k < m; ++k) {
            int p = k;
            for (int i = k + 1; i < m; ++i)
                if (fabs(w[i][k]) > fabs(w[p][k])) p = i;
            if (p != k) {
                sgn = -sgn;
                for (int j = 0; j < m; ++j) {
                    const double t = w[k][j];
                    w[k][j] = w[p][j];
                    w[p][j] = t;
                }
            }
            if (w[k][k] < 0.0) sgn = -sgn;
            for (int i = k + 1; i < m; ++i) {
                const double f = w[i][k] / w[k][k];
                for (int j = k; j < m; ++j) w[i][j] -= f * w[k][j];
            }
        }
        return sgn;
    };
    c.parity = det_sign(c.u) * det_sign(c.v);
}

__device__ void resolve_chart(const double* sact, int n, Chart& c) {
    double a[SB][SB];
    for (int i = 0; i < c.m; ++i)
        for (int j = 0; j < c.m; ++j) a[i][j] = sact[c.R[i] + n * c.L[j]];
    c.nullity = 0;
    c.lu_regular = c.m == 0;
    if (c.m == 0) {
        c.det = 1.0;
        return;
    }
    double w[SB][SB];
    for (int i = 0; i < c.m; ++i)
        for (int j = 0; j < c.m; ++j) w[i][j] = a[i][j];
    c.lu_regular = fullpiv_inverse(c.m, w, c.y, c.det);
    if (c.lu_regular) return;
    jacobi_svd(c.m, a, c);
    for (int q = 0; q < c.m; ++q)
        if (c.s[q] < kRankThreshold) ++c.nullity;
}

// Node `bits` of the chart (null singular values -> +-tau): D[r][l] = Y(l, r), returns det.
__device__ double chart_node(const Chart& c, unsigned bits, double (&D)[SB][SB]) {
    const int m = c.m;
    if (c.lu_regular) {
        for (int r = 0; r < m; ++r)
            for (int l = 0; l < m; ++l) D[r][l] = c.y[l][r];
        return c.det;
    }
    const double tau = c.s[0] >= kRankThreshold ? c.s[0] : 1.0;
    double sn[SB];
    double det = c.parity;
    int bit = 0;
    for (int q = 0; q < m; ++q) {
        sn[q] = c.s[q];
        if (q >= m - c.nullity) sn[q] = ((bits >> bit++) & 1u) ? -tau : tau;
        det *= sn[q];
    }
    for (int r = 0; r < m; ++r)
        for (int l = 0; l < m; ++l) {
            double t = 0.0;
            for (int q = 0; q < m; ++q) t += c.v[l][q] * c.u[r][q] / sn[q];
            D[r][l] = t;
        }
    return det;
}

__device__ __forceinline__ int pidx(const Chart& c, int r, int l) { return pair_index(c.R[r], c.L[l]); }

// Phi = sum_s D_s.h + sum_s sum_{r<r2,l<l2} [(rl|r2l2) - (rl2|r2l)] minor_s + D_0.J[D_1]
__device__ double pair_phi(const double* G2, const double* hact, int n, int np, const Chart* ch,
                           const double (*D)[SB][SB]) {
    double phi = 0.0;
    for (int s = 0; s < 2; ++s) {
        const Chart& c = ch[s];
        const int m = c.m;
        for (int r = 0; r < m; ++r)
            for (int l = 0; l < m; ++l) phi += D[s][r][l] * hact[c.R[r] + n * c.L[l]];
        for (int r = 0; r < m; ++r)
            for (int r2 = r + 1; r2 < m; ++r2)
                for (int l = 0; l < m; ++l)
                    for (int l2 = l + 1; l2 < m; ++l2) {
                        const double A = G2[pidx(c, r, l) * np + pidx(c, r2, l2)] - G2[pidx(c, r, l2) * np + pidx(c, r2, l)];
                        phi += A * (D[s][r][l] * D[s][r2][l2] - D[s][r][l2] * D[s][r2][l]);
                    }
    }
    for (int r = 0; r < ch[0].m; ++r)
        for (int l = 0; l < ch[0].m; ++l) {
            const double* g = G2 + pidx(ch[0], r, l) * np;
            double inner = 0.0;
            for (int r2 = 0; r2 < ch[1].m; ++r2)
                for (int l2 = 0; l2 < ch[1].m; ++l2) inner += g[pidx(ch[1], r2, l2)] * D[1][r2][l2];
            phi += D[0][r][l] * inner;
        }
    return phi;
}

// Adjoints of wH * H + wS * S at one regular node with S = c.
template <bool FullDg>
__device__ void pair_adjoint(const double* G2, const double* hact, int n, int np, const Chart* ch,
                             const double (*D)[SB][SB], double c, double wH, double wS, double* dg, double* ds,
                             double* dh) {
    double F[2][SB][SB];
    double phi_h = 0.0;
    for (int s = 0; s < 2; ++s)
        for (int r = 0; r < ch[s].m; ++r)
            for (int l = 0; l < ch[s].m; ++l) {
                F[s][r][l] = hact[ch[s].R[r] + n * ch[s].L[l]];
                phi_h += D[s][r][l] * F[s][r][l];
            }
    const double cw = c * wH;
    for (int s = 0; s < 2; ++s) {
        const Chart& b = ch[s];
        const int m = b.m;
        for (int r = 0; r < m; ++r)
            for (int r2 = r + 1; r2 < m; ++r2)
                for (int l = 0; l < m; ++l)
                    for (int l2 = l + 1; l2 < m; ++l2) {
                        const int iJ = pidx(b, r, l) * np + pidx(b, r2, l2);
                        const int iK = pidx(b, r, l2) * np + pidx(b, r2, l);
                        const double A = G2[iJ] - G2[iK];
                        F[s][r][l] += A * D[s][r2][l2];
                        F[s][r2][l2] += A * D[s][r][l];
                        F[s][r][l2] -= A * D[s][r2][l];
                        F[s][r2][l] -= A * D[s][r][l2];
                        const double v = cw * (D[s][r][l] * D[s][r2][l2] - D[s][r][l2] * D[s][r2][l]);
                        if (FullDg) {
                            atomicAdd(&dg[iJ], v);
                            atomicAdd(&dg[iK], -v);
                        } else {
                            atomicAdd(&dg[pair_index(pidx(b, r, l), pidx(b, r2, l2))], v);
                            atomicAdd(&dg[pair_index(pidx(b, r, l2), pidx(b, r2, l))], -v);
                        }
                    }
    }
    for (int r = 0; r < ch[0].m; ++r)
        for (int l = 0; l < ch[0].m; ++l) {
            const int P = pidx(ch[0], r, l);
            double acc = 0.0;
            for (int r2 = 0; r2 < ch[1].m; ++r2)
                for (int l2 = 0; l2 < ch[1].m; ++l2) {
                    const int Q = pidx(ch[1], r2, l2);
                    const double g = G2[P * np + Q];
                    acc += g * D[1][r2][l2];
                    F[1][r2][l2] += g * D[0][r][l];
                    if (FullDg) atomicAdd(&dg[P * np + Q], cw * D[0][r][l] * D[1][r2][l2]);
                    else atomicAdd(&dg[pair_index(P, Q)], cw * D[0][r][l] * D[1][r2][l2]);
                }
            F[0][r][l] += acc;
        }
    double phi_f = 0.0;
    for (int s = 0; s < 2; ++s)
        for (int r = 0; r < ch[s].m; ++r)
            for (int l = 0; l < ch[s].m; ++l) phi_f += D[s][r][l] * F[s][r][l];
    const double sc = wH * 0.5 * (phi_h + phi_f) + wS;
    // ds += c [ (wH Phi + wS) D - wH D F^T D ], dh += c wH D
    for (int s = 0; s < 2; ++s) {
        const Chart& b = ch[s];
        const int m = b.m;
        for (int l = 0; l < m; ++l) {
            double z[SB];
            for (int a2 = 0; a2 < m; ++a2) {
                double t = 0.0;
                for (int b2 = 0; b2 < m; ++b2) t += F[s][b2][a2] * D[s][b2][l];
                z[a2] = t;
            }
            for (int r = 0; r < m; ++r) {
                double dfd = 0.0;
                for (int a2 = 0; a2 < m; ++a2) dfd += D[s][r][a2] * z[a2];
                const int at = b.R[r] + n * b.L[l];
                atomicAdd(&ds[at], c * (sc * D[s][r][l] - wH * dfd));
                atomicAdd(&dh[at], cw * D[s][r][l]);
            }
        }
    }
}

template <bool Backward, bool FullDg>
__global__ void __launch_bounds__(256) k_pairs_singular(const SlotDesc* __restrict__ slots, TopoDesc topo,
                                                        int part_offset) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n = topo.n, ng = topo.ng, n_det = topo.n_det;
    const int np = n * (n + 1) / 2;
    const int count = to_global(sd.flags)[FLAG_SINGULAR_PAIRS];
    extern __shared__ double lds[];
    double* G2 = lds;
    double* sact = G2 + np * np;
    double* hact = sact + n * n;
    double* dg = hact + n * n;
    double* ds = dg + (FullDg ? np * np : ng);
    double* dh = ds + n * n;
    if (count > 0) {
        for (int t = threadIdx.x; t < np * np; t += blockDim.x) {
            const int P = t % np, Q = t / np;
            G2[t] = to_global(sd.G)[pair_index(P, Q)];
        }
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
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        int d, e;
        unpack_pair(sd.singular[i], d, e);
        double w = 1.0, wS = 0.0;
        if (Backward) {
            pair_weights(sd, topo, d, e, d, e, false, d == e ? 1.0 : 2.0, w, wS);
            if (w == 0.0 && wS == 0.0) continue;
        }
        Chart ch[2];
        const uint64_t mL[2] = {topo.amask[topo.det_a[d]], topo.bmask[topo.det_b[d]]};
        const uint64_t mR[2] = {topo.amask[topo.det_a[e]], topo.bmask[topo.det_b[e]]};
        for (int s = 0; s < 2; ++s) {
            ch[s].m = __builtin_popcountll(mL[s]);
            occupied_list(mL[s], ch[s].L);
            occupied_list(mR[s], ch[s].R);
            resolve_chart(sact, n, ch[s]);
        }
        const int k = ch[0].nullity + ch[1].nullity;
        double H = 0.0, S = 0.0;
        if (k <= (Backward ? 3 : 2)) {
            const unsigned nodes = 1u << k;
            const double scale = 1.0 / (double)nodes;
            for (unsigned node = 0; node < nodes; ++node) {
                double D[2][SB][SB];
                const double det0 = chart_node(ch[0], node, D[0]);
                const double det1 = chart_node(ch[1], node >> ch[0].nullity, D[1]);
                const double c = det0 * det1;
                if (!Backward) {
                    H += scale * c * pair_phi(G2, hact, n, np, ch, D);
                    S += scale * c;
                } else {
                    pair_adjoint<FullDg>(G2, hact, n, np, ch, D, c, w * scale, wS * scale, dg, ds, dh);
                }
            }
        }
        if (!Backward) {
            to_global(sd.Hdet)[(size_t)d + (size_t)n_det * e] = H;
            to_global(sd.Hdet)[(size_t)e + (size_t)n_det * d] = H;
            to_global(sd.Sdet)[(size_t)d + (size_t)n_det * e] = S;
            to_global(sd.Sdet)[(size_t)e + (size_t)n_det * d] = S;
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

}  // namespace

// G2 (np x np), S_act, h_act; backward adds the kernel adjoint (full np x np or packed) and ds, dh
std::size_t pairs_singular_lds_bytes(int n, int ng, bool backward) {
    const std::size_t np = (std::size_t)n * (n + 1) / 2, nn = (std::size_t)n * n;
    std::size_t doubles = np * np + 2 * nn;
    if (backward) doubles += (pairs_full_dg(n, ng) ? np * np : (std::size_t)ng) + 2 * nn;
    return doubles * sizeof(double);
}

hipError_t configure_pairs_singular() {
    const int limit = 160 * 1024;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_singular<false, false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, limit);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_singular<true, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, limit);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_singular<true, false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, limit);
}

void launch_pairs_singular(hipStream_t st, bool backward, int part_offset, const SlotDesc* slots, int count,
                           const TopoDesc& topo) {
    const dim3 grid(kSingularBlocks, 1, count);
    const size_t lds = pairs_singular_lds_bytes(topo.n, topo.ng, backward);
    // the backward accumulates with LDS atomics: one wavefront per block keeps their order that of
    // the wavefront's own instructions, hence the same bits on every run
    if (!backward)
        hipLaunchKernelGGL((k_pairs_singular<false, false>), grid, dim3(256), lds, st, slots, topo, part_offset);
    else if (pairs_full_dg(topo.n, topo.ng))
        hipLaunchKernelGGL((k_pairs_singular<true, true>), grid, dim3(64), lds, st, slots, topo, part_offset);
    else
        hipLaunchKernelGGL((k_pairs_singular<true, false>), grid, dim3(64), lds, st, slots, topo, part_offset);
}

}  // namespace xmvb_amd
