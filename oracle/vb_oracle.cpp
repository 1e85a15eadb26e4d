// FP64 CPU oracle for one VBSCF energy+gradient evaluation.
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker or as the
// timed CPU baseline. The product engine (xmvb-cpp_amd/) never links or calls it.
//
// This file restates, in plain C++17 with OpenMP and without Eigen/LAPACK, the
// reference algorithm of Tao-Xia/xmvb-cpp for the RI energy+gradient path.
// Every function names the reference file:line it follows. Pinning (what
// independent check fixes each part; see tests/):
//   topology            - reference count fixtures (tests/input/test_generated_structures.cpp:84-141),
//                         Catalan/2^(n/2) closed forms, S^2 = 0 per structure, product-vs-oracle bit-exact
//   resolve/cofactors   - brute-force Laplace minors (tests/test_oracle_pinning.py), nullity cases
//   same/opposite spin  - orthonormal-basis FCI projected through the exterior map (independent formula)
//   GEP                 - scipy.linalg.eigh on the same H, S
//   RI transform        - dense numpy einsum at small sizes
//   gradient            - central finite differences (step 1e-5/1e-6) incl. exactly singular pairs,
//                         per-orbital scale invariance <dE/dc_i, c_i> = 0
// Nothing here is "parity unpinned".
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using Vec = std::vector<double>;

// ---------------------------------------------------------------------------
// packing conventions (core/pair_index.hpp:15-55)
// ---------------------------------------------------------------------------
inline int64_t pp(int64_t a, int64_t b) {
    const int64_t hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

// ---------------------------------------------------------------------------
// small dense linear algebra, column-major
// ---------------------------------------------------------------------------

// Complete-pivoting LU with Eigen FullPivLU rank semantics (determinants/algebra/overlap.cpp:50-110):
// rank = #pivots with |p| > thr * max|pivot|; elimination stops at an exactly-zero corner.
struct LuResult {
    int rank = 0;
    double det = 0.0;
    Vec inverse;  // n x n when rank == n
};

LuResult fullpiv_lu(int n, const double* x, double thr) {
    LuResult res;
    Vec a(x, x + n * n);
    std::vector<int> rp(n), cp(n);
    std::iota(rp.begin(), rp.end(), 0);
    std::iota(cp.begin(), cp.end(), 0);
    int perm_sign = 1, nonzero = n;
    double maxpivot = 0.0;
    for (int k = 0; k < n; ++k) {
        int bi = k, bj = k;
        double big = -1.0;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i)
                if (std::fabs(a[i + n * j]) > big) {
                    big = std::fabs(a[i + n * j]);
                    bi = i;
                    bj = j;
                }
        if (big == 0.0) {
            nonzero = k;
            break;
        }
        maxpivot = std::max(maxpivot, big);
        if (bi != k) {
            for (int j = 0; j < n; ++j) std::swap(a[k + n * j], a[bi + n * j]);
            std::swap(rp[k], rp[bi]);
            perm_sign = -perm_sign;
        }
        if (bj != k) {
            for (int i = 0; i < n; ++i) std::swap(a[i + n * k], a[i + n * bj]);
            std::swap(cp[k], cp[bj]);
            perm_sign = -perm_sign;
        }
        const double piv = a[k + n * k];
        for (int i = k + 1; i < n; ++i) a[i + n * k] /= piv;
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) a[i + n * j] -= a[i + n * k] * a[k + n * j];
    }
    int rank = 0;
    for (int k = 0; k < nonzero; ++k)
        if (std::fabs(a[k + n * k]) > thr * maxpivot) ++rank;
    res.rank = rank;
    double det = perm_sign;
    for (int k = 0; k < n; ++k) det *= (k < nonzero ? a[k + n * k] : 0.0);
    res.det = det;
    if (rank == n) {
        // P X Q = L U  ->  X^{-1} = Q U^{-1} L^{-1} P
        res.inverse.assign(n * n, 0.0);
        Vec col(n);
        for (int e = 0; e < n; ++e) {
            // solve X z = e_e : L U (Q^T z) = P e_e
            for (int i = 0; i < n; ++i) col[i] = (rp[i] == e) ? 1.0 : 0.0;
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < i; ++k) col[i] -= a[i + n * k] * col[k];
            for (int i = n - 1; i >= 0; --i) {
                for (int k = i + 1; k < n; ++k) col[i] -= a[i + n * k] * col[k];
                col[i] /= a[i + n * i];
            }
            for (int i = 0; i < n; ++i) res.inverse[cp[i] + n * e] = col[i];
        }
    }
    return res;
}

double det_partial(int n, Vec a) {
    double det = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a[i + n * k]) > std::fabs(a[p + n * k])) p = i;
        if (a[p + n * k] == 0.0) return 0.0;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(a[k + n * j], a[p + n * j]);
            det = -det;
        }
        det *= a[k + n * k];
        for (int i = k + 1; i < n; ++i) {
            const double f = a[i + n * k] / a[k + n * k];
            for (int j = k + 1; j < n; ++j) a[i + n * j] -= f * a[k + n * j];
        }
    }
    return det;
}

// One-sided (Hestenes) Jacobi SVD: X = U diag(s) V^T, s descending, U and V orthogonal
// (null columns of U completed by Gram-Schmidt). Stands in for Eigen BDCSVD (overlap.cpp:82-107).
void svd_jacobi(int n, const double* x, Vec& u, Vec& s, Vec& v) {
    Vec a(x, x + n * n);
    v.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i) v[i + n * i] = 1.0;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n - 1; ++i)
            for (int j = i + 1; j < n; ++j) {
                double al = 0, be = 0, ga = 0;
                for (int k = 0; k < n; ++k) {
                    al += a[k + n * i] * a[k + n * i];
                    be += a[k + n * j] * a[k + n * j];
                    ga += a[k + n * i] * a[k + n * j];
                }
                if (ga == 0.0 || std::fabs(ga) <= 1e-17 * std::sqrt(al * be)) continue;
                off = std::max(off, std::fabs(ga) / std::sqrt(al * be));
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int k = 0; k < n; ++k) {
                    const double ai = a[k + n * i], aj = a[k + n * j];
                    a[k + n * i] = c * ai - sn * aj;
                    a[k + n * j] = sn * ai + c * aj;
                    const double vi = v[k + n * i], vj = v[k + n * j];
                    v[k + n * i] = c * vi - sn * vj;
                    v[k + n * j] = sn * vi + c * vj;
                }
            }
        if (off < 1e-16) break;
    }
    Vec sv(n);
    for (int j = 0; j < n; ++j) {
        double t = 0;
        for (int k = 0; k < n; ++k) t += a[k + n * j] * a[k + n * j];
        sv[j] = std::sqrt(t);
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return sv[p] > sv[q]; });
    s.assign(n, 0.0);
    u.assign(n * n, 0.0);
    Vec vv(n * n);
    const double smax = n ? sv[order[0]] : 0.0;
    std::vector<bool> have(n, false);
    for (int c = 0; c < n; ++c) {
        const int j = order[c];
        s[c] = sv[j];
        for (int k = 0; k < n; ++k) vv[k + n * c] = v[k + n * j];
        if (sv[j] > 1e-300 && sv[j] > 1e-15 * smax) {
            for (int k = 0; k < n; ++k) u[k + n * c] = a[k + n * j] / sv[j];
            have[c] = true;
        }
    }
    v = vv;
    // complete missing left vectors
    for (int c = 0; c < n; ++c) {
        if (have[c]) continue;
        int best_e = -1;
        double best_norm = -1;
        Vec best;
        for (int e = 0; e < n; ++e) {
            Vec w(n, 0.0);
            w[e] = 1.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int d = 0; d < n; ++d) {
                    if (!have[d]) continue;
                    double dot = 0;
                    for (int k = 0; k < n; ++k) dot += u[k + n * d] * w[k];
                    for (int k = 0; k < n; ++k) w[k] -= dot * u[k + n * d];
                }
            double nn = 0;
            for (int k = 0; k < n; ++k) nn += w[k] * w[k];
            if (nn > best_norm) {
                best_norm = nn;
                best = w;
                best_e = e;
            }
        }
        (void)best_e;
        const double inv = 1.0 / std::sqrt(best_norm);
        for (int k = 0; k < n; ++k) u[k + n * c] = best[k] * inv;
        have[c] = true;
    }
}

// Upper Cholesky S = R^T R (vbscf/orbitals/preparation/cholesky.*, linalg dsygvd step 1).
bool cholesky_upper(int n, const double* s, Vec& r) {
    r.assign(n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i <= j; ++i) {
            double t = s[i + n * j];
            for (int k = 0; k < i; ++k) t -= r[k + n * i] * r[k + n * j];
            if (i == j) {
                if (!(t > 0.0)) return false;
                r[j + n * j] = std::sqrt(t);
            } else {
                r[i + n * j] = t / r[i + n * i];
            }
        }
    }
    return true;
}

// inverse of an upper-triangular matrix
Vec upper_inverse(int n, const Vec& r) {
    Vec ri(n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        ri[j + n * j] = 1.0 / r[j + n * j];
        for (int i = j - 1; i >= 0; --i) {
            double t = 0;
            for (int k = i + 1; k <= j; ++k) t += r[i + n * k] * ri[k + n * j];
            ri[i + n * j] = -t / r[i + n * i];
        }
    }
    return ri;
}

// Householder tridiagonalisation + implicit QL with Wilkinson-type shifts
// (textbook symmetric eigensolver; stands in for LAPACK dsyevd inside dsygvd).
// On exit: w ascending eigenvalues, z (n x n, columns) orthonormal eigenvectors.
void sym_eigen(int n, Vec a, Vec& w, Vec& z) {
    z = a;  // will hold accumulated transforms
    Vec d(n), e(n);
    // tridiagonalise (z column-major, symmetric)
    auto Z = [&](int i, int j) -> double& { return z[i + n * j]; };
    for (int i = n - 1; i > 0; --i) {
        const int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l > 0) {
            for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
            if (scale == 0.0) {
                e[i] = Z(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    Z(i, k) /= scale;
                    h += Z(i, k) * Z(i, k);
                }
                double f = Z(i, l);
                const double g = f >= 0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                Z(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    Z(j, i) = Z(i, j) / h;
                    double gg = 0.0;
                    for (int k = 0; k <= j; ++k) gg += Z(j, k) * Z(i, k);
                    for (int k = j + 1; k <= l; ++k) gg += Z(k, j) * Z(i, k);
                    e[j] = gg / h;
                    f += e[j] * Z(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    const double ff = Z(i, j);
                    const double gg = e[j] - hh * ff;
                    e[j] = gg;
                    for (int k = 0; k <= j; ++k) Z(j, k) -= (ff * e[k] + gg * Z(i, k));
                }
            }
        } else {
            e[i] = Z(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += Z(i, k) * Z(k, j);
                for (int k = 0; k <= l; ++k) Z(k, j) -= g * Z(k, i);
            }
        }
        d[i] = Z(i, i);
        Z(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) Z(j, i) = Z(i, j) = 0.0;
    }
    // implicit QL on (d, e)
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m != l) {
                if (++iter > 60) throw std::runtime_error("QL iteration did not converge");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * f;
                        Z(k, i) = c * Z(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    // sort ascending
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return d[p] < d[q]; });
    w.assign(n, 0.0);
    Vec zs(n * n);
    for (int c = 0; c < n; ++c) {
        w[c] = d[order[c]];
        for (int k = 0; k < n; ++k) zs[k + n * c] = z[k + n * order[c]];
    }
    z = zs;
}

// Generalised symmetric-definite eigenproblem H x = lambda S x, itype 1, like LAPACKE dsygvd('V','U')
// (linalg/eigensolver.cpp:365-387). Eigenvectors satisfy x^T S x = 1.
int gep(int n, const double* h, const double* s, Vec& w, Vec& x) {
    Vec r;
    if (!cholesky_upper(n, s, r)) return 3;
    const Vec ri = upper_inverse(n, r);
    // M = R^{-T} H R^{-1}
    Vec t(n * n, 0.0), m(n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k <= j; ++k) {
            const double rk = ri[k + n * j];
            for (int i = 0; i < n; ++i) t[i + n * j] += h[i + n * k] * rk;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double acc = 0;
            for (int k = 0; k <= i; ++k) acc += ri[k + n * i] * t[k + n * j];
            m[i + n * j] = acc;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) m[i + n * j] = m[j + n * i] = 0.5 * (m[i + n * j] + m[j + n * i]);
    Vec z;
    sym_eigen(n, m, w, z);
    x.assign(n * n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i) {
            double acc = 0;
            for (int k = i; k < n; ++k) acc += ri[i + n * k] * z[k + n * c];
            x[i + n * c] = acc;
        }
    return 0;
}

// ---------------------------------------------------------------------------
// topology: Rumer/Young generator (input/loading/generated_structures.cpp:131-432,673-716)
// and StructureExpander::expand (vbscf/structures/expansion/expander.cpp:130-558), general path.
// ---------------------------------------------------------------------------
uint64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    k = std::min(k, n - k);
    unsigned __int128 r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return (uint64_t)r;
}

uint64_t tableau_count(int n, int spin) {  // generated_structures.cpp:131-143
    if (n < 0 || spin <= 0 || spin > n + 1) return 0;
    if (n == 0) return 1;
    if ((n - spin + 1) % 2) return 0;
    const int beta = (n - spin + 1) / 2;
    return binom(n, beta) - (beta ? binom(n, beta - 1) : 0);
}

uint64_t class_count(int ion, int ndocc, int nskip, int nact, int mult) {  // :145-167
    const int tab = nact - 2 * ion - nskip + ndocc;
    if (tab < 0) return 0;
    return binom(nact, ion) * binom(nact - ion, nskip + ion - ndocc) * tableau_count(tab, mult);
}

bool single_reference_case(int nact, int nel, int mult) {  // :169-178
    const int nb = (nel - mult + 1) / 2;
    return (nb == 0 && nel == mult - 1 && nel == nact) || (nb == nact && mult == 1);
}

std::vector<int> parse_int_list(const std::string& text) {  // :80-101 (ranges a:b, a-b, a*k)
    std::vector<int> out;
    std::string tok;
    auto flush = [&]() {
        if (tok.empty()) return;
        auto c = tok.find(':'), h = tok.find('-'), st = tok.find('*');
        if (c != std::string::npos || h != std::string::npos) {
            const auto sep = c != std::string::npos ? c : h;
            const int a = std::stoi(tok.substr(0, sep)), b = std::stoi(tok.substr(sep + 1));
            if (b < a) throw std::runtime_error("descending range");
            for (int v = a; v <= b; ++v) out.push_back(v);
        } else if (st != std::string::npos) {
            const int a = std::stoi(tok.substr(0, st)), k = std::stoi(tok.substr(st + 1));
            for (int i = 0; i < k; ++i) out.push_back(a);
        } else {
            out.push_back(std::stoi(tok));
        }
        tok.clear();
    };
    for (char ch : text) {
        if (ch == ',' || ch == ' ' || ch == '\t') flush();
        else tok.push_back(ch);
    }
    flush();
    return out;
}

std::vector<int> ionic_classes(const std::string& kw_in, int nact, int nel, int mult) {  // :180-244
    std::string kw = kw_in;
    for (auto& ch : kw) ch = (char)std::toupper((unsigned char)ch);
    const int ndocc = std::max(0, nel - nact);
    const int rb = (nel - 2 * ndocc - mult + 1) / 2;
    if (rb < 0) throw std::runtime_error("negative reduced beta count");
    std::vector<int> cls;
    if (kw.find("FULL") != std::string::npos) {
        for (int k = 0; k <= rb; ++k) cls.push_back(k);
    } else {
        if (kw.find("COV") != std::string::npos) cls.push_back(0);
        const auto io = kw.find("ION(");
        if (io != std::string::npos) {
            const auto e = kw.find(')', io + 4);
            if (e == std::string::npos) throw std::runtime_error("ION( without ')'");
            for (int v : parse_int_list(kw.substr(io + 4, e - io - 4))) cls.push_back(v);
        } else if (kw.find("ION") != std::string::npos) {
            for (int k = 1; k <= rb; ++k) cls.push_back(k);
        }
    }
    if (cls.empty()) throw std::runtime_error("empty structure class");
    std::sort(cls.begin(), cls.end());
    cls.erase(std::unique(cls.begin(), cls.end()), cls.end());
    if (cls.front() < 0 || cls.back() > rb) throw std::runtime_error("structure class out of range");
    for (auto& c : cls) c += ndocc;
    return cls;
}

// all increasing rows of length k (labels from 1) bounded by the Young maxima (:247-299),
// produced depth-first, which is the same lexicographic order as the incrementing walk.
void young_rows(int k, int nb, int norb, std::vector<std::vector<int>>& out) {
    if (k <= 0) return;
    if (norb <= 0 || norb < k) throw std::runtime_error("invalid orbital count");
    std::vector<int> mx(k, 0);
    if (k - nb + 1 == 1) {
        mx[nb - 1] = norb - 1;
    } else {
        mx[k - 1] = norb;
        for (int i = k - 1; i >= nb + 1; --i) mx[i - 1] = mx[i] - 1;
        if (nb > 0) mx[nb - 1] = mx[nb] - 2;
    }
    for (int i = nb - 1; i >= 1; --i) mx[i - 1] = mx[i] - 2;
    std::vector<int> row(k);
    // The maxima grow by 1 or 2 per position, so every bounded prefix extends and the
    // reference's "increment the rightmost free slot, reset the suffix" walk visits exactly
    // the rows a depth-first lexicographic enumeration visits, in the same order.
    struct Rec {
        static void go(int pos, int k, const std::vector<int>& mx, std::vector<int>& row,
                       std::vector<std::vector<int>>& out) {
            if (pos == k) {
                out.push_back(row);
                return;
            }
            const int lo = pos == 0 ? 1 : row[pos - 1] + 1;
            for (int v = lo; v <= mx[pos]; ++v) {
                row[pos] = v;
                go(pos + 1, k, mx, row, out);
            }
        }
    };
    Rec::go(0, k, mx, row, out);
}

// beta row compatible with an alpha Young row (:301-316)
bool beta_ok(const std::vector<int>& a, const std::vector<int>& b, int na, int nb) {
    for (int i = 0; i < nb; ++i) {
        if (b[i] < a[i]) return false;
        for (int j = i; j < na; ++j)
            if (b[i] == a[j]) return false;
    }
    return true;
}

// (alpha_row, beta_row) pairs of a covalent pattern, rows padded to `stride` (:334-391)
void covalent_pairs(int na, int nb, int norb, int stride,
                    std::vector<std::pair<std::vector<int>, std::vector<int>>>& out) {
    if (na <= 0) return;
    auto pad = [&](std::vector<int>& b) {  // :318-332 open-shell dummy partners
        for (int i = nb; i < na; ++i) b[i] = norb + na - i;
    };
    std::vector<std::vector<int>> arows;
    if (nb <= 0) {
        young_rows(na, 0, norb, arows);
        for (auto& a : arows) {
            a.resize(stride, 0);
            out.push_back({a, std::vector<int>(stride, 0)});
        }
        return;
    }
    young_rows(na, nb, norb, arows);
    if (na + nb == norb) {
        for (auto& a : arows) {
            std::vector<int> b(stride, 0);
            int next = 0;
            for (int lab = 1; lab <= norb; ++lab)
                if (std::find(a.begin(), a.begin() + na, lab) == a.begin() + na) b[next++] = lab;
            pad(b);
            a.resize(stride, 0);
            out.push_back({a, b});
        }
        return;
    }
    std::vector<std::vector<int>> brows;
    young_rows(nb, 0, norb, brows);
    for (auto& a : arows) {
        for (const auto& cand : brows) {
            if (!beta_ok(a, cand, na, nb)) continue;
            std::vector<int> b(stride, 0);
            for (int i = 0; i < nb; ++i) b[i] = cand[i];
            pad(b);
            std::vector<int> ar = a;
            ar.resize(stride, 0);
            out.push_back({ar, b});
        }
    }
}

// Rumer bond matching (:393-432): each alpha label, from the largest down, is matched to the
// first remaining beta label not below it; alpha is then rewritten in beta-partner order.
void reorder_bonds(std::vector<int>& alpha, const std::vector<int>& beta, int nvec, int nrow, int norb) {
    std::vector<int> partner_pos(3 * norb + 2, 0), beta_pos(3 * norb + 2, 0);
    std::vector<int> work(beta.begin(), beta.begin() + nrow);
    for (int v = 0; v < nvec; ++v) beta_pos[work[v] - 1] = v + 1;
    int remaining = nvec;
    for (int v = nvec; v >= 1; --v) {
        const int amax = alpha[remaining - 1];
        int m = 0;
        for (; m < remaining; ++m)
            if (amax <= work[m]) {
                partner_pos[amax - 1] = beta_pos[work[m] - 1];
                break;
            }
        if (m < remaining) work.erase(work.begin() + m);
        --remaining;
    }
    std::vector<int> out(nvec, 0);
    for (int v = 0; v < nvec; ++v) out[partner_pos[alpha[v] - 1] - 1] = alpha[v];
    for (int v = 0; v < nvec; ++v) alpha[v] = out[v];
}

struct RawStructures {
    int n_el = 0;
    std::vector<std::vector<int>> rows;  // 1-based orbital labels, bond pairs first
};

// build_generated_raw_structures_from_structure_class (:673-716), no inactive orbitals
RawStructures generate(const std::string& kw, int nact, int nel, int mult) {
    RawStructures raw;
    raw.n_el = nel;
    const int nb_act = (nel - mult + 1) / 2, na_act = nel - nb_act;
    if (nb_act < 0 || nb_act > nact) throw std::runtime_error("bad beta count");
    if (single_reference_case(nact, nel, mult)) {  // :446-468
        std::vector<int> row(nel, 0);
        for (int i = 0; i < nb_act; ++i) row[2 * i] = row[2 * i + 1] = i + 1;
        for (int o = 0; o < mult - 1; ++o) row[2 * nb_act + o] = nb_act + o + 1;
        raw.rows.push_back(row);
        return raw;
    }
    for (int ion : ionic_classes(kw, nact, nel, mult)) {  // :534-619
        const int nca = na_act - ion, ncb = nb_act - ion;
        if (ion < 0 || ion > nb_act || nca < 0 || ncb < 0) throw std::runtime_error("bad ionic class");
        auto emit = [&](const std::vector<int>& ions, const std::vector<int>& rem,
                        const std::vector<int>& a, const std::vector<int>& b) {  // :470-509
            std::vector<int> row(nel, 0);
            if (ion > 0) {
                for (int i = 0; i < ion; ++i) row[2 * i] = row[2 * i + 1] = ions[i];
                for (int i = 0; i < ncb; ++i) {
                    row[2 * ion + 2 * i] = rem[a[i] - 1];
                    row[2 * ion + 2 * i + 1] = rem[b[i] - 1];
                }
                for (int i = ncb; i < nca; ++i) row[2 * ion + ncb + i] = rem[a[i] - 1];
            } else {
                for (int i = 0; i < ncb; ++i) {
                    row[2 * i] = a[i];
                    row[2 * i + 1] = b[i];
                }
                for (int i = ncb; i < nca; ++i) row[ncb + i] = a[i];
            }
            raw.rows.push_back(row);
        };
        if (ion > 0) {
            std::vector<std::vector<int>> ion_rows;
            young_rows(ion, 0, nact, ion_rows);
            for (const auto& ir : ion_rows) {
                std::vector<int> used(nact + 1, 0), rem;
                for (int i = 0; i < ion; ++i) used[ir[i]] = 1;
                for (int lab = 1; lab <= nact; ++lab)
                    if (!used[lab]) rem.push_back(lab);
                if (nca <= 0) {
                    emit(ir, rem, {}, {});
                    continue;
                }
                std::vector<std::pair<std::vector<int>, std::vector<int>>> pairs;
                covalent_pairs(nca, ncb, nact - ion, na_act, pairs);
                for (auto& [a, b] : pairs) {
                    std::vector<int> ra = a;
                    if (ncb > 0) reorder_bonds(ra, b, nca, na_act, nact);
                    emit(ir, rem, ra, b);
                }
            }
        } else {
            std::vector<std::pair<std::vector<int>, std::vector<int>>> pairs;
            covalent_pairs(na_act, nb_act, nact, na_act, pairs);
            for (auto& [a, b] : pairs) {
                std::vector<int> ra = a;
                if (nb_act > 0) reorder_bonds(ra, b, na_act, na_act, nact);
                emit({}, {}, ra, b);
            }
        }
    }
    return raw;
}

uint64_t count_structures(const std::string& kw, int nact, int nel, int mult) {  // :621-658
    const int nb_act = (nel - mult + 1) / 2;
    if (nb_act < 0 || nb_act > nact) throw std::runtime_error("bad beta count");
    if (single_reference_case(nact, nel, mult)) return 1;
    const int nskip = std::max(0, nact - nel), ndocc = std::max(0, nel - nact);
    uint64_t total = 0;
    for (int ion : ionic_classes(kw, nact, nel, mult)) total += class_count(ion, ndocc, nskip, nact, mult);
    return total;
}

struct Expansion {
    int n_struct = 0;
    std::vector<uint64_t> amask, bmask;          // canonical (ascending mask) unique strings
    std::vector<int> det_a, det_b;                // per determinant, canonical order
    std::vector<int64_t> off;                     // CSR over determinants
    std::vector<int> term_s;
    std::vector<double> term_c;
};

// inversion parity of an occupied list relative to ascending order (expander.cpp:33-71)
int list_parity(const std::vector<int>& v) {
    int inv = 0;
    for (size_t i = 0; i < v.size(); ++i)
        for (size_t j = i + 1; j < v.size(); ++j)
            if (v[i] > v[j]) ++inv;
            else if (v[i] == v[j]) throw std::invalid_argument("duplicate orbital in an occupied spin string");
    return (inv & 1) ? -1 : 1;
}

uint64_t list_mask(const std::vector<int>& v) {
    uint64_t m = 0;
    for (int o : v) m |= (uint64_t)1 << o;
    return m;
}

// StructureExpander::expand (expander.cpp:287-558) via the general enumerate_structure path (:130-186)
Expansion expand(const RawStructures& raw, int mult) {
    const int nel = raw.n_el, nopen = mult - 1, npaired = nel - nopen;
    if (npaired < 0 || npaired % 2) throw std::invalid_argument("electron/spin counts inconsistent");
    const int npairs = npaired / 2;
    struct Emitted { uint64_t a, b; int s; double c; };
    std::vector<Emitted> all;
    for (size_t s = 0; s < raw.rows.size(); ++s) {
        std::vector<int> act(nel);
        for (int i = 0; i < nel; ++i) act[i] = raw.rows[s][i] - 1;
        std::vector<int> var;
        for (int p = 0; p < npairs; ++p)
            if (act[2 * p] != act[2 * p + 1]) var.push_back(p);
        const uint64_t nchoice = (uint64_t)1 << var.size();
        for (uint64_t ch = 0; ch < nchoice; ++ch) {  // binary counting, bit 0 fastest (:171-178)
            std::vector<int> al, be;
            size_t vi = 0;
            for (int p = 0; p < npairs; ++p) {
                bool sw = false;
                if (vi < var.size() && var[vi] == p) {
                    sw = (ch >> vi) & 1;
                    ++vi;
                }
                al.push_back(act[2 * p + (sw ? 1 : 0)]);
                be.push_back(act[2 * p + (sw ? 0 : 1)]);
            }
            for (int i = 2 * npairs; i < nel; ++i) al.push_back(act[i]);
            const int sg = list_parity(al) * list_parity(be);
            all.push_back({list_mask(al), list_mask(be), (int)s, (double)sg});
        }
    }
    // intern strings / pairs in discovery order, then canonicalise
    std::vector<uint64_t> ua, ub;
    std::unordered_map<uint64_t, int> ia, ib;
    std::vector<std::pair<int, int>> pairs;
    std::unordered_map<uint64_t, int> ipair;  // key = a_id * 2^32 + b_id
    std::vector<int> pair_of(all.size());
    for (size_t k = 0; k < all.size(); ++k) {
        auto fa = ia.find(all[k].a);
        int aid = fa == ia.end() ? (ia[all[k].a] = (int)ua.size(), ua.push_back(all[k].a), (int)ua.size() - 1) : fa->second;
        auto fb = ib.find(all[k].b);
        int bid = fb == ib.end() ? (ib[all[k].b] = (int)ub.size(), ub.push_back(all[k].b), (int)ub.size() - 1) : fb->second;
        const uint64_t key = ((uint64_t)aid << 32) | (uint32_t)bid;
        auto fp = ipair.find(key);
        if (fp == ipair.end()) {
            ipair[key] = (int)pairs.size();
            pair_of[k] = (int)pairs.size();
            pairs.push_back({aid, bid});
        } else {
            pair_of[k] = fp->second;
        }
    }
    auto canon = [](const std::vector<uint64_t>& m, std::vector<int>& rank, std::vector<uint64_t>& sorted) {
        std::vector<int> order(m.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int p, int q) { return m[p] < m[q]; });
        rank.assign(m.size(), 0);
        sorted.resize(m.size());
        for (size_t i = 0; i < order.size(); ++i) {
            rank[order[i]] = (int)i;
            sorted[i] = m[order[i]];
        }
    };
    std::vector<int> ra, rb;
    Expansion ex;
    ex.n_struct = (int)raw.rows.size();
    canon(ua, ra, ex.amask);
    canon(ub, rb, ex.bmask);
    std::vector<int> porder(pairs.size());
    std::iota(porder.begin(), porder.end(), 0);
    std::stable_sort(porder.begin(), porder.end(), [&](int p, int q) {
        const int a1 = ra[pairs[p].first], a2 = ra[pairs[q].first];
        if (a1 != a2) return a1 < a2;
        return rb[pairs[p].second] < rb[pairs[q].second];
    });
    std::vector<int> new_index(pairs.size());
    for (size_t i = 0; i < porder.size(); ++i) new_index[porder[i]] = (int)i;
    // per determinant: terms in (structure, emission) order, then coalesced (types.hpp:46-97)
    std::vector<std::vector<std::pair<int, double>>> rows(pairs.size());
    for (size_t k = 0; k < all.size(); ++k) rows[new_index[pair_of[k]]].push_back({all[k].s, all[k].c});
    ex.off.push_back(0);
    for (size_t d = 0; d < rows.size(); ++d) {
        ex.det_a.push_back(ra[pairs[porder[d]].first]);
        ex.det_b.push_back(rb[pairs[porder[d]].second]);
        auto& r = rows[d];
        std::stable_sort(r.begin(), r.end(), [](auto& x, auto& y) { return x.first < y.first; });
        size_t i = 0;
        while (i < r.size()) {
            const int st = r[i].first;
            double c = 0;
            while (i < r.size() && r[i].first == st) c += r[i++].second;
            if (c != 0.0) {
                ex.term_s.push_back(st);
                ex.term_c.push_back(c);
            }
        }
        ex.off.push_back((int64_t)ex.term_s.size());
    }
    return ex;
}


// ---------------------------------------------------------------------------
// determinant-pair algebra
// ---------------------------------------------------------------------------
constexpr double kThreshold = 1.0e-12;  // DeterminantOverlapResolver default (overlap.hpp:33)

struct SpinBlock {
    int n = 0;
    std::vector<int> occL, occR;
    Vec x;          // X(r,l) = S_act(R_r, L_l): rows right determinant, columns left (types.hpp:76-82)
    int nullity = 0;
    double det = 1.0;
    bool have_inv = false;
    Vec inv;        // Y = X^{-1}, Y(l,r)
    bool have_svd = false;
    Vec u, s, v;
    double parity = 1.0;
    Vec cof1;       // first-order cofactor matrix (rows right, columns left); zero when nullity >= 2
};

std::vector<int> occ_of(uint64_t m) {
    std::vector<int> o;
    for (int b = 0; b < 64; ++b)
        if (m >> b & 1) o.push_back(b);
    return o;
}

// DeterminantOverlapResolver::resolve_matrix (overlap.cpp:50-110)
void resolve(SpinBlock& b) {
    const int n = b.n;
    if (n == 0) {
        b.det = 1.0;
        b.nullity = 0;
        return;
    }
    LuResult lu = fullpiv_lu(n, b.x.data(), kThreshold);
    if (lu.rank == n) {
        b.nullity = 0;
        b.det = lu.det;
        b.inv = std::move(lu.inverse);
        b.have_inv = true;
    } else {
        svd_jacobi(n, b.x.data(), b.u, b.s, b.v);
        b.have_svd = true;
        b.nullity = 0;
        for (int i = 0; i < n; ++i)
            if (b.s[i] < kThreshold) ++b.nullity;
        b.parity = (det_partial(n, b.u) * det_partial(n, b.v) >= 0.0) ? 1.0 : -1.0;
        if (b.nullity == 0) {
            double p = b.parity;
            for (int i = 0; i < n; ++i) p *= b.s[i];
            b.det = p;
            b.inv.assign(n * n, 0.0);
            for (int l = 0; l < n; ++l)
                for (int r = 0; r < n; ++r) {
                    double t = 0;
                    for (int k = 0; k < n; ++k) t += b.v[l + n * k] * b.u[r + n * k] / b.s[k];
                    b.inv[l + n * r] = t;
                }
            b.have_inv = true;
        } else {
            b.det = 0.0;
        }
    }
    // build_first_order_cofactor_matrix (contractions.cpp:172-236)
    b.cof1.assign(n * n, 0.0);
    const auto& x = b.x;
    auto X = [&](int i, int j) { return x[i + n * j]; };
    if (n <= 3) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double value = 1.0;
                if (n == 2) {
                    value = X(1 - i, 1 - j);
                } else if (n == 3) {
                    const int r = i == 0 ? 1 : 0, s2 = i == 2 ? 1 : 2;
                    const int c = j == 0 ? 1 : 0, d = j == 2 ? 1 : 2;
                    const double product = X(r, d) * X(s2, c);
                    value = std::fma(X(r, c), X(s2, d), -product) - std::fma(X(r, d), X(s2, c), -product);
                }
                b.cof1[i + n * j] = ((i + j) % 2) ? -value : value;
            }
        return;
    }
    if (b.nullity == 0) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) b.cof1[i + n * j] = b.det * b.inv[j + n * i];
    } else if (b.nullity == 1) {
        double pref = b.parity;
        for (int k = 0; k < n - 1; ++k) pref *= b.s[k];
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) b.cof1[i + n * j] = pref * b.u[i + n * (n - 1)] * b.v[j + n * (n - 1)];
    }
}

double minor_det(int n, const Vec& x, int r1, int r2, int l1, int l2) {
    const int m = n - 2;
    if (m == 0) return 1.0;
    Vec a(m * m);
    int rr = 0;
    for (int r = 0; r < n; ++r) {
        if (r == r1 || r == r2) continue;
        int cc = 0;
        for (int l = 0; l < n; ++l) {
            if (l == l1 || l == l2) continue;
            a[rr + m * cc] = x[r + n * l];
            ++cc;
        }
        ++rr;
    }
    return det_partial(m, a);
}

// calc_second_order_cofactor (contractions.cpp:901-983)
double cof2(const SpinBlock& b, int r1, int r2, int l1, int l2) {
    const int n = b.n;
    if (r1 == r2 || l1 == l2 || n < 2) return 0.0;
    const double sign = ((r1 + r2 + l1 + l2) % 2) ? -1.0 : 1.0;
    if (n == 2) return sign;
    if (n == 3) return sign * b.x[(3 - r1 - r2) + n * (3 - l1 - l2)];
    if (b.nullity == 0 && b.det != 0.0) return sign * minor_det(n, b.x, r1, r2, l1, l2);
    const auto U = [&](int i, int k) { return b.u[i + n * k]; };
    const auto V = [&](int i, int k) { return b.v[i + n * k]; };
    if (b.nullity == 1) {
        const int z = n - 1;
        double pref = b.parity;
        for (int k = 0; k < n - 1; ++k) pref *= b.s[k];
        double acc = 0;
        for (int k = 0; k < n - 1; ++k) {
            const double rw = U(r1, k) * U(r2, z) - U(r2, k) * U(r1, z);
            const double lw = V(l1, k) * V(l2, z) - V(l2, k) * V(l1, z);
            acc += rw * lw / b.s[k];
        }
        return pref * acc;
    }
    if (b.nullity == 2) {
        const int z1 = n - 2, z2 = n - 1;
        double pref = b.parity;
        for (int k = 0; k < n - 2; ++k) pref *= b.s[k];
        const double rw = U(r1, z1) * U(r2, z2) - U(r2, z1) * U(r1, z2);
        const double lw = V(l1, z1) * V(l2, z2) - V(l2, z1) * V(l1, z2);
        return pref * rw * lw;
    }
    return 0.0;
}

struct Integrals {
    int n = 0;          // active orbitals
    Vec sact, hact;     // n x n raw frame
    Vec g;              // packed pair-of-pairs raw kernel
    double G(int64_t p, int64_t q) const { return g[pp(p, q)]; }
};

// calc_same_spin_hamiltonian_impl (determinants/algebra/hamiltonian.cpp:35-236)
double same_spin_h(const SpinBlock& b, const Integrals& I) {
    const int n = b.n, N = I.n;
    if (n == 0 || b.nullity >= 3) return 0.0;
    double h = 0.0;
    if (b.nullity <= 1)
        for (int l = 0; l < n; ++l)
            for (int r = 0; r < n; ++r) h += I.hact[b.occR[r] + N * b.occL[l]] * b.cof1[r + n * l];
    if (n < 2) return h;
    const auto& R = b.occR;
    const auto& L = b.occL;
    if (b.nullity == 0 && n > 3) {
        double two = 0.0;
        for (int l1 = 0; l1 < n - 1; ++l1)
            for (int r1 = 0; r1 < n - 1; ++r1) {
                const double c11 = b.cof1[r1 + n * l1];
                const int64_t p11 = pp(R[r1], L[l1]);
                for (int l2 = l1 + 1; l2 < n; ++l2) {
                    const double c12 = b.cof1[r1 + n * l2];
                    const int64_t p12 = pp(R[r1], L[l2]);
                    double acc = 0.0;
                    for (int r2 = r1 + 1; r2 < n; ++r2) {
                        const double m2 = c11 * b.cof1[r2 + n * l2] - c12 * b.cof1[r2 + n * l1];
                        acc += (I.G(p11, pp(R[r2], L[l2])) - I.G(p12, pp(R[r2], L[l1]))) * m2;
                    }
                    two += acc;
                }
            }
        return h + two / b.det;
    }
    for (int l1 = 0; l1 < n - 1; ++l1)
        for (int r1 = 0; r1 < n - 1; ++r1)
            for (int l2 = l1 + 1; l2 < n; ++l2)
                for (int r2 = r1 + 1; r2 < n; ++r2) {
                    const double c2 = cof2(b, r1, r2, l1, l2);
                    h += (I.G(pp(R[r1], L[l1]), pp(R[r2], L[l2])) - I.G(pp(R[r1], L[l2]), pp(R[r2], L[l1]))) * c2;
                }
    return h;
}

// evaluate_opposite_spin_coulomb_coupling (determinants/pairs/evaluator.cpp:95-148)
double opposite_k(const SpinBlock& a, const SpinBlock& b, const Integrals& I) {
    if (a.n == 0 || b.n == 0 || a.nullity >= 2 || b.nullity >= 2) return 0.0;
    double k = 0.0;
    for (int la = 0; la < a.n; ++la)
        for (int ra = 0; ra < a.n; ++ra) {
            const double ca = a.cof1[ra + a.n * la];
            const int64_t pa = pp(a.occR[ra], a.occL[la]);
            double acc = 0.0;
            for (int lb = 0; lb < b.n; ++lb)
                for (int rb = 0; rb < b.n; ++rb)
                    acc += b.cof1[rb + b.n * lb] * I.G(pp(b.occR[rb], b.occL[lb]), pa);
            k += ca * acc;
        }
    return k;
}

SpinBlock make_block(uint64_t mL, uint64_t mR, const Integrals& I) {
    SpinBlock b;
    b.occL = occ_of(mL);
    b.occR = occ_of(mR);
    b.n = (int)b.occL.size();
    if ((int)b.occR.size() != b.n) throw std::invalid_argument("determinant electron counts differ");
    b.x.assign(b.n * b.n, 0.0);
    for (int l = 0; l < b.n; ++l)
        for (int r = 0; r < b.n; ++r) b.x[r + b.n * l] = I.sact[b.occR[r] + I.n * b.occL[l]];
    resolve(b);
    return b;
}

// DeterminantPairEvaluator::combine_spin_pair_evaluations (evaluator.cpp:354-384)
void pair_forward(uint64_t aL, uint64_t aR, uint64_t bL, uint64_t bR, const Integrals& I, double& S, double& H) {
    const SpinBlock A = make_block(aL, aR, I), B = make_block(bL, bR, I);
    const double ha = same_spin_h(A, I), hb = same_spin_h(B, I);
    S = A.det * B.det;
    H = ha * B.det + hb * A.det + opposite_k(A, B, I);
}

// ---------------------------------------------------------------------------
// backward of one determinant pair: F = wH * H_dd' + wS * S_dd'
// regular pairs: compute_{same,opposite}_spin_original_phi (contractions.cpp:239-391) and
// build_regular_same_spin_overlap_hamiltonian_gradient (contractions.cpp:395-411);
// singular pairs: the pair function is multilinear in the singular values of each block taken in
// its SVD chart, so its exact derivatives at the null point are the mean over the 2^k chart nodes
// where each null singular value is replaced by +tau or -tau (regular there).
// ---------------------------------------------------------------------------
struct Grad {
    Vec ds, dh, dg;  // adjoints of sact (n x n, indexed [R + n*L]), hact, packed g
    void init(int n, size_t ng) {
        ds.assign(n * n, 0.0);
        dh.assign(n * n, 0.0);
        dg.assign(ng, 0.0);
    }
    void add(const Grad& o) {
        for (size_t i = 0; i < ds.size(); ++i) ds[i] += o.ds[i];
        for (size_t i = 0; i < dh.size(); ++i) dh[i] += o.dh[i];
        for (size_t i = 0; i < dg.size(); ++i) dg[i] += o.dg[i];
    }
};

// regular pair backward with explicit blocks (Y = inverse, det given)
void pair_backward_regular(const SpinBlock* blk[2], const Vec* Y[2], const double det[2], const Integrals& I,
                           double wH, double wS, double scale, Grad& g) {
    const int N = I.n;
    Vec Gm[2];
    double phi = 0.0;
    for (int sp = 0; sp < 2; ++sp) {
        const SpinBlock& b = *blk[sp];
        const int n = b.n;
        Gm[sp].assign(n * n, 0.0);
        if (n == 0) continue;
        const Vec& y = *Y[sp];
        const auto& R = b.occR;
        const auto& L = b.occL;
        for (int l = 0; l < n; ++l)
            for (int r = 0; r < n; ++r) {
                const double hv = I.hact[R[r] + N * L[l]];
                phi += hv * y[l + n * r];
                Gm[sp][l + n * r] += hv;
            }
        for (int l1 = 0; l1 < n - 1; ++l1)
            for (int r1 = 0; r1 < n - 1; ++r1)
                for (int l2 = l1 + 1; l2 < n; ++l2)
                    for (int r2 = r1 + 1; r2 < n; ++r2) {
                        const double v = I.G(pp(R[r1], L[l1]), pp(R[r2], L[l2])) - I.G(pp(R[r1], L[l2]), pp(R[r2], L[l1]));
                        const double x11 = y[l1 + n * r1], x22 = y[l2 + n * r2];
                        const double x12 = y[l2 + n * r1], x21 = y[l1 + n * r2];
                        phi += v * (x11 * x22 - x12 * x21);
                        Gm[sp][l1 + n * r1] += v * x22;
                        Gm[sp][l2 + n * r2] += v * x11;
                        Gm[sp][l2 + n * r1] -= v * x21;
                        Gm[sp][l1 + n * r2] -= v * x12;
                    }
    }
    const SpinBlock& A = *blk[0];
    const SpinBlock& B = *blk[1];
    if (A.n > 0 && B.n > 0) {
        const Vec& ya = *Y[0];
        const Vec& yb = *Y[1];
        for (int la = 0; la < A.n; ++la)
            for (int ra = 0; ra < A.n; ++ra) {
                const double yav = ya[la + A.n * ra];
                const int64_t pa = pp(A.occR[ra], A.occL[la]);
                for (int lb = 0; lb < B.n; ++lb)
                    for (int rb = 0; rb < B.n; ++rb) {
                        const double v = I.G(pp(B.occR[rb], B.occL[lb]), pa);
                        const double ybv = yb[lb + B.n * rb];
                        phi += v * yav * ybv;
                        Gm[0][la + A.n * ra] += v * ybv;
                        Gm[1][lb + B.n * rb] += v * yav;
                    }
            }
    }
    const double D = det[0] * det[1];
    const double c = scale * D;  // overall multiplier of dPhi terms
    // overlap-block gradient: D [ (wH*phi + wS) Y^T - wH Y^T Gm Y^T ]
    for (int sp = 0; sp < 2; ++sp) {
        const SpinBlock& b = *blk[sp];
        const int n = b.n;
        if (n == 0) continue;
        const Vec& y = *Y[sp];
        Vec t(n * n, 0.0), yt_g_yt(n * n, 0.0);
        // t = Gm Y^T  -> t(l, r') = sum_r Gm(l, r) Y(r', r)   [Y(l,r) indexing: y[l + n*r]]
        for (int l = 0; l < n; ++l)
            for (int rp = 0; rp < n; ++rp) {
                double acc = 0;
                for (int r = 0; r < n; ++r) acc += Gm[sp][l + n * r] * y[rp + n * r];
                t[l + n * rp] = acc;
            }
        // Y^T t -> (r, r') = sum_l Y(l, r) t(l, r')
        for (int r = 0; r < n; ++r)
            for (int rp = 0; rp < n; ++rp) {
                double acc = 0;
                for (int l = 0; l < n; ++l) acc += y[l + n * r] * t[l + n * rp];
                yt_g_yt[r + n * rp] = acc;
            }
        for (int r = 0; r < n; ++r)
            for (int l = 0; l < n; ++l) {
                // gradient wrt X(r, l); Y^T(r,l) = Y(l,r)
                const double gx = c * ((wH * phi + wS) * y[l + n * r] - wH * yt_g_yt[r + n * l]);
                g.ds[b.occR[r] + N * b.occL[l]] += gx;
            }
        // one-electron adjoint: wH * D * Y(l, r) at h(R_r, L_l)
        for (int l = 0; l < n; ++l)
            for (int r = 0; r < n; ++r) g.dh[b.occR[r] + N * b.occL[l]] += c * wH * y[l + n * r];
        // same-spin two-electron adjoint
        const auto& R = b.occR;
        const auto& L = b.occL;
        for (int l1 = 0; l1 < n - 1; ++l1)
            for (int r1 = 0; r1 < n - 1; ++r1)
                for (int l2 = l1 + 1; l2 < n; ++l2)
                    for (int r2 = r1 + 1; r2 < n; ++r2) {
                        const double m = y[l1 + n * r1] * y[l2 + n * r2] - y[l2 + n * r1] * y[l1 + n * r2];
                        g.dg[pp(pp(R[r1], L[l1]), pp(R[r2], L[l2]))] += c * wH * m;
                        g.dg[pp(pp(R[r1], L[l2]), pp(R[r2], L[l1]))] -= c * wH * m;
                    }
    }
    if (A.n > 0 && B.n > 0) {
        const Vec& ya = *Y[0];
        const Vec& yb = *Y[1];
        for (int la = 0; la < A.n; ++la)
            for (int ra = 0; ra < A.n; ++ra) {
                const double yav = ya[la + A.n * ra];
                const int64_t pa = pp(A.occR[ra], A.occL[la]);
                for (int lb = 0; lb < B.n; ++lb)
                    for (int rb = 0; rb < B.n; ++rb)
                        g.dg[pp(pp(B.occR[rb], B.occL[lb]), pa)] += c * wH * yav * yb[lb + B.n * rb];
            }
    }
}

void pair_backward(uint64_t aL, uint64_t aR, uint64_t bL, uint64_t bR, const Integrals& I, double wH, double wS,
                   Grad& g) {
    SpinBlock A = make_block(aL, aR, I), B = make_block(bL, bR, I);
    const SpinBlock* blk[2] = {&A, &B};
    int k[2] = {A.nullity, B.nullity};
    if (k[0] == 0 && k[1] == 0) {
        const Vec* Y[2] = {&A.inv, &B.inv};
        const double det[2] = {A.det, B.det};
        pair_backward_regular(blk, Y, det, I, wH, wS, 1.0, g);
        return;
    }
    if (k[0] >= 4 || k[1] >= 4) return;  // every derivative of the pair function vanishes there
    const int ktot = k[0] + k[1];
    const double scale = 1.0 / double(1 << ktot);
    Vec Ynode[2];
    for (int node = 0; node < (1 << ktot); ++node) {
        double det[2];
        int bit = 0;
        for (int sp = 0; sp < 2; ++sp) {
            const SpinBlock& b = *blk[sp];
            const int n = b.n;
            if (k[sp] == 0) {
                Ynode[sp] = b.inv;
                det[sp] = b.det;
                continue;
            }
            const double tau = b.s[0] >= kThreshold ? b.s[0] : 1.0;
            Vec sn = b.s;
            for (int i = n - k[sp]; i < n; ++i) sn[i] = ((node >> bit++) & 1) ? -tau : tau;
            double d = b.parity;
            for (int i = 0; i < n; ++i) d *= sn[i];
            det[sp] = d;
            Ynode[sp].assign(n * n, 0.0);  // Y = V diag(1/sn) U^T
            for (int l = 0; l < n; ++l)
                for (int r = 0; r < n; ++r) {
                    double t = 0;
                    for (int q = 0; q < n; ++q) t += b.v[l + n * q] * b.u[r + n * q] / sn[q];
                    Ynode[sp][l + n * r] = t;
                }
        }
        const Vec* Y[2] = {&Ynode[0], &Ynode[1]};
        pair_backward_regular(blk, Y, det, I, wH, wS, scale, g);
    }
}

// ---------------------------------------------------------------------------
// one full evaluation
// ---------------------------------------------------------------------------
struct Problem {
    int nbf, nact, naux;
    const double *S, *h, *B, *C;
    const uint8_t* support;
    int nstr, nastr, nbstr, ndet;
    const uint64_t *am, *bm;
    const int32_t *da, *db;
    const int64_t* off;
    const int32_t* ts;
    const double* tc;
    double enuc;
    // selected states and their weights (normalised here); n_states == 0: the lowest root alone
    int n_states = 0;
    const int32_t* sidx = nullptr;
    const double* sw = nullptr;
};

struct Out {
    double energy = 0;
    Vec grad, H, Sm, coef, sact, hact, g, cn, hdet, sdet;
    Vec eigenvalues;
    Vec state_energies;  // E11 + lambda_s + E_nuc per selected state
};

int evaluate(const Problem& p, Out& o, bool want_grad, std::string& err) {
    const int nbf = p.nbf, n = p.nact, naux = p.naux;
    const int64_t nbp = (int64_t)nbf * (nbf + 1) / 2;
    const int np = n * (n + 1) / 2;
    const size_t ng = (size_t)np * (np + 1) / 2;
    // normalize_sparse_orbitals (orbitals/preparation/preparer.cpp:45-94)
    Vec cn(p.C, p.C + (size_t)n * nbf), nu2(n);
    if (p.support)
        for (size_t i = 0; i < cn.size(); ++i)
            if (!p.support[i]) cn[i] = 0.0;
    for (int k = 0; k < n; ++k) {
        double q = 0;
        for (int a = 0; a < nbf; ++a) {
            const double ca = cn[k * nbf + a];
            if (ca == 0.0) continue;
            for (int b = 0; b < nbf; ++b) q += ca * cn[k * nbf + b] * p.S[a + (size_t)nbf * b];
        }
        if (!std::isfinite(q) || q <= std::numeric_limits<double>::epsilon()) {
            err = "orbital normalization failed for orbital " + std::to_string(k);
            return 2;
        }
        nu2[k] = q;
        const double f = std::sqrt(1.0 / q);
        for (int a = 0; a < nbf; ++a) cn[k * nbf + a] *= f;
    }
    auto Cn = [&](int mu, int k) { return cn[(size_t)k * nbf + mu]; };
    // S_act = Cn^T S Cn (preparer.cpp:371-375); h via the same congruence
    Vec sc((size_t)nbf * n, 0.0), hc((size_t)nbf * n, 0.0);
    for (int k = 0; k < n; ++k)
        for (int b = 0; b < nbf; ++b) {
            const double cb = Cn(b, k);
            if (cb == 0.0) continue;
            for (int a = 0; a < nbf; ++a) {
                sc[a + (size_t)nbf * k] += p.S[a + (size_t)nbf * b] * cb;
                hc[a + (size_t)nbf * k] += p.h[a + (size_t)nbf * b] * cb;
            }
        }
    Vec sact(n * n), hraw_direct(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s1 = 0, h1 = 0;
            for (int a = 0; a < nbf; ++a) {
                s1 += Cn(a, i) * sc[a + (size_t)nbf * j];
                h1 += Cn(a, i) * hc[a + (size_t)nbf * j];
            }
            sact[i + n * j] = s1;
            hraw_direct[i + n * j] = h1;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) sact[i + n * j] = sact[j + n * i] = 0.5 * (sact[i + n * j] + sact[j + n * i]);
    // active frame R = chol(S_act) upper, Cbar = Cn R^{-1} (vbscf/integrals/active/space.cpp:99-113)
    Vec R;
    if (!cholesky_upper(n, sact.data(), R)) {
        err = "active orbital overlap is not positive definite";
        return 2;
    }
    const Vec Ri = upper_inverse(n, R);
    Vec cbar((size_t)nbf * n, 0.0);  // cbar[mu + nbf*q]
    for (int q = 0; q < n; ++q)
        for (int k = 0; k <= q; ++k) {
            const double rk = Ri[k + n * q];
            for (int mu = 0; mu < nbf; ++mu) cbar[mu + (size_t)nbf * q] += Cn(mu, k) * rk;
        }
    // hbar = Cbar^T h Cbar (space.cpp:128-131), then raw h = R^T hbar R (integrals/active/raw.hpp:9-11)
    Vec hbar(n * n, 0.0);
    {
        Vec t((size_t)nbf * n, 0.0);
        for (int q = 0; q < n; ++q)
            for (int b = 0; b < nbf; ++b) {
                const double cb = cbar[b + (size_t)nbf * q];
                for (int a = 0; a < nbf; ++a) t[a + (size_t)nbf * q] += p.h[a + (size_t)nbf * b] * cb;
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double acc = 0;
                for (int a = 0; a < nbf; ++a) acc += cbar[a + (size_t)nbf * i] * t[a + (size_t)nbf * j];
                hbar[i + n * j] = acc;
            }
    }
    Vec hact(n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0;
            for (int a = 0; a <= i; ++a)
                for (int b = 0; b <= j; ++b) acc += R[a + n * i] * hbar[a + n * b] * R[b + n * j];
            hact[i + n * j] = acc;
        }
    // RI one-index T[A,q,mu] = sum_nu B[A,(mu nu)] Cbar[nu,q] (integrals/ao/ri_transform.cpp:10-37),
    // then Lbar[A,(p,q)] = sum_mu T[A,q,mu] Cbar[mu,p] (integrals/active/ri.cpp:34-96)
    Vec Tbar((size_t)naux * n * nbf, 0.0);  // index (A + naux*(q + n*mu))
#pragma omp parallel for schedule(static)
    for (int mu = 0; mu < nbf; ++mu) {
        for (int nu = 0; nu < nbf; ++nu) {
            const double* bcol = p.B + (size_t)naux * pp(mu, nu);
            for (int q = 0; q < n; ++q) {
                const double c = cbar[nu + (size_t)nbf * q];
                double* t = &Tbar[(size_t)naux * (q + (size_t)n * mu)];
                for (int A = 0; A < naux; ++A) t[A] += bcol[A] * c;
            }
        }
    }
    Vec Lbar((size_t)naux * np, 0.0);  // (A + naux*P)
#pragma omp parallel for schedule(static)
    for (int a = 0; a < n; ++a)
        for (int b = 0; b <= a; ++b) {
            double* l = &Lbar[(size_t)naux * pp(a, b)];
            for (int mu = 0; mu < nbf; ++mu) {
                const double c = cbar[mu + (size_t)nbf * b];
                const double* t = &Tbar[(size_t)naux * (a + (size_t)n * mu)];
                for (int A = 0; A < naux; ++A) l[A] += t[A] * c;
            }
        }
    // packed Gbar = Lbar^T Lbar (integrals/active/ri.cpp:14-31)
    Vec gbar(ng, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (int P = 0; P < np; ++P)
        for (int Q = 0; Q <= P; ++Q) {
            double acc = 0;
            for (int A = 0; A < naux; ++A) acc += Lbar[A + (size_t)naux * P] * Lbar[A + (size_t)naux * Q];
            gbar[pp(P, Q)] = acc;
        }
    // raw kernel: materialize_raw_pair_kernel (integrals/active/raw.hpp:13-23), four one-index passes
    Vec graw(ng, 0.0);
    {
        const size_t n4 = (size_t)n * n * n * n;
        Vec t1(n4), t2(n4);
        auto I4 = [&](int a, int b, int c, int d) { return (((size_t)a * n + b) * n + c) * n + d; };
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
                for (int c = 0; c < n; ++c)
                    for (int d = 0; d < n; ++d) t1[I4(a, b, c, d)] = gbar[pp(pp(a, b), pp(c, d))];
        for (int pass = 0; pass < 4; ++pass) {
            std::fill(t2.begin(), t2.end(), 0.0);
            // transform the leading index by R (x_new[p,...] = sum_a R[a,p] x[a,...]) and rotate indices
            for (int pnew = 0; pnew < n; ++pnew)
                for (int a = 0; a <= pnew; ++a) {
                    const double r = R[a + n * pnew];
                    for (int b = 0; b < n; ++b)
                        for (int c = 0; c < n; ++c)
                            for (int d = 0; d < n; ++d) t2[I4(b, c, d, pnew)] += r * t1[I4(a, b, c, d)];
                }
            std::swap(t1, t2);
        }
        for (int a = 0; a < n; ++a)
            for (int b = 0; b <= a; ++b)
                for (int c = 0; c < n; ++c)
                    for (int d = 0; d <= c; ++d) {
                        if (pp(a, b) < pp(c, d)) continue;
                        graw[pp(pp(a, b), pp(c, d))] = t1[I4(a, b, c, d)];
                    }
    }
    Integrals I;
    I.n = n;
    I.sact = sact;
    I.hact = hact;
    I.g = graw;
    (void)hraw_direct;
    // determinant-pair H/S over all ordered pairs (structures/assembly/hamiltonian_overlap.cpp:889-1021)
    const int nd = p.ndet;
    Vec hdet((size_t)nd * nd), sdet((size_t)nd * nd);
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < nd; ++d)
        for (int e = 0; e < nd; ++e) {
            double s, hv;
            pair_forward(p.am[p.da[d]], p.am[p.da[e]], p.bm[p.db[d]], p.bm[p.db[e]], I, s, hv);
            hdet[d + (size_t)nd * e] = hv;
            sdet[d + (size_t)nd * e] = s;
        }
    // structure matrices H_IJ = sum t_dI t_eJ H_de, then symmetrised
    const int ns = p.nstr;
    Vec Hs((size_t)ns * ns, 0.0), Ss((size_t)ns * ns, 0.0);
    {
        // Y = Hdet T (n_det x n_str), then T^T Y
        Vec yh((size_t)nd * ns, 0.0), ys((size_t)nd * ns, 0.0);
        for (int e = 0; e < nd; ++e)
            for (int64_t t = p.off[e]; t < p.off[e + 1]; ++t) {
                const int J = p.ts[t];
                const double c = p.tc[t];
                for (int d = 0; d < nd; ++d) {
                    yh[d + (size_t)nd * J] += hdet[d + (size_t)nd * e] * c;
                    ys[d + (size_t)nd * J] += sdet[d + (size_t)nd * e] * c;
                }
            }
        for (int d = 0; d < nd; ++d)
            for (int64_t t = p.off[d]; t < p.off[d + 1]; ++t) {
                const int Ii = p.ts[t];
                const double c = p.tc[t];
                for (int J = 0; J < ns; ++J) {
                    Hs[Ii + (size_t)ns * J] += c * yh[d + (size_t)nd * J];
                    Ss[Ii + (size_t)ns * J] += c * ys[d + (size_t)nd * J];
                }
            }
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < i; ++j) {
                Hs[i + (size_t)ns * j] = Hs[j + (size_t)ns * i] = 0.5 * (Hs[i + (size_t)ns * j] + Hs[j + (size_t)ns * i]);
                Ss[i + (size_t)ns * j] = Ss[j + (size_t)ns * i] = 0.5 * (Ss[i + (size_t)ns * j] + Ss[j + (size_t)ns * i]);
            }
    }
    Vec w, x;
    if (gep(ns, Hs.data(), Ss.data(), w, x) != 0) {
        err = "structure metric is not positive definite";
        return 3;
    }
    // E = E11 + sum_s w_s lambda_s + E_nuc over the selected states with normalised weights, no
    // inactive core (E11 = 0) (workflow/evaluator.cpp:22-32,65-111; mcscf/states/weights.cpp:18-33;
    // derivatives/gradient/active_space/evaluator.cpp:1243-1248). Without a selection: the lowest
    // root with weight 1.
    const bool single = p.n_states == 0;
    std::vector<int> sel = single ? std::vector<int>{0} : std::vector<int>(p.sidx, p.sidx + p.n_states);
    Vec omega(sel.size(), 1.0);
    if (!single) {
        double wsum = 0.0;
        for (int k = 0; k < p.n_states; ++k) {
            if (!(p.sw[k] >= 0.0)) {
                err = "state_average_weights must be non-negative";
                return 1;
            }
            if (sel[k] < 0 || sel[k] >= ns) {
                err = "selected state index is out of range";
                return 1;
            }
            wsum += p.sw[k];
        }
        if (!(wsum > 0.0)) {
            err = "state_average_weights must sum to a positive value";
            return 1;
        }
        for (int k = 0; k < p.n_states; ++k) omega[k] = p.sw[k] / wsum;
    }
    double eavg = 0.0;
    o.state_energies.assign(sel.size(), 0.0);
    for (size_t k = 0; k < sel.size(); ++k) {
        eavg += omega[k] * w[sel[k]];
        o.state_energies[k] = w[sel[k]] + p.enuc;
    }
    o.energy = single ? w[0] + p.enuc : eavg + p.enuc;
    o.eigenvalues = w;
    o.coef.assign(x.begin() + (size_t)ns * sel[0], x.begin() + (size_t)ns * (sel[0] + 1));
    o.H = Hs;
    o.Sm = Ss;
    o.sact = sact;
    o.hact = hact;
    o.g = graw;
    o.cn = cn;
    o.hdet = hdet;
    o.sdet = sdet;
    if (!want_grad) return 0;

    // ---- gradient: Hellmann-Feynman over determinant pairs, weights a_d a_e ----
    // (selected states: sum_s w_s a^s_d a^s_e for dH and sum_s w_s lambda_s a^s_d a^s_e for dS, the
    // state-averaged form of the selected-state determinant coefficient matrices,
    // structures/assembly/selected_coefficients.hpp:47-66)
    const double E = w[0];
    Vec a(nd, 0.0);
    for (int d = 0; d < nd; ++d)
        for (int64_t t = p.off[d]; t < p.off[d + 1]; ++t) a[d] += p.tc[t] * o.coef[p.ts[t]];
    const size_t nsel = sel.size();
    Vec as(single ? 0 : nsel * nd, 0.0);
    if (!single)
        for (size_t k = 0; k < nsel; ++k)
            for (int d = 0; d < nd; ++d)
                for (int64_t t = p.off[d]; t < p.off[d + 1]; ++t)
                    as[k * nd + d] += p.tc[t] * x[p.ts[t] + (size_t)ns * sel[k]];
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    std::vector<Grad> part(nthreads);
    for (auto& gg : part) gg.init(n, ng);
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 1)
        for (int d = 0; d < nd; ++d) {
            if (single) {
                if (a[d] == 0.0) continue;
                for (int e = 0; e < nd; ++e) {
                    const double wgt = a[d] * a[e];
                    if (wgt == 0.0) continue;
                    pair_backward(p.am[p.da[d]], p.am[p.da[e]], p.bm[p.db[d]], p.bm[p.db[e]], I, wgt, -E * wgt,
                                  part[tid]);
                }
                continue;
            }
            for (int e = 0; e < nd; ++e) {
                double wh = 0.0, wsv = 0.0;
                for (size_t k = 0; k < nsel; ++k) {
                    const double c2 = omega[k] * as[k * nd + d] * as[k * nd + e];
                    wh += c2;
                    wsv -= w[sel[k]] * c2;
                }
                if (wh == 0.0 && wsv == 0.0) continue;
                pair_backward(p.am[p.da[d]], p.am[p.da[e]], p.bm[p.db[d]], p.bm[p.db[e]], I, wh, wsv, part[tid]);
            }
        }
    }
    Grad tot;
    tot.init(n, ng);
    for (auto& gg : part) tot.add(gg);
    // chain to the normalised orbitals in the raw frame:
    //   dE/dCn = S Cn (dS + dS^T) + h Cn (dh + dh^T) + sum_A sum_q W_A[p,q] T[A,q,mu]
    Vec gcn((size_t)nbf * n, 0.0);  // gcn[mu + nbf*p]
    Vec ws(n * n), wh(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            ws[i + n * j] = tot.ds[i + n * j] + tot.ds[j + n * i];
            wh[i + n * j] = tot.dh[i + n * j] + tot.dh[j + n * i];
        }
    for (int pq = 0; pq < n; ++pq)
        for (int i = 0; i < n; ++i) {
            const double a1 = ws[i + n * pq], a2 = wh[i + n * pq];
            for (int mu = 0; mu < nbf; ++mu) gcn[mu + (size_t)nbf * pq] += sc[mu + (size_t)nbf * i] * a1 + hc[mu + (size_t)nbf * i] * a2;
        }
    // RI part. raw L[A,(p,q)] = sum_ab R[a,p] R[b,q] Lbar[A,(a,b)]; raw T = Tbar R.
    Vec Lraw((size_t)naux * np, 0.0);
    // L_raw[A] = R^T Lbar[A] R on the symmetric n x n block of each auxiliary row
    for (int A = 0; A < naux; ++A) {
        Vec lb(n * n), t(n * n, 0.0);
        for (int aa = 0; aa < n; ++aa)
            for (int bb = 0; bb < n; ++bb) lb[aa + n * bb] = Lbar[A + (size_t)naux * pp(aa, bb)];
        for (int aa = 0; aa < n; ++aa)
            for (int q = 0; q < n; ++q) {
                double acc = 0;
                for (int bb = 0; bb <= q; ++bb) acc += lb[aa + n * bb] * R[bb + n * q];
                t[aa + n * q] = acc;
            }
        for (int P = 0; P < n; ++P)
            for (int Q = 0; Q <= P; ++Q) {
                double acc = 0;
                for (int aa = 0; aa <= P; ++aa) acc += R[aa + n * P] * t[aa + n * Q];
                Lraw[A + (size_t)naux * pp(P, Q)] = acc;
            }
    }
    // M[P,Q] = dg[pp(P,Q)] (P != Q), 2 dg[pp(P,P)]; dL[A,P] = sum_Q M[P,Q] L[A,Q]
    Vec dL((size_t)naux * np, 0.0);
#pragma omp parallel for schedule(static)
    for (int P = 0; P < np; ++P)
        for (int Q = 0; Q < np; ++Q) {
            const double m = (P == Q) ? 2.0 * tot.dg[pp(P, P)] : tot.dg[pp(P, Q)];
            if (m == 0.0) continue;
            for (int A = 0; A < naux; ++A) dL[A + (size_t)naux * P] += m * Lraw[A + (size_t)naux * Q];
        }
    // W_A symmetric: off-diagonal dL, diagonal 2 dL; dE/dCn[mu,p] += sum_A sum_q W_A[p,q] T[A,q,mu]
    // T (raw frame) = Tbar R: T[A,q,mu] = sum_b Tbar[A,b,mu] R[b,q]
#pragma omp parallel for schedule(static)
    for (int mu = 0; mu < nbf; ++mu) {
        Vec traw((size_t)naux * n, 0.0);
        for (int q = 0; q < n; ++q)
            for (int bb = 0; bb <= q; ++bb) {
                const double r = R[bb + n * q];
                const double* tb = &Tbar[(size_t)naux * (bb + (size_t)n * mu)];
                for (int A = 0; A < naux; ++A) traw[A + (size_t)naux * q] += tb[A] * r;
            }
        for (int P = 0; P < n; ++P)
            for (int q = 0; q < n; ++q) {
                const double* dl = &dL[(size_t)naux * pp(P, q)];
                const double f = (P == q) ? 2.0 : 1.0;
                double acc = 0;
                for (int A = 0; A < naux; ++A) acc += dl[A] * traw[A + (size_t)naux * q];
                gcn[mu + (size_t)nbf * P] += f * acc;
            }
    }
    // normalization pullback to raw slots (orbitals/pullback/operator.cpp:139-215)
    o.grad.assign((size_t)n * nbf, 0.0);
    for (int k = 0; k < n; ++k) {
        const double f = std::sqrt(1.0 / nu2[k]);
        double dot = 0;
        for (int mu = 0; mu < nbf; ++mu) dot += gcn[mu + (size_t)nbf * k] * cn[(size_t)k * nbf + mu];
        for (int mu = 0; mu < nbf; ++mu) {
            if (p.support && !p.support[(size_t)k * nbf + mu]) continue;
            o.grad[(size_t)k * nbf + mu] = f * gcn[mu + (size_t)nbf * k] - f * dot * sc[mu + (size_t)nbf * k];
        }
    }
    return 0;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for tests and the CPU baseline
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_err;
orc::RawStructures g_raw;
orc::Expansion g_ex;
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

int oracle_count_structures(const char* kw, int nact, int nel, int mult, uint64_t* count) {
    try {
        *count = orc::count_structures(kw, nact, nel, mult);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// sizes[0..4] = n_structures, n_alpha_strings, n_beta_strings, n_determinants, n_terms
int oracle_topology(const char* kw, int nact, int nel, int mult, int64_t* sizes) {
    try {
        g_raw = orc::generate(kw, nact, nel, mult);
        g_ex = orc::expand(g_raw, mult);
        sizes[0] = g_ex.n_struct;
        sizes[1] = (int64_t)g_ex.amask.size();
        sizes[2] = (int64_t)g_ex.bmask.size();
        sizes[3] = (int64_t)g_ex.det_a.size();
        sizes[4] = (int64_t)g_ex.term_s.size();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int oracle_topology_fill(uint64_t* amask, uint64_t* bmask, int32_t* det_a, int32_t* det_b, int64_t* off,
                         int32_t* term_s, double* term_c, int32_t* raw) {
    std::copy(g_ex.amask.begin(), g_ex.amask.end(), amask);
    std::copy(g_ex.bmask.begin(), g_ex.bmask.end(), bmask);
    std::copy(g_ex.det_a.begin(), g_ex.det_a.end(), det_a);
    std::copy(g_ex.det_b.begin(), g_ex.det_b.end(), det_b);
    std::copy(g_ex.off.begin(), g_ex.off.end(), off);
    std::copy(g_ex.term_s.begin(), g_ex.term_s.end(), term_s);
    std::copy(g_ex.term_c.begin(), g_ex.term_c.end(), term_c);
    if (raw)
        for (size_t s = 0; s < g_raw.rows.size(); ++s)
            std::copy(g_raw.rows[s].begin(), g_raw.rows[s].end(), raw + s * g_raw.n_el);
    return 0;
}

// DeterminantOverlapResolver + cofactor caches for one block X (n x n, column-major, rows = right).
// cof2 (optional) is dense [r1 + n*(r2 + n*(l1 + n*l2))].
int oracle_resolve(int n, const double* x, double* det, int* nullity, double* cof1, double* cof2) {
    orc::SpinBlock b;
    b.n = n;
    b.x.assign(x, x + n * n);
    orc::resolve(b);
    *det = b.det;
    *nullity = b.nullity;
    if (cof1) std::copy(b.cof1.begin(), b.cof1.end(), cof1);
    if (cof2)
        for (int r1 = 0; r1 < n; ++r1)
            for (int r2 = 0; r2 < n; ++r2)
                for (int l1 = 0; l1 < n; ++l1)
                    for (int l2 = 0; l2 < n; ++l2) {
                        double v = 0.0;
                        if (r1 != r2 && l1 != l2) {
                            // calc_second_order_cofactor is defined for ordered index pairs
                            const int a1 = std::min(r1, r2), a2 = std::max(r1, r2);
                            const int c1 = std::min(l1, l2), c2 = std::max(l1, l2);
                            const double sgn = ((r1 > r2) != (l1 > l2)) ? -1.0 : 1.0;
                            v = sgn * orc::cof2(b, a1, a2, c1, c2);
                        }
                        cof2[r1 + n * (r2 + n * (l1 + n * l2))] = v;
                    }
    return 0;
}

// same-spin pair quantities for occupied masks (raw-frame integrals given)
int oracle_spin_pair(int nact, const double* sact, const double* hact, const double* gpacked, uint64_t mL,
                     uint64_t mR, double* det, double* hsig, int* nullity) {
    try {
        orc::Integrals I;
        I.n = nact;
        I.sact.assign(sact, sact + nact * nact);
        I.hact.assign(hact, hact + nact * nact);
        const int np = nact * (nact + 1) / 2;
        I.g.assign(gpacked, gpacked + (size_t)np * (np + 1) / 2);
        orc::SpinBlock b = orc::make_block(mL, mR, I);
        *det = b.det;
        *nullity = b.nullity;
        *hsig = orc::same_spin_h(b, I);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// full determinant pair (alpha and beta masks)
int oracle_det_pair(int nact, const double* sact, const double* hact, const double* gpacked, uint64_t aL,
                    uint64_t aR, uint64_t bL, uint64_t bR, double* s, double* h) {
    try {
        orc::Integrals I;
        I.n = nact;
        I.sact.assign(sact, sact + nact * nact);
        I.hact.assign(hact, hact + nact * nact);
        const int np = nact * (nact + 1) / 2;
        I.g.assign(gpacked, gpacked + (size_t)np * (np + 1) / 2);
        orc::pair_forward(aL, aR, bL, bR, I, *s, *h);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int oracle_gep(int n, const double* h, const double* s, double* w, double* x) {
    orc::Vec ww, xx;
    const int rc = orc::gep(n, h, s, ww, xx);
    if (rc) return rc;
    std::copy(ww.begin(), ww.end(), w);
    std::copy(xx.begin(), xx.end(), x);
    return 0;
}

// One energy(+gradient) evaluation. Optional outputs may be null.
//   C, grad, support: orbital-major [k*n_bf + mu]; S, h: n_bf x n_bf; B: naux x n_bf(n_bf+1)/2 column-major.
//   out_H, out_S: n_str^2; out_coef: n_str; out_sact/out_hact: n_act^2; out_g: packed kernel;
//   out_hdet/out_sdet: n_det^2.
int oracle_eval_states(int n_bf, int n_act, int naux, const double* S, const double* h, const double* B,
                       const double* C, const uint8_t* support, int n_str, int n_astr, const uint64_t* amask,
                       int n_bstr, const uint64_t* bmask, int n_det, const int32_t* det_a, const int32_t* det_b,
                       const int64_t* off, const int32_t* term_s, const double* term_c, double e_nuc, int want_grad,
                       int n_states, const int32_t* state_idx, const double* state_w, double* out_energy,
                       double* out_grad, double* out_H, double* out_S, double* out_coef, double* out_sact,
                       double* out_hact, double* out_g, double* out_hdet, double* out_sdet,
                       double* out_state_energies) {
    try {
        orc::Problem p{n_bf, n_act, naux, S, h, B, C, support, n_str, n_astr, n_bstr, n_det, amask, bmask,
                       det_a, det_b, off, term_s, term_c, e_nuc, n_states, state_idx, state_w};
        orc::Out o;
        std::string err;
        const int rc = orc::evaluate(p, o, want_grad != 0, err);
        if (rc) {
            g_err = err;
            return rc;
        }
        *out_energy = o.energy;
        auto put = [](const orc::Vec& v, double* dst) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        if (want_grad) put(o.grad, out_grad);
        put(o.H, out_H);
        put(o.Sm, out_S);
        put(o.coef, out_coef);
        put(o.sact, out_sact);
        put(o.hact, out_hact);
        put(o.g, out_g);
        put(o.hdet, out_hdet);
        put(o.sdet, out_sdet);
        put(o.state_energies, out_state_energies);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int oracle_eval(int n_bf, int n_act, int naux, const double* S, const double* h, const double* B, const double* C,
                const uint8_t* support, int n_str, int n_astr, const uint64_t* amask, int n_bstr, const uint64_t* bmask,
                int n_det, const int32_t* det_a, const int32_t* det_b, const int64_t* off, const int32_t* term_s,
                const double* term_c, double e_nuc, int want_grad, double* out_energy, double* out_grad, double* out_H,
                double* out_S, double* out_coef, double* out_sact, double* out_hact, double* out_g, double* out_hdet,
                double* out_sdet) {
    return oracle_eval_states(n_bf, n_act, naux, S, h, B, C, support, n_str, n_astr, amask, n_bstr, bmask, n_det,
                              det_a, det_b, off, term_s, term_c, e_nuc, want_grad, 0, nullptr, nullptr, out_energy,
                              out_grad, out_H, out_S, out_coef, out_sact, out_hact, out_g, out_hdet, out_sdet,
                              nullptr);
}

}  // extern "C"
