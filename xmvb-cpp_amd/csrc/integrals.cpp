// AO integrals over contracted Cartesian Gaussians for the small decks the engine runs end to end
// (H2 in 6-31G**): overlap, kinetic, nuclear attraction and two-electron repulsion integrals by the
// McMurchie-Davidson scheme (Hermite expansion coefficients E_t^{ij}, Hermite Coulomb integrals
// R_tuv from the Boys function), and a pivoted Cholesky factorisation of the two-electron matrix
// into the RI-shaped factors the engine consumes, (mu nu|la si) = sum_P B[P,(mu nu)] B[P,(la si)].
//
// The reference obtains these integrals from libcint (libcint/materialized_provider.cpp) with the
// basis read from basis/6-31gss.gbs ("cartesian" functions); libcint is not available here, so
// this is an independent implementation of the published method:
//   L. E. McMurchie and E. R. Davidson, J. Comput. Phys. 26, 218 (1978);
//   T. Helgaker, P. Jorgensen, J. Olsen, Molecular Electronic-Structure Theory, ch. 9 (2000).
// Basis data for hydrogen (6-31G, 6-31G*, 6-31G**): W. J. Hehre, R. Ditchfield, J. A. Pople,
// J. Chem. Phys. 56, 2257 (1972); P. C. Hariharan, J. A. Pople, Theor. Chim. Acta 28, 213 (1973).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xmvb_amd.h"
#include "common.hpp"

namespace {

struct Shell {
    int l = 0;                 // 0 = s, 1 = p, 2 = d
    std::vector<double> a, c;  // primitive exponents and contraction coefficients
    int atom = 0;
    double center[3] = {0, 0, 0};
};

struct BasisFunction {
    int shell = 0;
    int lx = 0, ly = 0, lz = 0;
    double norm = 1.0;  // contracted normalisation (applied on top of primitive normalisation)
};

// hydrogen, 6-31G family (basis/6-31gss.gbs lists the same numbers)
std::vector<Shell> hydrogen_shells(const std::string& basis) {
    std::vector<Shell> sh;
    Shell s1;
    s1.l = 0;
    s1.a = {18.7311370, 2.8253937, 0.6401217};
    s1.c = {0.03349460, 0.23472695, 0.81375733};
    sh.push_back(s1);
    Shell s2;
    s2.l = 0;
    s2.a = {0.1612778};
    s2.c = {1.0};
    sh.push_back(s2);
    if (basis == "6-31G**" || basis == "6-31GSS" || basis == "6-31G(D,P)") {
        Shell p;
        p.l = 1;
        p.a = {1.1};
        p.c = {1.0};
        sh.push_back(p);
    } else if (basis != "6-31G" && basis != "6-31G*" && basis != "6-31GS" && basis != "6-31G(D)") {
        throw std::invalid_argument("basis " + basis + " is not available (6-31G, 6-31G*, 6-31G** for H)");
    }
    return sh;
}

std::string upper(const char* s) {
    std::string r(s ? s : "");
    for (auto& ch : r) ch = (char)std::toupper((unsigned char)ch);
    return r;
}

double double_factorial(int n) {  // n!! with (-1)!! = 1
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

double primitive_norm(double a, int lx, int ly, int lz) {
    const int L = lx + ly + lz;
    return std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * L) /
           std::sqrt(double_factorial(2 * lx - 1) * double_factorial(2 * ly - 1) * double_factorial(2 * lz - 1));
}

// Hermite expansion coefficients E_t^{ij} for one Cartesian direction, t = 0..i+j
void hermite_e(int i, int j, double p, double xpa, double xpb, double eab, std::vector<double>& E, int ldt) {
    // E[(ii * (j + 1) + jj) * ldt + t]
    const int nj = j + 1;
    std::fill(E.begin(), E.end(), 0.0);
    auto at = [&](int ii, int jj, int t) -> double& { return E[(ii * nj + jj) * ldt + t]; };
    at(0, 0, 0) = eab;
    const double h = 0.5 / p;
    for (int ii = 0; ii <= i; ++ii)
        for (int jj = 0; jj <= j; ++jj) {
            if (ii == 0 && jj == 0) continue;
            const bool up_i = ii > 0;
            const int pi = up_i ? ii - 1 : ii, pj = up_i ? jj : jj - 1;
            const double x = up_i ? xpa : xpb;
            for (int t = 0; t <= ii + jj; ++t) {
                double v = x * (t <= pi + pj ? at(pi, pj, t) : 0.0);
                if (t > 0 && t - 1 <= pi + pj) v += h * at(pi, pj, t - 1);
                if (t + 1 <= pi + pj) v += (t + 1) * at(pi, pj, t + 1);
                at(ii, jj, t) = v;
            }
        }
}

// Boys function F_n(T), n = 0..nmax: series for small T, asymptotic form for large T, then
// downward recursion F_{n-1} = (2T F_n + e^-T) / (2n - 1)
void boys(int nmax, double T, double* F) {
    const double et = std::exp(-T);
    if (T < 35.0) {
        // F_nmax(T) = e^-T sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
        double term = 1.0 / (2 * nmax + 1), sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0 * T / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        F[nmax] = et * sum;
    } else {
        F[nmax] = double_factorial(2 * nmax - 1) / std::pow(2.0, nmax + 1) * std::sqrt(M_PI / std::pow(T, 2 * nmax + 1));
    }
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
}

// Hermite Coulomb integrals R_{tuv}^0 for t+u+v <= L, scaled: R^n_000 = (-2 alpha)^n F_n(alpha |PC|^2)
struct HermiteR {
    int L;
    std::vector<double> R;  // R[(t * (L+1) + u) * (L+1) + v], n = 0 only
    double& at(int t, int u, int v) { return R[((size_t)t * (L + 1) + u) * (L + 1) + v]; }
};

HermiteR hermite_r(int L, double alpha, const double* pc) {
    const double T = alpha * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
    std::vector<double> F(L + 1);
    boys(L, T, F.data());
    // R^n_{tuv} for all n <= L - (t+u+v)
    const int D = L + 1;
    std::vector<double> Rn((size_t)D * D * D * D, 0.0);
    auto r = [&](int n, int t, int u, int v) -> double& { return Rn[(((size_t)n * D + t) * D + u) * D + v]; };
    double m2a = 1.0;
    for (int n = 0; n <= L; ++n) {
        r(n, 0, 0, 0) = m2a * F[n];
        m2a *= -2.0 * alpha;
    }
    for (int tot = 1; tot <= L; ++tot)
        for (int t = 0; t <= tot; ++t)
            for (int u = 0; u <= tot - t; ++u) {
                const int v = tot - t - u;
                for (int n = 0; n + tot <= L; ++n) {
                    double val;
                    if (t > 0) {
                        val = pc[0] * r(n + 1, t - 1, u, v);
                        if (t > 1) val += (t - 1) * r(n + 1, t - 2, u, v);
                    } else if (u > 0) {
                        val = pc[1] * r(n + 1, t, u - 1, v);
                        if (u > 1) val += (u - 1) * r(n + 1, t, u - 2, v);
                    } else {
                        val = pc[2] * r(n + 1, t, u, v - 1);
                        if (v > 1) val += (v - 1) * r(n + 1, t, u, v - 2);
                    }
                    r(n, t, u, v) = val;
                }
            }
    HermiteR out{L, std::vector<double>((size_t)D * D * D, 0.0)};
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) out.at(t, u, v) = r(0, t, u, v);
    return out;
}

struct Molecule {
    std::vector<Shell> shells;
    std::vector<BasisFunction> bf;
    std::vector<int> Z;
    std::vector<double> xyz;  // bohr
};

Molecule build(const char* basis, int n_atoms, const int32_t* Z, const double* xyz) {
    if (n_atoms <= 0 || !Z || !xyz) throw std::invalid_argument("empty molecule");
    Molecule m;
    const std::string b = upper(basis);
    for (int at = 0; at < n_atoms; ++at) {
        if (Z[at] != 1) throw std::invalid_argument("only hydrogen atoms are available in the embedded basis data");
        for (Shell s : hydrogen_shells(b)) {
            s.atom = at;
            for (int k = 0; k < 3; ++k) s.center[k] = xyz[3 * at + k];
            m.shells.push_back(s);
        }
        m.Z.push_back(Z[at]);
    }
    m.xyz.assign(xyz, xyz + 3 * n_atoms);
    for (int si = 0; si < (int)m.shells.size(); ++si) {
        const int l = m.shells[si].l;
        // Cartesian components in the order x, y, z (l = 1); xx, xy, xz, yy, yz, zz (l = 2)
        for (int lx = l; lx >= 0; --lx)
            for (int ly = l - lx; ly >= 0; --ly) m.bf.push_back({si, lx, ly, l - lx - ly, 1.0});
    }
    return m;
}

// primitive-contracted overlap of two basis functions (without contracted norms)
double overlap_raw(const Molecule& m, const BasisFunction& f, const BasisFunction& g) {
    const Shell& A = m.shells[f.shell];
    const Shell& B = m.shells[g.shell];
    const int la[3] = {f.lx, f.ly, f.lz}, lb[3] = {g.lx, g.ly, g.lz};
    double s = 0.0;
    std::vector<double> E(64);
    for (size_t i = 0; i < A.a.size(); ++i)
        for (size_t j = 0; j < B.a.size(); ++j) {
            const double a = A.a[i], b = B.a[j], p = a + b, mu = a * b / p;
            double v = A.c[i] * B.c[j] * primitive_norm(a, f.lx, f.ly, f.lz) * primitive_norm(b, g.lx, g.ly, g.lz);
            for (int d = 0; d < 3; ++d) {
                const double X = A.center[d] - B.center[d];
                const double P = (a * A.center[d] + b * B.center[d]) / p;
                E.assign((size_t)(la[d] + 1) * (lb[d] + 1) * (la[d] + lb[d] + 1), 0.0);
                hermite_e(la[d], lb[d], p, P - A.center[d], P - B.center[d], std::exp(-mu * X * X), E,
                          la[d] + lb[d] + 1);
                v *= E[((size_t)la[d] * (lb[d] + 1) + lb[d]) * (la[d] + lb[d] + 1) + 0] * std::sqrt(M_PI / p);
            }
            s += v;
        }
    return s;
}

// 1D overlap of primitives with powers i, j (for the kinetic-energy formula)
double overlap_1d(int i, int j, double a, double b, double Ax, double Bx) {
    if (i < 0 || j < 0) return 0.0;
    const double p = a + b, mu = a * b / p, X = Ax - Bx, P = (a * Ax + b * Bx) / p;
    std::vector<double> E((size_t)(i + 1) * (j + 1) * (i + j + 1));
    hermite_e(i, j, p, P - Ax, P - Bx, std::exp(-mu * X * X), E, i + j + 1);
    return E[((size_t)i * (j + 1) + j) * (i + j + 1)] * std::sqrt(M_PI / p);
}

}  // namespace

extern "C" {

int vbe_gaussian_basis_size(const char* basis, int32_t n_atoms, const int32_t* Z, int32_t* n_bf) {
    try {
        if (!n_bf) throw std::invalid_argument("null argument");
        std::vector<double> xyz(3 * std::max(n_atoms, 0), 0.0);
        *n_bf = (int32_t)build(basis, n_atoms, Z, xyz.data()).bf.size();
        return VBE_OK;
    } catch (const std::exception& e) {
        xmvb_amd::set_global_error(e.what());
        return VBE_ERR_ARGUMENT;
    }
}

int vbe_ao_integrals(const char* basis, int32_t n_atoms, const int32_t* Z, const double* xyz_bohr, double* S,
                     double* T, double* V, double* eri_packed, int32_t* atom_of_function, double* e_nuc) {
    try {
        Molecule m = build(basis, n_atoms, Z, xyz_bohr);
        const int n = (int)m.bf.size();
        // contracted normalisation
        for (auto& f : m.bf) f.norm = 1.0 / std::sqrt(overlap_raw(m, f, f));
        if (atom_of_function)
            for (int i = 0; i < n; ++i) atom_of_function[i] = m.shells[m.bf[i].shell].atom;
        if (e_nuc) {
            double e = 0.0;
            for (int a = 0; a < n_atoms; ++a)
                for (int b = 0; b < a; ++b) {
                    const double dx = m.xyz[3 * a] - m.xyz[3 * b], dy = m.xyz[3 * a + 1] - m.xyz[3 * b + 1],
                                 dz = m.xyz[3 * a + 2] - m.xyz[3 * b + 2];
                    e += m.Z[a] * m.Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
                }
            *e_nuc = e;
        }
        // one-electron integrals
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const BasisFunction &f = m.bf[i], &g = m.bf[j];
                const Shell &A = m.shells[f.shell], &B = m.shells[g.shell];
                const int la[3] = {f.lx, f.ly, f.lz}, lb[3] = {g.lx, g.ly, g.lz};
                double s = 0.0, t = 0.0, v = 0.0;
                for (size_t pa = 0; pa < A.a.size(); ++pa)
                    for (size_t pb = 0; pb < B.a.size(); ++pb) {
                        const double a = A.a[pa], b = B.a[pb], p = a + b, mu = a * b / p;
                        const double cc = A.c[pa] * B.c[pb] * primitive_norm(a, f.lx, f.ly, f.lz) *
                                          primitive_norm(b, g.lx, g.ly, g.lz);
                        double s1[3], t1[3], P[3];
                        std::vector<std::vector<double>> E(3);
                        for (int d = 0; d < 3; ++d) {
                            const double Ax = A.center[d], Bx = B.center[d];
                            P[d] = (a * Ax + b * Bx) / p;
                            s1[d] = overlap_1d(la[d], lb[d], a, b, Ax, Bx);
                            // -1/2 d^2/dx^2 acting on the right primitive
                            t1[d] = -0.5 * (4.0 * b * b * overlap_1d(la[d], lb[d] + 2, a, b, Ax, Bx) -
                                            2.0 * b * (2 * lb[d] + 1) * overlap_1d(la[d], lb[d], a, b, Ax, Bx) +
                                            lb[d] * (lb[d] - 1) * overlap_1d(la[d], lb[d] - 2, a, b, Ax, Bx));
                            E[d].assign((size_t)(la[d] + 1) * (lb[d] + 1) * (la[d] + lb[d] + 1), 0.0);
                            const double X = Ax - Bx;
                            hermite_e(la[d], lb[d], p, P[d] - Ax, P[d] - Bx, std::exp(-mu * X * X), E[d],
                                      la[d] + lb[d] + 1);
                        }
                        s += cc * s1[0] * s1[1] * s1[2];
                        t += cc * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                        const int L = la[0] + lb[0] + la[1] + lb[1] + la[2] + lb[2];
                        for (int c = 0; c < n_atoms; ++c) {
                            const double pc[3] = {P[0] - m.xyz[3 * c], P[1] - m.xyz[3 * c + 1], P[2] - m.xyz[3 * c + 2]};
                            HermiteR R = hermite_r(L, p, pc);
                            double acc = 0.0;
                            for (int tx = 0; tx <= la[0] + lb[0]; ++tx)
                                for (int ty = 0; ty <= la[1] + lb[1]; ++ty)
                                    for (int tz = 0; tz <= la[2] + lb[2]; ++tz)
                                        acc += E[0][((size_t)la[0] * (lb[0] + 1) + lb[0]) * (la[0] + lb[0] + 1) + tx] *
                                               E[1][((size_t)la[1] * (lb[1] + 1) + lb[1]) * (la[1] + lb[1] + 1) + ty] *
                                               E[2][((size_t)la[2] * (lb[2] + 1) + lb[2]) * (la[2] + lb[2] + 1) + tz] *
                                               R.at(tx, ty, tz);
                            v += -m.Z[c] * cc * 2.0 * M_PI / p * acc;
                        }
                    }
                const double nn = f.norm * g.norm;
                if (S) S[i + (size_t)n * j] = s * nn;
                if (T) T[i + (size_t)n * j] = t * nn;
                if (V) V[i + (size_t)n * j] = v * nn;
            }
        if (!eri_packed) return VBE_OK;
        // two-electron integrals over packed pairs (mu >= nu): eri[PQ + npair * RS]
        const int npair = n * (n + 1) / 2;
        struct PairData {
            int i, j;
            std::vector<double> p, cc, P[3];
            std::vector<std::vector<double>> E[3];
        };
        std::vector<PairData> pairs(npair);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                PairData& pd = pairs[(size_t)i * (i + 1) / 2 + j];
                pd.i = i;
                pd.j = j;
                const BasisFunction &f = m.bf[i], &g = m.bf[j];
                const Shell &A = m.shells[f.shell], &B = m.shells[g.shell];
                const int la[3] = {f.lx, f.ly, f.lz}, lb[3] = {g.lx, g.ly, g.lz};
                for (size_t pa = 0; pa < A.a.size(); ++pa)
                    for (size_t pb = 0; pb < B.a.size(); ++pb) {
                        const double a = A.a[pa], b = B.a[pb], p = a + b, mu = a * b / p;
                        pd.p.push_back(p);
                        pd.cc.push_back(A.c[pa] * B.c[pb] * primitive_norm(a, f.lx, f.ly, f.lz) *
                                        primitive_norm(b, g.lx, g.ly, g.lz) * f.norm * g.norm);
                        for (int d = 0; d < 3; ++d) {
                            const double Ax = A.center[d], Bx = B.center[d], P = (a * Ax + b * Bx) / p, X = Ax - Bx;
                            pd.P[d].push_back(P);
                            std::vector<double> E((size_t)(la[d] + 1) * (lb[d] + 1) * (la[d] + lb[d] + 1));
                            hermite_e(la[d], lb[d], p, P - Ax, P - Bx, std::exp(-mu * X * X), E, la[d] + lb[d] + 1);
                            std::vector<double> row(la[d] + lb[d] + 1);
                            for (int t = 0; t <= la[d] + lb[d]; ++t)
                                row[t] = E[((size_t)la[d] * (lb[d] + 1) + lb[d]) * (la[d] + lb[d] + 1) + t];
                            pd.E[d].push_back(row);
                        }
                    }
            }
        for (int x = 0; x < npair; ++x)
            for (int y = 0; y <= x; ++y) {
                const PairData &L = pairs[x], &Rp = pairs[y];
                double val = 0.0;
                for (size_t u = 0; u < L.p.size(); ++u)
                    for (size_t w = 0; w < Rp.p.size(); ++w) {
                        const double p = L.p[u], q = Rp.p[w], alpha = p * q / (p + q);
                        const int lx1 = (int)L.E[0][u].size() - 1, ly1 = (int)L.E[1][u].size() - 1,
                                  lz1 = (int)L.E[2][u].size() - 1;
                        const int lx2 = (int)Rp.E[0][w].size() - 1, ly2 = (int)Rp.E[1][w].size() - 1,
                                  lz2 = (int)Rp.E[2][w].size() - 1;
                        const int Ltot = lx1 + ly1 + lz1 + lx2 + ly2 + lz2;
                        const double pq[3] = {L.P[0][u] - Rp.P[0][w], L.P[1][u] - Rp.P[1][w], L.P[2][u] - Rp.P[2][w]};
                        HermiteR R = hermite_r(Ltot, alpha, pq);
                        double acc = 0.0;
                        for (int t = 0; t <= lx1; ++t)
                            for (int uu = 0; uu <= ly1; ++uu)
                                for (int v = 0; v <= lz1; ++v) {
                                    const double e1 = L.E[0][u][t] * L.E[1][u][uu] * L.E[2][u][v];
                                    if (e1 == 0.0) continue;
                                    for (int tau = 0; tau <= lx2; ++tau)
                                        for (int nu = 0; nu <= ly2; ++nu)
                                            for (int phi = 0; phi <= lz2; ++phi) {
                                                const double sign = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                                acc += e1 * sign * Rp.E[0][w][tau] * Rp.E[1][w][nu] * Rp.E[2][w][phi] *
                                                       R.at(t + tau, uu + nu, v + phi);
                                            }
                                }
                        val += L.cc[u] * Rp.cc[w] * 2.0 * std::pow(M_PI, 2.5) / (p * q * std::sqrt(p + q)) * acc;
                    }
                eri_packed[x + (size_t)npair * y] = val;
                eri_packed[y + (size_t)npair * x] = val;
            }
        return VBE_OK;
    } catch (const std::exception& e) {
        xmvb_amd::set_global_error(e.what());
        return VBE_ERR_ARGUMENT;
    }
}

int vbe_pivoted_cholesky(int32_t n, const double* M, double tolerance, int32_t max_rank, int32_t* rank, double* B) {
    // M (n x n, symmetric positive semidefinite) ~= B^T B with B rank x n (column-major, leading
    // dimension max_rank); pivots are taken largest-diagonal-first until the largest remaining
    // diagonal falls below `tolerance`
    try {
        if (n <= 0 || !M || !rank || !B || max_rank <= 0) throw std::invalid_argument("invalid arguments");
        std::vector<double> d(n);
        for (int i = 0; i < n; ++i) d[i] = M[i + (size_t)n * i];
        std::fill(B, B + (size_t)max_rank * n, 0.0);
        int r = 0;
        while (r < max_rank) {
            int piv = (int)(std::max_element(d.begin(), d.end()) - d.begin());
            const double dmax = d[piv];
            if (!(dmax > tolerance)) break;
            const double s = std::sqrt(dmax);
            for (int i = 0; i < n; ++i) {
                double v = M[i + (size_t)n * piv];
                for (int k = 0; k < r; ++k) v -= B[k + (size_t)max_rank * i] * B[k + (size_t)max_rank * piv];
                B[r + (size_t)max_rank * i] = v / s;
            }
            for (int i = 0; i < n; ++i) {
                const double b = B[r + (size_t)max_rank * i];
                d[i] -= b * b;
            }
            d[piv] = 0.0;
            ++r;
        }
        *rank = r;
        return VBE_OK;
    } catch (const std::exception& e) {
        xmvb_amd::set_global_error(e.what());
        return VBE_ERR_ARGUMENT;
    }
}

}  // extern "C"
