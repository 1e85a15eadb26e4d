// hipvbscf: runs an .xmi deck whose $CTRL selects HIPVBSCF on one MI355X through the engine's C ABI.
//
// It takes the place of run_cuvbscf (src/cli/cuvbscf.cpp:31-210) for the new method token:
//   - the deck reader follows parse_input_deck_model (src/input/deck/model.cpp:432-560): $CTRL
//     flags and KEY=VALUE assignments, $GEO, $FRAG, $ORB, $GUS, with the same integer-token
//     expansion (a:b doubled, a-b, v*n) and '#' comments;
//   - orbital supports follow build_input_deck_orbital_support_chart
//     (src/input/deck/orbital_support.cpp:455-495) for ORBTYP=HAO with FRGTYP=ATOM and for OEO;
//   - GUESS=AUTO builds the block guess of build_hcore_block_guess (src/guess/initial.cpp:355-362,
//     src/guess/block.cpp:141-221): per support block, the generalized eigenvectors of the core
//     Hamiltonian on the support, in ascending order; GUESS=READ reads $GUS as
//     build_read_or_rdci_guess does (src/guess/initial.cpp:262-296);
//   - the .orb file is written in the layout of write_orbital_file (src/output/orbitals/writer.cpp:
//     14-49); the exit status is 0 when converged and 2 otherwise, and errors print
//     "XMVB-CPP error: ..." with status 2 (src/cli/main.cpp:8-20).
// Options HIPVBSCF does not implement are rejected (as run_cuvbscf rejects its own): the structure
// class must come from STR=, one state, no inactive electrons, ISCF=8 (P-LBFGS) or unset. Exact
// two-electron integrals enter as pivoted-Cholesky factors (largest dropped diagonal <= 1e-13).
//
// Usage: hipvbscf [--max-iterations N] [--eigensolver dense|davidson] [--dry-run] deck.xmi
// --dry-run stops after the deck, integrals, supports and guess are built (no device work).

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <locale>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xmvb_amd.h"

namespace {

constexpr double kBohrPerAngstrom = 1.0 / 0.529177249;  // src/input/loading/loader.cpp:42

std::string upper(std::string s) {
    for (auto& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

std::string trim(const std::string& s) {
    const auto a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return {};
    const auto b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::string strip_comment(const std::string& s) {
    const auto p = s.find('#');
    return p == std::string::npos ? s : s.substr(0, p);
}

std::vector<std::string> split(const std::string& s) {
    std::istringstream in(s);
    std::vector<std::string> out;
    std::string t;
    while (in >> t) out.push_back(t);
    return out;
}

double parse_double(std::string t) {
    for (auto& c : t)
        if (c == 'd' || c == 'D') c = 'E';
    size_t used = 0;
    const double v = std::stod(t, &used);
    if (used != t.size()) throw std::runtime_error("invalid floating-point token in input deck: " + t);
    return v;
}

int parse_int(const std::string& t) {
    size_t used = 0;
    const int v = std::stoi(t, &used);
    if (used != t.size()) throw std::runtime_error("invalid integer token in input deck: " + t);
    return v;
}

// a:b (each value twice when `doubled`), a-b, v*n, or a plain integer
void expand_token(const std::string& t, bool doubled, std::vector<int>& out) {
    const auto colon = t.find(':'), dash = t.find('-'), star = t.find('*');
    if (colon != std::string::npos) {
        const int a = parse_int(t.substr(0, colon)), b = parse_int(t.substr(colon + 1));
        if (b < a) throw std::runtime_error("descending ':' range is not supported in input deck");
        for (int v = a; v <= b; ++v) {
            out.push_back(v);
            if (doubled) out.push_back(v);
        }
    } else if (dash != std::string::npos && dash > 0) {
        const int a = parse_int(t.substr(0, dash)), b = parse_int(t.substr(dash + 1));
        if (b < a) throw std::runtime_error("descending '-' range is not supported in input deck");
        for (int v = a; v <= b; ++v) out.push_back(v);
    } else if (star != std::string::npos) {
        const int v = parse_int(t.substr(0, star)), n = parse_int(t.substr(star + 1));
        if (n < 0) throw std::runtime_error("negative repetition count is not supported in input deck");
        for (int k = 0; k < n; ++k) out.push_back(v);
    } else {
        out.push_back(parse_int(t));
    }
}

std::vector<int> expand(const std::vector<std::string>& tokens, bool doubled) {
    std::vector<int> out;
    for (const auto& t : tokens) expand_token(t, doubled, out);
    return out;
}

struct Atom {
    std::string symbol;
    double xyz[3];
};

struct Deck {
    std::string title;
    std::string ctrl_text;  // every $CTRL line, for the method parse
    std::map<std::string, std::string> assign;
    std::vector<Atom> atoms;
    bool frag_present = false;
    std::vector<int> frag_sizes;
    std::vector<std::vector<std::string>> frag_entries;
    bool orb_present = false, actorb_present = false;
    std::vector<int> orb_sizes;
    std::vector<std::vector<std::string>> orb_entries;
    bool gus_present = false;
    std::vector<std::string> gus_lines;
    bool str_present = false;
    bool molden = false;
};

Deck read_deck(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("failed to open input file: " + path);
    enum Block { None, Ctrl, Geo, Frag, Orb, Gus, Str } block = None;
    Deck d;
    bool frag_header = false, orb_header = false, first = true;
    std::string raw;
    while (std::getline(in, raw)) {
        const std::string norm = upper(trim(strip_comment(raw)));
        if (first && block == None && !norm.empty() && norm[0] != '$') {
            d.title = trim(raw);
            first = false;
            continue;
        }
        if (norm.empty()) {
            if (block == Gus) d.gus_lines.push_back(raw);
            continue;
        }
        first = false;
        if (norm == "$CTRL") { block = Ctrl; continue; }
        if (norm == "$GEO") { block = Geo; continue; }
        if (norm == "$FRAG") { block = Frag; d.frag_present = true; frag_header = false; continue; }
        if (norm == "$ORB" || norm == "$ACTORB") {
            if (d.orb_present || d.actorb_present)
                throw std::runtime_error("$ORB and $ACTORB cannot both appear in one input deck");
            (norm == "$ORB" ? d.orb_present : d.actorb_present) = true;
            block = Orb;
            orb_header = false;
            continue;
        }
        if (norm == "$GUS") { block = Gus; d.gus_present = true; continue; }
        if (norm == "$STR") { block = Str; d.str_present = true; continue; }
        if (norm == "$END") { block = None; continue; }
        const auto tokens = split(strip_comment(raw));
        switch (block) {
            case Ctrl:
                d.ctrl_text += strip_comment(raw) + " ";
                for (const auto& t : tokens) {
                    const auto eq = t.find('=');
                    if (eq == std::string::npos) {
                        if (upper(t) == "MOLDEN") d.molden = true;
                        continue;
                    }
                    d.assign[upper(t.substr(0, eq))] = t.substr(eq + 1);
                }
                break;
            case Geo: {
                if (tokens.size() < 4) throw std::runtime_error("invalid $GEO line: " + raw);
                Atom a;
                a.symbol = tokens[0];
                for (int k = 0; k < 3; ++k) a.xyz[k] = parse_double(tokens[1 + k]);
                d.atoms.push_back(a);
                break;
            }
            case Frag:
                if (!frag_header) {
                    d.frag_sizes = expand(tokens, true);
                    frag_header = true;
                } else {
                    d.frag_entries.push_back(tokens);
                }
                break;
            case Orb:
                if (d.orb_present && !orb_header) {
                    d.orb_sizes = expand(tokens, true);
                    orb_header = true;
                } else {
                    d.orb_entries.push_back(tokens);
                }
                break;
            case Gus: d.gus_lines.push_back(raw); break;
            case Str: break;
            case None: break;
        }
    }
    return d;
}

std::string get(const Deck& d, const std::string& key, const std::string& fallback) {
    const auto it = d.assign.find(key);
    return it == d.assign.end() ? fallback : it->second;
}

int element_z(const std::string& symbol) {
    static const char* table[] = {"H",  "HE", "LI", "BE", "B",  "C",  "N",  "O",  "F",
                                  "NE", "NA", "MG", "AL", "SI", "P",  "S",  "CL", "AR"};
    std::string s = upper(symbol);
    while (!s.empty() && std::isdigit((unsigned char)s.back())) s.pop_back();  // labels like H1
    for (int z = 0; z < (int)(sizeof(table) / sizeof(table[0])); ++z)
        if (s == table[z]) return z + 1;
    throw std::runtime_error("unknown element symbol in $GEO: " + symbol);
}

void check(int rc, const char* what) {
    if (rc != VBE_OK) throw std::runtime_error(std::string(what) + ": " + vbe_global_error());
}

void check_ctx(vbe_context* ctx, int rc, const char* what) {
    if (rc == VBE_OK) return;
    char buf[1024];
    vbe_last_error(ctx, buf, sizeof buf);
    throw std::runtime_error(std::string(what) + ": " + buf);
}

// Symmetric generalized eigenproblem A x = w B x (B positive definite, n small): Cholesky B = L L^T,
// cyclic Jacobi on L^-1 A L^-T, x = L^-T y; ascending order, x^T B x = 1.
void sygv(int n, std::vector<double> A, std::vector<double> B, std::vector<double>& w, std::vector<double>& X) {
    std::vector<double> L(n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double s = B[j + n * j];
        for (int k = 0; k < j; ++k) s -= L[j + n * k] * L[j + n * k];
        if (!(s > 0.0)) throw std::runtime_error("overlap block is not positive definite in the guess build");
        L[j + n * j] = std::sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double t = B[i + n * j];
            for (int k = 0; k < j; ++k) t -= L[i + n * k] * L[j + n * k];
            L[i + n * j] = t / L[j + n * j];
        }
    }
    // M = L^-1 A L^-T
    std::vector<double> T(A);
    for (int c = 0; c < n; ++c)  // T = L^-1 A (column by column forward substitution)
        for (int i = 0; i < n; ++i) {
            double s = T[i + n * c];
            for (int k = 0; k < i; ++k) s -= L[i + n * k] * T[k + n * c];
            T[i + n * c] = s / L[i + n * i];
        }
    std::vector<double> M(n * n);
    for (int r = 0; r < n; ++r)  // M = T L^-T: rows of T solved against L
        for (int j = 0; j < n; ++j) {
            double s = T[r + n * j];
            for (int k = 0; k < j; ++k) s -= M[r + n * k] * L[j + n * k];
            M[r + n * j] = s / L[j + n * j];
        }
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i + n * i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += M[p + n * q] * M[p + n * q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = M[p + n * q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (M[q + n * q] - M[p + n * p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double mkp = M[k + n * p], mkq = M[k + n * q];
                    M[k + n * p] = c * mkp - s * mkq;
                    M[k + n * q] = s * mkp + c * mkq;
                }
                for (int k = 0; k < n; ++k) {
                    const double mpk = M[p + n * k], mqk = M[q + n * k];
                    M[p + n * k] = c * mpk - s * mqk;
                    M[q + n * k] = s * mpk + c * mqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k + n * p], vkq = V[k + n * q];
                    V[k + n * p] = c * vkp - s * vkq;
                    V[k + n * q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return M[a + n * a] < M[b + n * b]; });
    w.assign(n, 0.0);
    X.assign(n * n, 0.0);
    for (int jj = 0; jj < n; ++jj) {
        const int j = order[jj];
        w[jj] = M[j + n * j];
        for (int i = n - 1; i >= 0; --i) {  // x = L^-T y
            double s = V[i + n * j];
            for (int k = i + 1; k < n; ++k) s -= L[k + n * i] * X[k + n * jj];
            X[i + n * jj] = s / L[i + n * i];
        }
    }
}

struct Problem {
    int n_bf = 0, n_act = 0, n_alpha = 0, n_beta = 0, n_aux = 0, n_atoms = 0;
    std::vector<int32_t> Z;
    std::vector<double> xyz_bohr;
    std::vector<double> S, h, B;
    std::vector<int32_t> atom_of;
    double e_nuc = 0.0;
    std::vector<std::vector<int>> support;  // 1-based AO indices per orbital, ascending within fragments
    std::vector<double> C0;                 // n_act x n_bf, orbital-major (dense rows)
    std::string basis, str_class, guess = "AUTO";
    int eigensolver = VBE_EIGENSOLVER_DAVIDSON;
    int max_iterations = 2000;
    int multiplicity = 1;
    int n_states = 1;
};

std::vector<std::vector<int>> build_supports(const Deck& d, const Problem& p, const std::string& orbtyp,
                                             const std::string& frgtyp) {
    const int n_orb = p.n_act;
    std::vector<std::vector<int>> atom_functions(p.n_atoms);
    for (int f = 0; f < p.n_bf; ++f) atom_functions[p.atom_of[f]].push_back(f + 1);
    auto atom_fragment = [&](const std::vector<int>& atoms) {
        std::vector<int> out;
        for (int a : atoms) {
            if (a <= 0 || a > p.n_atoms) throw std::runtime_error("fragment atom index is out of range");
            out.insert(out.end(), atom_functions[a - 1].begin(), atom_functions[a - 1].end());
        }
        std::sort(out.begin(), out.end());
        return out;
    };
    std::vector<std::vector<int>> sup(n_orb);
    if (orbtyp == "OEO") {
        for (auto& s : sup)
            for (int f = 1; f <= p.n_bf; ++f) s.push_back(f);
        return sup;
    }
    if (orbtyp != "HAO")
        throw std::invalid_argument("HIPVBSCF supports ORBTYP=HAO (with FRGTYP=ATOM) and ORBTYP=OEO; got ORBTYP=" +
                                    orbtyp);
    if (frgtyp != "ATOM")
        throw std::invalid_argument("HIPVBSCF supports FRGTYP=ATOM only (SAO labels need the libcint AO table)");
    std::vector<int> all;
    for (int a = 1; a <= p.n_atoms; ++a) all.push_back(a);
    if (d.actorb_present) {  // one line of atoms per active orbital (orbital_support.cpp:340-388)
        if ((int)d.orb_entries.size() != n_orb)
            throw std::runtime_error("$ACTORB must provide exactly one support line for each active orbital");
        for (int i = 0; i < n_orb; ++i) sup[i] = atom_fragment(expand(d.orb_entries[i], false));
        return sup;
    }
    std::vector<std::vector<int>> fragments;
    if (!d.frag_present) {
        for (int a = 1; a <= p.n_atoms; ++a) fragments.push_back(atom_fragment({a}));
    } else {
        std::vector<int> buffered;
        size_t next = 0;
        for (size_t f = 0; f < d.frag_sizes.size(); ++f) {
            while ((int)buffered.size() < d.frag_sizes[f] && next < d.frag_entries.size()) {
                const auto e = expand(d.frag_entries[next++], false);
                buffered.insert(buffered.end(), e.begin(), e.end());
            }
            if ((int)buffered.size() < d.frag_sizes[f]) throw std::runtime_error("not enough atom labels in $FRAG block");
            fragments.push_back(atom_fragment(std::vector<int>(buffered.begin(), buffered.begin() + d.frag_sizes[f])));
            buffered.erase(buffered.begin(), buffered.begin() + d.frag_sizes[f]);
        }
    }
    if (!d.orb_present) throw std::runtime_error("orbital support builder requires a $ORB block");
    if ((int)d.orb_sizes.size() < n_orb) throw std::runtime_error("orbital support header does not declare enough orbitals");
    std::vector<int> buffered;
    size_t next = 0;
    for (int i = 0; i < n_orb; ++i) {
        while ((int)buffered.size() < d.orb_sizes[i] && next < d.orb_entries.size()) {
            const auto e = expand(d.orb_entries[next++], false);
            buffered.insert(buffered.end(), e.begin(), e.end());
        }
        if ((int)buffered.size() < d.orb_sizes[i]) throw std::runtime_error("not enough fragment labels in $ORB block");
        for (int k = 0; k < d.orb_sizes[i]; ++k) {
            const int fl = buffered[k];
            if (fl <= 0 || fl > (int)fragments.size())
                throw std::runtime_error("orbital support references an out-of-range fragment");
            sup[i].insert(sup[i].end(), fragments[fl - 1].begin(), fragments[fl - 1].end());
        }
        buffered.erase(buffered.begin(), buffered.begin() + d.orb_sizes[i]);
    }
    return sup;
}

// per block of orbitals with the same support: ascending generalized eigenvectors of h on it
void hcore_block_guess(Problem& p) {
    const int n = p.n_bf;
    p.C0.assign((size_t)p.n_act * n, 0.0);
    std::vector<bool> done(p.n_act, false);
    for (int i = 0; i < p.n_act; ++i) {
        if (done[i]) continue;
        std::vector<int> block;
        for (int j = i; j < p.n_act; ++j)
            if (!done[j] && p.support[j] == p.support[i]) {
                block.push_back(j);
                done[j] = true;
            }
        const auto& idx = p.support[i];
        const int m = (int)idx.size();
        if ((int)block.size() > m) throw std::runtime_error("more orbitals than functions on one support block");
        std::vector<double> Hb(m * m), Sb(m * m), w, X;
        for (int a = 0; a < m; ++a)
            for (int b = 0; b < m; ++b) {
                Hb[a + m * b] = p.h[(idx[a] - 1) + (size_t)n * (idx[b] - 1)];
                Sb[a + m * b] = p.S[(idx[a] - 1) + (size_t)n * (idx[b] - 1)];
            }
        sygv(m, Hb, Sb, w, X);
        for (size_t k = 0; k < block.size(); ++k)
            for (int a = 0; a < m; ++a) p.C0[(size_t)block[k] * n + (idx[a] - 1)] = X[a + m * k];
    }
}

// $GUS: support counts header, then (coefficient, 1-based AO index) pairs per orbital, scattered
// into dense rows and gathered back onto each orbital's support
void read_guess(const Deck& d, Problem& p) {
    std::vector<int> counts;
    size_t first = d.gus_lines.size();
    for (size_t l = 0; l < d.gus_lines.size(); ++l) {
        const std::string line = strip_comment(d.gus_lines[l]);
        if (line.find('.') != std::string::npos) {
            first = l;
            break;
        }
        const auto t = split(line);
        if (!t.empty()) {
            const auto e = expand(t, true);
            counts.insert(counts.end(), e.begin(), e.end());
        }
    }
    if (first == d.gus_lines.size())
        throw std::runtime_error("GUESS=READ could not find coefficient data after the $GUS header");
    if ((int)counts.size() < p.n_act) throw std::runtime_error("GUESS=READ needs one $GUS header entry per VB orbital");
    const int n = p.n_bf;
    std::vector<double> dense((size_t)p.n_act * n, 0.0);
    std::vector<std::string> buf;
    size_t at = 0;
    int orb = 0;
    for (size_t l = first; l < d.gus_lines.size() && orb < p.n_act; ++l) {
        const auto t = split(strip_comment(d.gus_lines[l]));
        buf.insert(buf.end(), t.begin(), t.end());
        while (orb < p.n_act && buf.size() - at >= (size_t)counts[orb] * 2) {
            for (int k = 0; k < counts[orb]; ++k) {
                const double v = parse_double(buf[at + 2 * k]);
                const int f = parse_int(buf[at + 2 * k + 1]);
                if (f <= 0 || f > n) throw std::runtime_error("basis index in $GUS section is out of range");
                dense[(size_t)orb * n + (f - 1)] = v;
            }
            at += (size_t)counts[orb] * 2;
            ++orb;
        }
    }
    if (orb != p.n_act) throw std::runtime_error("GUESS=READ did not find enough orbital coefficient data in $GUS");
    p.C0.assign((size_t)p.n_act * n, 0.0);
    for (int i = 0; i < p.n_act; ++i)
        for (int f : p.support[i]) p.C0[(size_t)i * n + (f - 1)] = dense[(size_t)i * n + (f - 1)];
}

void normalize_guess(Problem& p) {
    const int n = p.n_bf;
    for (int i = 0; i < p.n_act; ++i) {
        double norm = 0.0;
        for (int a : p.support[i])
            for (int b : p.support[i])
                norm += p.C0[(size_t)i * n + a - 1] * p.C0[(size_t)i * n + b - 1] * p.S[(a - 1) + (size_t)n * (b - 1)];
        if (!(norm > 0.0)) throw std::runtime_error("encountered non-positive orbital norm during the guess build");
        const double s = 1.0 / std::sqrt(norm);
        for (int a : p.support[i]) p.C0[(size_t)i * n + a - 1] *= s;
    }
}

std::filesystem::path write_orbitals(const std::string& input_path, const Problem& p, const std::vector<double>& C) {
    auto path = std::filesystem::absolute(input_path);
    path.replace_extension(".orb");
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot open orbital output: " + path.string());
    out.imbue(std::locale::classic());
    const int n_orb = p.n_act;
    for (int i = 0; i < n_orb; ++i) {
        out << std::setw(6) << p.support[i].size();
        if ((i + 1) % 20 == 0 || i + 1 == n_orb) out << '\n';
    }
    out << std::fixed << std::setprecision(10);
    for (int i = 0; i < n_orb; ++i) {
        const int count = (int)p.support[i].size();
        out << "# ORBITAL " << i + 1 << " NAO = " << count << '\n';
        for (int j = 0; j < count; ++j) {
            const int f = p.support[i][j];
            out << std::setw(14) << C[(size_t)i * p.n_bf + f - 1] << ' ' << std::setw(6) << f;
            out << (((j + 1) % 4 == 0 || j + 1 == count) ? '\n' : ' ');
        }
    }
    out.close();
    if (!out) throw std::runtime_error("cannot write orbital output: " + path.string());
    return path;
}

struct Options {
    std::string input;
    int max_iterations = -1;
    int eigensolver = -1;
    bool dry_run = false;
};

Options parse_args(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--max-iterations" && i + 1 < argc) {
            o.max_iterations = parse_int(argv[++i]);
        } else if (a == "--eigensolver" && i + 1 < argc) {
            const std::string v = argv[++i];
            if (v == "dense") o.eigensolver = VBE_EIGENSOLVER_DENSE;
            else if (v == "davidson") o.eigensolver = VBE_EIGENSOLVER_DAVIDSON;
            else throw std::invalid_argument("invalid structure eigensolver: " + v);
        } else if (a == "--dry-run") {
            o.dry_run = true;
        } else if (!a.empty() && a[0] == '-') {
            throw std::invalid_argument("unknown option: " + a);
        } else {
            o.input = a;
        }
    }
    if (o.input.empty())
        throw std::invalid_argument("usage: hipvbscf [--max-iterations N] [--eigensolver dense|davidson] [--dry-run] deck.xmi");
    return o;
}

Problem build_problem(const Deck& d, const Options& o) {
    int32_t method = -1;
    check(vbe_parse_ctrl_method(d.ctrl_text.c_str(), &method), "$CTRL");
    if (method != VBE_METHOD_HIPVBSCF) throw std::invalid_argument("hipvbscf requires a HIPVBSCF input ($CTRL HIPVBSCF)");
    if (d.str_present) throw std::invalid_argument("HIPVBSCF takes its structures from STR=; explicit $STR is not supported");
    Problem p;
    p.basis = upper(get(d, "BASIS", ""));
    if (p.basis.empty()) throw std::invalid_argument("BASIS= is required");
    p.str_class = upper(get(d, "STR", ""));
    if (p.str_class.empty()) throw std::invalid_argument("STR= is required (structure class keyword)");
    p.n_act = parse_int(get(d, "NAO", "0"));
    const int nae = parse_int(get(d, "NAE", "0"));
    p.multiplicity = parse_int(get(d, "NMUL", "1"));
    const int charge = parse_int(get(d, "NCHARGE", "0"));
    if (p.n_act <= 0 || nae <= 0) throw std::invalid_argument("NAO and NAE must be positive");
    if ((nae + p.multiplicity - 1) % 2 != 0 || p.multiplicity < 1 || p.multiplicity - 1 > nae)
        throw std::invalid_argument("NAE and NMUL are inconsistent");
    p.n_alpha = (nae + p.multiplicity - 1) / 2;
    p.n_beta = nae - p.n_alpha;
    const std::string iscf = get(d, "ISCF", "");
    if (!iscf.empty() && parse_int(iscf) != 8) {
        const int v = parse_int(iscf);
        if (v != 7 && v != 9)
            throw std::invalid_argument("unsupported ISCF value " + iscf +
                                        "; use ISCF=7 for TNHVP, ISCF=8 for P-LBFGS, or ISCF=9 for CIAH");
        throw std::invalid_argument("HIPVBSCF supports P-LBFGS (ISCF=8) only");
    }
    const std::string es = upper(get(d, "EIGENSOLVER", "DAVIDSON"));
    if (es == "DAVIDSON") p.eigensolver = VBE_EIGENSOLVER_DAVIDSON;
    else if (es == "DENSE") p.eigensolver = VBE_EIGENSOLVER_DENSE;
    else throw std::invalid_argument("unsupported EIGENSOLVER value " + es + "; use EIGENSOLVER=DAVIDSON or EIGENSOLVER=DENSE");
    if (o.eigensolver >= 0) p.eigensolver = o.eigensolver;
    const std::string itmax = get(d, "ITMAX", "");
    if (!itmax.empty() && parse_int(itmax) > 0) p.max_iterations = parse_int(itmax);
    if (o.max_iterations >= 0) p.max_iterations = o.max_iterations;
    // NSTATE and WSTATE(1)=w,w,... (equal weights only), as apply_ctrl_assignment
    // (src/input/deck/model.cpp:167-209): the state count is the larger of the two
    int n_states = parse_int(get(d, "NSTATE", "1"));
    if (n_states <= 0) throw std::invalid_argument("NSTATE must be positive");
    for (const auto& [key, value] : d.assign) {
        if (key.rfind("WSTATE(", 0) != 0 || key.back() != ')') continue;
        if (parse_int(key.substr(7, key.size() - 8)) != 1)
            throw std::invalid_argument("equal-weight state averaging supports only WSTATE(1)");
        std::vector<double> w;
        size_t begin = 0;
        for (;;) {
            const size_t comma = value.find(',', begin);
            const std::string tok = value.substr(begin, comma == std::string::npos ? std::string::npos : comma - begin);
            if (tok.find_first_not_of(" \t") == std::string::npos) throw std::invalid_argument("WSTATE contains an empty weight");
            std::string f = tok;  // Fortran exponent letters D/d read as E
            for (char& ch : f)
                if (ch == 'D' || ch == 'd') ch = 'E';
            w.push_back(std::stod(f));
            if (comma == std::string::npos) break;
            begin = comma + 1;
        }
        if (!(w.front() > 0.0)) throw std::invalid_argument("WSTATE weights must be positive");
        for (double x : w)
            if (!(x > 0.0) || std::abs(x - w.front()) > 1e-12 * std::max({1.0, std::abs(w.front()), std::abs(x)}))
                throw std::invalid_argument("only equal positive WSTATE weights are supported");
        n_states = std::max(n_states, (int)w.size());
    }
    p.n_states = n_states;
    if (upper(get(d, "INT", "")) == "RI")
        throw std::invalid_argument("INT=RI needs an auxiliary basis, which this build does not carry; use exact integrals");
    // geometry
    if (d.atoms.empty()) throw std::invalid_argument("$GEO is empty");
    const bool bohr = upper(get(d, "UNIT", "ANGS")) == "BOHR";
    p.n_atoms = (int)d.atoms.size();
    int electrons = -charge;
    for (const auto& a : d.atoms) {
        p.Z.push_back(element_z(a.symbol));
        electrons += p.Z.back();
        for (int k = 0; k < 3; ++k) p.xyz_bohr.push_back(bohr ? a.xyz[k] : a.xyz[k] * kBohrPerAngstrom);
    }
    if (electrons < nae || (electrons - nae) % 2 != 0)
        throw std::invalid_argument("NAE is inconsistent with the molecule's electron count");
    check(vbe_check_hipvbscf_options(p.eigensolver, 1, n_states, electrons - nae), "HIPVBSCF options");
    // integrals: exact ERIs as pivoted-Cholesky factors
    int32_t nbf = 0;
    check(vbe_gaussian_basis_size(p.basis.c_str(), p.n_atoms, p.Z.data(), &nbf), "basis");
    p.n_bf = nbf;
    const int n = nbf, npair = n * (n + 1) / 2;
    std::vector<double> T((size_t)n * n), V((size_t)n * n), eri((size_t)npair * npair);
    p.S.resize((size_t)n * n);
    p.atom_of.resize(n);
    check(vbe_ao_integrals(p.basis.c_str(), p.n_atoms, p.Z.data(), p.xyz_bohr.data(), p.S.data(), T.data(), V.data(),
                           eri.data(), p.atom_of.data(), &p.e_nuc),
          "AO integrals");
    p.h.resize((size_t)n * n);
    for (size_t k = 0; k < p.h.size(); ++k) p.h[k] = T[k] + V[k];
    std::vector<double> Bfull((size_t)npair * npair);
    int32_t rank = 0;
    check(vbe_pivoted_cholesky(npair, eri.data(), 1e-13, npair, &rank, Bfull.data()), "pivoted Cholesky");
    p.n_aux = rank;
    p.B.assign((size_t)rank * npair, 0.0);  // rank x npair, column-major
    for (int c = 0; c < npair; ++c)
        for (int r = 0; r < rank; ++r) p.B[r + (size_t)rank * c] = Bfull[r + (size_t)npair * c];
    // supports and guess
    p.support = build_supports(d, p, upper(get(d, "ORBTYP", "GEN")), upper(get(d, "FRGTYP", "ATOM")));
    std::string guess = upper(get(d, "GUESS", d.gus_present ? "READ" : "AUTO"));
    if (guess == "RDCI") guess = "READ";
    p.guess = guess;
    if (guess == "AUTO") hcore_block_guess(p);
    else if (guess == "READ") {
        if (!d.gus_present) throw std::runtime_error("GUESS=READ requires a $GUS block");
        read_guess(d, p);
    } else throw std::invalid_argument("HIPVBSCF supports GUESS=AUTO and GUESS=READ; got GUESS=" + guess);
    normalize_guess(p);
    return p;
}

const char* termination_name(int t) {
    switch (t) {
        case VBE_TERM_MAX_ITERATIONS: return "max_iterations";
        case VBE_TERM_INITIAL_TOLERANCE: return "plbfgs_initial_tolerance";
        case VBE_TERM_CONVERGED: return "plbfgs_projected_gradient_tolerance";
        case VBE_TERM_LINE_SEARCH_FAILED: return "plbfgs_line_search_failed";
        case VBE_TERM_LINE_SEARCH_STALLED: return "plbfgs_line_search_stalled";
        case VBE_TERM_NON_DESCENT: return "plbfgs_non_descent_direction";
        case VBE_TERM_EMPTY_SPACE: return "nonredundant_space_empty";
    }
    return "unknown";
}

int run(const Options& o) {
    const auto start = std::chrono::steady_clock::now();
    const Deck d = read_deck(o.input);
    Problem p = build_problem(d, o);
    vbe_topology* topo = nullptr;
    int64_t sizes[5] = {0, 0, 0, 0, 0};
    check(vbe_topology_build(p.str_class.c_str(), p.n_act, p.n_alpha + p.n_beta, p.multiplicity, &topo, sizes),
          "structures");
    vbe_topology_view view{};
    check(vbe_topology_view_of(topo, &view), "structures");
    std::cout << std::setprecision(16);
    std::cout << " XMVB-CPP HIPVBSCF (MI355X engine)\n Input: " << o.input << "\n Title: " << d.title << "\n"
              << " Basis: " << p.basis << "  basis functions: " << p.n_bf << "  atoms: " << p.n_atoms << "\n"
              << " Active orbitals: " << p.n_act << "  electrons: " << p.n_alpha + p.n_beta << " (" << p.n_alpha
              << " alpha, " << p.n_beta << " beta)  structures (STR=" << p.str_class << "): " << sizes[0]
              << "  determinants: " << sizes[3] << "\n"
              << " Two-electron integrals: exact, as pivoted-Cholesky factors of rank " << p.n_aux << "\n"
              << " Optimizer: P-LBFGS (ISCF=8)  eigensolver: "
              << (p.eigensolver == VBE_EIGENSOLVER_DAVIDSON ? "DAVIDSON" : "DENSE") << "  guess: " << p.guess
              << "  itmax: " << p.max_iterations << "\n Nuclear repulsion: " << p.e_nuc << "\n";
    for (int i = 0; i < p.n_act; ++i) {
        std::cout << " Orbital " << i + 1 << " support:";
        for (int f : p.support[i]) std::cout << ' ' << f;
        std::cout << "\n";
    }
    if (o.dry_run) {
        std::cout << "\n Guess orbitals (coefficient, AO index):\n" << std::fixed << std::setprecision(10);
        for (int i = 0; i < p.n_act; ++i) {
            std::cout << " " << i + 1 << ":";
            for (int f : p.support[i]) std::cout << ' ' << p.C0[(size_t)i * p.n_bf + f - 1] << ' ' << f;
            std::cout << "\n";
        }
        vbe_topology_free(topo);
        return 0;
    }
    vbe_context* ctx = nullptr;
    check(vbe_context_create(0, 0, &ctx), "device context");
    check_ctx(ctx,
              vbe_set_eigensolver(ctx, p.eigensolver, 1e-7, 1e-3, 1.0, 1), "eigensolver");
    if (p.n_states > 1) check_ctx(ctx, vbe_set_states(ctx, p.n_states, nullptr, nullptr), "states");
    vbe_dims dims{p.n_bf, p.n_act, p.n_alpha, p.n_beta, p.n_aux};
    vbe_problem* prob = nullptr;
    check_ctx(ctx, vbe_problem_upload(ctx, &dims, p.S.data(), p.h.data(), p.B.data(), &view, p.e_nuc, &prob),
              "problem upload");
    std::vector<uint8_t> mask((size_t)p.n_act * p.n_bf, 0);
    for (int i = 0; i < p.n_act; ++i)
        for (int f : p.support[i]) mask[(size_t)i * p.n_bf + f - 1] = 1;
    vbe_optimizer_options opt;
    vbe_optimizer_defaults(&opt);
    opt.max_iterations = p.max_iterations;
    std::vector<double> C = p.C0, grad(C.size()), trace(p.max_iterations + 1);
    vbe_optimizer_result res{};
    const auto t0 = std::chrono::steady_clock::now();
    check_ctx(ctx, vbe_optimize(ctx, prob, C.data(), mask.data(), &opt, &res, grad.data(), trace.data(),
                                (int32_t)trace.size()),
              "P-LBFGS");
    const double scf_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "\n Iter          Energy (Eh)            Delta E\n";
    for (int k = 0; k <= res.iterations && k < (int)trace.size(); ++k) {
        std::cout << std::setw(5) << k << "  " << std::fixed << std::setprecision(12) << std::setw(22) << trace[k];
        if (k) std::cout << "  " << std::scientific << std::setprecision(3) << trace[k] - trace[k - 1];
        std::cout << "\n";
    }
    std::cout << std::defaultfloat << std::setprecision(16);
    std::cout << "\n SCF " << (res.converged ? "converged" : "NOT converged") << " after " << res.iterations
              << " iterations; total energy " << std::fixed << std::setprecision(12) << res.energy << " Eh\n"
              << std::defaultfloat << std::setprecision(16);
    std::string orbital_path;
    if (res.converged) {
        orbital_path = write_orbitals(o.input, p, C).string();
        // structure coefficients and Coulson-Chirgwin weights W_I = c_I (S c)_I at the final orbitals
        const int ns = (int)sizes[0];
        double e = 0.0;
        vbe_flags flags{VBE_PRECISION_FP64, p.eigensolver, 0};
        check_ctx(ctx, vbe_eval(ctx, prob, C.data(), mask.data(), flags, &e, nullptr, nullptr), "final evaluation");
        std::vector<double> Hs((size_t)ns * ns), Ss((size_t)ns * ns), c(ns);
        check_ctx(ctx, vbe_last_intermediates(ctx, prob, Hs.data(), Ss.data(), c.data(), nullptr, nullptr, nullptr),
                  "structure matrices");
        if (p.n_states > 1) {
            std::vector<double> es(p.n_states);
            check_ctx(ctx, vbe_last_state_energies(ctx, prob, es.data()), "state energies");
            std::cout << "\n State-averaged over " << p.n_states << " states (equal weights)\n";
            for (int k = 0; k < p.n_states; ++k)
                std::cout << " State " << k + 1 << " total energy " << std::fixed << std::setprecision(12) << es[k]
                          << " Eh\n";
            std::cout << " Coefficients below: state 1\n" << std::defaultfloat << std::setprecision(16);
        }
        std::cout << "\n Structure   Coefficient        Coulson-Chirgwin weight\n";
        for (int I = 0; I < ns; ++I) {
            double sc = 0.0;
            for (int J = 0; J < ns; ++J) sc += Ss[I + (size_t)ns * J] * c[J];
            std::cout << std::setw(10) << I + 1 << "  " << std::fixed << std::setprecision(10) << std::setw(16) << c[I]
                      << "  " << std::setw(16) << c[I] * sc << "\n";
        }
        std::cout << "\n Optimized orbitals (coefficient, AO index):\n";
        for (int i = 0; i < p.n_act; ++i) {
            std::cout << " " << i + 1 << ":";
            for (int f : p.support[i]) std::cout << ' ' << C[(size_t)i * p.n_bf + f - 1] << ' ' << f;
            std::cout << "\n";
        }
        std::cout << std::defaultfloat << std::setprecision(16);
    }
    std::cout << "\n XMVB-CPP RUN DIAGNOSTICS\n\n"
              << "method=hipvbscf_plbfgs\nstates=" << p.n_states << "\nstructures=" << sizes[0] << "\ndeterminants=" << sizes[3]
              << "\nbasis_functions=" << p.n_bf << "\nactive_orbitals=" << p.n_act << "\nfactor_rank=" << p.n_aux
              << "\n";
    if (!orbital_path.empty()) std::cout << "orbital_output=" << orbital_path << "\n";
    std::cout << "converged=" << res.converged << "\ntermination=" << termination_name(res.termination)
              << "\ntotal_energy_hartree=" << res.energy << "\nscf_iterations=" << res.iterations
              << "\nenergy_evaluations=" << res.energy_evaluations << "\ngradient_evaluations="
              << res.gradient_evaluations << "\ngradient_norm=" << res.gradient_norm << "\nscf_seconds=" << scf_seconds
              << "\nprocess_seconds="
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() << "\n";
    vbe_problem_release(prob);
    vbe_context_destroy(ctx);
    vbe_topology_free(topo);
    return res.converged ? 0 : 2;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(parse_args(argc, argv));
    } catch (const std::exception& e) {
        std::cerr << "XMVB-CPP error: " << e.what() << '\n';
        return 2;
    }
}
