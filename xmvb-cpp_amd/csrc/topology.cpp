// Structure generation and determinant expansion (host C++23). See topology.hpp.
#include "topology.hpp"

#include <algorithm>
#include <bit>
#include <cctype>
#include <numeric>
#include <stdexcept>
#include <unordered_map>
#include <utility>

namespace xmvb_amd {
namespace {

std::uint64_t choose(int n, int k) {
    if (k < 0 || k > n) return 0;
    k = std::min(k, n - k);
    unsigned __int128 acc = 1;
    for (int i = 1; i <= k; ++i) acc = acc * static_cast<unsigned>(n - k + i) / static_cast<unsigned>(i);
    return static_cast<std::uint64_t>(acc);
}

// number of spin tableaux (generated_structures.cpp:131-143)
std::uint64_t tableaux(int n, int multiplicity) {
    if (n < 0 || multiplicity <= 0 || multiplicity > n + 1) return 0;
    if (n == 0) return 1;
    if ((n - multiplicity + 1) % 2 != 0) return 0;
    const int beta = (n - multiplicity + 1) / 2;
    return choose(n, beta) - (beta > 0 ? choose(n, beta - 1) : 0);
}

struct Spins {
    int alpha, beta;
};

Spins active_spins(int n_electrons, int multiplicity) {
    const int beta = (n_electrons - multiplicity + 1) / 2;
    return {n_electrons - beta, beta};
}

bool is_single_reference(int n_orbitals, int n_electrons, int multiplicity) {
    const int beta = active_spins(n_electrons, multiplicity).beta;
    return (beta == 0 && n_electrons == multiplicity - 1 && n_electrons == n_orbitals) ||
           (beta == n_orbitals && multiplicity == 1);
}

std::string upper(std::string s) {
    for (auto& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

// "a:b", "a-b", "a*k" and plain integers separated by commas/blanks (generated_structures.cpp:35-101)
void expand_token(const std::string& tok, std::vector<int>& out) {
    auto to_int = [](const std::string& t) {
        std::size_t used = 0;
        const int v = std::stoi(t, &used);
        if (used != t.size()) throw std::runtime_error("invalid integer token in structure class: " + t);
        return v;
    };
    if (auto p = tok.find(':'); p != std::string::npos) {
        const int a = to_int(tok.substr(0, p)), b = to_int(tok.substr(p + 1));
        if (b < a) throw std::runtime_error("descending ':' range is not supported in structure class");
        for (int v = a; v <= b; ++v) out.push_back(v);
    } else if (auto q = tok.find('-'); q != std::string::npos) {
        const int a = to_int(tok.substr(0, q)), b = to_int(tok.substr(q + 1));
        if (b < a) throw std::runtime_error("descending '-' range is not supported in structure class");
        for (int v = a; v <= b; ++v) out.push_back(v);
    } else if (auto r = tok.find('*'); r != std::string::npos) {
        const int v = to_int(tok.substr(0, r)), k = to_int(tok.substr(r + 1));
        if (k < 0) throw std::runtime_error("negative repetition count is not supported in structure class");
        out.insert(out.end(), static_cast<std::size_t>(k), v);
    } else {
        out.push_back(to_int(tok));
    }
}

std::vector<int> class_list(const std::string& keyword, int n_orbitals, int n_electrons, int multiplicity) {
    const std::string kw = upper(keyword);
    const int forced_pairs = std::max(0, n_electrons - n_orbitals);
    const int reduced_beta = (n_electrons - 2 * forced_pairs - multiplicity + 1) / 2;
    if (reduced_beta < 0) throw std::runtime_error("structure class generation requires non-negative reduced beta count");
    std::vector<int> classes;
    if (kw.find("FULL") != std::string::npos) {
        for (int k = 0; k <= reduced_beta; ++k) classes.push_back(k);
    } else {
        if (kw.find("COV") != std::string::npos) classes.push_back(0);
        if (auto open = kw.find("ION("); open != std::string::npos) {
            const auto close = kw.find(')', open + 4);
            if (close == std::string::npos) throw std::runtime_error("structure class contains ION( without a closing ')'");
            std::string body = kw.substr(open + 4, close - open - 4), tok;
            for (char c : body + " ") {
                if (c == ',' || std::isspace(static_cast<unsigned char>(c))) {
                    if (!tok.empty()) expand_token(tok, classes);
                    tok.clear();
                } else {
                    tok.push_back(c);
                }
            }
        } else if (kw.find("ION") != std::string::npos) {
            for (int k = 1; k <= reduced_beta; ++k) classes.push_back(k);
        }
    }
    if (classes.empty()) throw std::runtime_error("unsupported or empty structure class specification: " + keyword);
    std::ranges::sort(classes);
    classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
    if (classes.front() < 0) throw std::runtime_error("structure class cannot be smaller than 0");
    if (classes.back() > reduced_beta) throw std::runtime_error("structure class exceeds the maximum ionic order");
    for (auto& c : classes) c += forced_pairs;
    return classes;
}

// Increasing label rows bounded by Young-diagram maxima, walked by "bump the rightmost slot that
// is below its maximum and refill the suffix consecutively" (generated_structures.cpp:246-299).
template <typename Visit>
void walk_young_rows(int length, int n_beta, int n_orbitals, Visit&& visit) {
    if (length <= 0) return;
    if (n_orbitals <= 0 || n_orbitals < length) throw std::runtime_error("invalid orbital count in structure generator");
    std::vector<int> ceiling(length, 0);
    if (length == n_beta) {
        ceiling[n_beta - 1] = n_orbitals - 1;
    } else {
        ceiling[length - 1] = n_orbitals;
        for (int i = length - 1; i > n_beta; --i) ceiling[i - 1] = ceiling[i] - 1;
        if (n_beta > 0) ceiling[n_beta - 1] = ceiling[n_beta] - 2;
    }
    for (int i = n_beta - 1; i >= 1; --i) ceiling[i - 1] = ceiling[i] - 2;
    std::vector<int> row(length);
    std::iota(row.begin(), row.end(), 1);
    for (;;) {
        visit(row);
        int slot = length - 1;
        while (slot >= 0 && row[slot] == ceiling[slot]) --slot;
        if (slot < 0) return;
        ++row[slot];
        for (int j = slot + 1; j < length; ++j) row[j] = row[j - 1] + 1;
    }
}

// Pair each alpha label, largest first, with the smallest unclaimed beta label that is not below
// it, then write the alpha labels into their partner's beta position (generated_structures.cpp:393-432).
// Beta labels may exceed 63 through open-shell padding, so the unclaimed set is a small vector.
void match_rumer_bonds(std::vector<int>& alpha, const std::vector<int>& beta, int n_vector, int n_row) {
    std::vector<int> unclaimed(beta.begin(), beta.begin() + n_row);
    std::vector<int> partner_slot(n_vector, -1);  // for alpha position v: beta slot of its partner
    for (int v = n_vector - 1; v >= 0; --v) {
        // the reference scans only the first (v+1) unclaimed entries
        const int a = alpha[v];
        const int window = v + 1;
        for (int j = 0; j < window && j < static_cast<int>(unclaimed.size()); ++j) {
            if (a <= unclaimed[j]) {
                const int label = unclaimed[j];
                for (int s = 0; s < n_vector; ++s)
                    if (beta[s] == label) {
                        partner_slot[v] = s;
                        break;
                    }
                unclaimed.erase(unclaimed.begin() + j);
                break;
            }
        }
    }
    std::vector<int> placed(n_vector, 0);
    for (int v = 0; v < n_vector; ++v) {
        if (partner_slot[v] < 0) throw std::runtime_error("Rumer bond matching left an alpha label unpaired");
        placed[partner_slot[v]] = alpha[v];
    }
    std::copy(placed.begin(), placed.end(), alpha.begin());
}

template <typename Visit>
void walk_covalent_pairs(int n_alpha, int n_beta, int n_orbitals, int stride, Visit&& visit) {
    if (n_alpha <= 0) return;
    auto pad_open_shells = [&](std::vector<int>& beta) {
        for (int i = n_beta; i < n_alpha; ++i) beta[i] = n_orbitals + n_alpha - i;
    };
    if (n_beta <= 0) {
        walk_young_rows(n_alpha, 0, n_orbitals, [&](const std::vector<int>& a) {
            std::vector<int> alpha(stride, 0), beta(stride, 0);
            std::copy(a.begin(), a.end(), alpha.begin());
            visit(alpha, beta);
        });
        return;
    }
    if (n_alpha + n_beta == n_orbitals) {
        walk_young_rows(n_alpha, n_beta, n_orbitals, [&](const std::vector<int>& a) {
            std::vector<int> alpha(stride, 0), beta(stride, 0);
            std::copy(a.begin(), a.end(), alpha.begin());
            std::uint64_t used = 0;
            for (int v : a) used |= std::uint64_t{1} << (v - 1);
            int next = 0;
            for (int label = 1; label <= n_orbitals; ++label)
                if (!(used >> (label - 1) & 1)) beta[next++] = label;
            pad_open_shells(beta);
            visit(alpha, beta);
        });
        return;
    }
    walk_young_rows(n_alpha, n_beta, n_orbitals, [&](const std::vector<int>& a) {
        walk_young_rows(n_beta, 0, n_orbitals, [&](const std::vector<int>& b) {
            for (int i = 0; i < n_beta; ++i) {  // generated_structures.cpp:301-316
                if (b[i] < a[i]) return;
                for (int j = i; j < n_alpha; ++j)
                    if (b[i] == a[j]) return;
            }
            std::vector<int> alpha(stride, 0), beta(stride, 0);
            std::copy(a.begin(), a.end(), alpha.begin());
            std::copy(b.begin(), b.end(), beta.begin());
            pad_open_shells(beta);
            visit(alpha, beta);
        });
    });
}

}  // namespace

std::uint64_t count_generated_structures(const std::string& structure_class, int n_orbitals, int n_electrons,
                                         int multiplicity) {
    const auto spins = active_spins(n_electrons, multiplicity);
    if (spins.beta < 0) throw std::runtime_error("active beta electron count is negative in structure generator");
    if (spins.beta > n_orbitals) throw std::runtime_error("active beta electron count exceeds the active orbital count");
    if (is_single_reference(n_orbitals, n_electrons, multiplicity)) return 1;
    const int skipped = std::max(0, n_orbitals - n_electrons);
    const int forced_pairs = std::max(0, n_electrons - n_orbitals);
    std::uint64_t total = 0;
    for (int ion : class_list(structure_class, n_orbitals, n_electrons, multiplicity)) {
        const int tab = n_orbitals - 2 * ion - skipped + forced_pairs;
        if (tab < 0) continue;
        total += choose(n_orbitals, ion) * choose(n_orbitals - ion, skipped + ion - forced_pairs) *
                 tableaux(tab, multiplicity);
    }
    if (total == 0) throw std::runtime_error("structure generator produced no raw structures");
    return total;
}

RawStructureSet generate_structures(const std::string& structure_class, int n_orbitals, int n_electrons,
                                    int multiplicity) {
    RawStructureSet out;
    out.n_active_orbitals = n_orbitals;
    out.n_active_electrons = n_electrons;
    out.spin_multiplicity = multiplicity;
    const auto spins = active_spins(n_electrons, multiplicity);
    const std::uint64_t expected = count_generated_structures(structure_class, n_orbitals, n_electrons, multiplicity);
    if (is_single_reference(n_orbitals, n_electrons, multiplicity)) {
        std::vector<int> row(n_electrons, 0);
        for (int i = 0; i < spins.beta; ++i) row[2 * i] = row[2 * i + 1] = i + 1;
        for (int o = 0; o < multiplicity - 1; ++o) row[2 * spins.beta + o] = spins.beta + o + 1;
        out.rows.push_back(std::move(row));
        return out;
    }
    for (int ion : class_list(structure_class, n_orbitals, n_electrons, multiplicity)) {
        const int cov_alpha = spins.alpha - ion, cov_beta = spins.beta - ion;
        if (ion < 0 || ion > spins.beta) throw std::runtime_error("structure class exceeds the active beta electron count");
        if (cov_alpha < 0 || cov_beta < 0) throw std::runtime_error("structure class produces negative covalent dimensions");
        // generated_structures.cpp:470-509
        auto write = [&](const std::vector<int>& ions, const std::vector<int>& relabel, const std::vector<int>& a,
                         const std::vector<int>& b) {
            std::vector<int> row(n_electrons, 0);
            auto map = [&](int label) { return ion > 0 ? relabel[label - 1] : label; };
            int at = 0;
            for (int i = 0; i < ion; ++i) {
                row[at++] = ions[i];
                row[at++] = ions[i];
            }
            for (int i = 0; i < cov_beta; ++i) {
                row[at++] = map(a[i]);
                row[at++] = map(b[i]);
            }
            for (int i = cov_beta; i < cov_alpha; ++i) row[at++] = map(a[i]);
            out.rows.push_back(std::move(row));
        };
        if (ion == 0) {
            walk_covalent_pairs(spins.alpha, spins.beta, n_orbitals, spins.alpha,
                                [&](std::vector<int> a, const std::vector<int>& b) {
                                    if (spins.beta > 0) match_rumer_bonds(a, b, spins.alpha, spins.alpha);
                                    write({}, {}, a, b);
                                });
            continue;
        }
        walk_young_rows(ion, 0, n_orbitals, [&](const std::vector<int>& ions) {
            std::uint64_t taken = 0;
            for (int i = 0; i < ion; ++i) taken |= std::uint64_t{1} << ions[i];
            std::vector<int> relabel;
            for (int label = 1; label <= n_orbitals; ++label)
                if (!(taken >> label & 1)) relabel.push_back(label);
            if (cov_alpha <= 0) {
                write(ions, relabel, {}, {});
                return;
            }
            walk_covalent_pairs(cov_alpha, cov_beta, n_orbitals - ion, spins.alpha,
                                [&](std::vector<int> a, const std::vector<int>& b) {
                                    if (cov_beta > 0) match_rumer_bonds(a, b, cov_alpha, spins.alpha);
                                    write(ions, relabel, a, b);
                                });
        });
    }
    if (out.rows.size() != expected) throw std::runtime_error("raw structure generation count mismatch");
    return out;
}

DeterminantTopology expand_structures(const RawStructureSet& raw) {
    const int n_el = raw.n_active_electrons;
    const int n_open = raw.spin_multiplicity - 1;
    const int n_paired = n_el - n_open;
    if (n_el <= 0 || n_open < 0 || n_paired < 0 || n_paired % 2 != 0)
        throw std::invalid_argument("raw structure electron and spin counts are inconsistent");
    if (raw.n_active_orbitals > 64) throw std::invalid_argument("more than 64 active orbitals are not supported");
    const int n_bonds = n_paired / 2;
    DeterminantTopology topo;
    topo.n_structures = static_cast<int>(raw.rows.size());
    topo.n_alpha_electrons = n_el - n_bonds;
    topo.n_beta_electrons = n_bonds;

    // Discovery pass over every structure's 2^(#covalent bonds) determinants. The sign of the
    // starting determinant is the product of the inversion parities of the bond-ordered alpha
    // and beta lists; swapping bond (a,b) between spins flips it exactly when an odd number of
    // singly occupied orbitals lie strictly between a and b (expander.cpp:188-283). Determinants
    // are emitted in binary-counting order of the swap choices, bond 0 fastest.
    struct Emission {
        std::uint64_t alpha, beta;
        std::int32_t structure;
        std::int8_t sign;
    };
    std::vector<Emission> emitted;
    auto inversion_sign = [](std::uint64_t& mask, int orbital, int sign) {
        const std::uint64_t flag = std::uint64_t{1} << orbital;
        if (mask & flag) throw std::invalid_argument("duplicate orbital in an occupied spin string");
        const std::uint64_t above = orbital == 63 ? 0 : (mask >> (orbital + 1));
        if (std::popcount(above) & 1) sign = -sign;
        mask |= flag;
        return sign;
    };
    for (std::size_t s = 0; s < raw.rows.size(); ++s) {
        const auto& row = raw.rows[s];
        std::uint64_t am = 0, bm = 0;
        int sign = 1;
        for (int p = 0; p < n_bonds; ++p) sign = inversion_sign(am, row[2 * p] - 1, sign);
        for (int i = 2 * n_bonds; i < n_el; ++i) sign = inversion_sign(am, row[i] - 1, sign);
        int bsign = 1;
        for (int p = 0; p < n_bonds; ++p) bsign = inversion_sign(bm, row[2 * p + 1] - 1, bsign);
        sign *= bsign;
        struct Swap {
            std::uint64_t exchange;
            bool flips;
        };
        std::vector<Swap> swaps;
        for (int p = 0; p < n_bonds; ++p) {
            const int a = row[2 * p] - 1, b = row[2 * p + 1] - 1;
            if (a == b) continue;
            const int lo = std::min(a, b), hi = std::max(a, b);
            const std::uint64_t between = ((std::uint64_t{1} << hi) - 1) & ~((std::uint64_t{1} << (lo + 1)) - 1);
            swaps.push_back({(std::uint64_t{1} << a) | (std::uint64_t{1} << b),
                             (std::popcount((am ^ bm) & between) & 1) != 0});
        }
        const std::uint64_t n_choices = std::uint64_t{1} << swaps.size();
        std::uint64_t state = 0;
        for (std::uint64_t c = 0;; ++c) {
            emitted.push_back({am, bm, static_cast<std::int32_t>(s), static_cast<std::int8_t>(sign)});
            if (c + 1 == n_choices) break;
            // binary increment: every bit that carries and the bit that is set each swap one bond
            const std::uint64_t next = state + 1, changed = state ^ next;
            for (std::uint64_t bits = changed; bits; bits &= bits - 1) {
                const auto& sw = swaps[std::countr_zero(bits)];
                am ^= sw.exchange;
                bm ^= sw.exchange;
                if (sw.flips) sign = -sign;
            }
            state = next;
        }
    }

    // Intern unique strings and (alpha, beta) pairs in discovery order.
    std::unordered_map<std::uint64_t, std::int32_t> alpha_id, beta_id;
    std::vector<std::uint64_t> alpha_found, beta_found;
    std::unordered_map<std::uint64_t, std::int32_t> pair_id;  // key: alpha_id << 32 | beta_id
    std::vector<std::pair<std::int32_t, std::int32_t>> pairs;
    std::vector<std::int32_t> emission_pair(emitted.size());
    auto intern = [](auto& map, auto& list, std::uint64_t key) {
        auto [it, fresh] = map.try_emplace(key, static_cast<std::int32_t>(list.size()));
        if (fresh) list.push_back(key);
        return it->second;
    };
    for (std::size_t k = 0; k < emitted.size(); ++k) {
        const auto a = intern(alpha_id, alpha_found, emitted[k].alpha);
        const auto b = intern(beta_id, beta_found, emitted[k].beta);
        const std::uint64_t key = (static_cast<std::uint64_t>(a) << 32) | static_cast<std::uint32_t>(b);
        auto [it, fresh] = pair_id.try_emplace(key, static_cast<std::int32_t>(pairs.size()));
        if (fresh) pairs.emplace_back(a, b);
        emission_pair[k] = it->second;
    }
    // Canonical order: strings ascending by mask; determinants by (alpha rank, beta rank).
    auto rank_of = [](const std::vector<std::uint64_t>& found, std::vector<std::uint64_t>& sorted) {
        sorted = found;
        std::ranges::sort(sorted);
        std::unordered_map<std::uint64_t, std::int32_t> pos;
        for (std::size_t i = 0; i < sorted.size(); ++i) pos[sorted[i]] = static_cast<std::int32_t>(i);
        std::vector<std::int32_t> rank(found.size());
        for (std::size_t i = 0; i < found.size(); ++i) rank[i] = pos[found[i]];
        return rank;
    };
    const auto alpha_rank = rank_of(alpha_found, topo.alpha_masks);
    const auto beta_rank = rank_of(beta_found, topo.beta_masks);
    std::vector<std::int32_t> order(pairs.size());
    std::iota(order.begin(), order.end(), 0);
    std::ranges::sort(order, [&](std::int32_t p, std::int32_t q) {
        const auto ka = std::pair{alpha_rank[pairs[p].first], beta_rank[pairs[p].second]};
        const auto kb = std::pair{alpha_rank[pairs[q].first], beta_rank[pairs[q].second]};
        return ka < kb;
    });
    std::vector<std::int32_t> slot(pairs.size());
    for (std::size_t i = 0; i < order.size(); ++i) slot[order[i]] = static_cast<std::int32_t>(i);
    // Counting sort of emissions into determinant rows keeps (structure, emission) order per row.
    std::vector<std::int64_t> counts(pairs.size() + 1, 0);
    for (auto p : emission_pair) ++counts[slot[p] + 1];
    for (std::size_t i = 1; i < counts.size(); ++i) counts[i] += counts[i - 1];
    std::vector<std::pair<std::int32_t, double>> staged(emitted.size());
    {
        auto cursor = counts;
        for (std::size_t k = 0; k < emitted.size(); ++k)
            staged[cursor[slot[emission_pair[k]]]++] = {emitted[k].structure, static_cast<double>(emitted[k].sign)};
    }
    // Coalesce equal structures within a row and drop zero coefficients (types.hpp:46-97).
    topo.term_offsets.assign(1, 0);
    for (std::size_t d = 0; d < pairs.size(); ++d) {
        topo.det_alpha.push_back(alpha_rank[pairs[order[d]].first]);
        topo.det_beta.push_back(beta_rank[pairs[order[d]].second]);
        std::int64_t i = counts[d];
        const std::int64_t end = counts[d + 1];
        std::int32_t previous = -1;
        while (i < end) {
            const std::int32_t st = staged[i].first;
            if (st <= previous) throw std::invalid_argument("invalid structure incidence ordering");
            previous = st;
            double c = 0.0;
            while (i < end && staged[i].first == st) c += staged[i++].second;
            if (c != 0.0) {
                topo.term_structure.push_back(st);
                topo.term_coefficient.push_back(c);
            }
        }
        topo.term_offsets.push_back(static_cast<std::int64_t>(topo.term_structure.size()));
    }
    return topo;
}

}  // namespace xmvb_amd
