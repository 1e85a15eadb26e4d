// Structure generation and determinant expansion for the VBSCF engine (host C++23).
//
// Restates the reference's structure-class generator
// (src/input/loading/generated_structures.cpp:131-716) and StructureExpander::expand
// (src/vbscf/structures/expansion/expander.cpp:188-558, bitmask path) so that structure order,
// unique-string order, determinant order and term signs are bit-identical.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace xmvb_amd {

struct RawStructureSet {
    int n_active_orbitals = 0;
    int n_active_electrons = 0;
    int spin_multiplicity = 1;
    // n_structures x n_active_electrons orbital labels (1-based), bond pairs first, then open shells
    std::vector<std::vector<int>> rows;
};

struct DeterminantTopology {
    int n_structures = 0;
    int n_alpha_electrons = 0;
    int n_beta_electrons = 0;
    std::vector<std::uint64_t> alpha_masks;  // ascending
    std::vector<std::uint64_t> beta_masks;   // ascending
    std::vector<std::int32_t> det_alpha;     // per determinant, ordered by (alpha rank, beta rank)
    std::vector<std::int32_t> det_beta;
    std::vector<std::int64_t> term_offsets;  // CSR over determinants
    std::vector<std::int32_t> term_structure;
    std::vector<double> term_coefficient;
};

// Closed-form structure count (generated_structures.cpp:621-658).
std::uint64_t count_generated_structures(const std::string& structure_class, int n_active_orbitals,
                                         int n_active_electrons, int spin_multiplicity);

// Raw structures of a structure class keyword ("COV", "FULL", "ION", "COV ION(1,2)", ...).
RawStructureSet generate_structures(const std::string& structure_class, int n_active_orbitals,
                                    int n_active_electrons, int spin_multiplicity);

// Signed determinant expansion of raw structures.
DeterminantTopology expand_structures(const RawStructureSet& raw);

}  // namespace xmvb_amd
