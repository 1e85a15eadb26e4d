// Host-only part of the C ABI: $CTRL method flag, option checks, topology and synthetic problems.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>
#include <random>
#include <sstream>
#include <string>

#include "../../include/xmvb_amd.h"
#include "common.hpp"
#include "topology.hpp"

struct vbe_topology {
    xmvb_amd::RawStructureSet raw;
    xmvb_amd::DeterminantTopology topo;
};

namespace xmvb_amd {

thread_local std::string g_global_error;

void set_global_error(const std::string& what) { g_global_error = what; }

namespace {

// Marsaglia polar normal sampler over mt19937_64; the draw order is fixed so that the same
// seed gives the same problem on any host (std::normal_distribution is implementation-defined).
class Normal {
  public:
    explicit Normal(std::uint64_t seed) : engine_(seed) {}
    double operator()() {
        if (has_spare_) {
            has_spare_ = false;
            return spare_;
        }
        for (;;) {
            const double u = 2.0 * unit() - 1.0, v = 2.0 * unit() - 1.0;
            const double s = u * u + v * v;
            if (s >= 1.0 || s == 0.0) continue;
            const double f = std::sqrt(-2.0 * std::log(s) / s);
            spare_ = v * f;
            has_spare_ = true;
            return u * f;
        }
    }

  private:
    double unit() { return static_cast<double>(engine_() >> 11) * 0x1.0p-53; }
    std::mt19937_64 engine_;
    double spare_ = 0.0;
    bool has_spare_ = false;
};

}  // namespace
}  // namespace xmvb_amd

using xmvb_amd::set_global_error;

extern "C" {

const char* vbe_global_error(void) { return xmvb_amd::g_global_error.c_str(); }

int vbe_parse_ctrl_method(const char* ctrl_line, int32_t* method) {
    if (!ctrl_line || !method) {
        set_global_error("vbe_parse_ctrl_method: null argument");
        return VBE_ERR_ARGUMENT;
    }
    std::istringstream in(ctrl_line);
    std::string tok;
    bool explicit_method = false;
    int32_t chosen = VBE_METHOD_VBSCF;
    while (in >> tok) {
        for (auto& c : tok) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
        if (tok.find('=') != std::string::npos) continue;  // KEY=VALUE pairs are not method tokens
        int32_t requested = -1;
        if (tok == "VBSCF") requested = VBE_METHOD_VBSCF;
        else if (tok == "CUVBSCF") requested = VBE_METHOD_CUVBSCF;
        else if (tok == "HIPVBSCF") requested = VBE_METHOD_HIPVBSCF;
        else if (tok == "CASSCF") requested = VBE_METHOD_CASSCF;
        else if (tok == "DFT") requested = VBE_METHOD_DFT;
        else if (tok == "TBVBSCF" || tok == "VBPDFT") {
            set_global_error(tok + " is no longer supported");
            return VBE_ERR_UNSUPPORTED;
        }
        if (requested < 0) continue;
        if (explicit_method && requested != chosen) {
            set_global_error("$CTRL must select only one electronic method");
            return VBE_ERR_ARGUMENT;
        }
        chosen = requested;
        explicit_method = true;
    }
    *method = chosen;
    return VBE_OK;
}

int vbe_check_hipvbscf_options(int32_t eigensolver, int32_t ri, int32_t n_states, int32_t n_inactive_electrons) {
    if (eigensolver != VBE_EIGENSOLVER_DENSE && eigensolver != VBE_EIGENSOLVER_DAVIDSON) {
        set_global_error("HIPVBSCF supports EIGENSOLVER=DENSE and EIGENSOLVER=DAVIDSON only");
        return VBE_ERR_UNSUPPORTED;
    }
    if (!ri) {
        set_global_error("HIPVBSCF takes two-electron integrals as RI-shaped factors (INT=RI, or exact integrals "
                         "factorised by vbe_pivoted_cholesky)");
        return VBE_ERR_UNSUPPORTED;
    }
    if (n_states < 1 || n_states > 8) {
        set_global_error("HIPVBSCF supports 1 to 8 averaged states (NSTATE)");
        return VBE_ERR_UNSUPPORTED;
    }
    if (n_states > 1 && eigensolver != VBE_EIGENSOLVER_DENSE) {
        set_global_error("state-averaged HIPVBSCF (NSTATE > 1) needs EIGENSOLVER=DENSE: the engine's Davidson finds "
                         "the lowest structure root only");
        return VBE_ERR_UNSUPPORTED;
    }
    if (n_inactive_electrons != 0) {
        set_global_error("HIPVBSCF currently requires all electrons to be active");
        return VBE_ERR_UNSUPPORTED;
    }
    return VBE_OK;
}

int vbe_count_structures(const char* structure_class, int32_t n_act, int32_t n_electrons, int32_t multiplicity,
                         uint64_t* count) {
    try {
        if (!structure_class || !count) throw std::invalid_argument("null argument");
        *count = xmvb_amd::count_generated_structures(structure_class, n_act, n_electrons, multiplicity);
        return VBE_OK;
    } catch (const std::exception& e) {
        set_global_error(e.what());
        return VBE_ERR_TOPOLOGY;
    }
}

int vbe_topology_build(const char* structure_class, int32_t n_act, int32_t n_electrons, int32_t multiplicity,
                       vbe_topology** out, int64_t sizes[5]) {
    try {
        if (!structure_class || !out || !sizes) throw std::invalid_argument("null argument");
        if (n_act <= 0 || n_act > 64 || n_electrons <= 0 || multiplicity <= 0)
            throw std::invalid_argument("invalid active-space dimensions");
        auto t = std::make_unique<vbe_topology>();
        t->raw = xmvb_amd::generate_structures(structure_class, n_act, n_electrons, multiplicity);
        t->topo = xmvb_amd::expand_structures(t->raw);
        sizes[0] = t->topo.n_structures;
        sizes[1] = static_cast<int64_t>(t->topo.alpha_masks.size());
        sizes[2] = static_cast<int64_t>(t->topo.beta_masks.size());
        sizes[3] = static_cast<int64_t>(t->topo.det_alpha.size());
        sizes[4] = static_cast<int64_t>(t->topo.term_structure.size());
        *out = t.release();
        return VBE_OK;
    } catch (const std::exception& e) {
        set_global_error(e.what());
        return VBE_ERR_TOPOLOGY;
    }
}

int vbe_topology_copy(const vbe_topology* t, uint64_t* alpha_masks, uint64_t* beta_masks, int32_t* det_alpha,
                      int32_t* det_beta, int64_t* term_offsets, int32_t* term_structure, double* term_coefficient,
                      int32_t* raw_structures) {
    if (!t) {
        set_global_error("vbe_topology_copy: null topology");
        return VBE_ERR_ARGUMENT;
    }
    auto put = [](const auto& v, auto* dst) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    put(t->topo.alpha_masks, alpha_masks);
    put(t->topo.beta_masks, beta_masks);
    put(t->topo.det_alpha, det_alpha);
    put(t->topo.det_beta, det_beta);
    put(t->topo.term_offsets, term_offsets);
    put(t->topo.term_structure, term_structure);
    put(t->topo.term_coefficient, term_coefficient);
    if (raw_structures)
        for (std::size_t s = 0; s < t->raw.rows.size(); ++s)
            std::copy(t->raw.rows[s].begin(), t->raw.rows[s].end(), raw_structures + s * t->raw.n_active_electrons);
    return VBE_OK;
}

int vbe_topology_view_of(const vbe_topology* t, vbe_topology_view* v) {
    if (!t || !v) {
        set_global_error("vbe_topology_view_of: null argument");
        return VBE_ERR_ARGUMENT;
    }
    v->n_structures = t->topo.n_structures;
    v->n_alpha_strings = static_cast<int32_t>(t->topo.alpha_masks.size());
    v->alpha_masks = t->topo.alpha_masks.data();
    v->n_beta_strings = static_cast<int32_t>(t->topo.beta_masks.size());
    v->beta_masks = t->topo.beta_masks.data();
    v->n_determinants = static_cast<int32_t>(t->topo.det_alpha.size());
    v->det_alpha = t->topo.det_alpha.data();
    v->det_beta = t->topo.det_beta.data();
    v->term_offsets = t->topo.term_offsets.data();
    v->term_structure = t->topo.term_structure.data();
    v->term_coefficient = t->topo.term_coefficient.data();
    return VBE_OK;
}

void vbe_topology_free(vbe_topology* t) { delete t; }

int vbe_synthetic_problem(uint64_t seed, int32_t n_bf, int32_t n_act, int32_t n_aux, double scan_r, double* S,
                          double* h, double* B, double* C) {
    if (n_bf <= 0 || n_act <= 0 || n_aux <= 0 || !S || !h || !B || !C) {
        set_global_error("vbe_synthetic_problem: invalid argument");
        return VBE_ERR_ARGUMENT;
    }
    xmvb_amd::Normal normal(seed);
    const std::size_t n = static_cast<std::size_t>(n_bf);
    std::vector<double> a(n * n);
    for (auto& x : a) x = normal();
    for (std::size_t j = 0; j < n; ++j)
        for (std::size_t i = 0; i < n; ++i) {
            double acc = 0.0;
            for (std::size_t k = 0; k < n; ++k) acc += a[i + n * k] * a[j + n * k];
            S[i + n * j] = (i == j ? 1.0 : 0.0) + 0.1 * acc / static_cast<double>(n_bf);
        }
    const double sd_h = std::sqrt(0.5);
    for (std::size_t k = 0; k < n * n; ++k) a[k] = sd_h * normal();
    const double coupling = scan_r >= 0.0 ? std::exp(-0.5 * scan_r) : 1.0;
    const std::size_t half = n / 2;
    auto crosses = [&](std::size_t i, std::size_t j) { return (i < half) != (j < half); };
    for (std::size_t j = 0; j < n; ++j)
        for (std::size_t i = 0; i < n; ++i) {
            double v = 0.5 * (a[i + n * j] + a[j + n * i]) - (i == j ? 2.0 : 0.0);
            if (crosses(i, j)) v *= coupling;
            h[i + n * j] = v;
        }
    const double sd_b = std::sqrt(1.0 / static_cast<double>(n_bf));
    const std::size_t npair = n * (n + 1) / 2;
    for (int32_t p = 0; p < n_aux; ++p) {
        for (std::size_t k = 0; k < n * n; ++k) a[k] = sd_b * normal();
        for (std::size_t hi = 0; hi < n; ++hi)
            for (std::size_t lo = 0; lo <= hi; ++lo) {
                double v = 0.5 * (a[hi + n * lo] + a[lo + n * hi]);
                if (crosses(hi, lo)) v *= coupling;
                B[static_cast<std::size_t>(p) + static_cast<std::size_t>(n_aux) * (hi * (hi + 1) / 2 + lo)] = v;
            }
    }
    (void)npair;
    const double sc = 1.0 / std::sqrt(static_cast<double>(n_bf));
    for (std::size_t k = 0; k < static_cast<std::size_t>(n_act) * n; ++k) C[k] = sc * normal();
    return VBE_OK;
}

}  // extern "C"
