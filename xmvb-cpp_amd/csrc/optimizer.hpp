// Projected L-BFGS orbital optimisation on device-resident state (internal interface).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>

#include "../../include/xmvb_amd.h"

namespace xmvb_amd {

// What the optimiser needs from one resident problem. `energy` runs a forward (energy-only)
// evaluation of device orbitals d_C and retains it, with d_grad as the gradient destination of
// the later `gradient` call; it returns NaN when the trial point is numerically invalid
// (normalisation or metric failure), which the line search treats as a rejected trial.
struct OptimizerHooks {
    hipStream_t stream = nullptr;
    int n_act = 0;
    int n_bf = 0;
    const double* S_ao = nullptr;  // host AO overlap, n_bf x n_bf
    std::function<double(const double* d_C, const uint8_t* d_support, double* d_grad)> energy;
    std::function<void(const double* d_C, const uint8_t* d_support, double* d_grad)> gradient;
    // forward and backward in one launch sequence (NaN as for `energy`)
    std::function<double(const double* d_C, const uint8_t* d_support, double* d_grad)> energy_gradient;
};

void optimizer_defaults(vbe_optimizer_options* o);

// Throws std::invalid_argument for bad options and std::runtime_error for failures.
void run_plbfgs(const OptimizerHooks& hooks, const uint8_t* support, double* C, const vbe_optimizer_options& opt,
                vbe_optimizer_result& result, double* grad, double* energy_trace, int trace_capacity);

}  // namespace xmvb_amd
