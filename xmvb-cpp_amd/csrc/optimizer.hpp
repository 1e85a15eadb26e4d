// Projected L-BFGS orbital optimisation on device-resident state (internal interface).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/xmvb_amd.h"

namespace xmvb_amd {

// One optimisation of a batch: host start orbitals in, converged orbitals out.
struct OptProblem {
    int n_bf = 0;
    const double* S_ao = nullptr;      // host AO overlap, n_bf x n_bf
    const uint8_t* support = nullptr;  // host n_act x n_bf mask or null
    double* C = nullptr;               // host n_act x n_bf, in/out
    double* grad = nullptr;            // host n_act x n_bf or null
    double* trace = nullptr;           // energy after each accepted iteration (optional)
    int trace_capacity = 0;
    vbe_optimizer_result* result = nullptr;
};

// What the optimiser needs from the engine. `evaluate` runs one launch sequence over the listed
// problems (indices into the batch) at device orbitals C[j] (support masks sup[j], device or null),
// writing energies (NaN where the point is numerically invalid: normalisation or metric failure,
// which the line search treats as a rejected trial) and, with with_grad, gradients into G[j].
// Without with_grad (only for a batch of one) the forward stays on the device and `complete`
// turns it into the gradient.
struct OptimizerHooks {
    hipStream_t stream = nullptr;
    int n_act = 0;
    std::function<void(const std::vector<int>& which, const std::vector<const double*>& C,
                       const std::vector<const uint8_t*>& sup, const std::vector<double*>& G, bool with_grad,
                       std::vector<double>& E)>
        evaluate;
    std::function<void(int which, const double* C, double* G)> complete;
    bool energy_only_trials = false;  // `complete` is available (a single problem, not a pool)
    // device memory of at least `bytes` for the optimizer's state, owned by the caller and valid
    // for the whole call; unset: the call allocates and frees its own
    std::function<void*(size_t bytes)> device_block;
    // pinned host memory of at least `bytes`, likewise (hipHostFree also waits for the device)
    std::function<void*(size_t bytes)> host_block;
};

void optimizer_defaults(vbe_optimizer_options* o);

// Converges every problem of the batch; problems advance in lockstep, one batched launch per
// phase. Throws std::invalid_argument for bad options and std::runtime_error for failures.
void run_plbfgs(const OptimizerHooks& hooks, std::vector<OptProblem>& problems, const vbe_optimizer_options& opt);

}  // namespace xmvb_amd
