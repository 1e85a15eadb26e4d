// Small device helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "gmem.hpp"
#include "kernels.hpp"

namespace xmvb_amd {


// packed index of the unordered pair (a, b): hi(hi+1)/2 + lo
__device__ __forceinline__ int pair_index(int a, int b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

// (x, z) of this workgroup in a grid (nx, 1, nz) of nx workgroups per problem z, renumbered so
// that all workgroups of one problem run on one XCD: workgroups go to the 8 XCDs in turn by linear
// id, so linear id L runs on XCD L % 8 and takes the (L / 8)-th place of that XCD's sequence,
// which is mapped to problem L % 8 + 8 (L / 8 / nx), workgroup (L / 8) % nx. A problem's
// workgroups then share one L2, and the data they all read (S, h, the determinant matrices) comes
// from HBM once. A bijection of the grid when nz is a multiple of 8; otherwise the identity.
__device__ __forceinline__ void problem_block(int& x, int& z) {
    const unsigned nx = gridDim.x, nz = gridDim.z;
    if (nz % 8u != 0u || gridDim.y != 1u) {
        x = (int)blockIdx.x;
        z = (int)blockIdx.z;
        return;
    }
    const unsigned L = blockIdx.x + nx * blockIdx.z, k = L / 8u;
    z = (int)(L % 8u + 8u * (k / nx));
    x = (int)(k % nx);
}

// unordered pair t = e(e+1)/2 + d with d <= e
__device__ __forceinline__ void unpack_pair(long long t, int& d, int& e) {
    long long x = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > t) --x;
    while ((x + 1) * (x + 2) / 2 <= t) ++x;
    d = (int)(t - x * (x + 1) / 2);
    e = (int)x;
}

// Weights of one unordered determinant pair {d, e} (and of its spin-flip partner {fd, fe} when
// flip) in the backward: the dH weight wH = sum_s w_s a^s_d a^s_e (x mult: 2 off the diagonal)
// and the dS weight wS = -sum_s w_s lambda_s a^s_d a^s_e, so the pair's adjoint is that of
// wH H_de + wS S_de. One state: wH = a_d a_e mult, wS = -lambda wH.
__device__ __forceinline__ void pair_weights(const SlotDesc& sd, const TopoDesc& topo, int d, int e, int fd, int fe,
                                             bool flip, double mult, double& wH, double& wS) {
    const auto dc = to_global(sd.dcoef);
    if (topo.n_states <= 1) {
        wH = (dc[d] * dc[e] + (flip ? dc[fd] * dc[fe] : 0.0)) * mult;
        wS = -to_global(sd.energy)[0] * wH;
        return;
    }
    double h = 0.0, s = 0.0;
    for (int k = 0; k < topo.n_states; ++k) {
        const auto a = dc + (size_t)k * topo.n_det;
        const double c = topo.state_w[k] * ((a[d] * a[e] + (flip ? a[fd] * a[fe] : 0.0)) * mult);
        h += c;
        s -= to_global(sd.state_energy)[k] * c;
    }
    wH = h;
    wS = s;
}

}  // namespace xmvb_amd
