// Small device helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace xmvb_amd {

// packed index of the unordered pair (a, b): hi(hi+1)/2 + lo
__device__ __forceinline__ int pair_index(int a, int b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

// unordered pair t = e(e+1)/2 + d with d <= e
__device__ __forceinline__ void unpack_pair(long long t, int& d, int& e) {
    long long x = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > t) --x;
    while ((x + 1) * (x + 2) / 2 <= t) ++x;
    d = (int)(t - x * (x + 1) / 2);
    e = (int)x;
}

}  // namespace xmvb_amd
