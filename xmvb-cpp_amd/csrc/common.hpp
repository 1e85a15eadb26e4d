// Shared host helpers of the engine.
#pragma once

#include <string>

namespace xmvb_amd {

// Text of the last error raised before a context existed (vbe_global_error).
void set_global_error(const std::string& what);

// Packed index of an unordered pair (core/pair_index.hpp:15-55).
inline long long packed_pair(long long a, long long b) {
    const long long hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

}  // namespace xmvb_amd
