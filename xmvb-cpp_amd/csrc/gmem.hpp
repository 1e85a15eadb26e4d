// Global-address-space views of device pointers (shared by the HIP translation units).
#pragma once

#include <hip/hip_runtime.h>

namespace xmvb_amd {

// Device-memory pointers read from descriptors are generic to the compiler, so their accesses
// become flat_* instructions, which count on both the vector-memory and the LDS counters: a wait
// for an LDS read then also waits for every outstanding global load, and prefetches overlap
// nothing. Every descriptor pointer addresses device memory; to_global states that, and accesses
// through the result are global_* instructions (vector-memory counter only).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}

// clang vector types for wide loads through global pointers (HIP's float4 / double2 are classes
// whose assignment does not bind to an address-space-qualified operand)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

}  // namespace xmvb_amd
