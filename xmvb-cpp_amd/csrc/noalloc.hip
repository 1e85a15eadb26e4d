// No-allocation evaluation: vbe_eval_device_noalloc runs one E+grad evaluation of a resident
// problem with its whole per-evaluation workspace (orbital intermediates, one-index factors, packed
// kernel, determinant H/S, structure matrices, eigensolver work, backward partial rows, pair lists)
// in a static device array instead of the problem's own workspace, so the call makes no device
// allocation: for callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The array is file-scope device memory of a fixed size; every GPU the library runs on holds its
// own copy until the process ends. Host code reaches a GPU's copy only through the address that
// hipGetSymbolAddress returns for it (looked up once per device, with that device current); the
// array's name in host code is a host-side stand-in and is never passed to a kernel or a copy.
// Every call on one GPU uses that GPU's copy, which keeps the last call's contents: calls that
// overlap on one GPU race. The kernels clear what they accumulate into before use.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/xmvb_amd.h"
#include "noalloc.hpp"

static __device__ __align__(256) unsigned char scratch[VBE_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
unsigned char* scratch_ptr[kMaxDevices] = {};

// Device address of `device`'s copy of the scratch array; the device must be current.
hipError_t scratch_on(int device, unsigned char** out) {
    std::lock_guard<std::mutex> g(scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int vbe_eval_device_noalloc(vbe_context* ctx, vbe_problem* prob, const double* d_C_raw,
                                       const uint8_t* d_support, vbe_flags flags, double* E_total,
                                       double* d_grad_raw) {
    if (!ctx) return VBE_ERR_ARGUMENT;
    const int device = xmvb_amd::context_device(ctx);
    hipError_t e = hipSetDevice(device);
    unsigned char* base = nullptr;
    if (e == hipSuccess) e = scratch_on(device, &base);
    if (e != hipSuccess) {
        const std::string what = std::string("static scratch lookup (hipGetSymbolAddress): ") + hipGetErrorString(e);
        return xmvb_amd::report_error(ctx, VBE_ERR_DEVICE, what.c_str());
    }
    return xmvb_amd::eval_device_in_scratch(ctx, prob, d_C_raw, d_support, flags, E_total, d_grad_raw, base,
                                            VBE_NOALLOC_SCRATCH_BYTES);
}
