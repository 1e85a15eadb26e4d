// Engine hooks of the no-allocation entry point (noalloc.hip): an evaluation of one resident
// problem whose per-evaluation workspace is a caller-supplied block of device memory.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/xmvb_amd.h"

namespace xmvb_amd {

// HIP device of a context (-1 for a null context).
int context_device(const vbe_context* ctx);
// Records `what` as the context's last error and returns status.
int report_error(vbe_context* ctx, int status, const char* what);
// vbe_eval_device with the workspace carved from [scratch, scratch + scratch_bytes) instead of the
// problem's own; VBE_ERR_MEMORY when the workspace does not fit. Allocates no device memory.
int eval_device_in_scratch(vbe_context* ctx, vbe_problem* prob, const double* d_C, const uint8_t* d_support,
                           vbe_flags flags, double* E_total, double* d_grad, unsigned char* scratch,
                           size_t scratch_bytes);

}  // namespace xmvb_amd
