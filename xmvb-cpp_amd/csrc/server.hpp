// Engine hooks of the request scheduler (server.cpp): launch a batch of resident slots of a pool
// from host orbitals on a free launch lane, poll it, collect its results. Every hook takes the
// pool's mutex, so the scheduler's dispatcher thread and the caller's thread (loading and
// unloading slots) never touch a pool's lanes or slots at the same time.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/xmvb_amd.h"

namespace xmvb_amd {

// HIP device of a pool's context.
int pool_device(const vbe_pool* p);
// The context's shared streams (created on first use); returns how many were written.
int context_streams(vbe_context* ctx, void** streams, int max_streams);
// n_act and the slot's n_bf; VBE_ERR_ARGUMENT (with a message) when the slot is out of range or
// not loaded.
int pool_slot_dims(vbe_pool* p, int slot, int* n_act, int* n_bf, std::string* err);
// True when an in-flight launch of the pool holds the slot.
bool pool_slot_busy(vbe_pool* p, int slot);
// Copies the orbitals C[i] (n_act x n_bf_i, host) of `count` distinct free loaded slots to the
// device and enqueues their forward (and backward) on a free lane; *lane receives the lane index,
// or -1 when every lane is busy (nothing is enqueued then). Errors come back as a status with
// *err set.
// stream: the stream to run it on (one of the context's; null: the lane's own). side: the stream
// that small launches may fork their adjoint kernels onto (one of the context's; null: none, the
// whole launch stays on `stream`); used only together with `stream`.
int pool_launch(vbe_pool* p, int count, const int* slots, const double* const* C, bool gradient, void* stream,
                void* side, int* lane, std::string* err);
// 1 when the lane's launch has completed, 0 while it runs.
int pool_launch_query(vbe_pool* p, int lane, std::string* err);
// Completes the lane's launch: energies[i] (E_total), grads[i] (n_act x n_bf_i, skipped when null),
// status[i] per slot, the launch's device time, frees the lane.
int pool_launch_collect(vbe_pool* p, int lane, double* energies, double* const* grads, int* status,
                        float* device_ms, std::string* err);

}  // namespace xmvb_amd
