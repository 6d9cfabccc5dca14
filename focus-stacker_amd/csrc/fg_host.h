// Host-side internals shared by the C-ABI layer (fg_api.cpp) and the stack server (fg_server.cpp).
#pragma once

#include <stddef.h>

#include "../../include/focus_gpu.h"

// Device bytes of a stack with these options when `budget` bytes are available; *cached tells
// whether every frame then stays on the device.  0 for invalid options.  host_frames: frames may
// come from host memory, so the caching decision leaves room for the upload ring.
size_t fg_plan_stack_bytes(const fg_stack_opts* o, double budget, int* cached, bool host_frames = true);
// Device memory source for a stack's buffers (the server's arena): alloc returns nullptr when it
// cannot serve the request; free takes back what alloc gave.
struct fg_dev_allocator {
    void* (*alloc)(void* user, size_t bytes);
    void (*free)(void* user, void* p);
    void* user;
};
// fg_stack_create with the caching decision taken against `budget` bytes (< 0: the options'
// memory fraction of free HBM) and its buffers from `allocator` (nullptr: hipMalloc).
int fg_stack_create_budget(fg_context* ctx, const fg_stack_opts* o, double budget, fg_stack** out,
                           const fg_dev_allocator* allocator = nullptr, bool host_frames = true);
// HIP device of a context, and its --memory-fraction.
int fg_context_device(const fg_context* ctx);
float fg_context_memory_fraction(const fg_context* ctx);
// The context's own stream (hipStream_t).
void* fg_context_stream(const fg_context* ctx);
namespace fg {
struct FrameDesc;
struct StackView;
}
// The buffers and pass arguments of a cached pyramid-contrast stack (FG_ERR_STATE otherwise).
int fg_stack_view(fg_stack* stack, fg::StackView* view);
// Validates frames 0..n-1 of a stack (device pointers) as fg_stack_add_frame_device does and writes
// their pass-1 descriptors to `out`, without running anything.
int fg_stack_describe_device_frames(fg_stack* stack, uint32_t n, const fg_image_desc* descs, const void* const* frames,
                                    const fg_pose* poses, const fg_rect* roi, fg::FrameDesc* out);
// Takes the device address of this GPU's scratch area for fg_fuse_stack_device_noalloc, once per
// device, with that device current (fg_noalloc.hip).
void fg_noalloc_setup(int device);
// fg_stack_destroy for a stack whose work is known complete: does not wait on its stream (which
// may be shared with, and busy running, another stack).
void fg_stack_destroy_idle(fg_stack* stack);
// A frame too large for the kernels' 32-bit byte offsets into its samples.
bool frame_too_large(const fg_image_desc& d);
// Sets the calling thread's fg_last_error() message and returns `code`.
int fg_set_error(int code, const char* msg);
