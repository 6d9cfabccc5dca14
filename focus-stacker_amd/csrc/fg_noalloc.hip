// fg_fuse_stack_device_noalloc: the one-shot stack fusion for callers that cannot allocate device
// memory.  Every working buffer of the stack (frame cache, weights, level buffers, norm sum, union
// support, accumulators, scratch of the small levels, hybrid images, output staging) is carved from
// `scratch`, a file-scope device array that each GPU holds from the library's load to the end of
// the process.  The call makes no hipMalloc / hipFree and creates no stream or event.
//
// `scratch` in host code is only a stand-in symbol, not a device address: its address on a GPU is
// taken once per device, with that device current, by hipGetSymbolAddress (fg_noalloc_setup, run
// by fg_context_create), and only that pointer is ever handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "fg_host.h"

namespace {

constexpr size_t kScratchBytes = FG_NOALLOC_SCRATCH_BYTES;
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_setup_mu;
void* g_scratch_ptr[kMaxDevices];     // the device address of `scratch` on each GPU
int g_scratch_state[kMaxDevices];     // 0: not taken yet, 1: taken, < 0: hipGetSymbolAddress failed
std::mutex g_call_mu[kMaxDevices];    // one call at a time per GPU: every call shares the area

// Bump allocation inside the scratch area; frees are no-ops (the whole area is the call's).
struct Carve {
    unsigned char* base;
    size_t size, used;
};
void* carve_alloc(void* u, size_t bytes) {
    Carve* c = (Carve*)u;
    const size_t b = (bytes + kAlign - 1) & ~(kAlign - 1);
    if (c->used + b > c->size) return nullptr;
    void* p = c->base + c->used;
    c->used += b;
    return p;
}
void carve_free(void*, void*) {}

}  // namespace

void fg_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (g_scratch_state[device] != 0) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_scratch_state[device] = -1;
        return;
    }
    g_scratch_ptr[device] = p;
    g_scratch_state[device] = 1;
}

extern "C" int fg_fuse_stack_device_noalloc(fg_context* ctx, uint32_t n, const fg_image_desc* descs,
                                            const void* const* d_frames, const fg_pose* poses, const fg_rect* roi,
                                            const fg_stack_opts* opts, void* d_out) {
    if (!ctx || !descs || !d_frames || !opts || !d_out || n == 0)
        return fg_set_error(FG_ERR_INVALID_ARG, "null argument");
    const int dev = fg_context_device(ctx);
    if (dev < 0 || dev >= kMaxDevices) return fg_set_error(FG_ERR_UNSUPPORTED, "device index beyond the scratch table");
    if (hipSetDevice(dev) != hipSuccess) return fg_set_error(FG_ERR_DEVICE, "hipSetDevice");
    {
        std::lock_guard<std::mutex> lk(g_setup_mu);
        if (g_scratch_state[dev] != 1)
            return fg_set_error(FG_ERR_DEVICE, "the scratch area's device address could not be taken on this GPU");
    }
    fg_stack_opts o = *opts;
    o.frame_count = n;
    o.debug_frame = -1;  // debug copies would need buffers of their own
    if (!o.stream) o.stream = fg_context_stream(ctx);
    // the working set under a budget of the whole area: the frame cache when it fits, else the
    // uncached second pass (fed from the same device frames); a stack whose fixed buffers alone
    // exceed the area is refused
    const double budget = (double)(kScratchBytes - 64 * kAlign);  // room for the carve's rounding
    int cached = 0;
    const size_t need = fg_plan_stack_bytes(&o, budget, &cached, false);
    if (need == 0) return fg_set_error(FG_ERR_INVALID_ARG, "invalid stack options");
    if ((double)need > budget)
        return fg_set_error(FG_ERR_INVALID_ARG,
                            ("stack needs " + std::to_string(need) + " bytes of working memory; the allocation-free "
                             "entry point's scratch area holds " + std::to_string(kScratchBytes) + " bytes")
                                .c_str());

    std::lock_guard<std::mutex> call(g_call_mu[dev]);
    Carve cv{(unsigned char*)g_scratch_ptr[dev], kScratchBytes, 0};
    const fg_dev_allocator alloc{carve_alloc, carve_free, &cv};
    fg_stack* s = nullptr;
    int rc = fg_stack_create_budget(ctx, &o, budget, &s, &alloc, false);
    if (rc) return rc;
    for (uint32_t k = 0; k < n && rc == FG_OK; k++)
        rc = fg_stack_add_frame_device(s, &descs[k], d_frames[k], poses ? &poses[k] : nullptr, roi);
    if (rc == FG_OK && fg_stack_needs_pass2_frames(s))
        for (uint32_t k = 0; k < n && rc == FG_OK; k++)
            rc = fg_stack_pass2_frame_device(s, &descs[k], d_frames[k], poses ? &poses[k] : nullptr, roi);
    if (rc == FG_OK) rc = fg_stack_finish_device(s, d_out);
    // the next call reuses the area: wait for this one's kernels before returning
    const hipError_t e = hipStreamSynchronize((hipStream_t)o.stream);
    fg_stack_destroy_idle(s);
    if (rc == FG_OK && e != hipSuccess) return fg_set_error(FG_ERR_DEVICE, hipGetErrorString(e));
    return rc;
}
