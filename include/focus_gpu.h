/*
 * focus_gpu.h — C ABI of the MI355X (gfx950) remap + pyramid-contrast fusion library.
 *
 * Drop-in for the per-frame hot path of asingleoat/focus-stacker: the Zig callers
 * keep their loops and swap the per-frame bodies for these calls, the same way they
 * already bind C libraries with @cImport (src/fft_backend.zig:5-7, build.zig:266-282).
 * Alignment (feature matching, phase correlation, the pose solve) stays on the CPU
 * and hands in poses.
 *
 * Conventions
 *   - C99, plain pointers and sizes.  Every call returns an int status: FG_OK (0) or a
 *     negative FG_ERR_*; fg_last_error() gives a thread-local message for the last failure.
 *   - Pixels are interleaved row-major, stride = width * channels samples, no padding,
 *     exactly as image_io.Image (src/image_io.zig:31-60).
 *   - Host-pointer calls copy their inputs before returning; `*_device` calls take
 *     device pointers and enqueue on the stack's (or the given) HIP stream.
 *   - One fg_stack is driven by one host thread at a time; a context may serve many.
 */
#ifndef FOCUS_GPU_H
#define FOCUS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fg_status {
    FG_OK = 0,
    FG_ERR_INVALID_ARG = -1,
    FG_ERR_OOM_HOST = -2,
    FG_ERR_OOM_DEVICE = -3,
    FG_ERR_SIZE_MISMATCH = -4,    /* error.MismatchedImageSizes  src/fuse/io.zig:6-9 */
    FG_ERR_FORMAT_MISMATCH = -5,  /* error.MismatchedImageFormats src/fuse/io.zig:6-9 */
    FG_ERR_UNSUPPORTED = -6,
    FG_ERR_DEVICE = -7,
    FG_ERR_STATE = -8             /* call out of order (e.g. finish before any frame) */
};

enum fg_sample_type { FG_SAMPLE_U8 = 0, FG_SAMPLE_U16 = 1 };   /* image_io.SampleType src/image_io.zig:25-28 */

/* fuse.config.Method (src/fuse/config.zig:10-15).  hardmask and softmask are the single-pass
 * methods (src/stack/pipeline.zig:245-312): every frame goes through fg_stack_add_frame once. */
enum fg_method {
    FG_METHOD_HARDMASK_CONTRAST = 0,
    FG_METHOD_SOFTMASK_CONTRAST = 1,
    FG_METHOD_PYRAMID_CONTRAST = 2,
    FG_METHOD_HYBRID_PYRAMID_CONTRAST = 3
};

/* optimize.ImagePose: the 15 f64 fields in declaration order (src/optimize.zig:32-48). */
typedef struct fg_pose {
    double yaw, pitch, roll, hfov_delta;
    double trans_x, trans_y, trans_z;
    double translation_plane_yaw, translation_plane_pitch;
    double radial_a, radial_b, radial_c;
    double center_shift_x, center_shift_y;
    double base_hfov_degrees;
} fg_pose;

/* remap.Rect (src/remap.zig:573-578): [left, right) x [top, bottom) in panorama pixels. */
typedef struct fg_rect { int32_t left, top, right, bottom; } fg_rect;

/* The parts of image_io.ImageInfo the path reads (src/image_io.zig:31-43). */
typedef struct fg_image_desc {
    uint32_t width, height;
    uint32_t color_channels;   /* 1 (grayscale) or 3 (rgb); fusion needs 3 */
    uint32_t extra_channels;   /* 0, or 1 = a support alpha follows the colour channels */
    uint32_t sample_type;      /* enum fg_sample_type */
} fg_image_desc;

typedef struct fg_context fg_context;
typedef struct fg_stack fg_stack;

typedef struct fg_context_opts {
    int32_t device;            /* HIP device ordinal */
    float memory_fraction;     /* --memory-fraction (src/memory_budget.zig:3): share of free HBM a
                                  stack may hold for its frame cache; 0 = no cap */
} fg_context_opts;

typedef struct fg_stack_opts {
    uint32_t method;           /* enum fg_method */
    uint32_t contrast_window_size; /* --contrast-window-size, odd, 3..31 (src/fuse/config.zig:295-299) */
    float hybrid_sharpness;    /* --hybrid-sharpness in [0,1] (default 0.35, src/fuse/pyramid.zig:11) */
    uint32_t frame_count;      /* N active frames (the 1/N fallback of normalizeWeightsInto) */
    uint32_t width, height;    /* fused size = ROI size (or the aligned frame size) */
    uint32_t sample_type;      /* enum fg_sample_type of the frames and of the output */
    uint32_t max_levels;       /* 0 = full depth down to 1x1 (reference); else cap (enfuse --levels) */
    float memory_fraction;     /* < 0: use the context's; else as fg_context_opts */
    int32_t debug_frame;       /* >= 0: keep that frame's mask/gaussian levels for fg_stack_debug_readback */
    void* stream;              /* hipStream_t to run on, or NULL for a stack-owned stream */
} fg_stack_opts;

enum fg_debug_buffer {
    FG_DEBUG_NORM_SUM = 0,     /* f32 W*H   : sum of weights (src/stack/pipeline.zig:469) */
    FG_DEBUG_UNION_SUPPORT = 1,/* f32 W*H   : 1 where any frame has alpha > 0 */
    FG_DEBUG_WEIGHTS = 2,      /* f32 W*H   : frame `index` contrast * alpha (cached frames only) */
    FG_DEBUG_ACC_LEVEL = 3,    /* f32 lw*lh*3 interleaved: accumulator level `index` (before finish) */
    FG_DEBUG_MASK_LEVEL = 4,   /* f32 lw*lh : debug_frame's normalised mask level `index` */
    FG_DEBUG_GAUSS_LEVEL = 5,  /* f32 lw*lh*3 interleaved: debug_frame's gaussian level `index` */
    FG_DEBUG_SOFT = 6,         /* samples W*H*3: hybrid soft (pyramid) image (after finish) */
    FG_DEBUG_HARD = 7          /* samples W*H*3: hybrid hard-mask image */
};

/* Page-locked host memory for frame pixels (image_io.loadImage's pixel buffers,
 * src/image_io.zig:160-205, can be allocated here): fg_stack_add_frame* then uploads at full PCIe
 * rate.  fg_host_free releases it.  Returns NULL on failure. */
void* fg_host_alloc(size_t bytes);
void fg_host_free(void* p);

/* Device memory on the context's GPU, for callers without a HIP binding (the Zig host) that keep
 * frames device-resident for the *_device entry points and the server.  The copies are
 * synchronous.  fg_device_alloc returns NULL on failure. */
void* fg_device_alloc(fg_context* ctx, size_t bytes);
void fg_device_free(fg_context* ctx, void* d_ptr);
int fg_copy_to_device(fg_context* ctx, void* d_dst, const void* src, size_t bytes);
int fg_copy_to_host(fg_context* ctx, void* dst, const void* d_src, size_t bytes);

/* Message of the last failed call on this thread ("" if none). */
const char* fg_last_error(void);
/* Library version string. */
const char* fg_version(void);

int fg_context_create(const fg_context_opts* opts, fg_context** out);
/* Destroy the context's stacks, batches and servers first: they use its device, stream and the
 * threads that copy host-fed frames. */
void fg_context_destroy(fg_context* ctx);

/* remap.remapRigidImage (src/remap.zig:166-204): remap `src` under `pose` into `roi`
 * (NULL = full frame).  dst receives roi_w*roi_h*(color_channels+1) samples, the last
 * channel being the support alpha.  The source must have extra_channels == 0. */
int fg_remap(fg_context* ctx, const fg_image_desc* src, const void* src_px, const fg_pose* pose,
             const fg_rect* roi, void* dst_px);
/* Same on device buffers, enqueued on `stream` (hipStream_t, NULL = the context's stream). */
int fg_remap_device(fg_context* ctx, const fg_image_desc* src, const void* d_src, const fg_pose* pose,
                    const fg_rect* roi, void* d_dst, void* stream);

/* remap.computeCommonOverlapRoi (src/remap.zig:585-633), host arithmetic.  active may be
 * NULL (all active).  Returns 1 and fills *out if the overlap is non-empty, 0 if not. */
int fg_common_overlap_roi(uint32_t n, const uint8_t* active, const uint32_t* widths, const uint32_t* heights,
                          const fg_pose* poses, fg_rect* out);

/* image_io.writeTiff (src/image_io.zig:175-200) with the tags of openTiffWriter (:100-143):
 * uncompressed, contiguous, 150 dpi, unassociated alpha for the extra channel.  The file is
 * byte-identical to what that libtiff call sequence writes.  Host only (no GPU work). */
int fg_write_tiff(const char* path, const fg_image_desc* desc, const void* px);
/* writeAlignedImageTask (src/remap.zig:134-164) for a decoded frame: fg_remap into `roi`
 * (NULL = full frame), then fg_write_tiff to "{prefix}_{output_index:04}.tif". */
int fg_write_aligned_tiff(fg_context* ctx, const fg_image_desc* src, const void* src_px, const fg_pose* pose,
                          const fg_rect* roi, const char* prefix, uint32_t output_index);

/* Pyramid-contrast stack: the two-pass loop of runPyramidStackFusion
 * (src/stack/pipeline.zig:361-605) / runPyramidPass (src/fuse/pipeline.zig:254-445). */
int fg_stack_create(fg_context* ctx, const fg_stack_opts* opts, fg_stack** out);
void fg_stack_destroy(fg_stack* stack);

/* Pass 1 for one frame, in stack order.  With a pose the frame is a source image
 * (extra_channels 0) remapped into `roi` (NULL = full frame); without a pose it is an
 * already aligned image of the stack size (src/fuse/pipeline.zig:289-336). */
int fg_stack_add_frame(fg_stack* stack, const fg_image_desc* desc, const void* px, const fg_pose* pose,
                       const fg_rect* roi);
int fg_stack_add_frame_device(fg_stack* stack, const fg_image_desc* desc, const void* d_px, const fg_pose* pose,
                              const fg_rect* roi);

/* 1 if the frame cache did not fit the memory budget: the caller must then feed every frame
 * again, in the same order, through fg_stack_pass2_frame* before fg_stack_finish (the
 * reference's uncached second loop, src/stack/pipeline.zig:561-597). */
int fg_stack_needs_pass2_frames(const fg_stack* stack);
int fg_stack_pass2_frame(fg_stack* stack, const fg_image_desc* desc, const void* px, const fg_pose* pose,
                         const fg_rect* roi);
int fg_stack_pass2_frame_device(fg_stack* stack, const fg_image_desc* desc, const void* d_px,
                                const fg_pose* pose, const fg_rect* roi);

/* Pass 2 (if cached), collapse, hybrid merge; writes W*H*3 samples (extra_channels = 0,
 * src/fuse/pyramid.zig:252-253).  The _device form leaves the result in device memory and
 * does not wait for it.  From a cached stack's second finish on (after fg_stack_reset), pass 2
 * and the collapse run as a HIP graph captured from the stack's stream on that finish; with
 * profiling on, or debug buffers kept, they are launched one by one. */
int fg_stack_finish(fg_stack* stack, void* out_px);
int fg_stack_finish_device(fg_stack* stack, void* d_out);

/* Debug readback (the --dump-masks-dir buffers, src/fuse/debug.zig). */
int fg_stack_debug_readback(fg_stack* stack, int which, uint32_t index, void* host_dst, size_t dst_bytes);
/* Number of pyramid levels of the stack and the size of level `level`. */
int fg_stack_levels(const fg_stack* stack, uint32_t* levels);
int fg_stack_level_size(const fg_stack* stack, uint32_t level, uint32_t* width, uint32_t* height);
/* The HIP stream the stack runs on (hipStream_t). */
void* fg_stack_stream(const fg_stack* stack);
/* Blocks until the stack's queued work is done. */
int fg_stack_sync(fg_stack* stack);

/* Starts a new stack of the same geometry and options on the same device buffers (no
 * allocation): clears the norm sum, union support and frame counters.  The accumulators start
 * over with the first pass-2 frame (softmask's sums are cleared here). */
int fg_stack_reset(fg_stack* stack);

/* Device time per phase, from HIP events recorded on the stack's stream while profiling is
 * on, accumulated over every stack run since fg_stack_set_profiling(stack, 1).  ms[k] is the
 * total and launches[k] the number of bracketed regions, for k < FG_PHASE_COUNT:
 *   INGEST        host-frame uploads (on the stack's upload stream, overlapping the others);
 *   PASS1         the pass-1 launch (remap + weights + norm sum), one per flush;
 *   PASS2         every pass-2 launch of a stack (all levels);
 *   FINISH        collapse, quantize and the hybrid merge;
 *   PASS2_LEVEL0  the level-0 pass-2 launch alone (inside PASS2). */
enum fg_phase {
    FG_PHASE_INGEST = 0, FG_PHASE_PASS1 = 1, FG_PHASE_PASS2 = 2, FG_PHASE_FINISH = 3, FG_PHASE_PASS2_LEVEL0 = 4,
    FG_PHASE_COUNT = 5
};
int fg_stack_set_profiling(fg_stack* stack, int enable);
int fg_stack_phase_times(fg_stack* stack, double* ms, uint32_t* launches);

/* One-shot: every frame through pass 1, pass 2 and collapse (poses may be NULL for
 * aligned frames, roi may be NULL). */
int fg_fuse_stack(fg_context* ctx, uint32_t n, const fg_image_desc* descs, const void* const* pxs,
                  const fg_pose* poses, const fg_rect* roi, const fg_stack_opts* opts, void* out_px);

/* ---- Allocation-free one-shot fusion ----
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  fg_fuse_stack for frames
 * already in device memory (d_frames[k], described by descs[k]) into a device buffer d_out of
 * W*H*3 samples, with every working buffer carved from a fixed scratch area of
 * FG_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's load until the process ends.  It
 * calls no hipMalloc or hipFree and creates no stream: the work runs on opts->stream, or the
 * context's stream when that is NULL, and the call returns when d_out is written.  The frame cache
 * is used when it fits the area, else the uncached second pass (fed again from d_frames); a stack
 * whose fixed buffers alone exceed the area (about 2.5 megapixels per frame) fails with
 * FG_ERR_INVALID_ARG and a message giving the bytes it needs.  Every call on one GPU shares the
 * area, which keeps the last call's contents: the library runs one such call per GPU at a time,
 * and work that overlaps a call on the same GPU must not touch the area (nothing outside the
 * library can address it). */
#define FG_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int fg_fuse_stack_device_noalloc(fg_context* ctx, uint32_t n, const fg_image_desc* descs, const void* const* d_frames,
                                 const fg_pose* poses, const fg_rect* roi, const fg_stack_opts* opts, void* d_out);

/* ---- Serving: continuous batching of whole stacks ----
 * No counterpart in the reference, which fuses one stack per process.  A server keeps up to
 * `slots` stacks in flight, each on its own HIP stream, so the kernels of uneven stacks overlap
 * on the device.  Queued requests are admitted, in submission order, whenever a slot and the
 * memory budget (memory_fraction x free HBM at creation, the --memory-fraction of
 * src/memory_budget.zig:3) allow.  Frames stay device-resident from submission to the fused
 * output.  Idle stacks are pooled per slot for same-geometry requests.  A stack's result does not
 * depend on what else runs: every request computes exactly what fg_fuse_stack computes. */
typedef struct fg_server fg_server;

typedef struct fg_server_opts {
    int32_t slots;             /* stacks in flight at once (0 = 4) */
    float memory_fraction;     /* share of free HBM the server may hold (0 = the context's; <= 0 there: all but 1 GB) */
    int32_t pool_per_slot;     /* idle stacks kept per slot for reuse (0 = 2, < 0 = none) */
} fg_server_opts;

typedef struct fg_request {
    uint32_t frame_count;
    const fg_image_desc* descs;    /* per frame */
    const void* const* frames;     /* per frame: device pointer, valid until the request is done */
    const fg_pose* poses;          /* per frame, or NULL for aligned frames */
    const fg_rect* roi;            /* NULL = full frame */
    uint32_t method;               /* enum fg_method */
    uint32_t contrast_window_size;
    float hybrid_sharpness;
    uint32_t max_levels;
    void* out;                     /* device buffer: W*H*3 samples of the stack's sample type */
} fg_request;

enum fg_request_state { FG_REQ_QUEUED = 0, FG_REQ_RUNNING = 1, FG_REQ_DONE = 2 };

typedef struct fg_request_status {
    int32_t state;                 /* enum fg_request_state, or a negative FG_ERR_* of a failed request */
    double queue_ms;               /* submission to admission */
    double run_ms;                 /* admission to completion on the device */
    double total_ms;               /* submission to completion */
} fg_request_status;

int fg_server_create(fg_context* ctx, const fg_server_opts* opts, fg_server** out);
/* Finishes every submitted request, then releases the server. */
void fg_server_destroy(fg_server* server);
/* Queues a copy of *req (the descriptor, pose and frame-pointer arrays are copied). */
int fg_server_submit(fg_server* server, const fg_request* req, uint64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for a request.  A done or failed request fills *st and
 * releases its ticket; on timeout *st holds QUEUED or RUNNING. */
int fg_server_wait(fg_server* server, uint64_t ticket, double timeout_ms, fg_request_status* st);

/* ---- Batched fusion: many stacks in one launch schedule ----
 * No counterpart in the reference, which fuses one stack per process.  fg_batch_fuse takes many
 * requests (device frames and outputs, as fg_server_submit does) and cuts them, in order, into
 * groups of pyramid-contrast stacks of one sample type and contrast window whose frame caches fit the
 * memory budget together.  A group runs as one persistent pass-1 launch over every stack's tiles, one
 * pass-2 launch per pyramid level and one collapse launch per level, captured as a HIP graph the second
 * time that group geometry (the ordered list of stack shapes) runs and replayed from then on (the
 * batch keeps up to plan_cache geometries with their buffers).  Stacks that fill the GPU alone (an
 * output of at least fg_tuning.batch_large_kpix thousand pixels), other methods and stacks too large
 * to cache run as stacks of their own (the plain stack calls), longest first, on side streams beside
 * the groups.  Every result equals the stack fused alone, and the call returns when every output is
 * written; status (optional) receives one entry per request, with run_ms the device time of its
 * group (or stack) and total_ms from the call's start to its end. */
typedef struct fg_batch fg_batch;
typedef struct fg_batch_opts {
    float memory_fraction;     /* share of free HBM the batch may hold (0 = the context's; <= 0 there: all but 1 GB) */
    int32_t plan_cache;        /* group geometries kept with their graphs (0 = 4, < 0 = none) */
} fg_batch_opts;
int fg_batch_create(fg_context* ctx, const fg_batch_opts* opts, fg_batch** out);
void fg_batch_destroy(fg_batch* batch);
int fg_batch_fuse(fg_batch* batch, uint32_t count, const fg_request* reqs, fg_request_status* status);

/* ---- Launch shapes ----
 * Process-wide knobs of the launch shapes, whose defaults the project's autotuner measured on MI355X
 * (focus-stacker_amd/autotune.py).  Results do not depend on them: every shape computes the same
 * bits.  fg_set_tuning applies to stacks and batch plans created afterwards. */
typedef struct fg_tuning {
    int32_t pass2_min_tiles;           /* single stacks: 64x16 tiles while a level has this many, then 16x16, then
                                          one frame per block */
    int32_t batch_wide_level;          /* batches: 64x16 level tiles when the mean level width is at least this */
    int32_t batch_pass1_blocks_per_cu; /* batches: persistent pass-1 blocks per CU (0 = as many as fit) */
    int32_t batch_group_mib;           /* batches: MiB of stack buffers a group may hold at most (0 = the budget) */
    int32_t batch_streams;             /* batches: streams of a batch (1..8; read by fg_batch_create): the groups
                                          run on the first, stacks of their own on the others */
    int32_t batch_large_kpix;          /* batches: output size (thousands of pixels) from which a pyramid stack
                                          runs as a stack of its own (0 = never) */
} fg_tuning;
int fg_get_tuning(fg_tuning* t);
int fg_set_tuning(const fg_tuning* t);

#ifdef __cplusplus
}
#endif
#endif
