/*
 * C-side conformance test of the focus_gpu ABI (include/focus_gpu.h, libfocus_gpu.so).
 *
 * The Zig host binds the library through @cImport, which sees exactly what a C compiler sees,
 * so this program plays the Zig glue's part: it is built with gcc as strict C99 against the
 * header and the shared library, and makes the calls in the order the reference's loops make
 * them (runPyramidStackFusion src/stack/pipeline.zig:361-605, runPyramidPass
 * src/fuse/pipeline.zig:254-445, writeAlignedImageTask src/remap.zig:134-164).
 *
 *   fg_conformance          layout, constants and every host-side path (no GPU needed)
 *   fg_conformance --gpu D  also the GPU paths, checked against the oracle (liboracle.so);
 *                           files are written under directory D
 *
 * Exit status 0 when every check passes.
 */
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "focus_gpu.h"
#include "focus_oracle.h"

static int g_checks, g_fail;

#define CHECK(cond, ...)                                                \
    do {                                                                \
        g_checks++;                                                     \
        if (!(cond)) {                                                  \
            g_fail++;                                                   \
            printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond);      \
            printf(__VA_ARGS__);                                        \
            printf(" [%s]\n", fg_last_error());                         \
            fflush(stdout);                                             \
        }                                                               \
    } while (0)

/* compile-time checks in C99: a negative array size fails the build */
#define STATIC_CHECK(cond, name) typedef char static_check_##name[(cond) ? 1 : -1]

/* optimize.ImagePose, field for field (src/optimize.zig:32-48): the glue copies by name, the
 * oracle's or_pose shares the layout */
STATIC_CHECK(sizeof(fg_pose) == 15 * sizeof(double), pose_size);
STATIC_CHECK(offsetof(fg_pose, yaw) == 0 * 8 && offsetof(fg_pose, pitch) == 1 * 8 && offsetof(fg_pose, roll) == 2 * 8 &&
                 offsetof(fg_pose, hfov_delta) == 3 * 8 && offsetof(fg_pose, trans_x) == 4 * 8 &&
                 offsetof(fg_pose, trans_y) == 5 * 8 && offsetof(fg_pose, trans_z) == 6 * 8 &&
                 offsetof(fg_pose, translation_plane_yaw) == 7 * 8 &&
                 offsetof(fg_pose, translation_plane_pitch) == 8 * 8 && offsetof(fg_pose, radial_a) == 9 * 8 &&
                 offsetof(fg_pose, radial_b) == 10 * 8 && offsetof(fg_pose, radial_c) == 11 * 8 &&
                 offsetof(fg_pose, center_shift_x) == 12 * 8 && offsetof(fg_pose, center_shift_y) == 13 * 8 &&
                 offsetof(fg_pose, base_hfov_degrees) == 14 * 8,
             pose_fields);
STATIC_CHECK(sizeof(fg_pose) == sizeof(or_pose), pose_as_oracle);
/* remap.Rect (src/remap.zig:573-578) */
STATIC_CHECK(sizeof(fg_rect) == 16 && offsetof(fg_rect, left) == 0 && offsetof(fg_rect, top) == 4 &&
                 offsetof(fg_rect, right) == 8 && offsetof(fg_rect, bottom) == 12,
             rect);
STATIC_CHECK(sizeof(fg_image_desc) == 20 && offsetof(fg_image_desc, sample_type) == 16, image_desc);
STATIC_CHECK(sizeof(fg_context_opts) == 8, context_opts);
STATIC_CHECK(offsetof(fg_stack_opts, memory_fraction) == 32 && offsetof(fg_stack_opts, debug_frame) == 36 &&
                 offsetof(fg_stack_opts, stream) == 40 && sizeof(fg_stack_opts) == 48,
             stack_opts);
STATIC_CHECK(sizeof(fg_server_opts) == 12, server_opts);
STATIC_CHECK(offsetof(fg_request, descs) == 8 && offsetof(fg_request, roi) == 32 && offsetof(fg_request, method) == 40 &&
                 offsetof(fg_request, out) == 56 && sizeof(fg_request) == 64,
             request);
STATIC_CHECK(sizeof(fg_request_status) == 32, request_status);
/* fuse.config.Method order (src/fuse/config.zig:10-15) and image_io.SampleType (src/image_io.zig:25-28) */
STATIC_CHECK(FG_METHOD_HARDMASK_CONTRAST == 0 && FG_METHOD_SOFTMASK_CONTRAST == 1 && FG_METHOD_PYRAMID_CONTRAST == 2 &&
                 FG_METHOD_HYBRID_PYRAMID_CONTRAST == 3,
             methods);
STATIC_CHECK(FG_SAMPLE_U8 == 0 && FG_SAMPLE_U16 == 1, sample_types);
STATIC_CHECK(FG_OK == 0 && FG_ERR_INVALID_ARG == -1 && FG_ERR_OOM_HOST == -2 && FG_ERR_OOM_DEVICE == -3 &&
                 FG_ERR_SIZE_MISMATCH == -4 && FG_ERR_FORMAT_MISMATCH == -5 && FG_ERR_UNSUPPORTED == -6 &&
                 FG_ERR_DEVICE == -7 && FG_ERR_STATE == -8,
             status_codes);
STATIC_CHECK(FG_PHASE_COUNT == 5, phases);

/* ------------------------------------------------------------------ synthetic stack */

enum { NF = 4, SW = 160, SH = 120 };

static uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

/* frame k: a texture whose fine detail is strongest in band k of the image (a focus sweep) */
static void make_frame(int k, int bytes, void* px) {
    for (int y = 0; y < SH; y++)
        for (int x = 0; x < SW; x++) {
            const int band = (x * NF) / SW;
            const int detail = band == k ? 90 : 20;
            for (int c = 0; c < 3; c++) {
                const uint32_t h = hash32((uint32_t)((y * SW + x) * 3 + c) * 2654435761U + 17U);
                int v = 60 + (x + 2 * y + 40 * c) % 120 + (int)(h % (uint32_t)detail) - detail / 2 + 8 * k;
                v = v < 0 ? 0 : (v > 255 ? 255 : v);
                const size_t i = ((size_t)y * SW + x) * 3 + c;
                if (bytes == 1)
                    ((uint8_t*)px)[i] = (uint8_t)v;
                else
                    ((uint16_t*)px)[i] = (uint16_t)(v * 256 + (int)(h >> 24));
            }
        }
}

static fg_pose make_pose(int k) {
    fg_pose p;
    memset(&p, 0, sizeof p);
    p.yaw = 0.002 * (k - 1);
    p.pitch = -0.0015 * k;
    p.roll = 0.001 * k;
    p.hfov_delta = 0.05 * k;
    p.base_hfov_degrees = 50.0;
    return p;
}

/* the final-image bar: at most one step, mean absolute difference below 0.05 */
static int images_close(const void* a, const void* b, size_t n, int bytes, int* maxd, double* meand) {
    double sum = 0.0;
    int mx = 0;
    for (size_t i = 0; i < n; i++) {
        const int va = bytes == 1 ? ((const uint8_t*)a)[i] : ((const uint16_t*)a)[i];
        const int vb = bytes == 1 ? ((const uint8_t*)b)[i] : ((const uint16_t*)b)[i];
        const int d = va > vb ? va - vb : vb - va;
        if (d > mx) mx = d;
        sum += d;
    }
    *maxd = mx;
    *meand = sum / (double)n;
    return mx <= 1 && *meand < 0.05;
}

/* share of the pixels whose rgb differs between two images (the hardmask winners must all agree) */
static double pixels_differing(const void* a, const void* b, size_t npx3, int bytes) {
    size_t n = 0;
    for (size_t p = 0; p < npx3 / 3; p++) {
        int d = 0;
        for (int c = 0; c < 3; c++) {
            const size_t i = p * 3 + c;
            d |= bytes == 1 ? ((const uint8_t*)a)[i] != ((const uint8_t*)b)[i]
                            : ((const uint16_t*)a)[i] != ((const uint16_t*)b)[i];
        }
        n += d != 0;
    }
    return (double)n / (double)(npx3 / 3);
}

static long file_size(const char* path) {
    FILE* f = fopen(path, "rb");
    long n;
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    n = ftell(f);
    fclose(f);
    return n;
}

/* ------------------------------------------------------------------ host-side paths */

static void* other_thread_fails(void* arg) {
    (void)arg;
    (void)fg_stack_create(NULL, NULL, NULL);
    return strstr(fg_last_error(), "null") ? (void*)1 : NULL;
}

static void host_checks(const char* dir) {
    const char* v = fg_version();
    CHECK(v && strstr(v, "gfx950"), "version '%s'", v ? v : "(null)");
    CHECK(strcmp(fg_last_error(), "") == 0, "no error yet");

    /* null arguments fail with FG_ERR_INVALID_ARG before anything reaches the GPU */
    CHECK(fg_context_create(NULL, NULL) == FG_ERR_INVALID_ARG, "context_create");
    CHECK(strlen(fg_last_error()) > 0, "message set");
    CHECK(fg_stack_create(NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "stack_create");
    CHECK(fg_stack_add_frame(NULL, NULL, NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "add_frame");
    CHECK(fg_stack_pass2_frame(NULL, NULL, NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "pass2_frame");
    CHECK(fg_stack_needs_pass2_frames(NULL) == 0, "needs_pass2");
    CHECK(fg_stack_finish(NULL, NULL) == FG_ERR_INVALID_ARG, "finish");
    CHECK(fg_stack_sync(NULL) == FG_ERR_INVALID_ARG, "sync");
    CHECK(fg_stack_stream(NULL) == NULL, "stream");
    CHECK(fg_remap(NULL, NULL, NULL, NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "remap");
    CHECK(fg_fuse_stack(NULL, 0, NULL, NULL, NULL, NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "fuse_stack");
    CHECK(fg_server_create(NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "server_create");
    CHECK(fg_server_submit(NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "server_submit");
    CHECK(fg_server_wait(NULL, 0, 0.0, NULL) == FG_ERR_INVALID_ARG, "server_wait");
    CHECK(fg_write_tiff(NULL, NULL, NULL) == FG_ERR_INVALID_ARG, "write_tiff");
    CHECK(fg_write_aligned_tiff(NULL, NULL, NULL, NULL, NULL, NULL, 0) == FG_ERR_INVALID_ARG, "write_aligned_tiff");
    fg_stack_destroy(NULL);
    fg_context_destroy(NULL);
    fg_server_destroy(NULL);

    /* the error message is per thread */
    {
        pthread_t t;
        void* ok = NULL;
        (void)fg_remap(NULL, NULL, NULL, NULL, NULL, NULL);
        const char* before = fg_last_error();
        char mine[256];
        snprintf(mine, sizeof mine, "%s", before);
        CHECK(pthread_create(&t, NULL, other_thread_fails, NULL) == 0, "thread");
        pthread_join(t, &ok);
        CHECK(ok != NULL, "other thread saw its own message");
        CHECK(strcmp(fg_last_error(), mine) == 0, "this thread's message unchanged");
    }

    /* computeCommonOverlapRoi: the library and the oracle agree on the synthetic poses */
    {
        fg_pose poses[NF];
        uint32_t ws[NF], hs[NF];
        fg_rect r;
        int32_t oroi[4];
        for (int k = 0; k < NF; k++) {
            poses[k] = make_pose(k);
            ws[k] = SW;
            hs[k] = SH;
        }
        const int a = fg_common_overlap_roi(NF, NULL, ws, hs, poses, &r);
        const int b = or_common_overlap_roi(NF, NULL, ws, hs, (const or_pose*)poses, oroi);
        CHECK(a == 1 && b == 1, "overlap found (%d, %d)", a, b);
        CHECK(r.left == oroi[0] && r.top == oroi[1] && r.right == oroi[2] && r.bottom == oroi[3],
              "roi (%d %d %d %d) vs oracle (%d %d %d %d)", r.left, r.top, r.right, r.bottom, oroi[0], oroi[1], oroi[2],
              oroi[3]);
        for (int k = 0; k < NF; k++) memset(&poses[k], 0, sizeof poses[k]), poses[k].base_hfov_degrees = 50.0;
        /* identity poses: the forward-mapped corners land within an ulp of the frame corners, and
         * ceil/floor+1 of them (src/remap.zig:626-631) is the frame up to that ulp */
        CHECK(fg_common_overlap_roi(NF, NULL, ws, hs, poses, &r) == 1 &&
                  or_common_overlap_roi(NF, NULL, ws, hs, (const or_pose*)poses, oroi) == 1 && r.left == oroi[0] &&
                  r.top == oroi[1] && r.right == oroi[2] && r.bottom == oroi[3] && r.left <= 1 && r.top <= 1 &&
                  r.right == SW && r.bottom == SH,
              "identity poses: (%d %d %d %d)", r.left, r.top, r.right, r.bottom);
    }

    /* writeTiff: an RGBA u8 image of 3x2 is one strip at offset 8, the directory after it */
    {
        uint8_t px[2 * 3 * 4];
        char path[1024];
        fg_image_desc d = {3, 2, 3, 1, FG_SAMPLE_U8};
        for (int i = 0; i < (int)sizeof px; i++) px[i] = (uint8_t)(i * 11);
        snprintf(path, sizeof path, "%s/fg_conformance_host.tif", dir);
        CHECK(fg_write_tiff(path, &d, px) == FG_OK, "write_tiff");
        FILE* f = fopen(path, "rb");
        uint8_t buf[512];
        size_t n = f ? fread(buf, 1, sizeof buf, f) : 0;
        if (f) fclose(f);
        /* header + 24 pixel bytes + directory of 18 entries + 32 bytes of out-of-line values */
        CHECK(n == 8 + 24 + (2 + 18 * 12 + 4) + 32, "file size %zu", n);
        CHECK(n >= 34 && buf[0] == 'I' && buf[1] == 'I' && buf[2] == 42 && buf[3] == 0 && buf[4] == 32 && buf[5] == 0,
              "header");
        CHECK(n >= 34 && memcmp(buf + 8, px, sizeof px) == 0 && buf[32] == 18, "pixels and entry count");
        fg_image_desc bad = {3, 2, 2, 0, FG_SAMPLE_U8};
        CHECK(fg_write_tiff(path, &bad, px) == FG_ERR_INVALID_ARG, "two colour channels rejected");
        remove(path);
    }
}

/* ------------------------------------------------------------------ GPU paths */

static int oracle_method(uint32_t m) {
    switch (m) {
    case FG_METHOD_HARDMASK_CONTRAST: return OR_METHOD_HARDMASK;
    case FG_METHOD_SOFTMASK_CONTRAST: return OR_METHOD_SOFTMASK;
    case FG_METHOD_HYBRID_PYRAMID_CONTRAST: return OR_METHOD_HYBRID;
    default: return OR_METHOD_PYRAMID;
    }
}

/* The Zig glue's loop: create once the ROI is known, one add_frame per decoded frame (freed
 * right after the call), the uncached second loop when asked for, finish. */
static int run_glue(fg_context* ctx, uint32_t method, int bytes, float memory_fraction, void* const* frames,
                    const fg_pose* poses, const fg_rect* roi, void* out, int* used_pass2) {
    fg_stack_opts o;
    fg_stack* s = NULL;
    const size_t fb = (size_t)SW * SH * 3 * bytes;
    fg_image_desc d = {SW, SH, 3, 0, bytes == 1 ? FG_SAMPLE_U8 : FG_SAMPLE_U16};
    memset(&o, 0, sizeof o); /* std.mem.zeroes */
    o.method = method;
    o.contrast_window_size = 5;
    o.hybrid_sharpness = 0.35f;
    o.frame_count = NF;
    o.width = (uint32_t)(roi->right - roi->left);
    o.height = (uint32_t)(roi->bottom - roi->top);
    o.sample_type = d.sample_type;
    o.memory_fraction = memory_fraction;
    o.debug_frame = -1;
    int rc = fg_stack_create(ctx, &o, &s);
    if (rc) return rc;
    for (int pass = 0; pass < 2 && rc == FG_OK; pass++) {
        if (pass == 1 && !fg_stack_needs_pass2_frames(s)) break;
        if (pass == 1) *used_pass2 = 1;
        for (int k = 0; k < NF && rc == FG_OK; k++) {
            void* tmp = malloc(fb); /* loadImage */
            memcpy(tmp, frames[k], fb);
            rc = pass == 0 ? fg_stack_add_frame(s, &d, tmp, &poses[k], roi) : fg_stack_pass2_frame(s, &d, tmp, &poses[k], roi);
            memset(tmp, 0xA5, fb); /* src.deinit: the library must have taken its copy */
            free(tmp);
        }
    }
    if (rc == FG_OK) rc = fg_stack_finish(s, out);
    fg_stack_destroy(s);
    return rc;
}

static void gpu_checks(const char* dir) {
    fg_context* ctx = NULL;
    fg_context_opts co = {0, 0.5f};
    CHECK(fg_context_create(&co, &ctx) == FG_OK && ctx, "context");
    if (!ctx) return;

    fg_pose poses[NF];
    uint32_t ws[NF], hs[NF];
    fg_rect roi;
    for (int k = 0; k < NF; k++) {
        poses[k] = make_pose(k);
        ws[k] = SW;
        hs[k] = SH;
    }
    CHECK(fg_common_overlap_roi(NF, NULL, ws, hs, poses, &roi) == 1, "roi");
    const int32_t oroi[4] = {roi.left, roi.top, roi.right, roi.bottom};
    const uint32_t W = (uint32_t)(roi.right - roi.left), H = (uint32_t)(roi.bottom - roi.top);
    const size_t npx = (size_t)W * H * 3;

    for (int bytes = 1; bytes <= 2; bytes++) {
        void* frames[NF];
        for (int k = 0; k < NF; k++) {
            frames[k] = malloc((size_t)SW * SH * 3 * bytes);
            make_frame(k, bytes, frames[k]);
        }
        void* out = malloc(npx * bytes);
        void* out2 = malloc(npx * bytes);
        void* ref = malloc(npx * bytes);
        const uint32_t methods[4] = {FG_METHOD_PYRAMID_CONTRAST, FG_METHOD_HYBRID_PYRAMID_CONTRAST,
                                     FG_METHOD_HARDMASK_CONTRAST, FG_METHOD_SOFTMASK_CONTRAST};
        for (int mi = 0; mi < (bytes == 1 ? 4 : 2); mi++) {
            const uint32_t m = methods[mi];
            int p2 = 0, maxd = 0;
            double meand = 0.0;
            or_fuse_opts oo = {oracle_method(m), 5, 0.35f, 0};
            printf("gpu: stack method %u, %d-byte samples, roi %ux%u\n", m, bytes, W, H);
            fflush(stdout);
            CHECK(run_glue(ctx, m, bytes, -1.0f, frames, poses, &roi, out, &p2) == FG_OK, "stack (method %u)", m);
            CHECK(or_stack(NF, (const void* const*)frames, SW, SH, bytes, (const or_pose*)poses, oroi, &oo, ref, NULL) == 0,
                  "oracle");
            if (m == FG_METHOD_HARDMASK_CONTRAST) {
                /* winners: a strict > over weights both sides round identically (8-bit frames) */
                const double frac = pixels_differing(out, ref, npx, bytes);
                CHECK(frac == 0.0, "hardmask vs oracle: %.2e of the pixels differ", frac);
            } else {
                CHECK(images_close(out, ref, npx, bytes, &maxd, &meand), "method %u vs oracle: max %d mean %.4f", m,
                      maxd, meand);
            }
            if (m == FG_METHOD_PYRAMID_CONTRAST || m == FG_METHOD_HYBRID_PYRAMID_CONTRAST) {
                /* a budget too small for the frame cache: the uncached second loop, same result */
                CHECK(run_glue(ctx, m, bytes, 1e-9f, frames, poses, &roi, out2, &p2) == FG_OK && p2 == 1,
                      "uncached stack (method %u)", m);
                CHECK(memcmp(out, out2, npx * bytes) == 0, "uncached == cached (method %u)", m);
            }
            /* the one-shot entry point computes the same */
            {
                fg_image_desc descs[NF];
                fg_stack_opts o;
                for (int k = 0; k < NF; k++) {
                    fg_image_desc d = {SW, SH, 3, 0, bytes == 1 ? FG_SAMPLE_U8 : FG_SAMPLE_U16};
                    descs[k] = d;
                }
                memset(&o, 0, sizeof o);
                o.method = m;
                o.contrast_window_size = 5;
                o.hybrid_sharpness = 0.35f;
                o.width = W;
                o.height = H;
                o.sample_type = descs[0].sample_type;
                o.memory_fraction = -1.0f;
                o.debug_frame = -1;
                CHECK(fg_fuse_stack(ctx, NF, descs, (const void* const*)frames, poses, &roi, &o, out2) == FG_OK,
                      "fuse_stack");
                CHECK(memcmp(out, out2, npx * bytes) == 0, "fuse_stack == stack loop (method %u)", m);
            }
        }

        /* focus_fuse_zig: already aligned RGBA frames, no pose (src/fuse/pipeline.zig:289-336) */
        {
            void* aligned[NF];
            const size_t ab = (size_t)W * H * 4 * bytes;
            fg_image_desc sd = {SW, SH, 3, 0, bytes == 1 ? FG_SAMPLE_U8 : FG_SAMPLE_U16};
            fg_image_desc ad = {W, H, 3, 1, sd.sample_type};
            fg_stack_opts o;
            fg_stack* s = NULL;
            int maxd = 0;
            double meand = 0.0;
            or_fuse_opts oo = {OR_METHOD_PYRAMID, 5, 0.35f, 0};
            printf("gpu: aligned frames, %d-byte samples\n", bytes);
            fflush(stdout);
            for (int k = 0; k < NF; k++) {
                aligned[k] = malloc(ab);
                CHECK(fg_remap(ctx, &sd, frames[k], &poses[k], &roi, aligned[k]) == FG_OK, "remap %d", k);
            }
            memset(&o, 0, sizeof o);
            o.method = FG_METHOD_PYRAMID_CONTRAST;
            o.contrast_window_size = 5;
            o.hybrid_sharpness = 0.35f;
            o.frame_count = NF;
            o.width = W;
            o.height = H;
            o.sample_type = ad.sample_type;
            o.memory_fraction = -1.0f;
            o.debug_frame = -1;
            CHECK(fg_stack_create(ctx, &o, &s) == FG_OK, "stack");
            if (s) {
                fg_image_desc wide = ad, other = ad;
                wide.width = W + 1;
                other.sample_type = bytes == 1 ? FG_SAMPLE_U16 : FG_SAMPLE_U8;
                CHECK(fg_stack_finish(s, out) == FG_ERR_STATE, "finish before the frames");
                CHECK(fg_stack_add_frame(s, &wide, aligned[0], NULL, NULL) == FG_ERR_SIZE_MISMATCH, "size mismatch");
                CHECK(fg_stack_add_frame(s, &other, aligned[0], NULL, NULL) == FG_ERR_FORMAT_MISMATCH, "format mismatch");
                for (int k = 0; k < NF; k++)
                    CHECK(fg_stack_add_frame(s, &ad, aligned[k], NULL, NULL) == FG_OK, "aligned frame %d", k);
                CHECK(fg_stack_add_frame(s, &ad, aligned[0], NULL, NULL) == FG_ERR_STATE, "one frame too many");
                CHECK(fg_stack_finish(s, out) == FG_OK, "finish");
                fg_stack_destroy(s);
            }
            CHECK(or_fuse_frames(NF, (const void* const*)aligned, W, H, bytes, 1, &oo, ref, NULL) == 0, "oracle");
            CHECK(images_close(out, ref, npx, bytes, &maxd, &meand), "aligned vs oracle: max %d mean %.4f", maxd, meand);
            for (int k = 0; k < NF; k++) free(aligned[k]);
        }

        /* remap + writeTiff: writeAlignedImageTask's file, with the remap checked against the oracle */
        {
            fg_image_desc sd = {SW, SH, 3, 0, bytes == 1 ? FG_SAMPLE_U8 : FG_SAMPLE_U16};
            const size_t rb = (size_t)W * H * 4 * bytes;
            void* got = malloc(rb);
            void* want = malloc(rb);
            char prefix[1024], path[1100];
            size_t differ = 0;
            int maxd = 0;
            printf("gpu: remap and aligned TIFF, %d-byte samples\n", bytes);
            fflush(stdout);
            CHECK(fg_remap(ctx, &sd, frames[2], &poses[2], &roi, got) == FG_OK, "remap");
            if (bytes == 1)
                or_remap_u8((const uint8_t*)frames[2], SW, SH, 3, (const or_pose*)&poses[2], oroi, (uint8_t*)want);
            else
                or_remap_u16((const uint16_t*)frames[2], SW, SH, 3, (const or_pose*)&poses[2], oroi, (uint16_t*)want);
            /* alpha always compared; colour where either side has support (a zero-support
             * sample's colour is not defined by the reference) */
            for (size_t i = 0; i < (size_t)W * H * 4; i++) {
                const size_t alpha = i - i % 4 + 3;
                const int a = bytes == 1 ? ((uint8_t*)got)[i] : ((uint16_t*)got)[i];
                const int b = bytes == 1 ? ((uint8_t*)want)[i] : ((uint16_t*)want)[i];
                const int ga = bytes == 1 ? ((uint8_t*)got)[alpha] : ((uint16_t*)got)[alpha];
                const int wa = bytes == 1 ? ((uint8_t*)want)[alpha] : ((uint16_t*)want)[alpha];
                const int d = (i != alpha && ga == 0 && wa == 0) ? 0 : (a > b ? a - b : b - a);
                differ += d != 0;
                if (d > maxd) maxd = d;
            }
            CHECK(differ == 0, "remap vs oracle: max %d, %zu samples differ", maxd, differ);
            snprintf(prefix, sizeof prefix, "%s/fg_conformance_%d", dir, bytes);
            CHECK(fg_write_aligned_tiff(ctx, &sd, frames[2], &poses[2], &roi, prefix, 2) == FG_OK, "write_aligned_tiff");
            snprintf(path, sizeof path, "%s_0002.tif", prefix);
            {
                FILE* f = fopen(path, "rb");
                uint8_t* buf = (uint8_t*)malloc(rb + 8);
                const size_t n = f ? fread(buf, 1, rb + 8, f) : 0;
                if (f) fclose(f);
                CHECK(n == rb + 8 && memcmp(buf + 8, got, rb) == 0, "pixels of %s", path);
                CHECK(file_size(path) == (long)(8 + rb + (2 + 18 * 12 + 4) + 32), "size of %s", path);
                free(buf);
            }
            remove(path);
            free(got);
            free(want);
        }

        /* the server: device-resident frames, same result as the stack loop */
        if (bytes == 1) {
            const size_t fb = (size_t)SW * SH * 3;
            const void* dframes[NF];
            fg_image_desc descs[NF];
            fg_server* srv = NULL;
            fg_server_opts so = {2, 0.0f, 0};
            fg_request q;
            fg_request_status st;
            uint64_t ticket = 0;
            void* dout = fg_device_alloc(ctx, npx);
            int p2 = 0;
            printf("gpu: server\n");
            fflush(stdout);
            CHECK(dout != NULL, "device output");
            for (int k = 0; k < NF; k++) {
                fg_image_desc d = {SW, SH, 3, 0, FG_SAMPLE_U8};
                void* p = fg_device_alloc(ctx, fb);
                descs[k] = d;
                CHECK(p && fg_copy_to_device(ctx, p, frames[k], fb) == FG_OK, "device frame %d", k);
                dframes[k] = p;
            }
            CHECK(run_glue(ctx, FG_METHOD_PYRAMID_CONTRAST, 1, -1.0f, frames, poses, &roi, ref, &p2) == FG_OK, "stack");
            CHECK(fg_server_create(ctx, &so, &srv) == FG_OK, "server_create");
            memset(&q, 0, sizeof q);
            q.frame_count = NF;
            q.descs = descs;
            q.frames = dframes;
            q.poses = poses;
            q.roi = &roi;
            q.method = FG_METHOD_PYRAMID_CONTRAST;
            q.contrast_window_size = 5;
            q.hybrid_sharpness = 0.35f;
            q.out = dout;
            CHECK(srv && fg_server_submit(srv, &q, &ticket) == FG_OK, "submit");
            CHECK(srv && fg_server_wait(srv, ticket, -1.0, &st) == FG_OK && st.state == FG_REQ_DONE, "wait");
            fg_server_destroy(srv);
            CHECK(fg_copy_to_host(ctx, out, dout, npx) == FG_OK, "copy out");
            CHECK(memcmp(out, ref, npx) == 0, "served == stack loop");
            for (int k = 0; k < NF; k++) fg_device_free(ctx, (void*)dframes[k]);
            fg_device_free(ctx, dout);
        }
        for (int k = 0; k < NF; k++) free(frames[k]);
        free(out);
        free(out2);
        free(ref);
    }
    fg_context_destroy(ctx);
}

int main(int argc, char** argv) {
    const int gpu = argc > 1 && strcmp(argv[1], "--gpu") == 0;
    const char* dir = argc > 2 ? argv[2] : (argc > 1 && !gpu ? argv[1] : "/tmp");
    host_checks(dir);
    if (gpu) gpu_checks(dir);
    printf("conformance: %d checks, %d failed%s\n", g_checks, g_fail, gpu ? " (with GPU)" : "");
    return g_fail ? 1 : 0;
}
