// C ABI of libfocus_gpu.so (include/focus_gpu.h): contexts, the stack sequence (pass 1 / pass 2 /
// collapse, src/stack/pipeline.zig:361-605 and src/fuse/pipeline.zig:254-445), the frame cache
// under the memory budget, and the host-side pose preparation (src/optimize.zig:3555-3586).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/focus_gpu.h"
#include "fg_host.h"
#include "fg_internal.h"
#include "fg_upload.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
int fail_hip(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? FG_ERR_OOM_DEVICE : FG_ERR_DEVICE;
}

#define HIP_TRY(expr, what)                      \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) return fail_hip(e_, what); \
    } while (0)

// ---- pose math (host, f64) ----
const double kPi = 3.14159265358979323846;

double focal_length_px(uint32_t width, double hfov) {  // focalLengthPixels src/optimize.zig:3705-3708
    double h = hfov < 1e-3 ? 1e-3 : (hfov > 179.0 ? 179.0 : hfov);
    return ((double)width * 0.5) / tan(h * kPi / 360.0);
}

void mat_mul(const double a[3][3], const double b[3][3], double out[3][3]) {  // matrixMatrixMul :3951-3959
    double r[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) r[i][k] = a[i][0] * b[0][k] + a[i][1] * b[1][k] + a[i][2] * b[2][k];
    memcpy(out, r, sizeof r);
}

void set_matrix(double a, double b, double c, double m[3][3]) {  // setMatrix(a, b, c, cl=true) :3901-3921
    double mx[3][3] = {{1, 0, 0}, {0, cos(a), sin(a)}, {0, -sin(a), cos(a)}};
    double my[3][3] = {{cos(b), 0, -sin(b)}, {0, 1, 0}, {sin(b), 0, cos(b)}};
    double mz[3][3] = {{cos(c), sin(c), 0}, {-sin(c), cos(c), 0}, {0, 0, 1}};
    double d[3][3];
    mat_mul(mz, mx, d);
    mat_mul(d, my, m);
}

bool has_translation(const fg_pose& p) {  // hasTranslation :3962-3964
    return fabs(p.trans_x) > 1e-12 || fabs(p.trans_y) > 1e-12 || fabs(p.trans_z) > 1e-12;
}
bool basic_rectilinear(const fg_pose& p) {  // hasBasicRectilinearPose :3074-3081
    return !has_translation(p) && fabs(p.radial_a) <= 1e-12 && fabs(p.radial_b) <= 1e-12 &&
           fabs(p.radial_c) <= 1e-12 && fabs(p.center_shift_x) <= 1e-12 && fabs(p.center_shift_y) <= 1e-12;
}

}  // namespace

namespace fg {

void make_levels(uint32_t w, uint32_t h, uint32_t max_levels, LevelTable* t) {  // computeLevelCount :163-176
    t->n = 0;
    size_t off = 0;
    for (;;) {
        if (t->n >= kMaxLevels) break;
        t->lv[t->n] = Level{w, h, off};
        t->n++;
        off += (size_t)w * h;
        if ((w == 1 && h == 1) || (max_levels && (uint32_t)t->n >= max_levels)) break;
        w = w > 1 ? (w + 1) / 2 : 1;
        h = h > 1 ? (h + 1) / 2 : 1;
    }
    t->total = off;
}

// initInverseTransformCache src/optimize.zig:3555-3586, folded into one 3x3 matrix per frame.
// Basic rectilinear poses: the cached legacy chain (:3618-3634) equals the pinhole homography
// src = c + f * (A p).xy / (A p).z with A = pitch_roll^T * yawMatrix(yaw), p = (x-cx, y-cy, d);
// it is written here in the lens path's form local = B (x-cx, y-cy, -d), B = D A D, D = diag(1,1,-1).
void make_remap_params(const double pose15[15], uint32_t w, uint32_t h, int32_t roi_left, int32_t roi_top,
                       RemapParams* p) {
    fg_pose pose;
    memcpy(&pose, pose15, sizeof pose);
    memset(p, 0, sizeof *p);
    double cx = ((double)w - 1.0) * 0.5, cy = ((double)h - 1.0) * 0.5;
    double pano = focal_length_px(w, pose.base_hfov_degrees);
    double eff = pose.base_hfov_degrees + pose.hfov_delta;
    double img_f = focal_length_px(w, eff < 1e-3 ? 1e-3 : (eff > 179.0 ? 179.0 : eff));
    double pr[3][3], ym[3][3], w2l[3][3];
    set_matrix(-pose.pitch, 0.0, pose.roll, pr);
    double cyw = cos(-pose.yaw), syw = sin(-pose.yaw);
    double ymt[3][3] = {{cyw, 0, -syw}, {0, 1, 0}, {syw, 0, cyw}};
    memcpy(ym, ymt, sizeof ym);
    mat_mul(pr, ym, w2l);
    p->neg_d = -pano;
    p->dcx = cx;
    p->dcy = cy;
    p->scx = cx;
    p->scy = cy;
    p->focal = img_f;
    p->roi_left = roi_left;
    p->roi_top = roi_top;
    double hw = ((double)w - 1.0) * 0.5, hh = ((double)h - 1.0) * 0.5;
    double den = hw * hw + hh * hh;
    p->inv_norm = den >= 1e-12 ? 1.0 / den : 1.0;  // radialNormalizer :4150-4156
    p->ra = pose.radial_a;
    p->rb = pose.radial_b;
    p->rc = pose.radial_c;
    if (basic_rectilinear(pose)) {
        double c = cos(pose.yaw), s = sin(pose.yaw);
        double ry[3][3] = {{c, 0, -s}, {0, 1, 0}, {s, 0, c}};
        double prt[3][3], a[3][3];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) prt[i][k] = pr[k][i];
        mat_mul(prt, ry, a);
        const double d[3] = {1.0, 1.0, -1.0};
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) p->B[i * 3 + k] = d[i] * a[i][k] * d[k];
        p->mode = REMAP_HOMOGRAPHY;
        p->legacy = 1;
        return;
    }
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) p->B[i * 3 + k] = w2l[i][k];
    p->csx = pose.center_shift_x;
    p->csy = pose.center_shift_y;
    if (has_translation(pose)) {
        // translationPlaneNormal :3966-3968 = cartErect(tp_yaw, -tp_pitch, 1) :4045-4053
        double phi = pose.translation_plane_yaw / 1.0;
        double tz = kPi * 0.5 - (-pose.translation_plane_pitch / 1.0);
        p->pn[0] = sin(tz) * sin(phi);
        p->pn[1] = cos(tz);
        p->pn[2] = sin(tz) * -cos(phi);
        p->tx = pose.trans_x;
        p->ty = pose.trans_y;
        p->tz = pose.trans_z;
        p->mode = REMAP_TRANSLATION;
        return;
    }
    p->mode = REMAP_LENS;
    p->radial = (pose.radial_a != 0.0 || pose.radial_b != 0.0 || pose.radial_c != 0.0) ? 1 : 0;
}

}  // namespace fg

// ---------------------------------------------------------------------------

struct fg_context {
    int device;
    float memory_fraction;
    hipStream_t stream;
    // fg_remap's device copies of the source and the output: grown on demand, kept until the
    // context ends (one hipFree per growth instead of a hipMalloc / hipFree pair per call)
    std::mutex remap_mu;
    void* remap_in;
    void* remap_out;
    size_t remap_in_cap, remap_out_cap;
    // threads that copy host-fed frames into the stacks' page-locked upload chunks, shared by the
    // context's stacks; made at the first such frame
    std::mutex pool_mu;
    fg::CopyPool* copy_pool;
    // fg_remap's page-locked chunk ring for pageable sources and outputs (under remap_mu)
    fg::HostUploader* remap_up;
    bool remap_up_tried;
};

// One stack: the pass-1 / pass-2 sequence.  Pass 1 runs over groups of frames (all pending frames
// at finish, or whenever the host staging slots are full); pass 2 runs over every cached frame
// at finish, or frame by frame when the frames are fed a second time.
struct fg_stack {
    fg_context* ctx;
    fg_stack_opts opts;
    hipStream_t stream;
    bool own_stream;
    uint32_t W, H;
    int sb;
    fg::LevelTable lt;
    bool cached;
    bool single;          // hardmask / softmask: one pass, no pyramids
    uint32_t added, p1_done, pass2_done;
    bool finished;
    uint32_t src_w, src_h, src_cc, src_extra;  // geometry of the first frame (later frames must match)
    bool have_src;
    std::vector<fg::FrameDesc> fdesc;
    // device memory
    std::vector<void*> allocs;
    fg_dev_allocator ext;   // buffers from an external allocator when ext.alloc is set
    fg::FrameDesc* d_fdesc;
    uint8_t* frames;      // RGBA cache slots
    float* weights;       // weight cache slots
    float4* lev;          // per-frame levels >= 1
    size_t lev_stride;
    uint32_t slots;
    float* S;
    uint8_t* U;
    float* acc;
    float* contrib;
    size_t contrib_floats;
    float* best;
    uint8_t* hard;
    uint8_t* soft;
    std::vector<uint8_t*> stage;  // device copies of host-fed source frames awaiting pass 1
    size_t stage_bytes;
    uint32_t stage_used;
    uint32_t stage_slots;         // slots of the ring (kStageSlots, fewer if device memory ran short)
    // uploads of host-fed frames run on their own stream: the compute stream waits for a slot's
    // upload, the next upload into a slot waits until pass 1 has read it
    hipStream_t cstream;
    hipEvent_t up_ev[8];
    hipEvent_t free_ev[8];
    bool free_rec[8];
    fg::HostUploader* up;  // pinned chunk ring for the uploads (nullptr: plain copies from the caller's memory)
    bool up_tried;
    float* accDbg;
    uint8_t* out_dev;
    float* merge_gs;
    float* merge_gh;
    // the cached stack's pass-2 + collapse launches, captured when the stack finishes a second time
    // (reset, pooled) and replayed by every later finish: the launches and their arguments are
    // fixed for the stack's life
    hipGraphExec_t tail;
    uint32_t tail_runs;   // cached finishes so far
    bool tail_eager;      // capture failed once: stay eager
    // profiling: event pairs per phase, resolved on query
    bool profiling;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[FG_PHASE_COUNT];
    std::vector<hipEvent_t> ev_pool;
};

namespace {

constexpr uint32_t kStageSlots = 8;

// The stack's page-locked chunk ring (and its context's copy threads), made at the first host-fed
// frame or host-memory output; nullptr where they are refused (plain copies then).
fg::HostUploader* make_uploader(fg_context* c) {
    fg::CopyPool* pool = nullptr;
    {
        std::lock_guard<std::mutex> lk(c->pool_mu);
        if (!c->copy_pool) c->copy_pool = new (std::nothrow) fg::CopyPool(fg::copy_pool_threads());
        pool = c->copy_pool;
    }
    fg::HostUploader* up = new (std::nothrow) fg::HostUploader();
    if (up && !up->init(pool)) {
        delete up;
        up = nullptr;
    }
    return up;
}

void ensure_uploader(fg_stack* s) {
    if (s->up_tried) return;
    s->up_tried = true;
    s->up = make_uploader(s->ctx);
}

// host memory that HIP allocated page-locked (the DMA engine reads it in place)
bool host_page_locked(const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory: an error on some runtimes, type 0 on others
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

hipEvent_t take_event(fg_stack* s) {
    hipEvent_t e = nullptr;
    if (!s->ev_pool.empty()) {
        e = s->ev_pool.back();
        s->ev_pool.pop_back();
        return e;
    }
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the work enqueued in its scope with two events on the stack's stream.
struct PhaseMark {
    fg_stack* s;
    int phase;
    hipStream_t st;
    hipEvent_t a;
    PhaseMark(fg_stack* sk, int ph, hipStream_t on = nullptr) : s(sk), phase(ph), st(on ? on : sk->stream), a(nullptr) {
        if (ph >= 0 && s->profiling && (a = take_event(s))) (void)hipEventRecord(a, st);
    }
    ~PhaseMark() {
        if (!a) return;
        hipEvent_t b = take_event(s);
        if (!b) return;
        (void)hipEventRecord(b, st);
        s->ev[phase].emplace_back(a, b);
    }
};

// Marks the staging slots [0, n) as read once the work queued so far on the compute stream is done.
int release_slots(fg_stack* s, uint32_t n) {
    if (!s->cstream) return FG_OK;
    for (uint32_t j = 0; j < n && j < s->stage_slots; j++) {
        HIP_TRY(hipEventRecord(s->free_ev[j], s->stream), "slot event");
        s->free_rec[j] = true;
    }
    return FG_OK;
}

void recycle_events(fg_stack* s) {
    for (int k = 0; k < FG_PHASE_COUNT; k++) {
        for (auto& pr : s->ev[k]) {
            s->ev_pool.push_back(pr.first);
            s->ev_pool.push_back(pr.second);
        }
        s->ev[k].clear();
    }
}

size_t frame_bytes(const fg_stack* s) { return (size_t)s->W * s->H * 4 * s->sb; }
size_t weight_bytes(const fg_stack* s) { return (size_t)s->W * s->H * sizeof(float); }

int dev_alloc(fg_stack* s, size_t bytes, void** out) {
    void* p = nullptr;
    if (s->ext.alloc) {
        p = s->ext.alloc(s->ext.user, bytes ? bytes : 256);
        if (!p) return fail(FG_ERR_OOM_DEVICE, "arena allocation of " + std::to_string(bytes) + " bytes failed");
        s->allocs.push_back(p);
        *out = p;
        return FG_OK;
    }
    hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(FG_ERR_OOM_DEVICE, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    s->allocs.push_back(p);
    *out = p;
    return FG_OK;
}

// `sync`: wait for the stack's streams first; a caller that knows the stack's work is complete
// (a pooled server stack, whose completion event fired) passes false, since the stream may be a
// shared one that already runs another stack's work.
void free_stack(fg_stack* s, bool sync = true) {
    if (!s) return;
    if (sync && s->cstream) (void)hipStreamSynchronize(s->cstream);
    if (sync && s->stream) (void)hipStreamSynchronize(s->stream);
    delete s->up;  // waits for transfers still reading its chunks
    s->up = nullptr;
    if (s->cstream) {
        for (int j = 0; j < 8; j++) {
            (void)hipEventDestroy(s->up_ev[j]);
            (void)hipEventDestroy(s->free_ev[j]);
        }
        (void)hipStreamDestroy(s->cstream);
    }
    recycle_events(s);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->tail) (void)hipGraphExecDestroy(s->tail);
    for (void* p : s->allocs) {
        if (s->ext.alloc)
            s->ext.free(s->ext.user, p);
        else
            (void)hipFree(p);
    }
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int check_desc(const fg_image_desc* d) {
    if (!d) return fail(FG_ERR_INVALID_ARG, "null image descriptor");
    if (d->width == 0 || d->height == 0) return fail(FG_ERR_INVALID_ARG, "empty image");
    if (d->color_channels != 1 && d->color_channels != 3)
        return fail(FG_ERR_UNSUPPORTED, "color_channels must be 1 or 3");
    if (d->extra_channels > 1) return fail(FG_ERR_UNSUPPORTED, "extra_channels must be 0 or 1");
    if (d->sample_type != FG_SAMPLE_U8 && d->sample_type != FG_SAMPLE_U16)
        return fail(FG_ERR_UNSUPPORTED, "sample_type must be u8 or u16");
    if (frame_too_large(*d)) return fail(FG_ERR_UNSUPPORTED, "frames must hold fewer than 2^32 - 16 bytes");
    return FG_OK;
}

size_t desc_bytes(const fg_image_desc* d) {
    return (size_t)d->width * d->height * (d->color_channels + d->extra_channels) *
           (d->sample_type == FG_SAMPLE_U16 ? 2 : 1);
}

int set_device(int dev) {
    HIP_TRY(hipSetDevice(dev), "hipSetDevice");
    return FG_OK;
}

// Validates frame k and fills its pass-1 descriptor; host pixels are copied to a staging slot.
int describe_frame(fg_stack* s, const fg_image_desc* d, const void* px, bool device_ptr, const fg_pose* pose,
                   const fg_rect* roi, fg::FrameDesc* f) {
    int rc = check_desc(d);
    if (rc) return rc;
    if (!px) return fail(FG_ERR_INVALID_ARG, "null pixel pointer");
    int sb = d->sample_type == FG_SAMPLE_U16 ? 2 : 1;
    if (sb != s->sb) return fail(FG_ERR_FORMAT_MISMATCH, "frame sample type differs from the stack's");
    if (d->color_channels != 3) return fail(FG_ERR_UNSUPPORTED, "pyramid fusion needs rgb frames");
    // every frame of a stack has the first frame's size and format (src/fuse/io.zig:31-42)
    if (s->have_src) {
        if (d->width != s->src_w || d->height != s->src_h)
            return fail(FG_ERR_SIZE_MISMATCH, "frame size differs from the first frame's");
        if (d->extra_channels != s->src_extra) return fail(FG_ERR_FORMAT_MISMATCH, "frame channels differ");
    }
    memset(f, 0, sizeof *f);
    f->sw = d->width;
    f->sh = d->height;
    if (pose) {
        if (d->extra_channels != 0) return fail(FG_ERR_UNSUPPORTED, "remap sources carry no alpha channel");
        fg_rect r = roi ? *roi : fg_rect{0, 0, (int32_t)d->width, (int32_t)d->height};
        if (r.right <= r.left || r.bottom <= r.top || (uint32_t)(r.right - r.left) != s->W ||
            (uint32_t)(r.bottom - r.top) != s->H)
            return fail(FG_ERR_SIZE_MISMATCH, "roi size differs from the stack size");
        fg::make_remap_params((const double*)pose, d->width, d->height, r.left, r.top, &f->p);
        f->kind = 0;
    } else {
        if (d->width != s->W || d->height != s->H)
            return fail(FG_ERR_SIZE_MISMATCH, "aligned frame size differs from the stack size");
        f->kind = 1;
        f->extra = (int32_t)d->extra_channels;
    }
    if (device_ptr) {
        if ((uintptr_t)px % (uintptr_t)sb != 0)
            return fail(FG_ERR_INVALID_ARG, "device frame pointer is not aligned to its sample size");
        f->src = px;
    } else {
        size_t nb = desc_bytes(d);
        if (s->stage.empty() || nb > s->stage_bytes) {
            if (!s->stage.empty()) return fail(FG_ERR_SIZE_MISMATCH, "frame larger than the first frame");
            if (!s->cstream) {
                HIP_TRY(hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking), "hipStreamCreate");
                for (uint32_t j = 0; j < kStageSlots; j++) {
                    HIP_TRY(hipEventCreateWithFlags(&s->up_ev[j], hipEventDisableTiming), "hipEventCreate");
                    HIP_TRY(hipEventCreateWithFlags(&s->free_ev[j], hipEventDisableTiming), "hipEventCreate");
                    s->free_rec[j] = false;
                }
            }
            // the ring is sized for overlap, not needed for correctness: when device memory runs
            // short it keeps the slots it got (at least one), and pass 1 then runs more often
            for (uint32_t j = 0; j < kStageSlots; j++) {
                void* p = nullptr;
                rc = dev_alloc(s, nb, &p);
                if (rc) {
                    if (j == 0) return rc;
                    g_last_error.clear();
                    break;
                }
                s->stage.push_back((uint8_t*)p);
            }
            s->stage_slots = (uint32_t)s->stage.size();
            s->stage_bytes = nb;
        }
        const uint32_t j = s->stage_used;
        uint8_t* slot = s->stage[j];
        if (s->free_rec[j]) HIP_TRY(hipStreamWaitEvent(s->cstream, s->free_ev[j], 0), "slot wait");
        ensure_uploader(s);
        // frames in page-locked memory (hipHostMalloc'd by the caller) go to the device directly
        const bool direct = !s->up || host_page_locked(px);
        {
            PhaseMark pm(s, FG_PHASE_INGEST, s->cstream);
            if (direct)
                HIP_TRY(hipMemcpyAsync(slot, px, nb, hipMemcpyHostToDevice, s->cstream), "upload frame");
            else
                HIP_TRY(s->up->upload(slot, px, nb, s->cstream), "upload frame");
        }
        HIP_TRY(hipEventRecord(s->up_ev[j], s->cstream), "upload event");
        HIP_TRY(hipStreamWaitEvent(s->stream, s->up_ev[j], 0), "upload wait");
        // the caller may release px on return: the uploader has read it by then; a direct copy is
        // waited for (pass 1 of earlier frames keeps running on the compute stream meanwhile)
        if (direct) HIP_TRY(hipEventSynchronize(s->up_ev[j]), "upload");
        f->src = slot;
    }
    if (!s->have_src) {
        s->have_src = true;
        s->src_w = d->width;
        s->src_h = d->height;
        s->src_cc = d->color_channels;
        s->src_extra = d->extra_channels;
    }
    return FG_OK;
}

// Runs pass 1 over the frames added since the last flush.
int flush_pass1(fg_stack* s) {
    if (s->p1_done >= s->added) return FG_OK;
    const uint32_t k0 = s->p1_done, k1 = s->added;
    HIP_TRY(hipMemcpyAsync(s->d_fdesc + k0, s->fdesc.data() + k0, (k1 - k0) * sizeof(fg::FrameDesc),
                           hipMemcpyHostToDevice, s->stream),
            "frame descriptors");
    fg::Pass1Args a{};
    a.frames = s->d_fdesc;
    a.k0 = (int)k0;
    a.k1 = (int)k1;
    a.slot_base = s->cached ? 0 : (int)k0;
    a.sample_bytes = s->sb;
    a.w = s->W;
    a.h = s->H;
    a.window = s->opts.contrast_window_size;
    a.rgba = s->frames;
    a.weights = s->weights;
    a.norm_sum = s->S;
    a.union_sup = s->U;
    a.best = s->best;
    a.hard = s->hard;
    a.lens = fg::frames_use_lens(s->fdesc.data() + k0, k1 - k0) ? 1 : 0;
    const bool softmask = s->opts.method == FG_METHOD_SOFTMASK_CONTRAST;
    if (s->single) {
        a.slot_base = (int)k0;
        a.norm_sum = nullptr;
        a.union_sup = nullptr;
        if (!softmask) {
            a.rgba = nullptr;
            a.weights = nullptr;
        }
    }
    {
        PhaseMark pm(s, FG_PHASE_PASS1);
        HIP_TRY(fg::launch_pass1(a, s->stream), "pass1 kernel");
    }
    if (softmask) {
        // blurFiveTapInto + accumulateSoft for the group (src/stack/pipeline.zig:304-308)
        PhaseMark pm(s, FG_PHASE_PASS2);
        fg::SoftArgs sa{};
        sa.sample_bytes = s->sb;
        sa.k0 = (int)k0;
        sa.k1 = (int)k1;
        sa.slot_base = (int)k0;
        sa.w = s->W;
        sa.h = s->H;
        sa.rgba = s->frames;
        sa.weights = s->weights;
        sa.wsum = s->S;
        sa.psum = s->acc;
        HIP_TRY(fg::launch_soft(sa, s->stream), "soft blend kernel");
    }
    s->p1_done = k1;
    int rc = release_slots(s, s->stage_used);
    s->stage_used = 0;
    return rc;
}

fg::Pass2Args pass2_args(fg_stack* s, int k0, int k1, int slot_base) {
    fg::Pass2Args a{};
    a.sample_bytes = s->sb;
    a.lt = &s->lt;
    a.k0 = k0;
    a.k1 = k1;
    a.slot_base = slot_base;
    a.rgba = s->frames;
    a.weights = s->weights;
    a.norm_sum = s->S;
    a.union_sup = s->U;
    a.frame_count = s->opts.frame_count;
    a.lev = s->lev;
    a.lev_stride = s->lev_stride;
    a.acc = s->acc;
    a.contrib = s->contrib;
    a.contrib_floats = s->contrib_floats;
    return a;
}

int run_pass2(fg_stack* s, int k0, int k1, int slot_base) {
    PhaseMark pm(s, FG_PHASE_PASS2);
    fg::Pass2Args a = pass2_args(s, k0, k1, slot_base);
    {
        PhaseMark pm0(s, FG_PHASE_PASS2_LEVEL0);
        a.level_begin = 0;
        a.level_end = 1;
        HIP_TRY(fg::launch_pass2(a, s->stream), "pass2 level-0 kernel");
    }
    a.level_begin = 1;
    a.level_end = 0;
    HIP_TRY(fg::launch_pass2(a, s->stream), "pass2 kernels");
    return FG_OK;
}

}  // namespace


extern "C" {

const char* fg_last_error(void) { return g_last_error.c_str(); }

void* fg_host_alloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail_hip(e, "hipHostMalloc");
        return nullptr;
    }
    return p;
}

void fg_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

void* fg_device_alloc(fg_context* ctx, size_t bytes) {
    if (!ctx || set_device(ctx->device)) return nullptr;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail_hip(e, "hipMalloc");
        return nullptr;
    }
    return p;
}

void fg_device_free(fg_context* ctx, void* p) {
    if (!ctx || !p || set_device(ctx->device)) return;
    (void)hipFree(p);
}

int fg_copy_to_device(fg_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || !dst || !src) return fail(FG_ERR_INVALID_ARG, "null argument");
    int rc = set_device(ctx->device);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "copy to device");
    return FG_OK;
}

int fg_copy_to_host(fg_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || !dst || !src) return fail(FG_ERR_INVALID_ARG, "null argument");
    int rc = set_device(ctx->device);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "copy to host");
    return FG_OK;
}
const char* fg_version(void) { return "focus_gpu 0.1 (gfx950)"; }

int fg_context_create(const fg_context_opts* opts, fg_context** out) {
    if (!out) return fail(FG_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int dev = opts ? opts->device : 0;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(FG_ERR_DEVICE, "no HIP device available");
    if (dev < 0 || dev >= count) return fail(FG_ERR_INVALID_ARG, "device index out of range");
    int rc = set_device(dev);
    if (rc) return rc;
    fg_context* c = new (std::nothrow) fg_context();
    if (!c) return fail(FG_ERR_OOM_HOST, "context allocation");
    c->device = dev;
    c->memory_fraction = opts ? opts->memory_fraction : 0.5f;
    c->remap_in = c->remap_out = nullptr;
    c->remap_in_cap = c->remap_out_cap = 0;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail_hip(e, "hipStreamCreate");
    }
    // the device address of this GPU's scratch area for fg_fuse_stack_device_noalloc (a failure
    // here only disables that entry point, which then reports it)
    fg_noalloc_setup(dev);
    *out = c;
    return FG_OK;
}

void fg_context_destroy(fg_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx->remap_up;
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->remap_in) (void)hipFree(ctx->remap_in);
    if (ctx->remap_out) (void)hipFree(ctx->remap_out);
    delete ctx->copy_pool;
    delete ctx;
}

int fg_remap_device(fg_context* ctx, const fg_image_desc* src, const void* d_src, const fg_pose* pose,
                    const fg_rect* roi, void* d_dst, void* stream) {
    if (!ctx || !pose || !d_src || !d_dst) return fail(FG_ERR_INVALID_ARG, "null argument");
    int rc = check_desc(src);
    if (rc) return rc;
    if (src->extra_channels != 0) return fail(FG_ERR_UNSUPPORTED, "remap sources carry no alpha channel");
    rc = set_device(ctx->device);
    if (rc) return rc;
    fg_rect r = roi ? *roi : fg_rect{0, 0, (int32_t)src->width, (int32_t)src->height};
    if (r.right <= r.left || r.bottom <= r.top) return fail(FG_ERR_INVALID_ARG, "empty roi");
    fg::RemapParams p;
    fg::make_remap_params((const double*)pose, src->width, src->height, r.left, r.top, &p);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(fg::launch_remap(d_src, src->sample_type == FG_SAMPLE_U16 ? 2 : 1, (int)src->color_channels, src->width,
                             src->height, p, (uint32_t)(r.right - r.left), (uint32_t)(r.bottom - r.top), d_dst, s),
            "remap kernel");
    return FG_OK;
}

int fg_remap(fg_context* ctx, const fg_image_desc* src, const void* src_px, const fg_pose* pose, const fg_rect* roi,
             void* dst_px) {
    if (!ctx || !pose || !src_px || !dst_px) return fail(FG_ERR_INVALID_ARG, "null argument");
    int rc = check_desc(src);
    if (rc) return rc;
    rc = set_device(ctx->device);
    if (rc) return rc;
    fg_rect r = roi ? *roi : fg_rect{0, 0, (int32_t)src->width, (int32_t)src->height};
    if (r.right <= r.left || r.bottom <= r.top) return fail(FG_ERR_INVALID_ARG, "empty roi");
    size_t sb = src->sample_type == FG_SAMPLE_U16 ? 2 : 1;
    size_t in_b = desc_bytes(src);
    size_t out_b = (size_t)(r.right - r.left) * (r.bottom - r.top) * (src->color_channels + 1) * sb;
    // the context's scratch copies, grown when a larger frame comes (calls on one context are
    // serialised here: they share the buffers and the context's stream)
    std::lock_guard<std::mutex> lk(ctx->remap_mu);
    auto grow = [&](void** buf, size_t* cap, size_t need) -> hipError_t {
        if (*cap >= need) return hipSuccess;
        if (*buf) (void)hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
        hipError_t e2 = hipMalloc(buf, need);
        if (e2 == hipSuccess) *cap = need;
        return e2;
    };
    hipError_t e = grow(&ctx->remap_in, &ctx->remap_in_cap, in_b);
    if (e == hipSuccess) e = grow(&ctx->remap_out, &ctx->remap_out_cap, out_b);
    if (e != hipSuccess) return fail_hip(e, "hipMalloc");
    rc = FG_OK;
    // pageable buffers through the context's page-locked chunks, page-locked ones directly
    if (!ctx->remap_up_tried) {
        ctx->remap_up_tried = true;
        ctx->remap_up = make_uploader(ctx);
    }
    fg::HostUploader* up = ctx->remap_up;
    const bool in_direct = !up || host_page_locked(src_px), out_direct = !up || host_page_locked(dst_px);
    e = in_direct ? hipMemcpyAsync(ctx->remap_in, src_px, in_b, hipMemcpyHostToDevice, ctx->stream)
                  : up->upload(ctx->remap_in, src_px, in_b, ctx->stream);
    if (e == hipSuccess) {
        rc = fg_remap_device(ctx, src, ctx->remap_in, pose, &r, ctx->remap_out, ctx->stream);
        if (rc == FG_OK) {
            if (out_direct) {
                e = hipMemcpyAsync(dst_px, ctx->remap_out, out_b, hipMemcpyDeviceToHost, ctx->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            } else {
                e = up->download(dst_px, ctx->remap_out, out_b, ctx->stream);
            }
        }
    }
    if (rc) return rc;
    if (e != hipSuccess) return fail_hip(e, "fg_remap");
    return FG_OK;
}

}  // extern "C"

// Sizes of a stack's device buffers; `budget` decides whether every frame is cached.
struct StackPlan {
    fg::LevelTable lt;
    size_t px, tot, lev_stride, contrib_floats;
    bool hybrid, hardmask, single, cached;
    uint32_t slots;
    size_t bytes;
};

static StackPlan plan_stack(const fg_stack_opts& o, double budget, bool host_frames = true) {
    StackPlan pl{};
    const size_t sb = o.sample_type == FG_SAMPLE_U16 ? 2 : 1;
    fg::make_levels(o.width, o.height, o.max_levels, &pl.lt);
    pl.hybrid = o.method == FG_METHOD_HYBRID_PYRAMID_CONTRAST;
    pl.hardmask = o.method == FG_METHOD_HARDMASK_CONTRAST;
    pl.single = pl.hardmask || o.method == FG_METHOD_SOFTMASK_CONTRAST;
    pl.px = (size_t)o.width * o.height;
    pl.tot = pl.lt.total;
    pl.lev_stride = pl.single ? 0 : pl.tot - pl.px;  // levels >= 1
    pl.contrib_floats = fg::pass2_contrib_floats(&pl.lt, o.frame_count);
    size_t fixed = pl.px * 4 + pl.px + pl.tot * 12 + pl.contrib_floats * 4 + pl.px * 3 * sb +
                   o.frame_count * sizeof(fg::FrameDesc);
    if (pl.hybrid || pl.hardmask) fixed += pl.px * 4 + pl.px * 3 * sb;
    if (pl.hybrid) fixed += pl.px * 3 * sb + 2 * (pl.tot > pl.px ? pl.tot - pl.px : 1) * 12;
    if (o.debug_frame >= 0) fixed += pl.tot * 12;
    const size_t per_frame = pl.px * 4 * sb + pl.px * 4 + pl.lev_stride * sizeof(float4);
    // frames fed from host memory go through an upload ring of kStageSlots source frames, which
    // describe_frame allocates at the first such frame; the caching decision leaves room for it
    // (sources the size of the fused area; a larger source makes the ring shrink, not fail)
    const double ring = host_frames ? (double)kStageSlots * (double)pl.px * 3.0 * (double)sb : 0.0;
    pl.cached = !pl.single && (double)fixed + ring + (double)o.frame_count * (double)per_frame <= budget;
    pl.slots = pl.cached ? o.frame_count : 1;
    // single-pass methods keep no frame past its group: softmask holds a group of frames
    // between pass 1 and the blend, hardmask only the winners
    if (pl.single) pl.slots = pl.hardmask ? 0 : (o.frame_count < kStageSlots ? o.frame_count : kStageSlots);
    pl.bytes = fixed + pl.slots * per_frame;
    return pl;
}

// The kernels address a frame's samples, and a stack's pixels at every level (16 bytes each for the
// level buffers), with 32-bit byte offsets.
bool frame_too_large(const fg_image_desc& d) {
    const uint64_t sb = d.sample_type == FG_SAMPLE_U16 ? 2 : 1;
    return (uint64_t)d.width * d.height * (d.color_channels + d.extra_channels) * sb >= (1ull << 32) - 16;
}
constexpr uint64_t kMaxStackPixels = 1ull << 28;

static int check_stack_opts(const fg_stack_opts* o) {
    if (o->method > FG_METHOD_HYBRID_PYRAMID_CONTRAST) return fail(FG_ERR_UNSUPPORTED, "unknown fusion method");
    if (o->contrast_window_size < 3 || (o->contrast_window_size & 1) == 0 || o->contrast_window_size > 31)
        return fail(FG_ERR_INVALID_ARG, "contrast_window_size must be odd and in [3, 31]");
    if (!(o->hybrid_sharpness >= 0.0f && o->hybrid_sharpness <= 1.0f))
        return fail(FG_ERR_INVALID_ARG, "hybrid_sharpness must be in [0, 1]");
    if (o->frame_count == 0 || o->width == 0 || o->height == 0) return fail(FG_ERR_INVALID_ARG, "empty stack");
    if ((uint64_t)o->width * o->height >= kMaxStackPixels)
        return fail(FG_ERR_UNSUPPORTED, "stacks must hold fewer than 2^28 pixels");
    if (o->sample_type != FG_SAMPLE_U8 && o->sample_type != FG_SAMPLE_U16)
        return fail(FG_ERR_UNSUPPORTED, "sample_type must be u8 or u16");
    return FG_OK;
}

int fg_context_device(const fg_context* ctx) { return ctx ? ctx->device : 0; }
int fg_set_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }

size_t fg_plan_stack_bytes(const fg_stack_opts* o, double budget, int* cached, bool host_frames) {
    if (check_stack_opts(o)) return 0;
    StackPlan pl = plan_stack(*o, budget, host_frames);
    if (cached) *cached = pl.cached ? 1 : 0;
    return pl.bytes;
}

int fg_stack_create_budget(fg_context* ctx, const fg_stack_opts* o, double budget, fg_stack** out,
                           const fg_dev_allocator* allocator, bool host_frames) {
    if (!ctx || !o || !out) return fail(FG_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int rc = check_stack_opts(o);
    if (rc) return rc;
    rc = set_device(ctx->device);
    if (rc) return rc;
    if (budget < 0.0) {
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        float frac = o->memory_fraction >= 0.0f ? o->memory_fraction : ctx->memory_fraction;
        budget = frac > 0.0f ? (double)frac * (double)free_b : (double)free_b - 1.0e9;
    }
    const StackPlan pl = plan_stack(*o, budget, host_frames);
    fg_stack* s = new (std::nothrow) fg_stack();
    if (!s) return fail(FG_ERR_OOM_HOST, "stack allocation");
    s->ctx = ctx;
    if (allocator) s->ext = *allocator;
    s->opts = *o;
    s->W = o->width;
    s->H = o->height;
    s->sb = o->sample_type == FG_SAMPLE_U16 ? 2 : 1;
    s->lt = pl.lt;
    s->fdesc.resize(o->frame_count);
    if (o->stream) {
        s->stream = (hipStream_t)o->stream;
        s->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete s;
            return fail_hip(e, "hipStreamCreate");
        }
        s->own_stream = true;
    }
    const size_t px = pl.px, tot = pl.tot;
    s->single = pl.single;
    s->lev_stride = pl.lev_stride;
    s->contrib_floats = pl.contrib_floats;
    s->cached = pl.cached;
    s->slots = pl.slots;
    void* p;
#define ALLOC(field, bytes)                      \
    do {                                         \
        rc = dev_alloc(s, (bytes), &p);          \
        if (rc) {                                \
            free_stack(s);                       \
            return rc;                           \
        }                                        \
        s->field = (decltype(s->field))p;        \
    } while (0)
    ALLOC(d_fdesc, o->frame_count * sizeof(fg::FrameDesc));
    ALLOC(frames, s->slots * frame_bytes(s));
    ALLOC(weights, s->slots * weight_bytes(s));
    ALLOC(lev, s->slots * s->lev_stride * sizeof(float4));
    ALLOC(S, px * sizeof(float));
    ALLOC(U, px);
    ALLOC(acc, tot * 3 * sizeof(float));
    ALLOC(contrib, s->contrib_floats * sizeof(float));
    ALLOC(out_dev, px * 3 * s->sb);
    if (pl.hybrid || pl.hardmask) {
        ALLOC(best, px * sizeof(float));
        ALLOC(hard, px * 3 * s->sb);
    }
    if (pl.hybrid) {
        ALLOC(soft, px * 3 * s->sb);
        // gaussian levels 1.. of the merge; level 0 is read from the images
        const size_t up = tot > px ? tot - px : 1;
        ALLOC(merge_gs, up * 3 * sizeof(float));
        ALLOC(merge_gh, up * 3 * sizeof(float));
    }
    if (o->debug_frame >= 0) ALLOC(accDbg, tot * 3 * sizeof(float));
#undef ALLOC
    rc = fg_stack_reset(s);
    if (rc) {
        free_stack(s);
        return rc;
    }
    *out = s;
    return FG_OK;
}

extern "C" {

int fg_stack_create(fg_context* ctx, const fg_stack_opts* o, fg_stack** out) {
    return fg_stack_create_budget(ctx, o, -1.0, out);
}

void fg_stack_destroy(fg_stack* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    free_stack(s);
}

}  // extern "C"

void fg_stack_destroy_idle(fg_stack* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    free_stack(s, false);
}

int fg_stack_view(fg_stack* s, fg::StackView* v) {
    if (!s || !v) return fail(FG_ERR_INVALID_ARG, "null argument");
    if (!s->cached || s->single || s->opts.method != FG_METHOD_PYRAMID_CONTRAST)
        return fail(FG_ERR_STATE, "only cached pyramid-contrast stacks run in a batch");
    memset(v, 0, sizeof *v);
    fg::Pass1Args& a = v->p1;
    a.frames = nullptr;
    a.k0 = 0;
    a.k1 = (int)s->opts.frame_count;
    a.slot_base = 0;
    a.sample_bytes = s->sb;
    a.w = s->W;
    a.h = s->H;
    a.window = s->opts.contrast_window_size;
    a.rgba = s->frames;
    a.weights = s->weights;
    a.norm_sum = s->S;
    a.union_sup = s->U;
    v->p2 = pass2_args(s, 0, (int)s->opts.frame_count, 0);
    v->lt = &s->lt;
    v->out_dev = s->out_dev;
    v->S = s->S;
    v->U = s->U;
    v->px = (size_t)s->W * s->H;
    v->sb = s->sb;
    return FG_OK;
}

int fg_stack_describe_device_frames(fg_stack* s, uint32_t n, const fg_image_desc* d, const void* const* px,
                                    const fg_pose* poses, const fg_rect* roi, fg::FrameDesc* out) {
    if (!s || !d || !px || !out) return fail(FG_ERR_INVALID_ARG, "null argument");
    if (n != s->opts.frame_count) return fail(FG_ERR_INVALID_ARG, "frame count differs from the stack's");
    s->have_src = false;
    for (uint32_t k = 0; k < n; k++) {
        const int rc = describe_frame(s, &d[k], px[k], true, poses ? &poses[k] : nullptr, roi, &out[k]);
        if (rc) return rc;
    }
    return FG_OK;
}

float fg_context_memory_fraction(const fg_context* ctx) { return ctx ? ctx->memory_fraction : 0.0f; }
void* fg_context_stream(const fg_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" {

static int add_frame(fg_stack* s, const fg_image_desc* d, const void* px, bool dev, const fg_pose* pose,
                     const fg_rect* roi) {
    if (!s) return fail(FG_ERR_INVALID_ARG, "null stack");
    if (s->finished || s->pass2_done) return fail(FG_ERR_STATE, "pass 1 is closed");
    if (s->added >= s->opts.frame_count) return fail(FG_ERR_STATE, "more frames than frame_count");
    int rc = set_device(s->ctx->device);
    if (rc) return rc;
    const uint32_t k = s->added;
    rc = describe_frame(s, d, px, dev, pose, roi, &s->fdesc[k]);
    if (rc) return rc;
    if (!dev) s->stage_used++;
    s->added++;
    // uncached stacks keep one frame slot: run pass 1 now; single-pass stacks when their group
    // of slots is full; otherwise when the host staging is full
    const bool ring_full = !dev && s->stage_used == s->stage_slots;
    if (s->single) {
        if (s->added - s->p1_done >= (s->slots ? s->slots : kStageSlots) || ring_full) return flush_pass1(s);
        return FG_OK;
    }
    if (!s->cached || ring_full) return flush_pass1(s);
    return FG_OK;
}

int fg_stack_add_frame(fg_stack* s, const fg_image_desc* d, const void* px, const fg_pose* pose, const fg_rect* roi) {
    return add_frame(s, d, px, false, pose, roi);
}
int fg_stack_add_frame_device(fg_stack* s, const fg_image_desc* d, const void* px, const fg_pose* pose,
                              const fg_rect* roi) {
    return add_frame(s, d, px, true, pose, roi);
}

int fg_stack_needs_pass2_frames(const fg_stack* s) { return s && !s->cached && !s->single ? 1 : 0; }

static int pass2_frame(fg_stack* s, const fg_image_desc* d, const void* px, bool dev, const fg_pose* pose,
                       const fg_rect* roi) {
    if (!s) return fail(FG_ERR_INVALID_ARG, "null stack");
    if (s->single) return fail(FG_ERR_STATE, "single-pass methods take every frame once; call fg_stack_finish");
    if (s->cached) return fail(FG_ERR_STATE, "frames are cached on the device; call fg_stack_finish");
    if (s->finished) return fail(FG_ERR_STATE, "stack already finished");
    if (s->added != s->opts.frame_count) return fail(FG_ERR_STATE, "pass 1 incomplete");
    if (s->pass2_done >= s->opts.frame_count) return fail(FG_ERR_STATE, "more pass-2 frames than frame_count");
    int rc = set_device(s->ctx->device);
    if (rc) return rc;
    const uint32_t k = s->pass2_done;
    rc = describe_frame(s, d, px, dev, pose, roi, &s->fdesc[k]);
    if (rc) return rc;
    // the frame's RGBA and weight map are rebuilt as in pass 1 (src/stack/pipeline.zig:566-572),
    // leaving the final norm sum and union support untouched
    HIP_TRY(hipMemcpyAsync(s->d_fdesc + k, &s->fdesc[k], sizeof(fg::FrameDesc), hipMemcpyHostToDevice, s->stream),
            "frame descriptor");
    fg::Pass1Args a{};
    a.frames = s->d_fdesc;
    a.k0 = (int)k;
    a.k1 = (int)k + 1;
    a.slot_base = (int)k;
    a.sample_bytes = s->sb;
    a.w = s->W;
    a.h = s->H;
    a.window = s->opts.contrast_window_size;
    a.rgba = s->frames;
    a.weights = s->weights;
    a.lens = fg::frames_use_lens(&s->fdesc[k], 1) ? 1 : 0;
    {
        PhaseMark pm(s, FG_PHASE_PASS1);
        HIP_TRY(fg::launch_pass1(a, s->stream), "pass1 kernel");
    }
    rc = release_slots(s, dev ? 0 : 1);
    if (rc) return rc;
    s->stage_used = 0;
    rc = run_pass2(s, (int)k, (int)k + 1, (int)k);
    if (rc) return rc;
    s->pass2_done++;
    return FG_OK;
}

int fg_stack_pass2_frame(fg_stack* s, const fg_image_desc* d, const void* px, const fg_pose* pose, const fg_rect* roi) {
    return pass2_frame(s, d, px, false, pose, roi);
}
int fg_stack_pass2_frame_device(fg_stack* s, const fg_image_desc* d, const void* px, const fg_pose* pose,
                                const fg_rect* roi) {
    return pass2_frame(s, d, px, true, pose, roi);
}

static int finish_tail(fg_stack* s);
static int tail_graph(fg_stack* s);

static int finish(fg_stack* s) {
    if (!s) return fail(FG_ERR_INVALID_ARG, "null stack");
    if (s->finished) return FG_OK;
    if (s->added != s->opts.frame_count) return fail(FG_ERR_STATE, "not every frame went through pass 1");
    int rc = set_device(s->ctx->device);
    if (rc) return rc;
    rc = flush_pass1(s);
    if (rc) return rc;
    if (s->single) {
        // src/stack/pipeline.zig:154-155: hardmask's output is the winners; softmask divides
        PhaseMark pm(s, FG_PHASE_FINISH);
        if (s->opts.method == FG_METHOD_HARDMASK_CONTRAST)
            HIP_TRY(hipMemcpyAsync(s->out_dev, s->hard, (size_t)s->W * s->H * 3 * s->sb, hipMemcpyDeviceToDevice,
                                   s->stream),
                    "winners");
        else
            HIP_TRY(fg::launch_soft_finalize(s->S, s->acc, s->W, s->H, s->sb, s->out_dev, s->stream), "soft finalize");
        s->finished = true;
        return FG_OK;
    }
    if (s->cached) {
        // without profiling events or debug copies the schedule is kernel launches only; a stack
        // finished once (the one-shot call) never pays for a capture
        const bool reuse = s->tail_runs++ > 0;
        if (reuse && !s->profiling && !s->accDbg && !s->tail_eager) {
            rc = tail_graph(s);
            if (rc) return rc;
            s->finished = true;
            return FG_OK;
        }
        rc = run_pass2(s, 0, (int)s->opts.frame_count, 0);
        if (rc) return rc;
    } else if (s->pass2_done != s->opts.frame_count) {
        return fail(FG_ERR_STATE, "uncached stack: feed every frame through fg_stack_pass2_frame first");
    }
    return finish_tail(s);
}

// collapse, quantize and the hybrid merge
static int finish_tail(fg_stack* s) {
    PhaseMark pm(s, FG_PHASE_FINISH);
    if (s->accDbg)
        HIP_TRY(hipMemcpyAsync(s->accDbg, s->acc, s->lt.total * 3 * sizeof(float), hipMemcpyDeviceToDevice, s->stream),
                "debug copy");
    const bool hybrid = s->opts.method == FG_METHOD_HYBRID_PYRAMID_CONTRAST;
    if (!hybrid) {
        HIP_TRY(fg::launch_collapse_quantize(s->acc, &s->lt, s->sb, s->out_dev, s->stream), "collapse");
    } else {
        // src/stack/pipeline.zig:177-236: soft = the pyramid result, hard = winners, then merge
        HIP_TRY(fg::launch_collapse_quantize(s->acc, &s->lt, s->sb, s->soft, s->stream), "collapse");
        HIP_TRY(fg::launch_hybrid_merge(s->soft, s->hard, s->sb, &s->lt, 2, 4, s->opts.hybrid_sharpness, s->merge_gs,
                                        s->merge_gh, s->acc, s->out_dev, s->stream),
                "hybrid merge");
    }
    s->finished = true;
    return FG_OK;
}

// The cached stack's pass 2 and collapse as one HIP graph: captured from the stack's stream on its
// second finish, instantiated, and launched by this and every later finish.  The captured launches
// are kernels only (clears are kernels too), on buffers the stack owns for its whole life.  If the
// runtime refuses the capture the stack stays on eager launches.
static int tail_graph(fg_stack* s) {
    if (!s->tail) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        int rc = FG_OK;
        if (e == hipSuccess) {
            rc = run_pass2(s, 0, (int)s->opts.frame_count, 0);
            if (rc == FG_OK) rc = finish_tail(s);
            s->finished = false;
            e = hipStreamEndCapture(s->stream, &g);
        }
        if (rc == FG_OK && e == hipSuccess && g) e = hipGraphInstantiate(&s->tail, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (rc != FG_OK || e != hipSuccess || !s->tail) {
            (void)hipGetLastError();
            if (s->tail) (void)hipGraphExecDestroy(s->tail);
            s->tail = nullptr;
            s->tail_eager = true;
            rc = run_pass2(s, 0, (int)s->opts.frame_count, 0);
            return rc ? rc : finish_tail(s);
        }
    }
    HIP_TRY(hipGraphLaunch(s->tail, s->stream), "graph launch");
    return FG_OK;
}

int fg_stack_finish_device(fg_stack* s, void* d_out) {
    int rc = finish(s);
    if (rc) return rc;
    if (d_out)
        HIP_TRY(hipMemcpyAsync(d_out, s->out_dev, (size_t)s->W * s->H * 3 * s->sb, hipMemcpyDeviceToDevice, s->stream),
                "copy out");
    return FG_OK;
}

int fg_stack_finish(fg_stack* s, void* out_px) {
    if (!out_px) return fail(FG_ERR_INVALID_ARG, "null output");
    int rc = finish(s);
    if (rc) return rc;
    const size_t bytes = (size_t)s->W * s->H * 3 * s->sb;
    ensure_uploader(s);
    if (s->up && !host_page_locked(out_px)) {  // pageable: through the chunks, every byte there on return
        HIP_TRY(s->up->download(out_px, s->out_dev, bytes, s->stream), "copy out");
        return FG_OK;
    }
    HIP_TRY(hipMemcpyAsync(out_px, s->out_dev, bytes, hipMemcpyDeviceToHost, s->stream), "copy out");
    HIP_TRY(hipStreamSynchronize(s->stream), "finish");
    return FG_OK;
}

int fg_stack_levels(const fg_stack* s, uint32_t* levels) {
    if (!s || !levels) return fail(FG_ERR_INVALID_ARG, "null argument");
    *levels = (uint32_t)s->lt.n;
    return FG_OK;
}

int fg_stack_level_size(const fg_stack* s, uint32_t level, uint32_t* w, uint32_t* h) {
    if (!s || !w || !h) return fail(FG_ERR_INVALID_ARG, "null argument");
    if (level >= (uint32_t)s->lt.n) return fail(FG_ERR_INVALID_ARG, "level out of range");
    *w = s->lt.lv[level].w;
    *h = s->lt.lv[level].h;
    return FG_OK;
}

void* fg_stack_stream(const fg_stack* s) { return s ? (void*)s->stream : nullptr; }

int fg_stack_sync(fg_stack* s) {
    if (!s) return fail(FG_ERR_INVALID_ARG, "null stack");
    if (s->cstream) HIP_TRY(hipStreamSynchronize(s->cstream), "sync");
    HIP_TRY(hipStreamSynchronize(s->stream), "sync");
    return FG_OK;
}

int fg_stack_debug_readback(fg_stack* s, int which, uint32_t index, void* dst, size_t dst_bytes) {
    if (!s || !dst) return fail(FG_ERR_INVALID_ARG, "null argument");
    int rc = set_device(s->ctx->device);
    if (rc) return rc;
    const size_t px = (size_t)s->W * s->H;
    auto copy = [&](const void* src, size_t bytes) -> int {
        if (bytes > dst_bytes) return fail(FG_ERR_INVALID_ARG, "destination too small");
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream), "readback");
        HIP_TRY(hipStreamSynchronize(s->stream), "readback");
        return FG_OK;
    };
    // pass 1 of pending frames has to have run for the weights / sums to be meaningful
    rc = flush_pass1(s);
    if (rc) return rc;
    switch (which) {
        case FG_DEBUG_NORM_SUM:
            return copy(s->S, px * 4);
        case FG_DEBUG_UNION_SUPPORT: {
            std::vector<uint8_t> u(px);
            HIP_TRY(hipMemcpyAsync(u.data(), s->U, px, hipMemcpyDeviceToHost, s->stream), "readback");
            HIP_TRY(hipStreamSynchronize(s->stream), "readback");
            if (px * 4 > dst_bytes) return fail(FG_ERR_INVALID_ARG, "destination too small");
            float* f = (float*)dst;
            for (size_t i = 0; i < px; i++) f[i] = u[i] ? 1.0f : 0.0f;
            return FG_OK;
        }
        case FG_DEBUG_WEIGHTS:
            if (!s->cached || index >= s->added) return fail(FG_ERR_INVALID_ARG, "weights of that frame not held");
            return copy(s->weights + index * px, px * 4);
        case FG_DEBUG_ACC_LEVEL:
        case FG_DEBUG_MASK_LEVEL:
        case FG_DEBUG_GAUSS_LEVEL: {
            if (index >= (uint32_t)s->lt.n) return fail(FG_ERR_INVALID_ARG, "level out of range");
            const fg::Level& l = s->lt.lv[index];
            const size_t n = (size_t)l.w * l.h;
            if (which == FG_DEBUG_ACC_LEVEL) {
                if (!s->single && !s->finished && (s->cached || s->pass2_done == 0)) {  // nothing summed yet
                    if (n * 12 > dst_bytes) return fail(FG_ERR_INVALID_ARG, "destination too small");
                    memset(dst, 0, n * 12);
                    return FG_OK;
                }
                const float* src = s->acc;
                if (s->finished) {
                    if (!s->accDbg) return fail(FG_ERR_STATE, "accumulator is collapsed in place by finish");
                    src = s->accDbg;
                }
                float* tmp = nullptr;
                HIP_TRY(hipMalloc(&tmp, n * 12), "hipMalloc");
                hipError_t e = fg::launch_planar_to_interleaved(src + 3 * l.off, n, tmp, s->stream);
                rc = e == hipSuccess ? copy(tmp, n * 12) : fail_hip(e, "readback");
                (void)hipFree(tmp);
                return rc;
            }
            const int f = s->opts.debug_frame;
            if (f < 0) return fail(FG_ERR_STATE, "stack was created without debug_frame");
            if (!s->cached) return fail(FG_ERR_STATE, "frame levels are kept for cached stacks only");
            if (!s->finished) return fail(FG_ERR_STATE, "frame levels exist after fg_stack_finish");
            std::vector<float4> v(n);
            if (index == 0) {
                float4* tmp = nullptr;
                HIP_TRY(hipMalloc(&tmp, n * sizeof(float4)), "hipMalloc");
                fg::Pass2Args a = pass2_args(s, 0, (int)s->opts.frame_count, 0);
                hipError_t e = fg::launch_level0_encode(a, f, tmp, s->stream);
                if (e == hipSuccess) e = hipMemcpyAsync(v.data(), tmp, n * sizeof(float4), hipMemcpyDeviceToHost, s->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
                (void)hipFree(tmp);
                if (e != hipSuccess) return fail_hip(e, "readback");
            } else {
                const float4* src = s->lev + (size_t)f * s->lev_stride + (l.off - s->lt.lv[1].off);
                HIP_TRY(hipMemcpyAsync(v.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost, s->stream), "readback");
                HIP_TRY(hipStreamSynchronize(s->stream), "readback");
            }
            const float m_unsup = index == 0 ? 1.0f / (float)s->opts.frame_count : 0.0f;
            if (which == FG_DEBUG_MASK_LEVEL) {
                if (n * 4 > dst_bytes) return fail(FG_ERR_INVALID_ARG, "destination too small");
                for (size_t i = 0; i < n; i++) ((float*)dst)[i] = v[i].w < 0.0f ? m_unsup : v[i].w;
                return FG_OK;
            }
            if (n * 12 > dst_bytes) return fail(FG_ERR_INVALID_ARG, "destination too small");
            for (size_t i = 0; i < n; i++) {
                ((float*)dst)[i * 3] = v[i].x < 0.0f ? -v[i].x - 1.0f : v[i].x;
                ((float*)dst)[i * 3 + 1] = v[i].y;
                ((float*)dst)[i * 3 + 2] = v[i].z;
            }
            return FG_OK;
        }
        case FG_DEBUG_SOFT:
        case FG_DEBUG_HARD:
            if (s->opts.method != FG_METHOD_HYBRID_PYRAMID_CONTRAST) return fail(FG_ERR_STATE, "not a hybrid stack");
            return copy(which == FG_DEBUG_SOFT ? s->soft : s->hard, px * 3 * s->sb);
        default:
            return fail(FG_ERR_INVALID_ARG, "unknown debug buffer");
    }
}

int fg_stack_reset(fg_stack* s) {
    if (!s) return fail(FG_ERR_INVALID_ARG, "null stack");
    int rc = set_device(s->ctx->device);
    if (rc) return rc;
    const size_t px = (size_t)s->W * s->H;
    HIP_TRY(fg::launch_zero(s->S, px * sizeof(float), s->stream), "reset");
    HIP_TRY(fg::launch_zero(s->U, px, s->stream), "reset");
    // pass 2 starts its sums at frame 0 without reading the accumulator; softmask adds into it
    if (s->single) HIP_TRY(fg::launch_zero(s->acc, s->lt.total * 3 * sizeof(float), s->stream), "reset");
    if (s->best) HIP_TRY(fg::launch_fill_f32(s->best, px, -INFINITY, s->stream), "reset");
    s->added = 0;
    s->p1_done = 0;
    s->pass2_done = 0;
    s->stage_used = 0;
    s->finished = false;
    s->have_src = false;
    return FG_OK;
}

int fg_stack_set_profiling(fg_stack* s, int enable) {
    if (!s) return fail(FG_ERR_INVALID_ARG, "null stack");
    if (enable) {
        HIP_TRY(hipStreamSynchronize(s->stream), "sync");
        recycle_events(s);
    }
    s->profiling = enable != 0;
    return FG_OK;
}

int fg_stack_phase_times(fg_stack* s, double* ms, uint32_t* launches) {
    if (!s || !ms) return fail(FG_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream), "sync");
    for (int k = 0; k < FG_PHASE_COUNT; k++) {
        double t = 0.0;
        for (auto& pr : s->ev[k]) {
            float f = 0.0f;
            HIP_TRY(hipEventElapsedTime(&f, pr.first, pr.second), "hipEventElapsedTime");
            t += f;
        }
        ms[k] = t;
        if (launches) launches[k] = (uint32_t)s->ev[k].size();
    }
    return FG_OK;
}

int fg_fuse_stack(fg_context* ctx, uint32_t n, const fg_image_desc* descs, const void* const* pxs,
                  const fg_pose* poses, const fg_rect* roi, const fg_stack_opts* opts, void* out_px) {
    if (!ctx || !descs || !pxs || !opts || !out_px || n == 0) return fail(FG_ERR_INVALID_ARG, "null argument");
    fg_stack_opts o = *opts;
    o.frame_count = n;
    fg_stack* s = nullptr;
    int rc = fg_stack_create(ctx, &o, &s);
    if (rc) return rc;
    for (uint32_t k = 0; k < n && rc == FG_OK; k++)
        rc = fg_stack_add_frame(s, &descs[k], pxs[k], poses ? &poses[k] : nullptr, roi);
    if (rc == FG_OK && fg_stack_needs_pass2_frames(s))
        for (uint32_t k = 0; k < n && rc == FG_OK; k++)
            rc = fg_stack_pass2_frame(s, &descs[k], pxs[k], poses ? &poses[k] : nullptr, roi);
    if (rc == FG_OK) rc = fg_stack_finish(s, out_px);
    fg_stack_destroy(s);
    return rc;
}


// computeCommonOverlapRoi src/remap.zig:585-633 with transformPoint src/optimize.zig:3506-3525.
static void transform_point(const fg_pose& p, double x, double y, uint32_t w, uint32_t h, double* ox, double* oy);

int fg_common_overlap_roi(uint32_t n, const uint8_t* active, const uint32_t* widths, const uint32_t* heights,
                          const fg_pose* poses, fg_rect* out) {
    if (!widths || !heights || !poses || !out) return fail(FG_ERR_INVALID_ARG, "null argument");
    struct V {
        double x, y;
    };
    auto inside = [](V p, V a, V b) { return ((b.x - a.x) * (p.y - a.y) - (b.y - a.y) * (p.x - a.x)) >= -1e-6; };
    auto isect = [](V p1, V p2, V q1, V q2) {
        double a1 = p2.y - p1.y, b1 = p1.x - p2.x, c1 = a1 * p1.x + b1 * p1.y;
        double a2 = q2.y - q1.y, b2 = q1.x - q2.x, c2 = a2 * q1.x + b2 * q1.y;
        double det = a1 * b2 - a2 * b1;
        if (fabs(det) < 1e-9) return p2;
        return V{(b2 * c1 - b1 * c2) / det, (a1 * c2 - a2 * c1) / det};
    };
    std::vector<V> poly, next;
    bool first = true;
    for (uint32_t i = 0; i < n; i++) {
        if (active && !active[i]) continue;
        uint32_t w = widths[i], h = heights[i];
        V q[4];
        double xs[4] = {0, (double)(w - 1), (double)(w - 1), 0}, ys[4] = {0, 0, (double)(h - 1), (double)(h - 1)};
        for (int k = 0; k < 4; k++) transform_point(poses[i], xs[k], ys[k], w, h, &q[k].x, &q[k].y);
        if (first) {
            poly.assign(q, q + 4);
            first = false;
            continue;
        }
        for (int e = 0; e < 4; e++) {
            if (poly.empty()) break;
            V a = q[e], b = q[(e + 1) % 4];
            next.clear();
            V prev = poly.back();
            bool prev_in = inside(prev, a, b);
            for (const V& cur : poly) {
                bool cur_in = inside(cur, a, b);
                if (cur_in != prev_in) next.push_back(isect(prev, cur, a, b));
                if (cur_in) next.push_back(cur);
                prev = cur;
                prev_in = cur_in;
            }
            poly.swap(next);
        }
        if (poly.empty()) return 0;
    }
    if (poly.empty()) return 0;
    double mnx = poly[0].x, mxx = poly[0].x, mny = poly[0].y, mxy = poly[0].y;
    for (const V& p : poly) {
        mnx = fmin(mnx, p.x);
        mxx = fmax(mxx, p.x);
        mny = fmin(mny, p.y);
        mxy = fmax(mxy, p.y);
    }
    fg_rect r{(int32_t)ceil(mnx), (int32_t)ceil(mny), (int32_t)floor(mxx) + 1, (int32_t)floor(mxy) + 1};
    if (r.right <= r.left || r.bottom <= r.top) return 0;
    *out = r;
    return 1;
}

}  // extern "C"

// ---- forward transform (host) ----
static void persp_sphere(double* x, double* y, const double m[3][3], double d) {  // perspSpherePoint :3272-3292
    double px = *x, py = *y;
    double r = sqrt(px * px + py * py);
    double th = r / d;
    double rs = r == 0.0 ? 0.0 : sin(th) / r;
    double vx = rs * px, vy = rs * py, vz = cos(th);
    double tx = m[0][0] * vx + m[1][0] * vy + m[2][0] * vz;
    double ty = m[0][1] * vx + m[1][1] * vy + m[2][1] * vz;
    double tz = m[0][2] * vx + m[1][2] * vy + m[2][2] * vz;
    double r2 = sqrt(tx * tx + ty * ty);
    double sc = r2 == 0.0 ? 0.0 : d * atan2(r2, tz) / r2;
    *x = sc * tx;
    *y = sc * ty;
}

static void transform_point(const fg_pose& p, double x, double y, uint32_t w, uint32_t h, double* ox, double* oy) {
    double cx = ((double)w - 1.0) * 0.5, cy = ((double)h - 1.0) * 0.5;
    double pano = focal_length_px(w, p.base_hfov_degrees);
    double eff = p.base_hfov_degrees + p.hfov_delta;
    double img_f = focal_length_px(w, eff < 1e-3 ? 1e-3 : (eff > 179.0 ? 179.0 : eff));
    if (basic_rectilinear(p)) {  // basicRectilinearTransformPoint :3662-3683
        double rs = pano / img_f;
        double px = (x - cx) * rs, py = (y - cy) * rs;
        double r = sqrt(px * px + py * py) / pano;  // sphereTpRectPoint :3238-3245
        double th = r == 0.0 ? 1.0 : atan(r) / r;
        px = th * px;
        py = th * py;
        double m[3][3];
        set_matrix(p.pitch, 0.0, p.roll, m);
        persp_sphere(&px, &py, m, pano);
        double rr = sqrt(px * px + py * py);  // erectSphereTpPoint :3294-3307
        double t = rr / pano;
        double s = t == 0.0 ? 1.0 / pano : sin(t) / rr;
        double v1 = s * px, v0 = cos(t);
        double ey = pano * atan((s * py) / sqrt(v0 * v0 + v1 * v1));
        double ex = pano * atan2(v1, v0);
        double half = pano * kPi, full = 2.0 * half;  // rotateErectPoint :3347-3362
        double xx = ex + p.yaw * pano;
        if (xx < -half) {
            xx += full;
            while (xx < -half) xx += full;
        } else if (xx > half) {
            xx -= full;
            while (xx > half) xx -= full;
        }
        double phi = xx / pano;  // rectErectPoint :3326-3335
        if (phi < -kPi * 0.5 || phi > kPi * 0.5) {
            *ox = -1.0;
            *oy = -1.0;
            return;
        }
        double theta = -(ey / pano) + kPi * 0.5;
        *ox = cx + pano * tan(phi);
        *oy = cy + pano / (tan(theta) * cos(phi));
        return;
    }
    // sourceWorldRay :3781-3793 (applyRadialDistortion :4164-4172, rectilinearRay :3883-3889)
    double dx = (x - cx) - p.center_shift_x, dy = (y - cy) - p.center_shift_y;
    double hw = ((double)w - 1.0) * 0.5, hh = ((double)h - 1.0) * 0.5;
    double dn = hw * hw + hh * hh;
    double nr2 = (dx * dx + dy * dy) * (dn >= 1e-12 ? 1.0 / dn : 1.0);
    double r4 = nr2 * nr2, r6 = r4 * nr2;
    double factor = 1.0 + p.radial_a * nr2 + p.radial_b * r4 + p.radial_c * r6;
    double v[3] = {dx * factor, dy * factor, -img_f};
    double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nrm < 1e-12) {
        v[0] = 0;
        v[1] = 0;
        v[2] = 1;
    } else {
        v[0] = v[0] / nrm;
        v[1] = v[1] / nrm;
        v[2] = v[2] / nrm;
    }
    double m[3][3];
    set_matrix(-p.pitch, 0.0, p.roll, m);
    double t[3] = {m[0][0] * v[0] + m[1][0] * v[1] + m[2][0] * v[2], m[0][1] * v[0] + m[1][1] * v[1] + m[2][1] * v[2],
                   m[0][2] * v[0] + m[1][2] * v[1] + m[2][2] * v[2]};
    double c = cos(p.yaw), s = sin(p.yaw);
    double ray[3] = {c * t[0] - s * t[2], t[1], s * t[0] + c * t[2]};
    if (has_translation(p)) {  // projectViaTranslationPlane :3970-3978
        double phi = p.translation_plane_yaw / 1.0;
        double tz = kPi * 0.5 - (-p.translation_plane_pitch / 1.0);
        double nv[3] = {sin(tz) * sin(phi), cos(tz), sin(tz) * -cos(phi)};
        double cam[3] = {p.trans_x, p.trans_y, p.trans_z};
        double tgt[3] = {cam[0] + ray[0], cam[1] + ray[1], cam[2] + ray[2]};
        double dir[3] = {tgt[0] - cam[0], tgt[1] - cam[1], tgt[2] - cam[2]};
        double num = 1.0 - (nv[0] * cam[0] + nv[1] * cam[1] + nv[2] * cam[2]);
        double den = nv[0] * dir[0] + nv[1] * dir[1] + nv[2] * dir[2];
        if (fabs(den) < 1e-12) {
            ray[0] = 0.0;
            ray[1] = 0.0;
            ray[2] = 1.0;
        } else {
            double u = num / den;
            for (int k = 0; k < 3; k++) ray[k] = cam[k] + u * dir[k];
        }
    }
    double denom = -ray[2];  // projectRectilinear :4073-4080
    if (denom <= 1e-12) {
        *ox = -1.0;
        *oy = -1.0;
        return;
    }
    *ox = cx + pano * (ray[0] / denom);
    *oy = cy + pano * (ray[1] / denom);
}
