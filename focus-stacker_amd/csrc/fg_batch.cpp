// Batched fusion of many stacks (fg_batch_*): no counterpart in the reference, which fuses one stack
// per process (src/stack/pipeline.zig:10).
//
// Requests are cut, in order, into groups of pyramid-contrast stacks of one sample type and contrast
// window whose frame caches fit the memory budget together.  A group runs as one launch schedule:
//   1. one launch clears every stack's norm sum, union support and the work-queue counter;
//   2. one persistent pass-1 launch: resident blocks pull (stack, tile) items, the tiles of the
//      stacks with the most frames first, each item running every frame of its stack;
//   3. one pass-2 launch per pyramid level, over the tiles of that level of every stack that has it;
//   4. one collapse launch per level (top down) and one that adds level 0 and quantizes.
// Each stack computes exactly what it computes alone: the same kernels run on the same tiles, with
// sums in frame order whatever the launch they share.  The schedule of a group geometry (the ordered
// list of stack shapes) is kept with the group's stacks (a plan) for later groups of the same
// geometry: launched one by one on its first run, captured from the batch's stream into a HIP graph
// on its second and replayed from then on; only the frame descriptors (which frames, which poses)
// are uploaded again per run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fg_arena.h"
#include "fg_host.h"
#include "fg_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

int bfail(int code, const std::string& msg) { return fg_set_error(code, msg.c_str()); }

// The shape of one stack of a group; a plan is reused by a group with the same list.
struct Geom {
    fg_stack_opts o;
    bool operator==(const Geom& g) const {
        return o.width == g.o.width && o.height == g.o.height && o.frame_count == g.o.frame_count &&
               o.sample_type == g.o.sample_type && o.contrast_window_size == g.o.contrast_window_size &&
               o.max_levels == g.o.max_levels;
    }
};

// A device table allocated from the batch's arena.
template <typename T>
struct DevTable {
    T* p = nullptr;
    size_t n = 0;
};

struct LevelLaunch {
    bool level0;
    int tw, th;
    DevTable<fg::Pass2Job> jobs;
    DevTable<fg::BatchItem> items;
};

struct CollapseLaunch {
    bool final_step;
    DevTable<fg::CollapseJob> jobs;
    DevTable<fg::BatchItem> items;
};

struct Plan {
    std::vector<Geom> key;
    std::vector<fg_stack*> stacks;
    std::vector<fg::StackView> views;
    std::vector<uint32_t> frame_off;       // first descriptor of each stack in the frame table
    size_t bytes = 0;                      // arena bytes of the stacks
    int sb = 1;
    uint32_t window = 5;
    DevTable<fg::FrameDesc> frames;        // every stack's frame descriptors, uploaded per run
    fg::FrameDesc* staging = nullptr;      // page-locked host copy of them
    hipEvent_t staged = nullptr;           // the upload of `staging` has been read
    bool staged_pending = false;
    DevTable<fg::Pass1Args> p1;
    DevTable<fg::BatchItem> items1;
    int blocks1 = 0;
    int* queue = nullptr;
    DevTable<fg::ZeroJob> zjobs;
    DevTable<fg::BatchItem> zitems;
    std::vector<LevelLaunch> levels;
    std::vector<CollapseLaunch> collapses;
    std::vector<int> single_level;         // stacks of one level: quantized directly
    hipGraphExec_t graph = nullptr;
    bool lens = false;                     // this run's frames include lens row-walk remaps
    bool graph_lens = false;               // ... as they did when the graph was captured
    bool eager = false;                    // capture failed once: launch one by one
    int runs = 0;                          // runs so far: the schedule is captured on the second
    hipStream_t stream = nullptr;          // the stream of its runs (a plan runs on one stream only)
    uint64_t last_use = 0;
    std::vector<void*> tables;             // arena blocks of the tables
};

}  // namespace

struct fg_batch {
    fg_context* ctx = nullptr;
    int device = 0;
    std::vector<hipStream_t> streams;      // groups alternate over them: one group's pass 2 overlaps the next one's pass 1
    double budget = 0.0;
    int plan_cache = 4;
    fg::Arena arena;
    fg_dev_allocator hook{fg::arena_alloc, fg::arena_free, nullptr};
    std::vector<std::unique_ptr<Plan>> plans;
    double held = 0.0;                     // arena bytes of the cached plans
    uint64_t clock = 0;
    std::mutex mu;
};

namespace {

template <typename T>
int upload_table(fg_batch* b, Plan* p, const std::vector<T>& host, DevTable<T>* t) {
    t->n = host.size();
    if (host.empty()) return FG_OK;
    void* d = b->arena.alloc(host.size() * sizeof(T));
    if (!d) return bfail(FG_ERR_OOM_DEVICE, "batch arena: work tables");
    p->tables.push_back(d);
    if (hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(FG_ERR_DEVICE, "batch work-table upload");
    t->p = (T*)d;
    return FG_OK;
}

void free_plan(fg_batch* b, Plan* p) {
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->staged_pending) (void)hipEventSynchronize(p->staged);
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->staged) (void)hipEventDestroy(p->staged);
    if (p->staging) (void)hipHostFree(p->staging);
    for (fg_stack* s : p->stacks) fg_stack_destroy_idle(s);
    for (void* t : p->tables) b->arena.release(t);
    if (p->queue) b->arena.release(p->queue);
}

// Frees the least recently used plan (after its stream has finished with it).
void evict_lru(fg_batch* b) {
    size_t lru = 0;
    for (size_t i = 1; i < b->plans.size(); i++)
        if (b->plans[i]->last_use < b->plans[lru]->last_use) lru = i;
    b->held -= (double)b->plans[lru]->bytes;
    free_plan(b, b->plans[lru].get());
    b->plans.erase(b->plans.begin() + (long)lru);
}

// The stacks, buffers and work tables of a group geometry.
int build_plan(fg_batch* b, const std::vector<Geom>& key, Plan* p, hipStream_t stream) {
    p->key = key;
    p->stream = stream;
    p->sb = key[0].o.sample_type == FG_SAMPLE_U16 ? 2 : 1;
    p->window = key[0].o.contrast_window_size;
    const fg::Tuning& tu = fg::tuning();
    int rc = FG_OK;
    for (const Geom& g : key) {
        fg_stack_opts o = g.o;
        o.stream = stream;
        fg_stack* s = nullptr;
        rc = fg_stack_create_budget(b->ctx, &o, 1e300, &s, &b->hook, false);
        if (rc) return rc;
        p->stacks.push_back(s);
        fg::StackView v;
        rc = fg_stack_view(s, &v);
        if (rc) return rc;
        p->views.push_back(v);
        int cached = 0;
        p->bytes += fg_plan_stack_bytes(&o, 1e300, &cached, false);
    }
    const int n = (int)key.size();
    // frame table and per-stack pass-1 arguments
    uint32_t total = 0;
    for (int i = 0; i < n; i++) {
        p->frame_off.push_back(total);
        total += key[(size_t)i].o.frame_count;
    }
    void* fd = b->arena.alloc((size_t)total * sizeof(fg::FrameDesc));
    if (!fd) return bfail(FG_ERR_OOM_DEVICE, "batch arena: frame table");
    p->tables.push_back(fd);
    p->frames.p = (fg::FrameDesc*)fd;
    p->frames.n = total;
    if (hipHostMalloc((void**)&p->staging, (size_t)total * sizeof(fg::FrameDesc), hipHostMallocDefault) != hipSuccess)
        return bfail(FG_ERR_OOM_HOST, "batch frame staging");
    if (hipEventCreateWithFlags(&p->staged, hipEventDisableTiming) != hipSuccess)
        return bfail(FG_ERR_DEVICE, "hipEventCreate");
    std::vector<fg::Pass1Args> p1(n);
    for (int i = 0; i < n; i++) {
        p1[(size_t)i] = p->views[(size_t)i].p1;
        p1[(size_t)i].frames = p->frames.p + p->frame_off[(size_t)i];
    }
    if ((rc = upload_table(b, p, p1, &p->p1))) return rc;
    // longest stacks first: their items take the most time, so the queue ends with short ones
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return key[(size_t)a].o.frame_count > key[(size_t)c].o.frame_count;
    });
    std::vector<fg::BatchItem> items;
    for (int i : order) {
        uint32_t gx = 0, gy = 0;
        fg::pass1_grid(key[(size_t)i].o.width, key[(size_t)i].o.height, p->window, &gx, &gy);
        for (uint32_t y = 0; y < gy; y++)
            for (uint32_t x = 0; x < gx; x++) items.push_back(fg::BatchItem{(uint32_t)i, (uint16_t)x, (uint16_t)y});
    }
    if ((rc = upload_table(b, p, items, &p->items1))) return rc;
    const int resident = fg::pass1_resident_blocks(p->sb, fg_context_device(b->ctx));
    const int per_cu_cap = tu.batch_pass1_blocks_per_cu;
    int blocks = resident;
    if (per_cu_cap > 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, fg_context_device(b->ctx)) == hipSuccess)
            blocks = std::min(blocks, per_cu_cap * cus);
    }
    p->blocks1 = std::max(1, std::min(blocks, (int)items.size()));
    p->queue = (int*)b->arena.alloc(sizeof(int));
    if (!p->queue) return bfail(FG_ERR_OOM_DEVICE, "batch arena: queue");
    // clears: norm sum and union support of every stack, and the queue counter
    std::vector<fg::ZeroJob> zj;
    for (const fg::StackView& v : p->views) {
        zj.push_back(fg::ZeroJob{(uint32_t*)v.S, v.px});
        zj.push_back(fg::ZeroJob{(uint32_t*)v.U, (v.px + 3) / 4});
    }
    zj.push_back(fg::ZeroJob{(uint32_t*)p->queue, 1});
    std::vector<fg::BatchItem> zi;
    for (size_t j = 0; j < zj.size(); j++)
        for (size_t c = 0; c * fg::kZeroChunkWords < zj[j].words; c++)
            zi.push_back(fg::BatchItem{(uint32_t)j, (uint16_t)c, 0});
    if (zi.size() && (zi.back().bx + 1) * fg::kZeroChunkWords < zj[zi.back().job].words)
        return bfail(FG_ERR_UNSUPPORTED, "batch clear: range too long");
    if ((rc = upload_table(b, p, zj, &p->zjobs))) return rc;
    if ((rc = upload_table(b, p, zi, &p->zitems))) return rc;
    // pass 2: one launch per level, 64x16 tiles while the level is wide on average, else 16x16
    int max_levels = 0;
    for (const fg::StackView& v : p->views) max_levels = std::max(max_levels, v.lt->n);
    for (int l = 0; l < max_levels; l++) {
        LevelLaunch ll{};
        ll.level0 = l == 0;
        double wsum = 0.0;
        int cnt = 0;
        for (const fg::StackView& v : p->views)
            if (l < v.lt->n) {
                wsum += v.lt->lv[l].w;
                cnt++;
            }
        ll.tw = cnt && wsum / cnt >= (double)tu.batch_wide_level ? 64 : 16;
        ll.th = 16;
        std::vector<fg::Pass2Job> jobs;
        std::vector<fg::BatchItem> its;
        for (int i : order) {
            const fg::StackView& v = p->views[(size_t)i];
            if (l >= v.lt->n) continue;
            const uint32_t j = (uint32_t)jobs.size();
            jobs.push_back(fg::Pass2Job{v.p2, fg::level_args(*v.lt, l)});
            int gx = 0, gy = 0;
            fg::pass2_tiles((int)v.lt->lv[l].w, (int)v.lt->lv[l].h, ll.tw, ll.th, &gx, &gy);
            for (int y = 0; y < gy; y++)
                for (int x = 0; x < gx; x++) its.push_back(fg::BatchItem{j, (uint16_t)x, (uint16_t)y});
        }
        if ((rc = upload_table(b, p, jobs, &ll.jobs))) return rc;
        if ((rc = upload_table(b, p, its, &ll.items))) return rc;
        p->levels.push_back(ll);
    }
    // collapse: level l-1 += EXPAND(level l) from the top down; the step into level 0 quantizes
    for (int l = max_levels - 1; l >= 1; l--) {
        CollapseLaunch cl{};
        cl.final_step = l == 1;
        std::vector<fg::CollapseJob> jobs;
        std::vector<fg::BatchItem> its;
        for (int i = 0; i < n; i++) {
            const fg::StackView& v = p->views[(size_t)i];
            if (l >= v.lt->n) continue;
            const fg::Level& pl = v.lt->lv[l - 1];
            const fg::Level& ch = v.lt->lv[l];
            const uint32_t j = (uint32_t)jobs.size();
            jobs.push_back(fg::CollapseJob{v.p2.acc + 3 * pl.off, v.p2.acc + 3 * ch.off, (int)pl.w, (int)pl.h,
                                           (int)ch.w, (int)ch.h, cl.final_step ? v.out_dev : nullptr});
            int gx = 0, gy = 0;
            fg::collapse_tiles((int)pl.w, (int)pl.h, &gx, &gy);
            for (int y = 0; y < gy; y++)
                for (int x = 0; x < gx; x++) its.push_back(fg::BatchItem{j, (uint16_t)x, (uint16_t)y});
        }
        if ((rc = upload_table(b, p, jobs, &cl.jobs))) return rc;
        if ((rc = upload_table(b, p, its, &cl.items))) return rc;
        p->collapses.push_back(cl);
    }
    for (int i = 0; i < n; i++)
        if (p->views[(size_t)i].lt->n == 1) p->single_level.push_back(i);
    return FG_OK;
}

// The schedule of a plan, enqueued on its stream (eagerly, or under capture).
int enqueue_schedule(Plan* p) {
    hipStream_t s = p->stream;
    if (fg::launch_zero_batch(p->zjobs.p, p->zitems.p, (int)p->zitems.n, s) != hipSuccess)
        return bfail(FG_ERR_DEVICE, "batch clear launch");
    if (fg::launch_pass1_batch(p->sb, p->window, p->lens, p->p1.p, p->items1.p, (int)p->items1.n, p->queue,
                               p->blocks1, s) != hipSuccess)
        return bfail(FG_ERR_DEVICE, "batch pass-1 launch");
    for (const LevelLaunch& ll : p->levels)
        if (fg::launch_pass2_batch(p->sb, ll.level0, ll.tw, ll.th, ll.jobs.p, ll.items.p, (int)ll.items.n, s) !=
            hipSuccess)
            return bfail(FG_ERR_DEVICE, "batch pass-2 launch");
    for (const CollapseLaunch& cl : p->collapses)
        if (fg::launch_collapse_batch(cl.jobs.p, cl.items.p, (int)cl.items.n, cl.final_step, p->sb, s) != hipSuccess)
            return bfail(FG_ERR_DEVICE, "batch collapse launch");
    for (int i : p->single_level) {
        const fg::StackView& v = p->views[(size_t)i];
        if (fg::launch_quantize(v.p2.acc, v.lt->lv[0].w, v.lt->lv[0].h, p->sb, v.out_dev, s) != hipSuccess)
            return bfail(FG_ERR_DEVICE, "batch quantize launch");
    }
    return FG_OK;
}

// Runs a plan on the group's requests: frame descriptors up, the schedule (launched one by one on the
// plan's first run, captured on its second, its graph replayed after), the outputs copied to the
// requests' buffers.
int run_plan(Plan* p, const std::vector<const fg_request*>& reqs) {
    hipStream_t stream = p->stream;
    if (p->staged_pending) {  // the previous run's upload has to have read the staging
        if (hipEventSynchronize(p->staged) != hipSuccess) return bfail(FG_ERR_DEVICE, "batch staging");
        p->staged_pending = false;
    }
    for (size_t i = 0; i < reqs.size(); i++) {
        const fg_request* q = reqs[i];
        const int rc = fg_stack_describe_device_frames(p->stacks[i], q->frame_count, q->descs, q->frames, q->poses,
                                                       q->roi, p->staging + p->frame_off[i]);
        if (rc) return rc;
    }
    if (hipMemcpyAsync(p->frames.p, p->staging, p->frames.n * sizeof(fg::FrameDesc), hipMemcpyHostToDevice,
                       stream) != hipSuccess ||
        hipEventRecord(p->staged, stream) != hipSuccess)
        return bfail(FG_ERR_DEVICE, "batch frame upload");
    p->staged_pending = true;
    // the pass-1 kernel carries the lens-path remap only for runs that have such frames: a graph
    // captured with the other kernel is captured again
    p->lens = fg::frames_use_lens(p->staging, p->frames.n);
    if (p->graph && p->graph_lens != p->lens) {
        (void)hipGraphExecDestroy(p->graph);
        p->graph = nullptr;
    }
    // a geometry's first run launches its schedule one by one (a batch that never recurs pays for
    // no capture); the second captures it, and it is replayed from then on
    if (!p->graph && !p->eager && p->runs > 0) {
        hipGraph_t g = nullptr;
        int rc = FG_OK;
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = enqueue_schedule(p);
            e = hipStreamEndCapture(stream, &g);
        }
        if (rc == FG_OK && e == hipSuccess && g) e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
        p->graph_lens = p->lens;
        if (g) (void)hipGraphDestroy(g);
        if (rc != FG_OK || e != hipSuccess || !p->graph) {
            (void)hipGetLastError();
            if (p->graph) (void)hipGraphExecDestroy(p->graph);
            p->graph = nullptr;
            p->eager = true;
        }
    }
    if (p->graph) {
        if (hipGraphLaunch(p->graph, stream) != hipSuccess) return bfail(FG_ERR_DEVICE, "batch graph launch");
    } else {
        const int rc = enqueue_schedule(p);
        if (rc) return rc;
    }
    p->runs++;
    for (size_t i = 0; i < reqs.size(); i++) {
        const fg::StackView& v = p->views[i];
        const size_t bytes = v.px * 3 * (size_t)v.sb;
        if (hipMemcpyAsync(reqs[i]->out, v.out_dev, bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
            return bfail(FG_ERR_DEVICE, "batch output copy");
    }
    return FG_OK;
}

// The stack options of a request (as fg_server_submit forms them).
fg_stack_opts request_opts(const fg_request* q) {
    fg_stack_opts o{};
    const fg_image_desc& d0 = q->descs[0];
    o.method = q->method;
    o.contrast_window_size = q->contrast_window_size ? q->contrast_window_size : 5;
    o.hybrid_sharpness = q->hybrid_sharpness;
    o.frame_count = q->frame_count;
    o.width = q->roi && q->poses ? (uint32_t)(q->roi->right - q->roi->left) : d0.width;
    o.height = q->roi && q->poses ? (uint32_t)(q->roi->bottom - q->roi->top) : d0.height;
    o.sample_type = d0.sample_type;
    o.max_levels = q->max_levels;
    o.memory_fraction = 0.0f;
    o.debug_frame = -1;
    return o;
}

// The frame checks of fg_stack_add_frame_device (src/fuse/io.zig:20-45: every frame of the first
// frame's size and format), made before a request joins a group so that a bad one fails alone.
int validate_request(const fg_request* q, const fg_stack_opts& o) {
    const fg_image_desc& d0 = q->descs[0];
    const uint32_t sb = d0.sample_type == FG_SAMPLE_U16 ? 2 : 1;
    for (uint32_t k = 0; k < q->frame_count; k++) {
        const fg_image_desc& d = q->descs[k];
        if (d.width == 0 || d.height == 0) return bfail(FG_ERR_INVALID_ARG, "empty image");
        if (d.color_channels != 3) return bfail(FG_ERR_UNSUPPORTED, "pyramid fusion needs rgb frames");
        if (d.extra_channels > 1) return bfail(FG_ERR_UNSUPPORTED, "extra_channels must be 0 or 1");
        if (d.sample_type != d0.sample_type) return bfail(FG_ERR_FORMAT_MISMATCH, "frame sample types differ");
        if (frame_too_large(d)) return bfail(FG_ERR_UNSUPPORTED, "frames must hold fewer than 2^32 - 16 bytes");
        if (d.width != d0.width || d.height != d0.height)
            return bfail(FG_ERR_SIZE_MISMATCH, "frame size differs from the first frame's");
        if (d.extra_channels != d0.extra_channels) return bfail(FG_ERR_FORMAT_MISMATCH, "frame channels differ");
        if (!q->frames[k]) return bfail(FG_ERR_INVALID_ARG, "null frame pointer");
        if ((uintptr_t)q->frames[k] % sb != 0)
            return bfail(FG_ERR_INVALID_ARG, "device frame pointer is not aligned to its sample size");
        if (q->poses) {
            if (d.extra_channels != 0) return bfail(FG_ERR_UNSUPPORTED, "remap sources carry no alpha channel");
        } else if (d.width != o.width || d.height != o.height) {
            return bfail(FG_ERR_SIZE_MISMATCH, "aligned frame size differs from the stack size");
        }
    }
    if (q->poses && q->roi && (q->roi->right <= q->roi->left || q->roi->bottom <= q->roi->top))
        return bfail(FG_ERR_SIZE_MISMATCH, "empty roi");
    return FG_OK;
}

// A request that runs as a stack of its own (too large to gain from sharing launches, another method,
// or too large to cache): the plain stack calls, enqueued on `stream` without waiting.
int start_single(fg_batch* b, const fg_request* q, const fg_stack_opts& o0, hipStream_t stream, double budget,
                 fg_stack** out) {
    fg_stack_opts o = o0;
    o.stream = stream;
    fg_stack* s = nullptr;
    int rc = fg_stack_create_budget(b->ctx, &o, budget, &s, &b->hook, false);
    if (rc) return rc;
    *out = s;
    for (uint32_t k = 0; rc == FG_OK && k < q->frame_count; k++)
        rc = fg_stack_add_frame_device(s, &q->descs[k], q->frames[k], q->poses ? &q->poses[k] : nullptr, q->roi);
    if (rc == FG_OK && fg_stack_needs_pass2_frames(s))
        for (uint32_t k = 0; rc == FG_OK && k < q->frame_count; k++)
            rc = fg_stack_pass2_frame_device(s, &q->descs[k], q->frames[k], q->poses ? &q->poses[k] : nullptr, q->roi);
    if (rc == FG_OK) rc = fg_stack_finish_device(s, q->out);
    return rc;
}

}  // namespace

extern "C" {

int fg_batch_create(fg_context* ctx, const fg_batch_opts* opts, fg_batch** out) {
    if (!ctx || !out) return bfail(FG_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    fg_batch* b = new (std::nothrow) fg_batch();
    if (!b) return bfail(FG_ERR_OOM_HOST, "batch allocation");
    b->ctx = ctx;
    b->device = fg_context_device(ctx);
    if (hipSetDevice(b->device) != hipSuccess) {
        delete b;
        return bfail(FG_ERR_DEVICE, "hipSetDevice");
    }
    if (opts && opts->plan_cache != 0) b->plan_cache = opts->plan_cache < 0 ? 0 : opts->plan_cache;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    float frac = opts && opts->memory_fraction > 0.0f ? opts->memory_fraction : 0.0f;
    if (frac <= 0.0f) frac = fg_context_memory_fraction(ctx);
    b->budget = frac > 0.0f ? (double)frac * (double)free_b : (double)free_b - 1.0e9;
    b->arena.limit = b->budget > 0.0 ? (size_t)b->budget : 0;
    b->hook.user = &b->arena;
    const int ns = std::max(1, std::min(8, fg::tuning().batch_streams));
    for (int i = 0; i < ns; i++) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : b->streams) (void)hipStreamDestroy(x);
            delete b;
            return bfail(FG_ERR_DEVICE, "hipStreamCreate");
        }
        b->streams.push_back(st);
    }
    *out = b;
    return FG_OK;
}

void fg_batch_destroy(fg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (hipStream_t st : b->streams) (void)hipStreamSynchronize(st);
    for (auto& p : b->plans) free_plan(b, p.get());
    b->plans.clear();
    b->arena.destroy();
    for (hipStream_t st : b->streams) (void)hipStreamDestroy(st);
    delete b;
}

int fg_batch_fuse(fg_batch* b, uint32_t count, const fg_request* reqs, fg_request_status* status) {
    if (!b || (!reqs && count)) return bfail(FG_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    if (hipSetDevice(b->device) != hipSuccess) return bfail(FG_ERR_DEVICE, "hipSetDevice");
    const auto t0 = Clock::now();
    hipStream_t gstream = b->streams[0];  // the groups' stream; stacks of their own use the others
    hipEvent_t ev0 = nullptr;
    if (hipEventCreate(&ev0) != hipSuccess || hipEventRecord(ev0, gstream) != hipSuccess)
        return bfail(FG_ERR_DEVICE, "batch events");
    // one entry per unit of work: a group, or a stack of its own
    struct Track {
        std::vector<uint32_t> idx;
        hipEvent_t start = nullptr, done = nullptr;
        double issued_ms = 0.0;
    };
    std::vector<Track> tracks;
    auto open_track = [&](hipStream_t st) -> Track& {
        tracks.emplace_back();
        Track& t = tracks.back();
        (void)hipEventCreate(&t.start);
        (void)hipEventCreate(&t.done);
        (void)hipEventRecord(t.start, st);
        t.issued_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        return t;
    };
    std::vector<int> state(count, FG_REQ_QUEUED);
    int first_error = FG_OK;
    std::string first_msg;
    auto note_error = [&](uint32_t i, int rc) {
        state[i] = rc;
        if (first_error == FG_OK) {
            first_error = rc;
            first_msg = fg_last_error();
        }
    };
    // stacks of their own in flight, oldest first; their buffers go back once their work is done
    struct Single {
        fg_stack* s;
        hipEvent_t done;
        double bytes;
    };
    std::deque<Single> inflight;
    double inflight_bytes = 0.0;
    auto retire_oldest = [&]() {
        Single& x = inflight.front();
        (void)hipEventSynchronize(x.done);
        fg_stack_destroy_idle(x.s);
        (void)hipEventDestroy(x.done);
        inflight_bytes -= x.bytes;
        inflight.pop_front();
    };
    // room for `need` more bytes: cached plans first (least recently used), then finished singles
    auto make_room = [&](double need) {
        while (b->held + inflight_bytes + need > b->budget) {
            if (!b->plans.empty())
                evict_lru(b);
            else if (!inflight.empty())
                retire_oldest();
            else
                break;
        }
    };
    const size_t nsingle = b->streams.size() > 1 ? b->streams.size() - 1 : 1;
    const double large_px = (double)fg::tuning().batch_large_kpix * 1000.0;
    const double group_cap = fg::tuning().batch_group_mib > 0
                                 ? std::min(b->budget, (double)fg::tuning().batch_group_mib * 1048576.0)
                                 : b->budget;

    std::vector<fg_stack_opts> opts(count);
    std::vector<Geom> cur_key;
    std::vector<uint32_t> cur;
    double cur_bytes = 0.0;
    auto flush = [&]() -> void {
        if (cur.empty()) return;
        Plan* p = nullptr;
        for (auto& q : b->plans)
            if (q->key == cur_key) p = q.get();
        int rc = FG_OK;
        if (!p) {
            while (!b->plans.empty() && (int)b->plans.size() >= std::max(1, b->plan_cache)) evict_lru(b);
            make_room(cur_bytes);
            auto np = std::make_unique<Plan>();
            rc = build_plan(b, cur_key, np.get(), gstream);
            if (rc == FG_ERR_OOM_DEVICE) {
                // the arena is full or fragmented: everything else gives its memory back, and a group of
                // one stack may then grow the arena past the budget, as a request served alone does
                free_plan(b, np.get());
                while (!b->plans.empty()) evict_lru(b);
                while (!inflight.empty()) retire_oldest();
                np = std::make_unique<Plan>();
                b->arena.overflow = cur.size() == 1;
                rc = build_plan(b, cur_key, np.get(), gstream);
                b->arena.overflow = false;
            }
            if (rc == FG_OK) {
                b->held += (double)np->bytes;
                p = np.get();
                b->plans.push_back(std::move(np));
            } else {
                const std::string msg = fg_last_error();
                free_plan(b, np.get());
                fg_set_error(rc, msg.c_str());
            }
        }
        std::vector<const fg_request*> rq;
        for (uint32_t i : cur) rq.push_back(&reqs[i]);
        Track& tr = open_track(gstream);
        tr.idx = cur;
        if (rc == FG_OK) {
            p->last_use = ++b->clock;
            rc = run_plan(p, rq);
        }
        (void)hipEventRecord(tr.done, gstream);
        for (uint32_t i : cur) {
            if (rc == FG_OK)
                state[i] = FG_REQ_DONE;
            else
                note_error(i, rc);
        }
        if (b->plan_cache == 0 && p) {  // nothing kept between groups
            for (size_t k = 0; k < b->plans.size(); k++)
                if (b->plans[k].get() == p) {
                    b->held -= (double)p->bytes;
                    free_plan(b, p);
                    b->plans.erase(b->plans.begin() + (long)k);
                    break;
                }
        }
        cur.clear();
        cur_key.clear();
        cur_bytes = 0.0;
    };
    // classify: stacks of their own, and the batchable ones (grouped in request order)
    struct SingleReq {
        uint32_t i;
        double cost;  // frame-pixels: both passes scale with them
    };
    std::vector<SingleReq> singles;
    std::vector<uint32_t> batched;
    std::vector<size_t> needs(count, 0);
    for (uint32_t i = 0; i < count; i++) {
        const fg_request* q = &reqs[i];
        if (!q->descs || !q->frames || !q->out || q->frame_count == 0) {
            fg_set_error(FG_ERR_INVALID_ARG, "null argument or empty request");
            note_error(i, FG_ERR_INVALID_ARG);
            continue;
        }
        opts[i] = request_opts(q);
        int cached = 0;
        const size_t need = fg_plan_stack_bytes(&opts[i], 1e300, &cached, false);
        if (need == 0) {
            fg_set_error(FG_ERR_INVALID_ARG, "invalid request options");
            note_error(i, FG_ERR_INVALID_ARG);
            continue;
        }
        const int vrc = validate_request(q, opts[i]);
        if (vrc) {
            note_error(i, vrc);
            continue;
        }
        needs[i] = need;
        const double roi_px = (double)opts[i].width * (double)opts[i].height;
        const bool batchable = opts[i].method == FG_METHOD_PYRAMID_CONTRAST && (double)need <= group_cap &&
                               !(large_px > 0.0 && roi_px >= large_px);
        if (batchable)
            batched.push_back(i);
        else
            singles.push_back(SingleReq{i, roi_px * (double)q->frame_count});
    }
    // stacks of their own first, longest first, each on the side stream with the least work queued
    // (longest-processing-time order: the batch ends with short stacks, not with one long one alone)
    std::stable_sort(singles.begin(), singles.end(),
                     [](const SingleReq& x, const SingleReq& y) { return x.cost > y.cost; });
    std::vector<double> load(nsingle, 0.0);
    for (const SingleReq& sr : singles) {
        const uint32_t i = sr.i;
        const fg_request* q = &reqs[i];
        size_t j = 0;
        for (size_t k = 1; k < nsingle; k++)
            if (load[k] < load[j]) j = k;
        load[j] += sr.cost;
        hipStream_t st = b->streams.size() > 1 ? b->streams[1 + j] : gstream;
        make_room(std::min((double)needs[i], b->budget));
        const double left = b->budget - b->held - inflight_bytes;
        int c2 = 0;
        const double sneed = (double)fg_plan_stack_bytes(&opts[i], left, &c2, false);
        Track& tr = open_track(st);
        tr.idx = {i};
        fg_stack* s = nullptr;
        int rc = start_single(b, q, opts[i], st, left, &s);
        if (rc == FG_ERR_OOM_DEVICE && !s) {  // alone on the device, as the reference runs it: uncached
            while (!b->plans.empty()) evict_lru(b);
            while (!inflight.empty()) retire_oldest();
            b->arena.overflow = true;
            rc = start_single(b, q, opts[i], st, b->budget, &s);
            b->arena.overflow = false;
        }
        (void)hipEventRecord(tr.done, st);
        if (s) {
            hipEvent_t done = nullptr;
            (void)hipEventCreateWithFlags(&done, hipEventDisableTiming);
            (void)hipEventRecord(done, st);
            inflight.push_back(Single{s, done, sneed});
            inflight_bytes += sneed;
        }
        if (rc == FG_OK)
            state[i] = FG_REQ_DONE;
        else
            note_error(i, rc);
    }
    // the groups, in request order, on the group stream beside them
    for (uint32_t i : batched) {
        const bool fits = !cur.empty() && opts[i].sample_type == cur_key[0].o.sample_type &&
                          opts[i].contrast_window_size == cur_key[0].o.contrast_window_size &&
                          cur_bytes + (double)needs[i] <= group_cap;
        if (!cur.empty() && !fits) flush();
        cur.push_back(i);
        cur_key.push_back(Geom{opts[i]});
        cur_bytes += (double)needs[i];
    }
    flush();
    hipError_t e = hipSuccess;
    for (hipStream_t st : b->streams) {
        const hipError_t e2 = hipStreamSynchronize(st);
        if (e == hipSuccess) e = e2;
    }
    while (!inflight.empty()) retire_oldest();
    const double host_total = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    if (status) {
        for (uint32_t i = 0; i < count; i++) status[i] = fg_request_status{state[i], 0.0, 0.0, 0.0};
        for (const Track& t : tracks) {
            float t_start = 0.0f, t_done = 0.0f;
            if (t.start && t.done && hipEventElapsedTime(&t_start, ev0, t.start) == hipSuccess &&
                hipEventElapsedTime(&t_done, ev0, t.done) == hipSuccess) {
                for (uint32_t i : t.idx) {
                    status[i].queue_ms = std::max((double)t_start, t.issued_ms);
                    status[i].run_ms = (double)(t_done - t_start);
                    status[i].total_ms = std::min((double)t_done, host_total);
                }
            }
        }
    }
    for (Track& t : tracks) {
        if (t.start) (void)hipEventDestroy(t.start);
        if (t.done) (void)hipEventDestroy(t.done);
    }
    (void)hipEventDestroy(ev0);
    if (e != hipSuccess) return bfail(FG_ERR_DEVICE, hipGetErrorString(e));
    if (first_error != FG_OK) return bfail(first_error, first_msg);
    return FG_OK;
}

int fg_get_tuning(fg_tuning* t) {
    if (!t) return bfail(FG_ERR_INVALID_ARG, "null argument");
    const fg::Tuning& c = fg::tuning();
    t->pass2_min_tiles = c.pass2_min_tiles;
    t->batch_wide_level = c.batch_wide_level;
    t->batch_pass1_blocks_per_cu = c.batch_pass1_blocks_per_cu;
    t->batch_group_mib = c.batch_group_mib;
    t->batch_streams = c.batch_streams;
    t->batch_large_kpix = c.batch_large_kpix;
    return FG_OK;
}

int fg_set_tuning(const fg_tuning* t) {
    if (!t) return bfail(FG_ERR_INVALID_ARG, "null argument");
    if (t->pass2_min_tiles < 1 || t->batch_wide_level < 1 || t->batch_pass1_blocks_per_cu < 0 || t->batch_group_mib < 0 ||
        t->batch_streams < 1 || t->batch_streams > 8 || t->batch_large_kpix < 0)
        return bfail(FG_ERR_INVALID_ARG, "tuning values out of range");
    fg::Tuning& c = fg::tuning();
    c.pass2_min_tiles = t->pass2_min_tiles;
    c.batch_wide_level = t->batch_wide_level;
    c.batch_pass1_blocks_per_cu = t->batch_pass1_blocks_per_cu;
    c.batch_group_mib = t->batch_group_mib;
    c.batch_streams = t->batch_streams;
    c.batch_large_kpix = t->batch_large_kpix;
    return FG_OK;
}

}  // extern "C"
