// Continuous batching of whole stacks: a scheduler thread admits queued requests into slots
// (one HIP stream each) as slots and the device-memory budget free up, runs every request as the
// plain stack calls (fg_stack_add_frame_device ... fg_stack_finish_device) on its slot's stream,
// and retires it when the slot's completion event fires.  Idle stacks are pooled per slot and
// reused by requests of the same geometry (fg_stack_reset).  Requests share nothing but the
// device, so each result equals the request run alone.
//
// The scheduler sleeps on a doorbell rung by fg_server_submit and by a host function queued on a
// slot's stream behind each request's work, so completions are seen as they happen (no polling).
// The doorbell has a mutex of its own: the host function never takes the server mutex, which the
// scheduler may hold across HIP calls that wait for the device.  A request's launches are issued
// with the server mutex released, so submit and wait never queue behind them; the slots, pools,
// byte counts and arena belong to the scheduler thread alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fg_arena.h"
#include "fg_host.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    std::vector<fg_image_desc> descs;
    std::vector<const void*> frames;
    std::vector<fg_pose> poses;
    bool has_poses = false;
    fg_rect roi{};
    bool has_roi = false;
    fg_stack_opts opts{};
    void* out = nullptr;
    int state = FG_REQ_QUEUED;
    std::string error;
    Clock::time_point t_submit, t_start, t_done;
    size_t bytes = 0;
    fg_stack* stack = nullptr;
};

struct Pooled {
    fg_stack* stack;
    fg_stack_opts key;
    size_t bytes;
    uint64_t last_use;
};

bool same_geometry(const fg_stack_opts& a, const fg_stack_opts& b) {
    return a.method == b.method && a.contrast_window_size == b.contrast_window_size &&
           a.hybrid_sharpness == b.hybrid_sharpness && a.frame_count == b.frame_count && a.width == b.width &&
           a.height == b.height && a.sample_type == b.sample_type && a.max_levels == b.max_levels;
}

}  // namespace

struct fg_server {
    fg_context* ctx = nullptr;
    int device = 0;
    int nslots = 4;
    int pool_per_slot = 2;
    double budget = 0.0;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done_ev;
    std::vector<std::shared_ptr<Request>> running;     // per slot
    std::vector<std::vector<Pooled>> idle;             // per slot
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> by_id;
    double in_use = 0.0;                               // bytes of running and pooled stacks
    uint64_t next_id = 1, use_clock = 0;
    std::mutex mu;
    std::condition_variable cv_done;
    std::mutex door_mu;                                // doorbell: submissions and completions
    std::condition_variable door_cv;
    uint64_t rings = 0;
    bool host_fn = true;                               // completions ring the doorbell
    std::thread worker;
    bool stop = false;
    fg::Arena arena;
    fg_dev_allocator hook{fg::arena_alloc, fg::arena_free, nullptr};

    void run();
    void ring();
    void drop_idle();
    bool admit(std::unique_lock<std::mutex>& lk);
    void start(int slot, const std::shared_ptr<Request>& r, std::unique_lock<std::mutex>& lk);
    void retire(int slot, int state);
};

void fg_server::ring() {
    {
        std::lock_guard<std::mutex> dl(door_mu);
        rings++;
    }
    door_cv.notify_all();
}

namespace {

int server_fail(int code, const char* msg) { return fg_set_error(code, msg); }

// Runs on the HIP runtime's callback thread when a slot's stream reaches the end of a request: no
// HIP calls, only the doorbell.
void ring_cb(void* user) { static_cast<fg_server*>(user)->ring(); }

}  // namespace

// Runs a request's stack calls on its slot's stream (called with the lock held; the launches are
// issued with it released).
void fg_server::start(int slot, const std::shared_ptr<Request>& r, std::unique_lock<std::mutex>& lk) {
    r->t_start = Clock::now();
    r->state = FG_REQ_RUNNING;
    running[slot] = r;
    fg_stack* st = nullptr;
    // a pooled stack of the same geometry, else a new one
    auto& pool = idle[slot];
    for (size_t i = 0; i < pool.size(); i++) {
        if (same_geometry(pool[i].key, r->opts)) {
            st = pool[i].stack;
            r->bytes = pool[i].bytes;
            pool.erase(pool.begin() + (long)i);
            break;
        }
    }
    int rc = FG_OK;
    if (st) {
        rc = fg_stack_reset(st);
    } else {
        fg_stack_opts o = r->opts;
        o.stream = streams[slot];
        o.debug_frame = -1;
        o.memory_fraction = 0.0f;
        const double left = budget > in_use ? budget - in_use : 0.0;
        rc = fg_stack_create_budget(ctx, &o, left, &st, &hook, false);
        if (rc == FG_ERR_OOM_DEVICE) {  // a fragmented arena: give back the idle stacks and retry
            drop_idle();
            rc = fg_stack_create_budget(ctx, &o, left, &st, &hook, false);
        }
        if (rc == FG_ERR_OOM_DEVICE) {
            bool others = false;
            for (int i = 0; i < nslots; i++) others = others || (i != slot && running[i]);
            if (others) {  // wait for running stacks to return their memory
                r->state = FG_REQ_QUEUED;
                running[slot] = nullptr;
                queue.push_front(r);
                return;
            }
            // alone: a stack larger than the whole budget still runs (uncached, as the
            // reference does when its frame cache does not fit)
            arena.overflow = true;
            rc = fg_stack_create_budget(ctx, &o, left, &st, &hook, false);
            arena.overflow = false;
        }
        if (rc == FG_OK) in_use += (double)r->bytes;
    }
    r->stack = st;
    std::string err;
    lk.unlock();
    const fg_pose* pz = r->has_poses ? r->poses.data() : nullptr;
    const fg_rect* roi = r->has_roi ? &r->roi : nullptr;
    for (uint32_t k = 0; rc == FG_OK && k < r->opts.frame_count; k++)
        rc = fg_stack_add_frame_device(st, &r->descs[k], r->frames[k], pz ? &pz[k] : nullptr, roi);
    if (rc == FG_OK && fg_stack_needs_pass2_frames(st))
        for (uint32_t k = 0; rc == FG_OK && k < r->opts.frame_count; k++)
            rc = fg_stack_pass2_frame_device(st, &r->descs[k], r->frames[k], pz ? &pz[k] : nullptr, roi);
    if (rc == FG_OK) rc = fg_stack_finish_device(st, r->out);
    if (rc == FG_OK && hipEventRecord(done_ev[slot], streams[slot]) != hipSuccess) rc = FG_ERR_DEVICE;
    if (rc == FG_OK && host_fn && hipLaunchHostFunc(streams[slot], ring_cb, this) != hipSuccess) {
        (void)hipGetLastError();
        host_fn = false;  // the scheduler polls the events instead
    }
    if (rc != FG_OK) {
        err = fg_last_error();
        (void)hipStreamSynchronize(streams[slot]);
    }
    lk.lock();
    if (rc != FG_OK) {
        r->error = err;
        retire(slot, rc);
    }
}

// Destroys every pooled stack (lock held).
void fg_server::drop_idle() {
    for (auto& pool : idle) {
        for (Pooled& p : pool) {
            fg_stack_destroy_idle(p.stack);
            in_use -= (double)p.bytes;
        }
        pool.clear();
    }
}

// Completes the request of a slot and pools (or frees) its stack (lock held).
void fg_server::retire(int slot, int state) {
    std::shared_ptr<Request> r = running[slot];
    running[slot] = nullptr;
    r->t_done = Clock::now();
    r->state = state;
    if (r->stack) {
        if (state == FG_REQ_DONE && pool_per_slot > 0) {
            auto& pool = idle[slot];
            if ((int)pool.size() >= pool_per_slot) {  // drop the least recently used
                size_t lru = 0;
                for (size_t i = 1; i < pool.size(); i++)
                    if (pool[i].last_use < pool[lru].last_use) lru = i;
                fg_stack_destroy_idle(pool[lru].stack);
                in_use -= (double)pool[lru].bytes;
                pool.erase(pool.begin() + (long)lru);
            }
            pool.push_back(Pooled{r->stack, r->opts, r->bytes, ++use_clock});
        } else {
            fg_stack_destroy(r->stack);
            in_use -= (double)r->bytes;
        }
        r->stack = nullptr;
    }
    cv_done.notify_all();
}

// Admits queued requests in order while a slot is free and the budget allows (lock held).
// Returns whether anything was admitted.
bool fg_server::admit(std::unique_lock<std::mutex>& lk) {
    bool any = false;
    while (!queue.empty()) {
        int slot = -1;
        for (int i = 0; i < nslots; i++)
            if (!running[i]) {
                slot = i;
                break;
            }
        if (slot < 0) break;
        std::shared_ptr<Request> r = queue.front();
        // reuse needs no new memory; otherwise the stack's size under what is left
        bool pooled = false;
        for (const Pooled& p : idle[slot]) pooled = pooled || same_geometry(p.key, r->opts);
        if (!pooled) {
            int cached = 0;
            size_t need = fg_plan_stack_bytes(&r->opts, budget > in_use ? budget - in_use : 0.0, &cached, false);
            // free idle stacks (least recently used first) until it fits
            while (in_use + (double)need > budget) {
                int bs = -1;
                size_t bi = 0;
                for (int s = 0; s < nslots; s++)
                    for (size_t i = 0; i < idle[s].size(); i++)
                        if (bs < 0 || idle[s][i].last_use < idle[bs][bi].last_use) {
                            bs = s;
                            bi = i;
                        }
                if (bs < 0) break;
                fg_stack_destroy_idle(idle[bs][bi].stack);
                in_use -= (double)idle[bs][bi].bytes;
                idle[bs].erase(idle[bs].begin() + (long)bi);
                need = fg_plan_stack_bytes(&r->opts, budget > in_use ? budget - in_use : 0.0, &cached, false);
            }
            bool busy = false;
            for (int i = 0; i < nslots; i++) busy = busy || running[i];
            // a request that does not fit waits for running ones; alone it runs uncached
            if (in_use + (double)need > budget && busy) break;
            r->bytes = need;
        }
        queue.pop_front();
        start(slot, r, lk);
        if (r->state == FG_REQ_QUEUED) break;  // requeued: memory comes back as others finish
        any = true;
    }
    return any;
}

void fg_server::run() {
    (void)hipSetDevice(device);
    for (;;) {
        uint64_t seen;
        {  // rings after this point bring the scheduler round again
            std::lock_guard<std::mutex> dl(door_mu);
            seen = rings;
        }
        bool busy = false, done = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (int i = 0; i < nslots; i++) {
                if (!running[i]) continue;
                hipError_t e = hipEventQuery(done_ev[i]);
                if (e == hipSuccess) {
                    retire(i, FG_REQ_DONE);
                } else if (e != hipErrorNotReady) {
                    running[i]->error = hipGetErrorString(e);
                    retire(i, FG_ERR_DEVICE);
                } else {
                    busy = true;
                }
            }
            if (admit(lk)) busy = true;
            done = stop && queue.empty() && !busy;
        }
        if (done) break;
        std::unique_lock<std::mutex> dl(door_mu);
        if (busy && !host_fn)
            door_cv.wait_for(dl, std::chrono::microseconds(50), [&] { return rings != seen; });
        else
            door_cv.wait(dl, [&] { return rings != seen; });
    }
}

extern "C" {

int fg_server_create(fg_context* ctx, const fg_server_opts* opts, fg_server** out) {
    if (!ctx || !out) return server_fail(FG_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    fg_server* s = new (std::nothrow) fg_server();
    if (!s) return server_fail(FG_ERR_OOM_HOST, "server allocation");
    s->ctx = ctx;
    s->device = fg_context_device(ctx);
    if (opts && opts->slots > 0) s->nslots = opts->slots > 64 ? 64 : opts->slots;
    if (opts && opts->pool_per_slot != 0) s->pool_per_slot = opts->pool_per_slot < 0 ? 0 : opts->pool_per_slot;
    if (hipSetDevice(s->device) != hipSuccess) {
        delete s;
        return server_fail(FG_ERR_DEVICE, "hipSetDevice");
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    // 0: the context's --memory-fraction; <= 0 there too: all free HBM but 1 GB
    float frac = opts && opts->memory_fraction > 0.0f ? opts->memory_fraction : 0.0f;
    if (frac <= 0.0f) frac = fg_context_memory_fraction(ctx);
    s->budget = frac > 0.0f ? (double)frac * (double)free_b : (double)free_b - 1.0e9;
    s->arena.limit = s->budget > 0.0 ? (size_t)s->budget : 0;
    s->hook.user = &s->arena;
    s->streams.resize(s->nslots, nullptr);
    s->done_ev.resize(s->nslots, nullptr);
    s->running.resize(s->nslots);
    s->idle.resize(s->nslots);
    for (int i = 0; i < s->nslots; i++) {
        if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->done_ev[i], hipEventDisableTiming) != hipSuccess) {
            fg_server_destroy(s);
            return server_fail(FG_ERR_DEVICE, "stream or event creation");
        }
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return FG_OK;
}

void fg_server_destroy(fg_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->ring();
    if (s->worker.joinable()) s->worker.join();
    // the host functions queued behind the last requests may not have run yet: they touch `s`
    for (hipStream_t st : s->streams)
        if (st) (void)hipStreamSynchronize(st);
    for (auto& pool : s->idle)
        for (Pooled& p : pool) fg_stack_destroy(p.stack);
    s->arena.destroy();
    for (hipEvent_t e : s->done_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t st : s->streams)
        if (st) (void)hipStreamDestroy(st);
    delete s;
}

int fg_server_submit(fg_server* s, const fg_request* q, uint64_t* ticket) {
    if (!s || !q || !ticket || !q->descs || !q->frames || !q->out || q->frame_count == 0)
        return server_fail(FG_ERR_INVALID_ARG, "null argument or empty request");
    auto r = std::make_shared<Request>();
    r->descs.assign(q->descs, q->descs + q->frame_count);
    r->frames.assign(q->frames, q->frames + q->frame_count);
    if (q->poses) {
        r->poses.assign(q->poses, q->poses + q->frame_count);
        r->has_poses = true;
    }
    if (q->roi) {
        r->roi = *q->roi;
        r->has_roi = true;
    }
    const fg_image_desc& d0 = q->descs[0];
    fg_stack_opts& o = r->opts;
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
    if (fg_plan_stack_bytes(&o, 1e300, nullptr) == 0) return server_fail(FG_ERR_INVALID_ARG, "invalid request options");
    r->out = q->out;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->by_id[r->id] = r;
        s->queue.push_back(r);
    }
    s->ring();
    *ticket = r->id;
    return FG_OK;
}

int fg_server_wait(fg_server* s, uint64_t ticket, double timeout_ms, fg_request_status* st) {
    if (!s || !st) return server_fail(FG_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->by_id.find(ticket);
    if (it == s->by_id.end()) return server_fail(FG_ERR_INVALID_ARG, "unknown ticket");
    std::shared_ptr<Request> r = it->second;
    auto finished = [&] { return r->state == FG_REQ_DONE || r->state < 0; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, finished);
    else
        s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished);
    st->state = r->state;
    st->queue_ms = st->run_ms = st->total_ms = 0.0;
    if (finished()) {
        st->queue_ms = ms_between(r->t_submit, r->t_start);
        st->run_ms = ms_between(r->t_start, r->t_done);
        st->total_ms = ms_between(r->t_submit, r->t_done);
        s->by_id.erase(it);
        if (r->state < 0) return server_fail(r->state, r->error.c_str());
    }
    return FG_OK;
}

}  // extern "C"
