// Uploads of host-fed frames from pageable memory (the buffers a CPU decoder fills), and the copy of a
// stack's output back into pageable memory: the frame goes to
// the device through a ring of pinned chunks.  Each chunk is filled by a few threads at once (one
// CPU copy stream does not keep up with the DMA engine), and its DMA is queued as soon as it is
// filled, so the copy of one chunk overlaps the transfer of the ones before it.  upload() returns
// once the caller's buffer has been read: the last chunks may still be in flight on the stream, and
// work that needs the frame waits on that stream.  A plain hipMemcpyAsync from pageable memory goes
// through the runtime's own staging at about half the pinned rate and has to be waited for before
// the caller may reuse the buffer.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace fg {

// memcpy split over `parts` threads: the caller's thread copies part 0 and waits for the others.
// Callers on several threads may share one pool: their jobs run one after another.
class CopyPool {
  public:
    explicit CopyPool(int parts) : parts_(std::max(1, parts)) {
        for (int i = 1; i < parts_; i++) th_.emplace_back([this, i] { worker(i); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : th_) t.join();
    }
    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;

    void copy(uint8_t* d, const uint8_t* s, size_t n) {
        if (parts_ == 1 || n < ((size_t)1 << 20)) {
            memcpy(d, s, n);
            return;
        }
        std::lock_guard<std::mutex> job(call_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            d_ = d;
            s_ = s;
            n_ = n;
            remaining_ = parts_ - 1;
            gen_++;
        }
        cv_.notify_all();
        piece(0, d, s, n);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return remaining_ == 0; });
    }

  private:
    // part i of n bytes: 4 KiB-aligned ranges, the last one shorter
    void piece(int i, uint8_t* d, const uint8_t* s, size_t n) const {
        const size_t per = ((n + (size_t)parts_ - 1) / (size_t)parts_ + 4095) & ~(size_t)4095;
        const size_t a = (size_t)i * per;
        if (a >= n) return;
        memcpy(d + a, s + a, std::min(n, a + per) - a);
    }
    void worker(int i) {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            uint8_t* d = d_;
            const uint8_t* s = s_;
            const size_t n = n_;
            lk.unlock();
            piece(i, d, s, n);
            lk.lock();
            if (--remaining_ == 0) done_.notify_one();
        }
    }

    const int parts_;
    std::vector<std::thread> th_;
    std::mutex call_mu_;  // one job at a time
    std::mutex mu_;
    std::condition_variable cv_, done_;
    uint64_t gen_ = 0;  // one job per generation; the next is posted only after every part is done
    int remaining_ = 0;
    bool stop_ = false;
    uint8_t* d_ = nullptr;
    const uint8_t* s_ = nullptr;
    size_t n_ = 0;
};

class HostUploader {
  public:
    static constexpr int kChunks = 4;
    static constexpr size_t kChunkBytes = (size_t)16 << 20;

    HostUploader() = default;
    ~HostUploader() { release(); }
    HostUploader(const HostUploader&) = delete;
    HostUploader& operator=(const HostUploader&) = delete;

    // Pinned chunks and their events, filled through `pool` (not owned); false (nothing held) if
    // any is refused.
    bool init(CopyPool* pool) {
        if (!pool) return false;
        pool_ = pool;
        for (int b = 0; b < kChunks; b++) {
            if (hipHostMalloc((void**)&pin_[b], kChunkBytes, hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                release();
                return false;
            }
        }
        return true;
    }

    // n bytes from host memory px to device memory dst, queued on `st`.  Every copy of px is done on
    // return; the transfers complete in stream order on `st`.
    hipError_t upload(void* dst, const void* px, size_t n, hipStream_t st) {
        uint8_t* d = (uint8_t*)dst;
        const uint8_t* s = (const uint8_t*)px;
        for (size_t off = 0; off < n; off += kChunkBytes) {
            const int b = next_;
            next_ = (next_ + 1) % kChunks;
            if (pending_[b]) {  // the chunk's previous transfer has read it
                const hipError_t e = hipEventSynchronize(ev_[b]);
                if (e != hipSuccess) return e;
                pending_[b] = false;
            }
            const size_t len = std::min(kChunkBytes, n - off);
            pool_->copy(pin_[b], s + off, len);
            hipError_t e = hipMemcpyAsync(d + off, pin_[b], len, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipEventRecord(ev_[b], st);
            if (e != hipSuccess) return e;
            pending_[b] = true;
        }
        return hipSuccess;
    }

    // n bytes from device memory src (complete in stream order on `st`) to host memory dst, the other
    // way through the same chunks: each chunk's transfer is queued ahead, then copied out by the
    // threads once it has landed.  Every byte is in dst on return.
    hipError_t download(void* dst, const void* src, size_t n, hipStream_t st) {
        uint8_t* d = (uint8_t*)dst;
        const uint8_t* s = (const uint8_t*)src;
        const size_t nc = (n + kChunkBytes - 1) / kChunkBytes;
        auto issue = [&](size_t c) -> hipError_t {
            const int b = (int)(c % kChunks);
            if (pending_[b]) {  // an upload's transfer still reading the chunk
                const hipError_t e = hipEventSynchronize(ev_[b]);
                if (e != hipSuccess) return e;
                pending_[b] = false;
            }
            const size_t off = c * kChunkBytes, len = std::min(kChunkBytes, n - off);
            hipError_t e = hipMemcpyAsync(pin_[b], s + off, len, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipEventRecord(ev_[b], st);
            if (e == hipSuccess) pending_[b] = true;
            return e;
        };
        for (size_t c = 0; c < nc && c < (size_t)kChunks; c++) {
            const hipError_t e = issue(c);
            if (e != hipSuccess) return e;
        }
        for (size_t c = 0; c < nc; c++) {
            const int b = (int)(c % kChunks);
            hipError_t e = hipEventSynchronize(ev_[b]);
            if (e != hipSuccess) return e;
            pending_[b] = false;
            const size_t off = c * kChunkBytes;
            pool_->copy(d + off, pin_[b], std::min(kChunkBytes, n - off));
            if (c + kChunks < nc && (e = issue(c + kChunks)) != hipSuccess) return e;
        }
        return hipSuccess;
    }

  private:
    // waits for transfers still reading the chunks, then frees them
    void release() {
        for (int b = 0; b < kChunks; b++) {
            if (pending_[b]) (void)hipEventSynchronize(ev_[b]);
            pending_[b] = false;
            if (ev_[b]) (void)hipEventDestroy(ev_[b]);
            if (pin_[b]) (void)hipHostFree(pin_[b]);
            ev_[b] = nullptr;
            pin_[b] = nullptr;
        }
        pool_ = nullptr;
    }

    uint8_t* pin_[kChunks] = {};
    hipEvent_t ev_[kChunks] = {};
    bool pending_[kChunks] = {};
    int next_ = 0;
    CopyPool* pool_ = nullptr;
};

// Threads for a context's copy pool: up to 8, half the host's hardware threads.
inline int copy_pool_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::min(8u, std::max(1u, hw / 2));
}

}  // namespace fg
