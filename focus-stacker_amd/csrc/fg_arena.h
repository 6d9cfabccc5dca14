// Device arena shared by the stack server (fg_server.cpp) and the batch engine (fg_batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <unordered_map>
#include <vector>

namespace fg {

// Device memory of a server or a batch: chunks from hipMalloc, added on demand up to the budget and
// kept until the owner ends (or, when the budget is reached and a request fits nowhere, returned once
// wholly free), so steady-state work never calls hipMalloc or hipFree (hipFree waits for the whole
// device); stacks take 256-byte aligned ranges, first fit, with coalescing frees.
struct Arena {
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::vector<Chunk> chunks;
    std::unordered_map<void*, std::pair<size_t, size_t>> live;  // pointer -> (chunk, length)
    size_t total = 0, limit = 0;
    bool overflow = false;  // a request running alone may grow the arena past the limit
    static constexpr size_t kMinChunk = (size_t)8 << 30;

    void* take(size_t c, size_t bytes) {
        auto& fl = chunks[c].free;
        for (auto it = fl.begin(); it != fl.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, rem = it->second - bytes;
            fl.erase(it);
            if (rem) fl[off + bytes] = rem;
            void* p = chunks[c].base + off;
            live[p] = {c, bytes};
            return p;
        }
        return nullptr;
    }
    // Returns chunks that hold nothing to the device (hipFree waits for the device, so this runs only
    // when an allocation can neither fit nor grow the arena).
    bool trim() {
        bool any = false;
        for (Chunk& c : chunks) {
            if (!c.base || c.free.size() != 1 || c.free.begin()->first != 0 || c.free.begin()->second != c.size)
                continue;
            (void)hipFree(c.base);
            total -= c.size;
            c.base = nullptr;
            c.size = 0;
            c.free.clear();
            any = true;
        }
        return any;
    }
    void* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (size_t c = 0; c < chunks.size(); c++)
            if (chunks[c].base)
                if (void* p = take(c, bytes)) return p;
        size_t sz = std::max(bytes, kMinChunk);
        if (total + sz > limit) sz = bytes;
        if (total + sz > limit && !overflow && trim()) {
            sz = std::max(bytes, kMinChunk);
            if (total + sz > limit) sz = bytes;
        }
        if (total + sz > limit && !overflow) return nullptr;
        void* base = nullptr;
        if (hipMalloc(&base, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        chunks.push_back(Chunk{(char*)base, sz, {}});
        chunks.back().free[0] = sz;
        total += sz;
        return take(chunks.size() - 1, bytes);
    }
    void release(void* p) {
        auto it = live.find(p);
        if (it == live.end()) return;
        const size_t c = it->second.first, len = it->second.second;
        live.erase(it);
        auto& fl = chunks[c].free;
        size_t off = (size_t)((char*)p - chunks[c].base), n = len;
        auto nx = fl.lower_bound(off);
        if (nx != fl.end() && off + n == nx->first) {  // merge with the following range
            n += nx->second;
            nx = fl.erase(nx);
        }
        if (nx != fl.begin()) {  // and with the preceding one
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                pv->second += n;
                return;
            }
        }
        fl[off] = n;
    }
    void destroy() {
        for (Chunk& c : chunks)
            if (c.base) (void)hipFree(c.base);
        chunks.clear();
        live.clear();
        total = 0;
    }
};

inline void* arena_alloc(void* u, size_t bytes) { return ((Arena*)u)->alloc(bytes); }
inline void arena_free(void* u, void* p) { ((Arena*)u)->release(p); }

}  // namespace fg
