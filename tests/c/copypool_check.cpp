// Checks fg::CopyPool (focus-stacker_amd/csrc/fg_upload.h), the threaded copy that fills the pinned
// upload chunks of host-fed frames: every byte lands once and in place for sizes that do not divide
// by the thread count or the 4 KiB piece alignment, for copies below the threading threshold, and
// for thousands of jobs in a row through one pool (each job is posted only after the previous one
// has finished on every thread), and two callers sharing one pool.  Host code only; prints the
// first mismatches and exits 1 on any.
#include <stdio.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../focus-stacker_amd/csrc/fg_upload.h"

int main() {
    std::mt19937_64 rng(7);
    long bad = 0, jobs = 0;
    const size_t cap = (size_t)24 << 20;
    std::vector<uint8_t> src(cap + 64), dst(cap + 64);
    for (size_t i = 0; i < src.size(); i++) src[i] = (uint8_t)(rng() >> 56);
    for (int parts : {1, 2, 3, 5, 8}) {
        fg::CopyPool pool(parts);
        std::vector<size_t> sizes = {0, 1, 4095, 4096, 4097, ((size_t)1 << 20) - 1, (size_t)1 << 20,
                                     ((size_t)1 << 20) + 1, (size_t)16 << 20, cap};
        for (int i = 0; i < 40; i++) sizes.push_back((size_t)(rng() % cap));
        for (int rep = 0; rep < (parts == 8 ? 400 : 1); rep++) sizes.push_back(((size_t)1 << 20) + rng() % 4000000);
        for (size_t n : sizes) {
            const size_t so = rng() % 64, dof = rng() % 64;  // unaligned ends
            memset(dst.data(), 0xA5, dst.size());
            pool.copy(dst.data() + dof, src.data() + so, n);
            jobs++;
            const bool inside = memcmp(dst.data() + dof, src.data() + so, n) == 0;
            bool outside = true;
            for (size_t i = 0; i < dof; i++) outside = outside && dst[i] == 0xA5;
            for (size_t i = dof + n; i < dst.size(); i++) outside = outside && dst[i] == 0xA5;
            if (!inside || !outside) {
                if (bad < 10) printf("mismatch: parts=%d n=%zu src+%zu dst+%zu\n", parts, n, so, dof);
                bad++;
            }
        }
    }
    {  // two callers sharing one pool (stacks of one context): their jobs run one after another
        fg::CopyPool pool(4);
        long tbad[2] = {0, 0};
        auto caller = [&](int id) {
            std::mt19937_64 r(100 + id);
            const size_t n0 = (size_t)3 << 20;
            std::vector<uint8_t> a(n0 + 8192), b(n0 + 8192);
            for (int it = 0; it < 150; it++) {
                const size_t n = n0 + r() % 8192;
                for (size_t i = 0; i < n; i += 997) a[i] = (uint8_t)(it * 31 + id + i);
                memset(b.data(), 0, b.size());
                pool.copy(b.data(), a.data(), n);
                if (memcmp(a.data(), b.data(), n) != 0) tbad[id]++;
            }
        };
        std::thread t0(caller, 0), t1(caller, 1);
        t0.join();
        t1.join();
        jobs += 300;
        if (tbad[0] || tbad[1]) printf("mismatch: shared pool, %ld + %ld jobs\n", tbad[0], tbad[1]);
        bad += tbad[0] + tbad[1];
    }
    printf("copy pool: %ld jobs, %ld mismatches\n", jobs, bad);
    return bad ? 1 : 0;
}
