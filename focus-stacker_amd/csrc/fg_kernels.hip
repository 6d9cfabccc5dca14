// HIP kernels (gfx950, wave64) of the remap + pyramid-contrast path.
//
// Arithmetic follows the reference operation by operation so that, given the same
// inputs, every f32 stage rounds the same way: the library is compiled with
// -ffp-contract=off (the reference's Zig float mode is strict, no FMA), sums keep the
// reference's association order, and divisions/sqrt are IEEE correctly rounded.
// Where the reference's own order is not a function of the data (the contrast sliding
// sums depend on its thread split) the kernel uses a fixed direct order in the same
// precision (f64).
#include "fg_device.h"

#include <math.h>

namespace fg {

template <typename T, int CH>
__global__ void __launch_bounds__(256) k_remap(const T* __restrict__ src, int sw, int sh, RemapParams p, int ow,
                                               int oh, T* __restrict__ dst) {
    int x = blockIdx.x * 64 + (threadIdx.x & 63);
    int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= ow || y >= oh) return;
    uint32_t o[CH + 1];
    remap_sample<T, CH>(p, src, sw, sh, x, y, o);
    T* d = dst + ((size_t)y * ow + x) * (CH + 1);
#pragma unroll
    for (int c = 0; c <= CH; c++) d[c] = (T)o[c];
}

hipError_t launch_remap(const void* src, int sb, int ch, uint32_t sw, uint32_t sh, const RemapParams& p, uint32_t ow,
                        uint32_t oh, void* dst, hipStream_t s) {
    dim3 grid((ow + 63) / 64, (oh + 3) / 4);
    if (sb == 1 && ch == 3)
        hipLaunchKernelGGL((k_remap<uint8_t, 3>), grid, dim3(256), 0, s, (const uint8_t*)src, (int)sw, (int)sh, p,
                           (int)ow, (int)oh, (uint8_t*)dst);
    else if (sb == 2 && ch == 3)
        hipLaunchKernelGGL((k_remap<uint16_t, 3>), grid, dim3(256), 0, s, (const uint16_t*)src, (int)sw, (int)sh, p,
                           (int)ow, (int)oh, (uint16_t*)dst);
    else if (sb == 1 && ch == 1)
        hipLaunchKernelGGL((k_remap<uint8_t, 1>), grid, dim3(256), 0, s, (const uint8_t*)src, (int)sw, (int)sh, p,
                           (int)ow, (int)oh, (uint8_t*)dst);
    else if (sb == 2 && ch == 1)
        hipLaunchKernelGGL((k_remap<uint16_t, 1>), grid, dim3(256), 0, s, (const uint16_t*)src, (int)sw, (int)sh, p,
                           (int)ow, (int)oh, (uint16_t*)dst);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// collapse (collapseToImageWithJobsAndDebug src/fuse/pyramid.zig:223-269)
// ---------------------------------------------------------------------------

// A 64 x 4 tile of a level expands from a 34 x 4 window of the level above it (taps i-1..i+1 and
// j-1..j+1 around i = x/2, j = y/2).  The window is staged in LDS, planar by channel, with the
// clamp applied on load, so every coordinate expand_at asks for (clamped, or checked against the
// level's bounds first) lies inside it and reads the value the level holds there.
constexpr int kExW = 64, kExH = 4, kExCols = kExW / 2 + 2, kExRows = kExH / 2 + 2, kExPlane = kExCols * kExRows;

struct TileAcc {
    const float* p;  // [3][kExRows][kExCols] from level coordinates (cx0, ry0)
    int cx0, ry0;
    __device__ __forceinline__ f3 operator()(int x, int y) const {
        const float* q = p + (y - ry0) * kExCols + (x - cx0);
        return mk3(q[0], q[kExPlane], q[2 * kExPlane]);
    }
};

// The window of planar level `src` (cw x ch) for the block's tile, loaded into registers (every
// load issued, unconditionally: spare lanes repeat the last element), then stored to LDS; the
// caller synchronises.  Loading several windows before storing any keeps all their loads in flight.
constexpr int kExIt = (3 * kExPlane + 255) / 256;
struct WindowRegs {
    float v[kExIt];
};

__device__ __forceinline__ WindowRegs load_window(const float* __restrict__ src, int cw, int ch, int bx, int by) {
    const int cx0 = bx * (kExW / 2) - 1, ry0 = by * (kExH / 2) - 1;
    const size_t cn = (size_t)cw * ch;
    WindowRegs w;
#pragma unroll
    for (int it = 0; it < kExIt; it++) {
        int k = (int)threadIdx.x + it * 256;
        k = k < 3 * kExPlane ? k : 3 * kExPlane - 1;
        const int c = k / kExPlane, rem = k - c * kExPlane;
        const int r = rem / kExCols, col = rem - r * kExCols;
        w.v[it] = src[c * cn + (size_t)clamp_coord(ry0 + r, ch) * cw + clamp_coord(cx0 + col, cw)];
    }
    return w;
}

__device__ __forceinline__ TileAcc store_window(float* t, const WindowRegs& w, int bx, int by) {
#pragma unroll
    for (int it = 0; it < kExIt; it++) {
        const int k = (int)threadIdx.x + it * 256;
        if (k < 3 * kExPlane) t[k] = w.v[it];
    }
    return TileAcc{t, bx * (kExW / 2) - 1, by * (kExH / 2) - 1};
}

__device__ __forceinline__ TileAcc stage_window(float* t, const float* __restrict__ src, int cw, int ch, int bx,
                                                int by) {
    return store_window(t, load_window(src, cw, ch, bx, by), bx, by);
}

static inline dim3 expand_grid(int w, int h) {
    return dim3((unsigned)((w + kExW - 1) / kExW), (unsigned)((h + kExH - 1) / kExH));
}

__device__ __forceinline__ void collapse_tile(float* __restrict__ parent, int pw, int ph, const float* __restrict__ child,
                                              int cw, int ch, int bx, int by) {
    __shared__ float win[3 * kExPlane];
    const int x = bx * kExW + (threadIdx.x & (kExW - 1)), y = by * kExH + threadIdx.x / kExW;
    const bool in = x < pw && y < ph;
    const size_t n = (size_t)pw * ph, i = in ? (size_t)y * pw + x : 0;
    const float p0 = parent[i], p1 = parent[i + n], p2 = parent[i + 2 * n];  // in flight with the window
    const TileAcc acc = stage_window(win, child, cw, ch, bx, by);
    __syncthreads();
    if (!in) return;
    const f3 e = expand_at(acc, cw, ch, pw, ph, x, y);
    parent[i] = p0 + e.x;
    parent[i + n] = p1 + e.y;
    parent[i + 2 * n] = p2 + e.z;
}

__global__ void __launch_bounds__(256) k_collapse(float* __restrict__ parent, int pw, int ph,
                                                  const float* __restrict__ child, int cw, int ch) {
    collapse_tile(parent, pw, ph, child, cw, ch, (int)blockIdx.x, (int)blockIdx.y);
}

template <typename T>
__global__ void k_quantize(const float* __restrict__ acc0, size_t n, T* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = acc0[i + c * n] + 0.5f;
        v = v < 0.0f ? 0.0f : (v > SampleMax<T>::f ? SampleMax<T>::f : v);
        out[i * 3 + c] = (T)(uint32_t)v;
    }
}

hipError_t launch_quantize(const float* acc0, uint32_t w, uint32_t h, int sb, void* out, hipStream_t s) {
    size_t n = (size_t)w * h;
    dim3 grid((unsigned)((n + 255) / 256));
    if (sb == 1)
        hipLaunchKernelGGL((k_quantize<uint8_t>), grid, dim3(256), 0, s, acc0, n, (uint8_t*)out);
    else
        hipLaunchKernelGGL((k_quantize<uint16_t>), grid, dim3(256), 0, s, acc0, n, (uint16_t*)out);
    return hipGetLastError();
}

// The last collapse step fused with the quantize: level 0 + EXPAND(level 1), then (int)clamp(v + 0.5)
// (src/fuse/pyramid.zig:223-269), the same f32 operations without storing the collapsed level 0.
template <typename T>
__device__ __forceinline__ void collapse_quantize_tile(const float* __restrict__ parent, int pw, int ph,
                                                       const float* __restrict__ child, int cw, int ch,
                                                       T* __restrict__ out, int bx, int by) {
    __shared__ float win[3 * kExPlane];
    const int x = bx * kExW + (threadIdx.x & (kExW - 1)), y = by * kExH + threadIdx.x / kExW;
    const bool in = x < pw && y < ph;
    const size_t n = (size_t)pw * ph, i = in ? (size_t)y * pw + x : 0;
    const float p0 = parent[i], p1 = parent[i + n], p2 = parent[i + 2 * n];  // in flight with the window
    const TileAcc acc = stage_window(win, child, cw, ch, bx, by);
    __syncthreads();
    if (!in) return;
    const f3 e = expand_at(acc, cw, ch, pw, ph, x, y);
    const float v[3] = {p0 + e.x, p1 + e.y, p2 + e.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float q = v[c] + 0.5f;
        q = q < 0.0f ? 0.0f : (q > SampleMax<T>::f ? SampleMax<T>::f : q);
        out[i * 3 + c] = (T)(uint32_t)q;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_collapse_quantize(const float* __restrict__ parent, int pw, int ph,
                                                           const float* __restrict__ child, int cw, int ch,
                                                           T* __restrict__ out) {
    collapse_quantize_tile<T>(parent, pw, ph, child, cw, ch, out, (int)blockIdx.x, (int)blockIdx.y);
}

// One collapse step of many stacks: block b runs item b (a stack's parent / child levels, a tile).
__global__ void __launch_bounds__(256) k_collapse_batch(const CollapseJob* __restrict__ jobs,
                                                        const BatchItem* __restrict__ items) {
    const BatchItem w = items[blockIdx.x];
    const CollapseJob j = jobs[w.job];
    collapse_tile(j.parent, j.pw, j.ph, j.child, j.cw, j.ch, (int)w.bx, (int)w.by);
}

template <typename T>
__global__ void __launch_bounds__(256) k_collapse_quantize_batch(const CollapseJob* __restrict__ jobs,
                                                                 const BatchItem* __restrict__ items) {
    const BatchItem w = items[blockIdx.x];
    const CollapseJob j = jobs[w.job];
    collapse_quantize_tile<T>(j.parent, j.pw, j.ph, j.child, j.cw, j.ch, (T*)j.out, (int)w.bx, (int)w.by);
}

hipError_t launch_collapse_batch(const CollapseJob* jobs, const BatchItem* items, int n_items, bool final_step, int sb,
                                 hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (!final_step)
        hipLaunchKernelGGL(k_collapse_batch, dim3((unsigned)n_items), dim3(256), 0, s, jobs, items);
    else if (sb == 1)
        hipLaunchKernelGGL((k_collapse_quantize_batch<uint8_t>), dim3((unsigned)n_items), dim3(256), 0, s, jobs, items);
    else
        hipLaunchKernelGGL((k_collapse_quantize_batch<uint16_t>), dim3((unsigned)n_items), dim3(256), 0, s, jobs, items);
    return hipGetLastError();
}

// Zeroes chunk `bx` of range `job` (words [bx * kZeroChunkWords, +kZeroChunkWords) clipped to it).
__global__ void __launch_bounds__(256) k_zero_batch(const ZeroJob* __restrict__ jobs, const BatchItem* __restrict__ items) {
    const BatchItem w = items[blockIdx.x];
    const ZeroJob j = jobs[w.job];
    const size_t base = (size_t)w.bx * kZeroChunkWords;
    for (size_t i = base + threadIdx.x; i < base + kZeroChunkWords && i < j.words; i += 256) j.p[i] = 0u;
}

hipError_t launch_zero_batch(const ZeroJob* jobs, const BatchItem* items, int n_items, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_batch, dim3((unsigned)n_items), dim3(256), 0, s, jobs, items);
    return hipGetLastError();
}

// collapses levels L-1 .. 1 in place: level l-1 += EXPAND(level l) for l = L-1 down to 2
static void launch_collapse_levels(float* acc, const LevelTable* lt, hipStream_t s) {
    for (int l = lt->n - 1; l >= 2; l--) {
        const Level& p = lt->lv[l - 1];
        const Level& c = lt->lv[l];
        hipLaunchKernelGGL(k_collapse, expand_grid((int)p.w, (int)p.h), dim3(256), 0, s, acc + 3 * p.off, (int)p.w,
                           (int)p.h, (const float*)(acc + 3 * c.off), (int)c.w, (int)c.h);
    }
}

hipError_t launch_collapse_quantize(float* acc, const LevelTable* lt, int sb, void* out, hipStream_t s) {
    const Level& l0 = lt->lv[0];
    if (lt->n == 1) return launch_quantize(acc, l0.w, l0.h, sb, out, s);
    launch_collapse_levels(acc, lt, s);
    const Level& c = lt->lv[1];
    const dim3 g = expand_grid((int)l0.w, (int)l0.h);
    if (sb == 1)
        hipLaunchKernelGGL((k_collapse_quantize<uint8_t>), g, dim3(256), 0, s, (const float*)(acc + 3 * l0.off),
                           (int)l0.w, (int)l0.h, (const float*)(acc + 3 * c.off), (int)c.w, (int)c.h, (uint8_t*)out);
    else
        hipLaunchKernelGGL((k_collapse_quantize<uint16_t>), g, dim3(256), 0, s, (const float*)(acc + 3 * l0.off),
                           (int)l0.w, (int)l0.h, (const float*)(acc + 3 * c.off), (int)c.w, (int)c.h, (uint16_t*)out);
    return hipGetLastError();
}

__global__ void k_zero(uint32_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void k_zero_bytes(uint8_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Clears with a kernel (not hipMemsetAsync) so the same code can be captured in a HIP graph.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if (((uintptr_t)p % 4) == 0 && bytes % 4 == 0) {
        size_t n = bytes / 4;
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (uint32_t*)p, n);
    } else {
        hipLaunchKernelGGL(k_zero_bytes, dim3((unsigned)((bytes + 255) / 256)), dim3(256), 0, s, (uint8_t*)p, bytes);
    }
    return hipGetLastError();
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

__global__ void k_planar_to_interleaved(const float* __restrict__ p, size_t n, float* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[i * 3] = p[i];
    out[i * 3 + 1] = p[i + n];
    out[i * 3 + 2] = p[i + 2 * n];
}
hipError_t launch_planar_to_interleaved(const float* planar, size_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_planar_to_interleaved, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, planar, n, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// hybrid merge (mergeHybridImages src/fuse/pyramid.zig:271-301 via buildImageLaplacianPyramid :418-487)
// ---------------------------------------------------------------------------

// reduceRgbRowsSequential :952-995 with horizontalReduceRgbRow :1726-1746 (clamped, no support);
// tap(c, row, col) is the level's sample of channel c
template <class Tap>
__device__ __forceinline__ void reduce_clamped_px(const Tap& tap, int sw, int sh, int x, int y, float out[3]) {
    const int sx = 2 * x;
    const int xs[5] = {clamp_coord(sx - 2, sw), clamp_coord(sx - 1, sw), clamp_coord(sx, sw), clamp_coord(sx + 1, sw),
                       clamp_coord(sx + 2, sw)};
    for (int c = 0; c < 3; c++) {
        float rowv[5];
        for (int ky = 0; ky < 5; ky++) {
            const int r = clamp_coord(2 * y + ky - 2, sh);
            rowv[ky] = (tap(c, r, xs[0]) + 4.0f * tap(c, r, xs[1]) + 6.0f * tap(c, r, xs[2]) +
                        4.0f * tap(c, r, xs[3]) + tap(c, r, xs[4])) *
                       (1.0f / 16.0f);
        }
        out[c] = (rowv[0] + 4.0f * rowv[1] + 6.0f * rowv[2] + 4.0f * rowv[3] + rowv[4]) * (1.0f / 16.0f);
    }
}

__global__ void k_reduce_clamped(const float* __restrict__ src, int sw, int sh, float* __restrict__ dst, int dw,
                                 int dh) {
    const size_t n = (size_t)dw * dh, sn = (size_t)sw * sh;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float o[3];
    reduce_clamped_px([&](int c, int r, int x) { return src[c * sn + (size_t)r * sw + x]; }, sw, sh, (int)(i % dw),
                      (int)(i / dw), o);
    dst[i] = o[0];
    dst[i + n] = o[1];
    dst[i + 2 * n] = o[2];
}

// The first reduce, read straight from the interleaved image: its f32 copy (fillRgbBaseRows
// :606-639) is exact, so the taps hold the same values.  A block makes 64 x 4 outputs: it stages
// the clamped 11 x 131 source window (one row of taps per tile row, the reference's clamps applied
// on load) in LDS, forms each tile row's horizontal 5-tap sums once, then the vertical sums; every
// sum has the operands and order of reduce_clamped_px.  blockIdx.z picks soft (0) or hard (1).
constexpr int kRbW = 64, kRbH = 4, kRbRows = 2 * kRbH + 3, kRbCols = 2 * kRbW + 3;
template <typename T>
__global__ void __launch_bounds__(256) k_reduce_base(const T* __restrict__ soft, const T* __restrict__ hard, int sw,
                                                     int sh, float* __restrict__ dsoft, float* __restrict__ dhard,
                                                     int dw, int dh) {
    __shared__ float raw[kRbRows][3][kRbCols];
    __shared__ float hsum[kRbRows][3][kRbW];
    const T* img = blockIdx.z ? hard : soft;
    float* dst = blockIdx.z ? dhard : dsoft;
    const int x0 = blockIdx.x * kRbW, y0 = blockIdx.y * kRbH, tid = threadIdx.x;
    constexpr int kN = kRbRows * kRbCols * 3, kIt = (kN + 255) / 256;
    float v[kIt];  // every load issued, unconditionally (spare lanes repeat the last), before the LDS stores
#pragma unroll
    for (int it = 0; it < kIt; it++) {
        int k = tid + it * 256;
        k = k < kN ? k : kN - 1;
        const int t = k / (kRbCols * 3), rem = k - t * (kRbCols * 3);
        const int col = rem / 3, c = rem - col * 3;
        const int gr = clamp_coord(2 * y0 - 2 + t, sh), gc = clamp_coord(2 * x0 - 2 + col, sw);
        v[it] = (float)img[((size_t)gr * sw + gc) * 3 + c];
    }
#pragma unroll
    for (int it = 0; it < kIt; it++) {
        const int k = tid + it * 256;
        if (k < kN) {
            const int t = k / (kRbCols * 3), rem = k - t * (kRbCols * 3);
            const int col = rem / 3, c = rem - col * 3;
            raw[t][c][col] = v[it];
        }
    }
    __syncthreads();
    for (int k = tid; k < kRbRows * 3 * kRbW; k += 256) {
        const int t = k / (3 * kRbW), rem = k - t * (3 * kRbW);
        const int c = rem / kRbW, ox = rem - c * kRbW;
        const float* r = &raw[t][c][2 * ox];
        hsum[t][c][ox] = (r[0] + 4.0f * r[1] + 6.0f * r[2] + 4.0f * r[3] + r[4]) * (1.0f / 16.0f);
    }
    __syncthreads();
    const int tx = tid & (kRbW - 1), ty = tid / kRbW;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= dw || y >= dh) return;
    const size_t n = (size_t)dw * dh, i = (size_t)y * dw + x;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int t = 2 * ty;
        dst[i + c * n] = (hsum[t][c][tx] + 4.0f * hsum[t + 1][c][tx] + 6.0f * hsum[t + 2][c][tx] +
                          4.0f * hsum[t + 3][c][tx] + hsum[t + 4][c][tx]) *
                         (1.0f / 16.0f);
    }
}

// out[l] = lap_soft + (lap_hard - lap_soft) * mix for l in [hs, he), lap_soft otherwise
// (mergeHybridImages src/fuse/pyramid.zig:271-301); the top level's laplacian is the gaussian.
__global__ void k_lap_mix(const float* __restrict__ gs, const float* __restrict__ gsn, const float* __restrict__ gh,
                          const float* __restrict__ ghn, int w, int h, int nw, int nh, int top, int mixed, float mix,
                          float* __restrict__ out) {
    const size_t n = (size_t)w * h;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % w), y = (int)(i / w);
    f3 s = mk3(gs[i], gs[i + n], gs[i + 2 * n]);
    f3 hv = mk3(0.0f, 0.0f, 0.0f);
    if (mixed) hv = mk3(gh[i], gh[i + n], gh[i + 2 * n]);
    // the top level is the gaussian itself (buildImageLaplacianPyramid src/fuse/pyramid.zig:477-482);
    // it is mixed like the others when it falls in the hard range, i.e. when the pyramid has at
    // most start + count levels
    f3 ls = s;
    if (!top) {
        ls = sub3(s, expand_at(PlanarAcc{gsn, (size_t)nw * nh, nw}, nw, nh, w, h, x, y));
        if (mixed) hv = sub3(hv, expand_at(PlanarAcc{ghn, (size_t)nw * nh, nw}, nw, nh, w, h, x, y));
    }
    if (mixed) ls = add3(ls, mul3r(sub3(hv, ls), mix));  // soft + (hard - soft) * mix
    out[i] = ls.x;
    out[i + n] = ls.y;
    out[i + 2 * n] = ls.z;
}

// The merge's last step: the level-0 laplacian (k_lap_mix's operations on the images' exact f32
// samples) plus EXPAND(collapsed level 1), quantized (collapseToImageWithJobs :223-269); the level-0
// laplacian is never stored.  With one level (child == 0) the level is the top, the image itself.
template <typename T>
__global__ void __launch_bounds__(256) k_merge_level0(const T* __restrict__ soft, const T* __restrict__ hard,
                                                      const float* __restrict__ gs1, const float* __restrict__ gh1,
                                                      const float* __restrict__ col1, int w, int h, int cw, int ch,
                                                      int child, int mixed, float mix, T* __restrict__ out) {
    __shared__ float win[3][3 * kExPlane];
    const int x = blockIdx.x * kExW + (threadIdx.x & (kExW - 1)), y = blockIdx.y * kExH + threadIdx.x / kExW;
    const bool in = x < w && y < h;
    const size_t i = in ? (size_t)y * w + x : 0;
    f3 ls = mk3((float)soft[i * 3], (float)soft[i * 3 + 1], (float)soft[i * 3 + 2]);  // in flight with the windows
    TileAcc as{}, ah{}, ac{};
    if (child) {
        const int bx = (int)blockIdx.x, by = (int)blockIdx.y;
        const WindowRegs rs = load_window(gs1, cw, ch, bx, by), rc = load_window(col1, cw, ch, bx, by);
        if (mixed) ah = stage_window(win[1], gh1, cw, ch, bx, by);
        as = store_window(win[0], rs, bx, by);
        ac = store_window(win[2], rc, bx, by);
    }
    __syncthreads();
    if (!in) return;
    if (child) ls = sub3(ls, expand_at(as, cw, ch, w, h, x, y));
    if (mixed) {
        f3 hv = mk3((float)hard[i * 3], (float)hard[i * 3 + 1], (float)hard[i * 3 + 2]);
        if (child) hv = sub3(hv, expand_at(ah, cw, ch, w, h, x, y));
        ls = add3(ls, mul3r(sub3(hv, ls), mix));  // soft + (hard - soft) * mix
    }
    if (child) ls = add3(ls, expand_at(ac, cw, ch, w, h, x, y));
    const float v[3] = {ls.x, ls.y, ls.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float q = v[c] + 0.5f;
        q = q < 0.0f ? 0.0f : (q > SampleMax<T>::f ? SampleMax<T>::f : q);
        out[i * 3 + c] = (T)(uint32_t)q;
    }
}

// gs/gh hold levels 1..L-1 of the soft and hard gaussians (level l at 3 * (off_l - off_1) floats);
// out_pyr holds the merged laplacian at the accumulator's offsets, level 0 unused.
hipError_t launch_hybrid_merge(const void* soft, const void* hard, int sb, const LevelTable* lt, uint32_t hard_start,
                               uint32_t hard_count, float mix, float* gs, float* gh, float* out_pyr, void* out,
                               hipStream_t s) {
    const int L = lt->n;
    const Level& l0 = lt->lv[0];
    const size_t base = L > 1 ? lt->lv[1].off : 0;
    auto at = [&](float* b, int l) { return b + 3 * (lt->lv[l].off - base); };
    const uint32_t hs = hard_start < (uint32_t)L ? hard_start : (uint32_t)L;
    const uint32_t he = hs + hard_count < (uint32_t)L ? hs + hard_count : (uint32_t)L;
    if (L > 1) {
        const Level& c1 = lt->lv[1];
        const dim3 g1((c1.w + kRbW - 1) / kRbW, (c1.h + kRbH - 1) / kRbH, 2);
        if (sb == 1)
            hipLaunchKernelGGL((k_reduce_base<uint8_t>), g1, dim3(256), 0, s, (const uint8_t*)soft,
                               (const uint8_t*)hard, (int)l0.w, (int)l0.h, at(gs, 1), at(gh, 1), (int)c1.w, (int)c1.h);
        else
            hipLaunchKernelGGL((k_reduce_base<uint16_t>), g1, dim3(256), 0, s, (const uint16_t*)soft,
                               (const uint16_t*)hard, (int)l0.w, (int)l0.h, at(gs, 1), at(gh, 1), (int)c1.w,
                               (int)c1.h);
        for (int l = 2; l < L; l++) {
            const Level& p = lt->lv[l - 1];
            const Level& c = lt->lv[l];
            const size_t n = (size_t)c.w * c.h;
            const dim3 g((unsigned)((n + 255) / 256));
            hipLaunchKernelGGL(k_reduce_clamped, g, dim3(256), 0, s, (const float*)at(gs, l - 1), (int)p.w, (int)p.h,
                               at(gs, l), (int)c.w, (int)c.h);
            hipLaunchKernelGGL(k_reduce_clamped, g, dim3(256), 0, s, (const float*)at(gh, l - 1), (int)p.w, (int)p.h,
                               at(gh, l), (int)c.w, (int)c.h);
        }
        for (int l = 1; l < L; l++) {
            const Level& c = lt->lv[l];
            const int top = l == L - 1;
            const int nl = top ? l : l + 1;
            const Level& nx = lt->lv[nl];
            const size_t n = (size_t)c.w * c.h;
            hipLaunchKernelGGL(k_lap_mix, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               (const float*)at(gs, l), (const float*)at(gs, nl), (const float*)at(gh, l),
                               (const float*)at(gh, nl), (int)c.w, (int)c.h, (int)nx.w, (int)nx.h, top,
                               (int)((uint32_t)l >= hs && (uint32_t)l < he), mix, out_pyr + 3 * c.off);
        }
        launch_collapse_levels(out_pyr, lt, s);
    }
    const Level& c1 = lt->lv[L > 1 ? 1 : 0];
    const dim3 g0 = expand_grid((int)l0.w, (int)l0.h);
    const int child = L > 1, mixed = 0 >= hs && 0 < he;
    const float* g1s = child ? at(gs, 1) : gs;
    const float* g1h = child ? at(gh, 1) : gh;
    const float* col1 = out_pyr + 3 * c1.off;
    if (sb == 1)
        hipLaunchKernelGGL((k_merge_level0<uint8_t>), g0, dim3(256), 0, s, (const uint8_t*)soft, (const uint8_t*)hard,
                           g1s, g1h, col1, (int)l0.w, (int)l0.h, (int)c1.w, (int)c1.h, child, mixed, mix,
                           (uint8_t*)out);
    else
        hipLaunchKernelGGL((k_merge_level0<uint16_t>), g0, dim3(256), 0, s, (const uint16_t*)soft,
                           (const uint16_t*)hard, g1s, g1h, col1, (int)l0.w, (int)l0.h, (int)c1.w, (int)c1.h, child,
                           mixed, mix, (uint16_t*)out);
    return hipGetLastError();
}

}  // namespace fg
