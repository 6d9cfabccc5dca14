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
    double sx, sy;
    if (map_pixel(p, x, y, &sx, &sy)) {
        sample_bilinear<T, CH>(src, sw, sh, sx, sy, o);
    } else {
#pragma unroll
        for (int c = 0; c <= CH; c++) o[c] = 0;
    }
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

__global__ void __launch_bounds__(256) k_collapse(float* __restrict__ parent, int pw, int ph,
                                                  const float* __restrict__ child, int cw, int ch) {
    const size_t n = (size_t)pw * ph;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % pw), y = (int)(i / pw);
    f3 e = expand_at(PlanarAcc{child, (size_t)cw * ch, cw}, cw, ch, pw, ph, x, y);
    parent[i] = parent[i] + e.x;
    parent[i + n] = parent[i + n] + e.y;
    parent[i + 2 * n] = parent[i + 2 * n] + e.z;
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
__global__ void __launch_bounds__(256) k_collapse_quantize(const float* __restrict__ parent, int pw, int ph,
                                                           const float* __restrict__ child, int cw, int ch,
                                                           T* __restrict__ out) {
    const size_t n = (size_t)pw * ph;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % pw), y = (int)(i / pw);
    const f3 e = expand_at(PlanarAcc{child, (size_t)cw * ch, cw}, cw, ch, pw, ph, x, y);
    const float v[3] = {parent[i] + e.x, parent[i + n] + e.y, parent[i + 2 * n] + e.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float q = v[c] + 0.5f;
        q = q < 0.0f ? 0.0f : (q > SampleMax<T>::f ? SampleMax<T>::f : q);
        out[i * 3 + c] = (T)(uint32_t)q;
    }
}

hipError_t launch_collapse_quantize(float* acc, const LevelTable* lt, int sb, void* out, hipStream_t s) {
    const Level& l0 = lt->lv[0];
    if (lt->n == 1) return launch_quantize(acc, l0.w, l0.h, sb, out, s);
    for (int l = lt->n - 1; l >= 2; l--) {
        const Level& p = lt->lv[l - 1];
        const Level& c = lt->lv[l];
        size_t n = (size_t)p.w * p.h;
        hipLaunchKernelGGL(k_collapse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, acc + 3 * p.off, (int)p.w,
                           (int)p.h, (const float*)(acc + 3 * c.off), (int)c.w, (int)c.h);
    }
    const Level& c = lt->lv[1];
    const size_t n = (size_t)l0.w * l0.h;
    const dim3 g((unsigned)((n + 255) / 256));
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

template <typename T>
__global__ void k_fill_base(const T* __restrict__ img, size_t n, float* __restrict__ g) {  // fillRgbBaseRows :606-639
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    g[i] = (float)img[i * 3];
    g[i + n] = (float)img[i * 3 + 1];
    g[i + 2 * n] = (float)img[i * 3 + 2];
}

// reduceRgbRowsSequential :952-995 with horizontalReduceRgbRow :1726-1746 (clamped, no support)
__global__ void k_reduce_clamped(const float* __restrict__ src, int sw, int sh, float* __restrict__ dst, int dw,
                                 int dh) {
    const size_t n = (size_t)dw * dh, sn = (size_t)sw * sh;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % dw), y = (int)(i / dw);
    const int sx = 2 * x;
    const int xs[5] = {clamp_coord(sx - 2, sw), clamp_coord(sx - 1, sw), clamp_coord(sx, sw), clamp_coord(sx + 1, sw),
                       clamp_coord(sx + 2, sw)};
    for (int c = 0; c < 3; c++) {
        const float* p = src + c * sn;
        float rowv[5];
        for (int ky = 0; ky < 5; ky++) {
            const float* r = p + (size_t)clamp_coord(2 * y + ky - 2, sh) * sw;
            rowv[ky] = (r[xs[0]] + 4.0f * r[xs[1]] + 6.0f * r[xs[2]] + 4.0f * r[xs[3]] + r[xs[4]]) * (1.0f / 16.0f);
        }
        dst[i + c * n] = (rowv[0] + 4.0f * rowv[1] + 6.0f * rowv[2] + 4.0f * rowv[3] + rowv[4]) * (1.0f / 16.0f);
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

hipError_t launch_hybrid_merge(const void* soft, const void* hard, int sb, const LevelTable* lt, uint32_t hard_start,
                               uint32_t hard_count, float mix, float* gs, float* gh, float* out_pyr, void* out,
                               hipStream_t s) {
    const int L = lt->n;
    const Level& l0 = lt->lv[0];
    size_t n0 = (size_t)l0.w * l0.h;
    dim3 g0((unsigned)((n0 + 255) / 256));
    if (sb == 1) {
        hipLaunchKernelGGL((k_fill_base<uint8_t>), g0, dim3(256), 0, s, (const uint8_t*)soft, n0, gs);
        hipLaunchKernelGGL((k_fill_base<uint8_t>), g0, dim3(256), 0, s, (const uint8_t*)hard, n0, gh);
    } else {
        hipLaunchKernelGGL((k_fill_base<uint16_t>), g0, dim3(256), 0, s, (const uint16_t*)soft, n0, gs);
        hipLaunchKernelGGL((k_fill_base<uint16_t>), g0, dim3(256), 0, s, (const uint16_t*)hard, n0, gh);
    }
    for (int l = 1; l < L; l++) {
        const Level& p = lt->lv[l - 1];
        const Level& c = lt->lv[l];
        size_t n = (size_t)c.w * c.h;
        dim3 g((unsigned)((n + 255) / 256));
        hipLaunchKernelGGL(k_reduce_clamped, g, dim3(256), 0, s, (const float*)(gs + 3 * p.off), (int)p.w, (int)p.h,
                           gs + 3 * c.off, (int)c.w, (int)c.h);
        hipLaunchKernelGGL(k_reduce_clamped, g, dim3(256), 0, s, (const float*)(gh + 3 * p.off), (int)p.w, (int)p.h,
                           gh + 3 * c.off, (int)c.w, (int)c.h);
    }
    uint32_t hs = hard_start < (uint32_t)L ? hard_start : (uint32_t)L;
    uint32_t he = hs + hard_count < (uint32_t)L ? hs + hard_count : (uint32_t)L;
    for (int l = 0; l < L; l++) {
        const Level& c = lt->lv[l];
        const int top = l == L - 1;
        const Level& nx = lt->lv[top ? l : l + 1];
        size_t n = (size_t)c.w * c.h;
        hipLaunchKernelGGL(k_lap_mix, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           (const float*)(gs + 3 * c.off), (const float*)(gs + 3 * nx.off),
                           (const float*)(gh + 3 * c.off), (const float*)(gh + 3 * nx.off), (int)c.w, (int)c.h,
                           (int)nx.w, (int)nx.h, top, (int)((uint32_t)l >= hs && (uint32_t)l < he), mix,
                           out_pyr + 3 * c.off);
    }
    return launch_collapse_quantize(out_pyr, lt, sb, out, s);
}

}  // namespace fg
