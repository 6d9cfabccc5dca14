// The single-pass methods (src/stack/pipeline.zig:245-312, src/fuse/pipeline.zig:189-252).
//
// hardmask-contrast is pass 1's winner update alone (k_pass1 with only `best` / `hard`), so it
// needs no kernel here.  softmask-contrast blurs each frame's support-scaled weight map
// (blurFiveTapInto, src/fuse/masks.zig:86-207), adds weight and weight * rgb where the blurred
// weight is positive (accumulateRange, src/fuse/blend.zig:97-148) and finally divides
// (finalizeSoft, :150-186).  k_soft runs a group of frames per launch with the sums of its
// 64x16 tile in registers, in frame order.
#include <type_traits>

#include "fg_device.h"

namespace fg {

namespace {

constexpr int SW = 64, SH = 16;                 // tile
constexpr int SRW = SW + 4, SRH = SH + 4;       // weight region with the blur's two-pixel halo

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// (((a + 4b) + 6c) + 4d) + e, then / 16: the reference's left-to-right sum; the division by 16 is
// exact, so it is a multiplication
__device__ __forceinline__ float five_tap(float a, float b, float c, float d, float e) {
    return ((((a + 4.0f * b) + 6.0f * c) + 4.0f * d) + e) * 0.0625f;
}

template <typename T>
__global__ void __launch_bounds__(256) k_soft(SoftArgs a) {
    using PT = typename std::conditional<sizeof(T) == 1, uint32_t, uint2>::type;
    __shared__ float reg[SRH * SRW];
    __shared__ float hb[SRH * SW];
    const int W = (int)a.w, H = (int)a.h;
    const int X0 = blockIdx.x * SW, Y0 = blockIdx.y * SH;
    const int tx = threadIdx.x % SW, ty = threadIdx.x / SW;
    const size_t n = (size_t)W * H;
    const int x = X0 + tx;

    float ws[4], p0[4], p1[4], p2[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int y = Y0 + ty + 4 * j;
        ws[j] = p0[j] = p1[j] = p2[j] = 0.0f;
        if (x < W && y < H) {
            const size_t i = (size_t)y * W + x;
            ws[j] = a.wsum[i];
            p0[j] = a.psum[i];
            p1[j] = a.psum[i + n];
            p2[j] = a.psum[i + 2 * n];
        }
    }
    for (int k = a.k0; k < a.k1; k++) {
        const size_t slot = (size_t)(k - a.slot_base);
        const float* wk = a.weights + slot * n;
        const PT* px = (const PT*)a.rgba + slot * n;
        // region with clamped coordinates: the clamp of the blur at the image edges
        // (horizontalPassRange :151-177, verticalPassRange :179-207)
        for (int idx = threadIdx.x; idx < SRH * SRW; idx += 256) {
            const int gy = clampi(Y0 - 2 + idx / SRW, H - 1), gx = clampi(X0 - 2 + idx % SRW, W - 1);
            reg[idx] = wk[(size_t)gy * W + gx];
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < SRH * SW; idx += 256) {
            const float* r = &reg[(idx / SW) * SRW + idx % SW];
            hb[idx] = five_tap(r[0], r[1], r[2], r[3], r[4]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int y = Y0 + ty + 4 * j;
            if (x >= W || y >= H) continue;
            const int t = (ty + 4 * j) * SW + tx;  // region row ty + 4j is output row - 2
            const float bw = five_tap(hb[t], hb[t + SW], hb[t + 2 * SW], hb[t + 3 * SW], hb[t + 4 * SW]);
            if (bw > 0.0f) {
                const PT q = px[(size_t)y * W + x];
                float r, g, b;
                if (sizeof(T) == 1) {
                    const uint32_t v = *(const uint32_t*)&q;
                    r = (float)(v & 0xffu);
                    g = (float)((v >> 8) & 0xffu);
                    b = (float)((v >> 16) & 0xffu);
                } else {
                    const uint2 v = *(const uint2*)&q;
                    r = (float)(v.x & 0xffffu);
                    g = (float)(v.x >> 16);
                    b = (float)(v.y & 0xffffu);
                }
                ws[j] = ws[j] + bw;
                p0[j] = p0[j] + bw * r;
                p1[j] = p1[j] + bw * g;
                p2[j] = p2[j] + bw * b;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int y = Y0 + ty + 4 * j;
        if (x >= W || y >= H) continue;
        const size_t i = (size_t)y * W + x;
        a.wsum[i] = ws[j];
        a.psum[i] = p0[j];
        a.psum[i + n] = p1[j];
        a.psum[i + 2 * n] = p2[j];
    }
}

// finalizeSoft src/fuse/blend.zig:150-186
template <typename T>
__global__ void k_soft_finalize(const float* __restrict__ wsum, const float* __restrict__ psum, size_t n,
                                T* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float wt = wsum[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        uint32_t q = 0;
        if (wt > 1e-12f) {
            const float inv = 1.0f / wt;
            float v = psum[i + c * n] * inv + 0.5f;
            v = v < 0.0f ? 0.0f : (v > SampleMax<T>::f ? SampleMax<T>::f : v);
            q = (uint32_t)v;
        }
        out[i * 3 + c] = (T)q;
    }
}

}  // namespace

hipError_t launch_soft(const SoftArgs& a, hipStream_t s) {
    if (a.k1 <= a.k0) return hipSuccess;
    dim3 grid((a.w + SW - 1) / SW, (a.h + SH - 1) / SH);
    if (a.sample_bytes == 1)
        hipLaunchKernelGGL(k_soft<uint8_t>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_soft<uint16_t>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_soft_finalize(const float* wsum, const float* psum, uint32_t w, uint32_t h, int sample_bytes,
                                void* out, hipStream_t s) {
    const size_t n = (size_t)w * h;
    dim3 grid((unsigned)((n + 255) / 256));
    if (sample_bytes == 1)
        hipLaunchKernelGGL(k_soft_finalize<uint8_t>, grid, dim3(256), 0, s, wsum, psum, n, (uint8_t*)out);
    else
        hipLaunchKernelGGL(k_soft_finalize<uint16_t>, grid, dim3(256), 0, s, wsum, psum, n, (uint16_t*)out);
    return hipGetLastError();
}

}  // namespace fg
