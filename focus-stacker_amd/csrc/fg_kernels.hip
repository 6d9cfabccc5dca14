// HIP kernels (gfx950, wave64) of the remap + pyramid-contrast path.
//
// Arithmetic follows the reference operation by operation so that, given the same
// inputs, every f32 stage rounds the same way: the library is compiled with
// -ffp-contract=off (the reference's Zig float mode is strict, no FMA), sums keep the
// reference's association order, and divisions/sqrt are IEEE correctly rounded.
// Where the reference's own order is not a function of the data (the contrast sliding
// sums depend on its thread split) the kernel uses a fixed direct order in the same
// precision (f64).
#include "fg_internal.h"

#include <math.h>

namespace fg {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float a, float b, float c) { return f3{a, b, c}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul3(float k, f3 a) { return f3{k * a.x, k * a.y, k * a.z}; }
__device__ __forceinline__ f3 mul3r(f3 a, float k) { return f3{a.x * k, a.y * k, a.z * k}; }
__device__ __forceinline__ f3 div3(f3 a, float k) { return f3{a.x / k, a.y / k, a.z / k}; }
__device__ __forceinline__ f3 zero3() { return f3{0.0f, 0.0f, 0.0f}; }

template <typename T>
struct SampleMax;
template <>
struct SampleMax<uint8_t> {
    static constexpr uint32_t i = 255u;
    static constexpr double d = 255.0;
    static constexpr float f = 255.0f;
    static constexpr float gray3 = 765.0f;  // 3.0 * 255.0 (src/fuse/grayscale.zig:34)
};
template <>
struct SampleMax<uint16_t> {
    static constexpr uint32_t i = 65535u;
    static constexpr double d = 65535.0;
    static constexpr float f = 65535.0f;
    static constexpr float gray3 = 196605.0f;  // 3.0 * 65535.0 (src/fuse/grayscale.zig:53)
};

__device__ __forceinline__ int clamp_coord(int c, int limit) {  // clampCoord src/fuse/pyramid.zig:1837-1842
    if (c <= 0) return 0;
    int mx = limit - 1;
    return c >= mx ? mx : c;
}

// ---------------------------------------------------------------------------
// remap (src/remap.zig:343-541, 710-994; transform src/optimize.zig:3593-3634, 3859-3881, 4204-4235)
// ---------------------------------------------------------------------------

// Newton inverse of the radial model (invertRadialDistortionCached src/optimize.zig:4204-4235,
// undistortSourcePoint src/remap.zig:755-786); returns the undistorted offset.
__device__ __forceinline__ void invert_radial(const RemapParams& p, double dx, double dy, double* ux, double* uy) {
    double q = sqrt(dx * dx + dy * dy);
    if (q < 1e-12) {
        *ux = dx;
        *uy = dy;
        return;
    }
    double radius = q;
    for (int it = 0; it < 6; it++) {
        double nr2 = radius * radius * p.inv_norm;
        double nr4 = nr2 * nr2;
        double factor = 1.0 + p.ra * nr2 + p.rb * nr4 + p.rc * nr4 * nr2;
        double fd = p.ra + 2.0 * p.rb * nr2 + 3.0 * p.rc * nr4;
        double f = radius * factor - q;
        if (fabs(f) <= 1e-9) break;
        double df = factor + 2.0 * radius * radius * p.inv_norm * fd;
        if (fabs(df) < 1e-12) break;
        radius -= f / df;
        radius = radius > 0.0 ? radius : 0.0;
    }
    double scale = radius / q;
    *ux = dx * scale;
    *uy = dy * scale;
}

// Source coordinate of output pixel (x, y).  Returns false when the reference writes zeros
// without sampling (|denom| < 1e-12 on the lens path, src/remap.zig:420-422; rays behind
// the camera on the basic path, whose legacy chain maps them 1.6e16 px away).
__device__ __forceinline__ bool map_pixel(const RemapParams& p, int x, int y, double* sx, double* sy) {
    double px = (double)(p.roi_left + x) - p.dcx;
    double py = (double)(p.roi_top + y) - p.dcy;
    double rz = p.neg_d;
    if (p.mode == REMAP_TRANSLATION) {
        double den = p.pn[0] * px + p.pn[1] * py + p.pn[2] * rz;
        if (fabs(den) < 1e-12) {
            *sx = -1.0;
            *sy = -1.0;
            return true;
        }
        double u = 1.0 / den;
        double hx = 0.0 + u * px, hy = 0.0 + u * py, hz = 0.0 + u * rz;
        px = hx - p.tx;
        py = hy - p.ty;
        rz = hz - p.tz;
    }
    double lx = p.B[0] * px + p.B[1] * py + p.B[2] * rz;
    double ly = p.B[3] * px + p.B[4] * py + p.B[5] * rz;
    double lz = p.B[6] * px + p.B[7] * py + p.B[8] * rz;
    double denom = -lz;
    double rx, ry;
    if (p.mode == REMAP_TRANSLATION) {
        if (fabs(denom) < 1e-12) {
            rx = -1.0;
            ry = -1.0;
        } else {
            rx = p.focal * (lx / denom) + p.csx;
            ry = p.focal * (ly / denom) + p.csy;
        }
        double ux, uy;
        invert_radial(p, rx, ry, &ux, &uy);
        *sx = p.scx + ux;
        *sy = p.scy + uy;
        return true;
    }
    if (p.legacy ? !(denom > 0.0) : fabs(denom) < 1e-12) return false;
    rx = p.focal * (lx / denom) + p.csx;
    ry = p.focal * (ly / denom) + p.csy;
    if (p.mode == REMAP_LENS && p.radial) {
        double ux, uy;
        invert_radial(p, rx, ry, &ux, &uy);
        *sx = p.scx + ux;
        *sy = p.scy + uy;
    } else {
        *sx = p.scx + rx;
        *sy = p.scy + ry;
    }
    return true;
}

// samplePixelBilinear{U8,U16} with interior and support-renormalised border cases.
template <typename T, int CH>
__device__ __forceinline__ void sample_bilinear(const T* __restrict__ px, int w, int h, double x, double y,
                                                uint32_t out[CH + 1]) {
#pragma unroll
    for (int c = 0; c <= CH; c++) out[c] = 0;
    if (!(x >= -1.0 && x < (double)w && y >= -1.0 && y < (double)h)) return;
    int x0 = (int)floor(x), y0 = (int)floor(y);
    int x1 = x0 + 1, y1 = y0 + 1;
    double fx = x - (double)x0, fy = y - (double)y0;
    double omx = 1.0 - fx, omy = 1.0 - fy;
    double w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
    size_t rs = (size_t)w * CH;
    if (x0 >= 0 && y0 >= 0 && x1 < w && y1 < h) {
        size_t b00 = (size_t)y0 * rs + (size_t)x0 * CH;
        size_t b10 = b00 + CH, b01 = b00 + rs, b11 = b01 + CH;
        double sup = w00 + w10 + w01 + w11;
#pragma unroll
        for (int c = 0; c < CH; c++) {
            double v = (double)px[b00 + c] * w00 + (double)px[b10 + c] * w10 + (double)px[b01 + c] * w01 +
                       (double)px[b11 + c] * w11;
            out[c] = (uint32_t)round(v / sup);
        }
        out[CH] = SampleMax<T>::i;
        return;
    }
    bool i00 = x0 >= 0 && y0 >= 0 && x0 < w && y0 < h;
    bool i10 = x1 >= 0 && y0 >= 0 && x1 < w && y0 < h;
    bool i01 = x0 >= 0 && y1 >= 0 && x0 < w && y1 < h;
    bool i11 = x1 >= 0 && y1 >= 0 && x1 < w && y1 < h;
    double sup = 0.0;
    if (i00) sup += w00;
    if (i10) sup += w10;
    if (i01) sup += w01;
    if (i11) sup += w11;
    if (sup <= 0.0) return;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        double v = 0.0;
        if (i00) v += (double)px[((size_t)y0 * w + x0) * CH + c] * w00;
        if (i10) v += (double)px[((size_t)y0 * w + x1) * CH + c] * w10;
        if (i01) v += (double)px[((size_t)y1 * w + x0) * CH + c] * w01;
        if (i11) v += (double)px[((size_t)y1 * w + x1) * CH + c] * w11;
        out[c] = (uint32_t)round(v / sup);
    }
    out[CH] = (uint32_t)round(sup * SampleMax<T>::d);
}

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

template <typename T>
__global__ void k_to_rgba(const T* __restrict__ src, int extra, size_t n, T* __restrict__ dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int st = 3 + extra;
    T r = src[i * st], g = src[i * st + 1], b = src[i * st + 2];
    T a = extra ? src[i * st + 3] : (T)SampleMax<T>::i;
    dst[i * 4] = r;
    dst[i * 4 + 1] = g;
    dst[i * 4 + 2] = b;
    dst[i * 4 + 3] = a;
}

hipError_t launch_to_rgba(const void* src, int sb, int extra, uint32_t w, uint32_t h, void* dst, hipStream_t s) {
    size_t n = (size_t)w * h;
    dim3 grid((unsigned)((n + 255) / 256));
    if (sb == 1)
        hipLaunchKernelGGL((k_to_rgba<uint8_t>), grid, dim3(256), 0, s, (const uint8_t*)src, extra, n, (uint8_t*)dst);
    else
        hipLaunchKernelGGL((k_to_rgba<uint16_t>), grid, dim3(256), 0, s, (const uint16_t*)src, extra, n,
                           (uint16_t*)dst);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pass 1: gray (fillAverage src/fuse/grayscale.zig:21-57), binary support (src/fuse/masks.zig:32-57),
// local contrast (src/fuse/contrast.zig:96-154, 172-290), applySupport (src/fuse/masks.zig:6-30),
// union support (:59-84), norm sum (src/stack/pipeline.zig:469); hybrid winners
// (updateRange src/fuse/blend.zig:237-314) fold in here since the effective winner weight
// is exactly the support-scaled contrast weight.
// ---------------------------------------------------------------------------

constexpr int P1_TW = 64, P1_TH = 4;

template <typename T>
__global__ void __launch_bounds__(256) k_pass1(const T* __restrict__ rgba, int W, int H, int win,
                                               float* __restrict__ wout, float* __restrict__ S,
                                               uint8_t* __restrict__ U, float* __restrict__ best,
                                               T* __restrict__ hard) {
    extern __shared__ float g_lds[];  // gray of the tile + halo; < 0 marks unsupported / outside
    const int b = win / 2;
    const int cols = P1_TW + 2 * b, rows = P1_TH + 2 * b;
    const int ox = blockIdx.x * P1_TW, oy = blockIdx.y * P1_TH;
    for (int k = threadIdx.x; k < cols * rows; k += 256) {
        int gx = ox - b + k % cols, gy = oy - b + k / cols;
        float v = -1.0f;
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            const T* q = rgba + ((size_t)gy * W + gx) * 4;
            if (q[3] > 0) {
                uint32_t sum = (uint32_t)q[0] + (uint32_t)q[1] + (uint32_t)q[2];
                v = (float)sum / SampleMax<T>::gray3;
            }
        }
        g_lds[k] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x = ox + tx, y = oy + ty;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    float std_w = 0.0f;
    const float* c0 = g_lds + (ty + b) * cols + (tx + b);
    if (W >= win && H >= win && x >= b && x < W - b && y >= b && y < H - b && *c0 >= 0.0f) {
        double sum = 0.0, sq = 0.0;
        uint32_t n = 0;
        for (int dy = -b; dy <= b; dy++) {
            const float* r = c0 + dy * cols;
            for (int dx = -b; dx <= b; dx++) {
                float v = r[dx];
                if (v >= 0.0f) {
                    sum += (double)v;
                    sq += (double)(v * v);
                    n++;
                }
            }
        }
        if (n > 1) {
            double denom = (double)(n - 1), nf = (double)n;
            double var = (sq - ((sum * sum) / nf)) / denom;
            if (!(var > 0.0)) var = 0.0;
            std_w = (float)sqrt(var);
        }
    }
    const T* q = rgba + i * 4;
    float wv = std_w * ((float)q[3] / SampleMax<T>::f);
    wout[i] = wv;
    if (S) {
        S[i] = S[i] + wv;
        if (q[3] > 0) U[i] = 1;
    }
    if (best) {
        if (wv > best[i]) {
            best[i] = wv;
            hard[i * 3] = q[0];
            hard[i * 3 + 1] = q[1];
            hard[i * 3 + 2] = q[2];
        }
    }
}

hipError_t launch_pass1(const Pass1Args& a, hipStream_t s) {
    int b = (int)a.window / 2;
    size_t lds = sizeof(float) * (size_t)(P1_TW + 2 * b) * (P1_TH + 2 * b);
    dim3 grid((a.w + P1_TW - 1) / P1_TW, (a.h + P1_TH - 1) / P1_TH);
    if (a.sample_bytes == 1)
        hipLaunchKernelGGL((k_pass1<uint8_t>), grid, dim3(256), lds, s, (const uint8_t*)a.rgba, (int)a.w, (int)a.h,
                           (int)a.window, a.weights, a.norm_sum, a.union_sup, a.best, (uint8_t*)a.hard);
    else
        hipLaunchKernelGGL((k_pass1<uint16_t>), grid, dim3(256), lds, s, (const uint16_t*)a.rgba, (int)a.w, (int)a.h,
                           (int)a.window, a.weights, a.norm_sum, a.union_sup, a.best, (uint16_t*)a.hard);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pass 2 (accumulateImageWithWorkspace src/fuse/pyramid.zig:190-204)
// ---------------------------------------------------------------------------

// normalizeWeightsInto src/fuse/pyramid.zig:303-316
__device__ __forceinline__ float normalized_mask(float w, float s, float def) { return s > 1e-12f ? w / s : def; }

// Level-0 sources: the cached RGBA frame and the weight / norm-sum / union maps.
template <typename T>
struct Level0 {
    const T* rgba;
    const float* wts;
    const float* S;
    const uint8_t* U;
    float def;
    int w, h;
    __device__ __forceinline__ float4 g(int x, int y) const {
        const T* q = rgba + ((size_t)y * w + x) * 4;
        return make_float4((float)q[0], (float)q[1], (float)q[2], q[3] > 0 ? 1.0f : 0.0f);
    }
    __device__ __forceinline__ float2 m(int x, int y) const {
        size_t i = (size_t)y * w + x;
        return make_float2(normalized_mask(wts[i], S[i], def), U[i] ? 1.0f : 0.0f);
    }
};
struct LevelN {
    const float4* G;
    const float2* M;
    int w, h;
    __device__ __forceinline__ float4 g(int x, int y) const { return G[(size_t)y * w + x]; }
    __device__ __forceinline__ float2 m(int x, int y) const { return M[(size_t)y * w + x]; }
};

// Support-aware 5x5 reduce of RGB (own support) and mask (union support) together:
// reduceRgbWithSupportRowsSequential :1044-1099 and reduceScalarWithSupportRowsSequential
// :997-1042 — taps skipped when outside or unsupported, weight (wy*wx)*support, sums in ky,kx order.
constexpr int RD_T = 16;             // 16x16 outputs per block
constexpr int RD_IN = 2 * RD_T + 3;  // 35x35 input tile

template <class Src>
__global__ void __launch_bounds__(256) k_reduce(Src src, int dw, int dh, float4* __restrict__ G,
                                                float2* __restrict__ M) {
    __shared__ float4 gt[RD_IN * RD_IN];
    __shared__ float2 mt[RD_IN * RD_IN];
    const int ox = blockIdx.x * RD_T, oy = blockIdx.y * RD_T;
    const int ix0 = 2 * ox - 2, iy0 = 2 * oy - 2;
    for (int k = threadIdx.x; k < RD_IN * RD_IN; k += 256) {
        int gx = ix0 + k % RD_IN, gy = iy0 + k / RD_IN;
        if (gx >= 0 && gy >= 0 && gx < src.w && gy < src.h) {
            gt[k] = src.g(gx, gy);
            mt[k] = src.m(gx, gy);
        } else {
            gt[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            mt[k] = make_float2(0.0f, 0.0f);
        }
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x = ox + tx, y = oy + ty;
    if (x >= dw || y >= dh) return;
    const float KW[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, ws = 0.0f, mv = 0.0f, mws = 0.0f;
#pragma unroll
    for (int ky = 0; ky < 5; ky++) {
        const int row = (2 * ty + ky) * RD_IN + 2 * tx;
#pragma unroll
        for (int kx = 0; kx < 5; kx++) {
            const float4 gv = gt[row + kx];
            const float2 m = mt[row + kx];
            if (gv.w > 0.0f) {
                float wt = KW[ky] * KW[kx] * gv.w;
                s0 += gv.x * wt;
                s1 += gv.y * wt;
                s2 += gv.z * wt;
                ws += wt;
            }
            if (m.y > 0.0f) {
                float wt = KW[ky] * KW[kx] * m.y;
                mv += m.x * wt;
                mws += wt;
            }
        }
    }
    const size_t o = (size_t)y * dw + x;
    G[o] = ws > 0.0f ? make_float4(s0 / ws, s1 / ws, s2 / ws, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    M[o] = mws > 0.0f ? make_float2(mv / mws, 1.0f) : make_float2(0.0f, 0.0f);
}

// EXPAND of level (sw, sh) to (dw, dh) at output (x, y): expandRgb src/fuse/pyramid.zig:849-909.
//  * max(dw, dh) > 512: clamp-to-edge separable expand (expandRgbRowsSequential :1101-1149 with
//    horizontalExpandRgbRow :1748-1775).  Odd output rows use source row y/2 alone (the reference
//    fetches clamp(y/2) for both of its rows, :1127-1128).
//  * otherwise: truncated kernel renormalised per expandRgbExactNoWrap :1167-1354, including its
//    divisors for one-pixel-wide or one-pixel-tall source levels.
template <class Acc>
__device__ __forceinline__ f3 expand_at(const Acc& s, int sw, int sh, int dw, int dh, int x, int y) {
    const int i = x >> 1, j = y >> 1;
    const bool ev_x = (x & 1) == 0, ev_y = (y & 1) == 0;
    if ((dw > dh ? dw : dh) > 512) {
        auto hrow = [&](int r) -> f3 {
            if (ev_x) {
                f3 a = s(clamp_coord(i - 1, sw), r), b = s(clamp_coord(i, sw), r), c = s(clamp_coord(i + 1, sw), r);
                return mul3r(add3(add3(a, mul3(6.0f, b)), c), 0.125f);
            }
            f3 a = s(clamp_coord(i, sw), r), b = s(clamp_coord(i + 1, sw), r);
            return mul3r(add3(a, b), 0.5f);
        };
        if (ev_y) {
            f3 h0 = hrow(clamp_coord(j - 1, sh)), h1 = hrow(clamp_coord(j, sh)), h2 = hrow(clamp_coord(j + 1, sh));
            return mul3r(add3(add3(h0, mul3(6.0f, h1)), h2), 0.125f);
        }
        f3 h = hrow(clamp_coord(j, sh));
        return mul3r(add3(h, h), 0.5f);
    }
    auto tap = [&](int xx, int r) -> f3 { return (xx >= 0 && xx < sw) ? s(xx, r) : zero3(); };
    auto H = [&](int r) -> f3 {
        if (r < 0 || r >= sh) return zero3();
        if (ev_x) return add3(add3(tap(i - 1, r), mul3(6.0f, tap(i, r))), tap(i + 1, r));
        return mul3(4.0f, add3(tap(i, r), tap(i + 1, r)));
    };
    f3 num = ev_y ? add3(add3(H(j - 1), mul3(6.0f, H(j))), H(j + 1)) : add3(H(j), H(j + 1));
    float hw = ev_x ? (float)((i > 0 ? 1 : 0) + 6 + (i + 1 < sw ? 1 : 0)) : (float)(4 + (i + 1 < sw ? 4 : 0));
    float vw = ev_y ? (float)((j > 0 ? 1 : 0) + 6 + (j + 1 < sh ? 1 : 0)) : (float)(1 + (j + 1 < sh ? 1 : 0));
    float div = hw * vw;
    if (sw == 1 && sh > 1 && j <= sh - 2 && ev_x) div = ev_y ? (j == 0 ? 49.0f : 56.0f) : 14.0f;
    if (sh == 1 && sw > 1 && y == 0 && i == sw - 1) div = ev_x ? 49.0f : 28.0f;
    return div3(num, div);
}

struct GAcc {  // float4 gaussian level
    const float4* G;
    int w;
    __device__ __forceinline__ f3 operator()(int x, int y) const {
        float4 v = G[(size_t)y * w + x];
        return mk3(v.x, v.y, v.z);
    }
};
struct PlanarAcc {  // planar R,G,B planes of n pixels
    const float* p;
    size_t n;
    int w;
    __device__ __forceinline__ f3 operator()(int x, int y) const {
        size_t i = (size_t)y * w + x;
        return mk3(p[i], p[i + n], p[i + 2 * n]);
    }
};

// acc[l] += (G[l] - EXPAND(G[l+1])) * M[l]   (accumulateLapRowsSequential :1506-1531)
// top level: acc += G * M                    (accumulateGaussianRowsSequential :1533-1554)
template <class Src>
__global__ void __launch_bounds__(256) k_expand_acc(Src src, const float4* __restrict__ Gn, int nw, int nh,
                                                    int top, float* __restrict__ acc) {
    const size_t n = (size_t)src.w * src.h;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % src.w), y = (int)(i / src.w);
    const float4 g = src.g(x, y);
    const float m = src.m(x, y).x;
    f3 v = mk3(g.x, g.y, g.z);
    if (!top) v = sub3(v, expand_at(GAcc{Gn, nw}, nw, nh, src.w, src.h, x, y));
    acc[i] = acc[i] + v.x * m;
    acc[i + n] = acc[i + n] + v.y * m;
    acc[i + 2 * n] = acc[i + 2 * n] + v.z * m;
}

template <typename T>
static hipError_t pass2_frame_t(const Pass2Args& a, hipStream_t s) {
    const LevelTable& lt = *a.lt;
    Level0<T> l0{(const T*)a.rgba, a.weights, a.norm_sum, a.union_sup, 1.0f / (float)a.frame_count,
                 (int)lt.lv[0].w, (int)lt.lv[0].h};
    auto Gl = [&](int l) { return a.G + lt.lv[l].off; };
    auto Ml = [&](int l) { return a.M + lt.lv[l].off; };
    auto accl = [&](int l) { return a.acc + 3 * lt.lv[l].off; };
    auto blocks1d = [](size_t n) { return dim3((unsigned)((n + 255) / 256)); };
    const int L = lt.n;
    if (L == 1) {
        hipLaunchKernelGGL((k_expand_acc<Level0<T>>), blocks1d((size_t)l0.w * l0.h), dim3(256), 0, s, l0,
                           (const float4*)nullptr, 0, 0, 1, accl(0));
        return hipGetLastError();
    }
    // gaussian + mask pyramids (buildMaskGaussianPyramidInto :384-416, buildAndAccumulate :489-522)
    for (int l = 1; l < L; l++) {
        dim3 grid((lt.lv[l].w + RD_T - 1) / RD_T, (lt.lv[l].h + RD_T - 1) / RD_T);
        if (l == 1)
            hipLaunchKernelGGL((k_reduce<Level0<T>>), grid, dim3(256), 0, s, l0, (int)lt.lv[1].w, (int)lt.lv[1].h,
                               Gl(1), Ml(1));
        else
            hipLaunchKernelGGL((k_reduce<LevelN>), grid, dim3(256), 0, s,
                               LevelN{Gl(l - 1), Ml(l - 1), (int)lt.lv[l - 1].w, (int)lt.lv[l - 1].h},
                               (int)lt.lv[l].w, (int)lt.lv[l].h, Gl(l), Ml(l));
    }
    // laplacian accumulation (:527-541)
    for (int l = 0; l < L; l++) {
        const int top = l == L - 1;
        const float4* gn = top ? nullptr : Gl(l + 1);
        const int nw = top ? 0 : (int)lt.lv[l + 1].w, nh = top ? 0 : (int)lt.lv[l + 1].h;
        const size_t n = (size_t)lt.lv[l].w * lt.lv[l].h;
        if (l == 0)
            hipLaunchKernelGGL((k_expand_acc<Level0<T>>), blocks1d(n), dim3(256), 0, s, l0, gn, nw, nh, top, accl(0));
        else
            hipLaunchKernelGGL((k_expand_acc<LevelN>), blocks1d(n), dim3(256), 0, s,
                               LevelN{Gl(l), Ml(l), (int)lt.lv[l].w, (int)lt.lv[l].h}, gn, nw, nh, top, accl(l));
    }
    return hipGetLastError();
}

hipError_t launch_pass2_frame(const Pass2Args& a, hipStream_t s) {
    return a.sample_bytes == 1 ? pass2_frame_t<uint8_t>(a, s) : pass2_frame_t<uint16_t>(a, s);
}

template <class Src>
__global__ void k_level0_debug(Src src, float4* G0, float2* M0) {
    const size_t n = (size_t)src.w * src.h;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % src.w), y = (int)(i / src.w);
    G0[i] = src.g(x, y);
    M0[i] = src.m(x, y);
}

hipError_t launch_level0_debug(const Pass2Args& a, float4* G0, float2* M0, hipStream_t s) {
    const LevelTable& lt = *a.lt;
    size_t n = (size_t)lt.lv[0].w * lt.lv[0].h;
    dim3 grid((unsigned)((n + 255) / 256));
    float def = 1.0f / (float)a.frame_count;
    if (a.sample_bytes == 1)
        hipLaunchKernelGGL((k_level0_debug<Level0<uint8_t>>), grid, dim3(256), 0, s,
                           Level0<uint8_t>{(const uint8_t*)a.rgba, a.weights, a.norm_sum, a.union_sup, def,
                                           (int)lt.lv[0].w, (int)lt.lv[0].h},
                           G0, M0);
    else
        hipLaunchKernelGGL((k_level0_debug<Level0<uint16_t>>), grid, dim3(256), 0, s,
                           Level0<uint16_t>{(const uint16_t*)a.rgba, a.weights, a.norm_sum, a.union_sup, def,
                                            (int)lt.lv[0].w, (int)lt.lv[0].h},
                           G0, M0);
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

hipError_t launch_collapse(float* acc, const LevelTable* lt, hipStream_t s) {
    for (int l = lt->n - 1; l >= 1; l--) {
        const Level& p = lt->lv[l - 1];
        const Level& c = lt->lv[l];
        size_t n = (size_t)p.w * p.h;
        hipLaunchKernelGGL(k_collapse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, acc + 3 * p.off, (int)p.w,
                           (int)p.h, (const float*)(acc + 3 * c.off), (int)c.w, (int)c.h);
    }
    return hipGetLastError();
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

// out[l] = lap_soft + (lap_hard - lap_soft) * mix for l in [hs, he), lap_soft otherwise; top = gaussian soft.
__global__ void k_lap_mix(const float* __restrict__ gs, const float* __restrict__ gsn, const float* __restrict__ gh,
                          const float* __restrict__ ghn, int w, int h, int nw, int nh, int top, int mixed, float mix,
                          float* __restrict__ out) {
    const size_t n = (size_t)w * h;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % w), y = (int)(i / w);
    f3 s = mk3(gs[i], gs[i + n], gs[i + 2 * n]);
    if (top) {
        out[i] = s.x;
        out[i + n] = s.y;
        out[i + 2 * n] = s.z;
        return;
    }
    f3 ls = sub3(s, expand_at(PlanarAcc{gsn, (size_t)nw * nh, nw}, nw, nh, w, h, x, y));
    if (mixed) {
        f3 hv = mk3(gh[i], gh[i + n], gh[i + 2 * n]);
        f3 lh = sub3(hv, expand_at(PlanarAcc{ghn, (size_t)nw * nh, nw}, nw, nh, w, h, x, y));
        ls = add3(ls, mul3r(sub3(lh, ls), mix));
    }
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
    hipError_t e = launch_collapse(out_pyr, lt, s);
    if (e != hipSuccess) return e;
    return launch_quantize(out_pyr, l0.w, l0.h, sb, out, s);
}

}  // namespace fg
