// Device helpers shared by the HIP kernels: pixel mapping and bilinear sampling of remap,
// the reference's two expand regimes, and small float3 arithmetic.  Compiled with
// -ffp-contract=off: every operation rounds where the reference rounds.
#pragma once

#include <math.h>

#include "fg_internal.h"
#include "fg_seqsum.h"

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
// `sens` (when given): how far the undistorted point moves per pixel that (dx, dy) moves, to first
// order: the larger of the radial derivative 1 / (d (r factor(r)) / dr) and the tangential scale
// radius / q at the solution.
__device__ __forceinline__ void invert_radial(const RemapParams& p, double dx, double dy, double* ux, double* uy,
                                              double* sens = nullptr) {
    double q = sqrt(dx * dx + dy * dy);
    if (q < 1e-12) {
        *ux = dx;
        *uy = dy;
        if (sens) *sens = 1.0;
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
    if (sens) {
        const double nr2 = radius * radius * p.inv_norm, nr4 = nr2 * nr2;
        const double factor = 1.0 + p.ra * nr2 + p.rb * nr4 + p.rc * nr4 * nr2;
        const double fd = p.ra + 2.0 * p.rb * nr2 + 3.0 * p.rc * nr4;
        const double df = fabs(factor + 2.0 * radius * radius * p.inv_norm * fd);
        *sens = fmax(scale, df > 1e-3 ? 1.0 / df : 1e3);
    }
}

// Source coordinate of output pixel (x, y).  Returns false when the reference writes zeros
// without sampling (|denom| < 1e-12 on the lens path, src/remap.zig:420-422; rays behind
// the camera on the basic path, whose legacy chain maps them 1.6e16 px away).
//
// The lens path: the reference walks each output row from the ROI's left edge, adding the
// homography's first column once per pixel (remapU8RowsNoTranslationLens / remapU16RowsNoTranslationLens,
// src/remap.zig:397-437, 497-537), so pixel x sees x roundings.  WALK forms the rows that way
// (seq_add); otherwise they are the products B (x - cx, y - cy, -d), and `err` (when given) receives
// a bound on |sx - sx_walk| + |sy - sy_walk| in source pixels.
template <bool WALK = false>
__device__ __forceinline__ bool map_pixel(const RemapParams& p, int x, int y, double* sx, double* sy,
                                          double* err = nullptr) {
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
    double lx, ly, lz;
    if (WALK && p.mode == REMAP_LENS) {
        const double bx = (double)p.roi_left - p.dcx;
        lx = seq_add(p.B[0] * bx + p.B[1] * py + p.B[2] * rz, p.B[0], x);
        ly = seq_add(p.B[3] * bx + p.B[4] * py + p.B[5] * rz, p.B[3], x);
        lz = seq_add(p.B[6] * bx + p.B[7] * py + p.B[8] * rz, p.B[6], x);
    } else {
        lx = p.B[0] * px + p.B[1] * py + p.B[2] * rz;
        ly = p.B[3] * px + p.B[4] * py + p.B[5] * rz;
        lz = p.B[6] * px + p.B[7] * py + p.B[8] * rz;
    }
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
    double dr = 0.0;  // bound on |rx - rx_walk| + |ry - ry_walk|
    if (!WALK && err && p.mode == REMAP_LENS) {
        // each walked row value differs from its product form by at most half an ulp of the row's
        // largest |value| per addition (x of them: the values are monotone along the row, so the
        // largest sits at an end) plus a few ulp of the terms of the start's and the products' sums
        const double apx = fabs(px), apy = fabs(py), arz = fabs(rz);
        auto dl = [&](double l, const double* b) {
            double m = fmax(fabs(l), fabs(l - (double)x * b[0]));
            m = fmax(m, fmax(fabs(b[0]) * (apx + (double)x), fmax(fabs(b[1]) * apy, fabs(b[2]) * arz)));
            return ((double)x * 0.5 + 8.0) * ldexp(1.0, ilogb(fmax(m, 1e-300)) - 52);
        };
        const double dlx = dl(lx, &p.B[0]), dly = dl(ly, &p.B[3]), dlz = dl(lz, &p.B[6]);
        const double ad = fabs(denom);
        dr = p.focal * (dlx + fabs(lx / denom) * dlz) / ad + p.focal * (dly + fabs(ly / denom) * dlz) / ad;
    }
    double sens = 1.0;
    if (p.mode == REMAP_LENS && p.radial) {
        double ux, uy;
        invert_radial(p, rx, ry, &ux, &uy, !WALK && err ? &sens : nullptr);
        *sx = p.scx + ux;
        *sy = p.scy + uy;
    } else {
        *sx = p.scx + rx;
        *sy = p.scy + ry;
    }
    if (!WALK && err && p.mode == REMAP_LENS) {
        // the radial inverse scales the difference by its local sensitivity (first order; 1.25 for
        // what that leaves out) and stops its Newton steps within 1e-9 of the radius equation, so
        // one more or one fewer step moves the point by up to 1e-9 sens on each run
        *err = 1.25 * sens * dr + 2e-9 * fmax(sens, 1.0);
    }
    return true;
}

// round(v / sup) with @round's half-away-from-zero, as the reference forms it from the correctly
// rounded f64 quotient.  v * (1/sup) is within a few ulp of that quotient (|v| < 2^17), so the
// rounded integer is the same unless the quotient sits within 1e-9 of a half; there the exact
// quotient is formed.
__device__ __forceinline__ uint32_t round_div(double v, double sup, double inv) {
    const double q = v * inv;
    const double f = q - floor(q);
    if (fabs(f - 0.5) > 1e-9) return (uint32_t)round(q);
    return (uint32_t)round(v / sup);
}

// Global-address-space view of a source pointer read from memory (frame descriptors), so its
// loads are global rather than flat.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}

// samplePixelBilinear{U8,U16} with interior and support-renormalised border cases.
//
// `band` >= 0: false when a rounded quantity (a sample's quotient, or the border alpha) lies
// within `band` of its rounding tie (for the border quotients v / sup, 2 band / sup: v and sup both
// move with the coordinate), or the border support is below 1/4.
template <typename T, int CH>
__device__ __forceinline__ bool sample_bilinear(const T* __restrict__ src, int w, int h, double x, double y,
                                                uint32_t out[CH + 1], double band = -1.0) {
    const __attribute__((address_space(1))) T* px = as_global(src);
#pragma unroll
    for (int c = 0; c <= CH; c++) out[c] = 0;
    // a coordinate within band of the sampled area's edge may fall on either side of it
    if (!(x >= -1.0 && x < (double)w && y >= -1.0 && y < (double)h))
        return !(band >= 0.0 && x >= -1.0 - band && x < (double)w + band && y >= -1.0 - band && y < (double)h + band);
    bool ok = true;
    auto decided = [&](double q, double bd) {
        const double f = q - floor(q);
        ok = ok && !(fabs(f - 0.5) <= bd);
    };
    const int x0 = (int)floor(x), y0 = (int)floor(y);
    const int x1 = x0 + 1, y1 = y0 + 1;
    const double fx = x - (double)x0, fy = y - (double)y0;
    const double omx = 1.0 - fx, omy = 1.0 - fy;
    const double w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
    const int rs = w * CH;
    if (x0 >= 0 && y0 >= 0 && x1 < w && y1 < h) {
        const int b00 = y0 * rs + x0 * CH;
        const int b10 = b00 + CH, b01 = b00 + rs, b11 = b01 + CH;
        const double sup = w00 + w10 + w01 + w11;
        const double inv = 1.0 / sup;
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const double v = (double)px[b00 + c] * w00 + (double)px[b10 + c] * w10 + (double)px[b01 + c] * w01 +
                             (double)px[b11 + c] * w11;
            out[c] = round_div(v, sup, inv);
            if (band >= 0.0) decided(v * inv, band);
        }
        out[CH] = SampleMax<T>::i;
        return ok;
    }
    const bool i00 = x0 >= 0 && y0 >= 0 && x0 < w && y0 < h;
    const bool i10 = x1 >= 0 && y0 >= 0 && x1 < w && y0 < h;
    const bool i01 = x0 >= 0 && y1 >= 0 && x0 < w && y1 < h;
    const bool i11 = x1 >= 0 && y1 >= 0 && x1 < w && y1 < h;
    double sup = 0.0;
    if (i00) sup += w00;
    if (i10) sup += w10;
    if (i01) sup += w01;
    if (i11) sup += w11;
    if (sup <= 0.0) return !(band >= 0.0);
    const double inv = 1.0 / sup;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        double v = 0.0;
        if (i00) v += (double)px[(y0 * w + x0) * CH + c] * w00;
        if (i10) v += (double)px[(y0 * w + x1) * CH + c] * w10;
        if (i01) v += (double)px[(y1 * w + x0) * CH + c] * w01;
        if (i11) v += (double)px[(y1 * w + x1) * CH + c] * w11;
        out[c] = round_div(v, sup, inv);
        if (band >= 0.0) decided(v * inv, 2.0 * band / sup);
    }
    out[CH] = (uint32_t)round(sup * SampleMax<T>::d);
    if (band >= 0.0) {
        decided(sup * SampleMax<T>::d, band);
        ok = ok && sup >= 0.25;
    }
    return ok;
}

// remapRigidImage's sample for output pixel (x, y): RGB(A) into o.  The lens path takes its row
// values as products, and walks the row (map_pixel<true>) only where a rounding tie lies within
// reach of their difference: a quotient moves by at most the sample range times the coordinate
// change (|dv/dx|, |dv/dy| <= max over the taps' differences, the support weights summing to 1),
// and so does the border alpha.
template <typename T, int CH>
__device__ __forceinline__ void remap_sample(const RemapParams& p, const T* __restrict__ src, int sw, int sh, int x,
                                             int y, uint32_t o[CH + 1]) {
    double sx, sy, err = 0.0;
    if (map_pixel(p, x, y, &sx, &sy, &err)) {
        if (p.mode != REMAP_LENS) {
            sample_bilinear<T, CH>(src, sw, sh, sx, sy, o);
            return;
        }
        if (sample_bilinear<T, CH>(src, sw, sh, sx, sy, o, 1.25 * SampleMax<T>::d * err + 1e-9)) return;
    } else if (p.mode != REMAP_LENS) {
#pragma unroll
        for (int c = 0; c <= CH; c++) o[c] = 0;
        return;
    }
    if (map_pixel<true>(p, x, y, &sx, &sy)) {
        sample_bilinear<T, CH>(src, sw, sh, sx, sy, o);
    } else {
#pragma unroll
        for (int c = 0; c <= CH; c++) o[c] = 0;
    }
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

// normalizeWeightsInto src/fuse/pyramid.zig:303-316
__device__ __forceinline__ float normalized_mask(float w, float s, float def) { return s > 1e-12f ? w / s : def; }

struct PlanarAcc {  // planar R,G,B planes of n pixels
    const float* p;
    size_t n;
    int w;
    __device__ __forceinline__ f3 operator()(int x, int y) const {
        size_t i = (size_t)y * w + x;
        return mk3(p[i], p[i + n], p[i + 2 * n]);
    }
};

}  // namespace fg
