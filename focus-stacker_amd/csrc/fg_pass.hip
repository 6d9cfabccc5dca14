// Pass 1 and pass 2 of a stack as frame-grouped tile kernels (gfx950, wave64, 256 threads).
//
// Pass 1 (k_pass1): for each 64x16 output tile, frames k0..k1-1 in order: the tile plus the
// contrast halo is remapped (f64) into LDS, the RGBA tile goes to the frame cache, gray and the
// local-contrast standard deviation are formed from LDS with f64 column sums, and the weight
// (contrast * alpha) goes to the weight cache.  The norm sum, union support and hybrid winners
// of the tile stay in registers across the frames and are written once.
//
// Pass 2 (k_pass2): for each level l and each tile of level l, frames k0..k1-1 in order: the
// level-l region (tile + 4-pixel halo) is loaded into LDS, the support-aware 5x5 reduce builds
// the level-(l+1) gaussian and mask of the tile (+1 halo) in LDS and writes their interior for
// the next level, and the laplacian G[l] - EXPAND(G[l+1]) times the mask is added to the
// accumulator held in registers.  The accumulator tile is read and written once per launch.
// Levels too small to fill the GPU with tiles run one frame per block (grid.z) and write each
// frame's product to scratch; k_ordered_sum then adds them to the accumulator in frame order.
//
// Every f32 value is rounded where the reference rounds and sums keep its order (frame order
// for the norm sum and the accumulators, ky-then-kx order in the reduce), so results match the
// reference's single-threaded arithmetic; see fg_device.h for the expand and remap formulas.
#include <type_traits>

#include "fg_device.h"
#include "fg_tuning.h"

namespace fg {

// ---------------------------------------------------------------------------
// packed RGBA samples in LDS
// ---------------------------------------------------------------------------

template <typename T>
struct Pack;
template <>
struct Pack<uint8_t> {
    using type = uint32_t;
    __device__ static __forceinline__ type make(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
        return r | (g << 8) | (b << 16) | (a << 24);
    }
    __device__ static __forceinline__ uint32_t ch(type v, int c) { return (v >> (8 * c)) & 0xffu; }
};
template <>
struct Pack<uint16_t> {
    using type = uint2;
    __device__ static __forceinline__ type make(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
        return make_uint2(r | (g << 16), b | (a << 16));
    }
    __device__ static __forceinline__ uint32_t ch(type v, int c) {
        return c < 2 ? ((v.x >> (16 * c)) & 0xffffu) : ((v.y >> (16 * (c - 2))) & 0xffffu);
    }
};

// RGBA of ROI pixel (x, y) of a frame: remapped source (remapRigidImage) or aligned frame
// (alpha = full support when the frame has none, which leaves applySupportInto a no-op).
// LENS: the launch holds frames of the lens row walk, whose exact remap (the walk redone where the
// direct products land near a rounding tie) is compiled in only then: its registers would otherwise
// count against every pass-1 launch through the out-of-line exact path below.
template <typename T, bool LENS>
__device__ __forceinline__ void frame_rgba(const FrameDesc& f, int x, int y, uint32_t o[4]) {
    if (f.kind == 0) {
        if constexpr (LENS) {
            remap_sample<T, 3>(f.p, (const T*)f.src, (int)f.sw, (int)f.sh, x, y, o);
        } else {
            double sx, sy;
            if (map_pixel(f.p, x, y, &sx, &sy)) {
                sample_bilinear<T, 3>((const T*)f.src, (int)f.sw, (int)f.sh, sx, sy, o);
            } else {
                o[0] = o[1] = o[2] = o[3] = 0;
            }
        }
    } else {
        const int st = 3 + f.extra;
        const T* q = (const T*)f.src + ((size_t)y * f.sw + x) * st;
        o[0] = q[0];
        o[1] = q[1];
        o[2] = q[2];
        o[3] = f.extra ? (uint32_t)q[3] : SampleMax<T>::i;
    }
}

// frame_rgba out of line, for the rare pixels the fast remap path hands back
template <typename T, bool LENS>
__device__ __attribute__((noinline)) uint64_t frame_rgba_exact(const FrameDesc& f, int x, int y) {
    uint32_t o[4];
    frame_rgba<T, LENS>(f, x, y, o);
    return (uint64_t)o[0] | ((uint64_t)o[1] << 16) | ((uint64_t)o[2] << 32) | ((uint64_t)o[3] << 48);
}

// ---------------------------------------------------------------------------
// pass 1
// ---------------------------------------------------------------------------
// A block owns a region of 64 columns x RH rows of the ROI: an output tile of (64 - 2b) x
// (RH - 2b) pixels plus the contrast halo b = window / 2.  Wave w remaps region rows w, w+4, ...
// (lane = column), so every LDS row is one wave and every global store is one contiguous run.

constexpr int P1C = 64;

int pass1_rows(int window) { return window / 2 <= 4 ? 32 : 64; }

// dynamic LDS of k_pass1 (byte offsets of every array in Pass1Lds): region gray and alpha (one
// byte per pixel for u8 frames), the tile's running norm sum and union support (a byte per pixel;
// for the u8 hybrid one 64-bit lane mask per tile row; neither for the single-pass methods) and, for
// the hybrid and hardmask methods, the winning weight and either the winner's frame slot (frames
// cached: its colour is read back from the frame cache at the end) or the region colour (frames
// not cached: a new winner's colour is written to the hard image right away).  At u8 and windows
// 5 to 9 the cached hybrid's arrays fit 40 KB, four blocks per CU.
struct Pass1Lds {
    size_t gray, alpha, colour, stage, tS, tU, tB, tK, total;
};
// staged source box: 4 waves x StageGeom::NR rows of 64 x StageGeom::NC dwords
constexpr size_t kStageBytesU8 = 4 * 12 * 64 * 1 * 4, kStageBytesU16 = 4 * 10 * 64 * 2 * 4;
__host__ __device__ inline Pass1Lds pass1_lds(int rh, int th, int sb, bool hybrid, bool cached, bool acc, bool pack) {
    Pass1Lds o;
    const size_t reg = (size_t)rh * P1C, til = (size_t)th * P1C, pt = sb == 1 ? 4 : 8;
    const bool colour = hybrid && !cached;
    size_t at = 0;
    o.stage = at;  at += sb == 1 ? kStageBytesU8 : kStageBytesU16;
    o.gray = at;   at += reg * 4;
    o.tS = at;     at += acc ? til * 4 : 0;
    o.tB = at;     at += hybrid ? til * 4 : 0;
    o.colour = at; at += colour ? reg * pt : 0;
    o.tK = at;     at += hybrid && cached ? til * 2 : 0;
    o.alpha = at;  at += reg * (pack ? 1 : 2);
    o.tU = at;     at += acc ? (pack || sb == 2 ? (size_t)th * 8 : til) : 0;
    o.total = (at + 15) & ~(size_t)15;
    return o;
}

// Hybrid winners are kept as 16-bit frame slots (0xffff = none) when the frames of the launch are
// cached; the kernel and its launch decide this the same way.
__host__ __device__ inline bool pass1_slot_winners(const Pass1Args& a) {
    return a.rgba != nullptr && a.k1 - a.slot_base < 0xffff;
}

size_t pass1_lds_bytes(int window, int sb, bool hybrid, bool cached, bool acc, bool pack) {
    const int b = window / 2, rh = pass1_rows(window);
    return pass1_lds(rh, rh - 2 * b, sb, hybrid, cached, acc, pack).total;
}

// 1 / d to within a few ulp: hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp_f64(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// a / b correctly rounded from y = RN(1/b) by one Markstein correction step (P. Markstein,
// "Computation of elementary functions on the IBM RISC System/6000 processor", IBM J. R&D 34,
// 1990): q = a*y, r = a - q*b (exact with an fma), RN(q + r*y).
__device__ __forceinline__ double div_markstein(double a, double b, double y) {
    const double q = a * y;
    const double r = fma(-q, b, a);
    return fma(r, y, q);
}
__device__ __forceinline__ float div_markstein(float a, float b, float y) {
    const float q = a * y;
    const float r = fmaf(-q, b, a);
    return fmaf(r, y, q);
}

// sqrt(x), correctly rounded, for normal x >= 2^-767: an rsq estimate refined by one Goldschmidt
// step (s ~ sqrt(x), h ~ 1/(2 sqrt(x))) and two Newton corrections of s by its residual x - s^2.
// These are the operations the compiler's own f64 sqrt runs, which first rescales inputs below
// 2^-767 by 2^256 (and the result back) and passes 0 and inf through; for x >= 2^-767 the rescale is
// by 2^0, so the bits are the same.
__device__ __forceinline__ double sqrt_normal(double x) {
    const double g = __builtin_amdgcn_rsq(x);
    double s = x * g, h = g * 0.5;
    const double r = fma(-h, s, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    double d = fma(-s, s, x);
    s = fma(d, h, s);
    d = fma(-s, s, x);
    return fma(d, h, s);
}

// LDS-only workgroup barrier: orders LDS accesses across the block without waiting for the
// block's outstanding global stores (no thread reads another thread's global stores in a launch).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- source staging: the footprint of a block's region in one frame's source ----
// Homography frames have a convex footprint (the image of the region's corners), so its
// bounding box, read row by row with coalesced dword loads, holds every interior tap.  The box
// is prefetched into registers one frame ahead and written to LDS over the window-column sums.
template <typename T>
struct StageGeom;
template <>
struct StageGeom<uint8_t> {
    static constexpr int NR = 12, NC = 1;  // row slots per wave, dword slots per lane
};
template <>
struct StageGeom<uint16_t> {
    static constexpr int NR = 10, NC = 2;
};

// Staged rows sit at a fixed pitch of 64 * NC dwords (the widest box), whatever the box's width:
// the second tap row of a pixel is then its first plus a constant.
template <typename T>
constexpr int kStagePitch = 64 * StageGeom<T>::NC;
static_assert(kStageBytesU8 == 4 * StageGeom<uint8_t>::NR * 64 * StageGeom<uint8_t>::NC * 4, "staging size");
static_assert(kStageBytesU16 == 4 * StageGeom<uint16_t>::NR * 64 * StageGeom<uint16_t>::NC * 4, "staging size");

struct StageBox {
    int ok;          // frame staged
    int x_lo, y_lo;  // first staged source column and row
    int rows, dw;    // staged rows, dwords per staged row
    int interior;    // every staged dword lies inside the source buffer (no byte-wise assembly)
    int inside;      // every region pixel maps inside the source interior and the staged box
};

// Column constants of the fast-path homography for region column gx: B * (x, 0, neg_d), with the
// first two rows scaled by the focal length (the remap loop's form; fast_coord keeps the plain one).
struct ColHomog {
    double cx, cy, cz;
};
__device__ __forceinline__ ColHomog col_homog(const RemapParams& p, int gx) {
    const double pxd = (double)(p.roi_left + gx) - p.dcx;
    return ColHomog{p.B[0] * pxd + p.B[2] * p.neg_d, p.B[3] * pxd + p.B[5] * p.neg_d, p.B[6] * pxd + p.B[8] * p.neg_d};
}
__device__ __forceinline__ ColHomog col_homog_focal(const RemapParams& p, int gx) {
    const ColHomog c = col_homog(p, gx);
    return ColHomog{p.focal * c.cx, p.focal * c.cy, c.cz};
}

// Fast-path source coordinate of region pixel (gx, gy) from its column constants; false where the
// exact path decides.
__device__ __forceinline__ bool fast_coord(const RemapParams& p, const ColHomog& ch, int gy, double* x, double* y,
                                           double* den = nullptr) {
    const double pyd = (double)(p.roi_top + gy) - p.dcy;
    const double lx = ch.cx + p.B[1] * pyd;
    const double ly = ch.cy + p.B[4] * pyd;
    const double lz = ch.cz + p.B[7] * pyd;
    const double denom = -lz;
    if (den) *den = denom;
    if (!(p.legacy ? denom > 1e-6 : fabs(denom) > 1e-6)) return false;
    const double inv = rcp_f64(denom);
    *x = p.scx + (p.focal * (lx * inv) + p.csx);
    *y = p.scy + (p.focal * (ly * inv) + p.csy);
    return true;
}

// Bounding box of the four corner coordinates in `cb` (NaN marks an invalid corner) in frame f's
// source; `cd` holds the homography denominators at the corners.
//
// `inside`: the region's footprint lies in the source interior, [m, sw-1-m] x [m, sh-1-m] for a
// margin m far above the few-ulp difference between the corner formula and the remap loop's.  The
// denominator is affine in the pixel position, so one sign and a magnitude above the fast path's
// limit at the four corners hold for the whole region, whose image is then the convex hull of the
// corner images: every region pixel's taps lie in the box (staged with a one-pixel margin) and in
// the source, and the remap loop needs no per-pixel checks but the rounding-tie band.
template <typename T>
__device__ __forceinline__ StageBox stage_box(const double* cb, const double* cd, const FrameDesc& f) {
    using G = StageGeom<T>;
    const int sw = (int)f.sw, sh = (int)f.sh;
    StageBox s{0, 0, 0, 0, 0, 0, 0};
    double mnx = cb[0], mxx = cb[0], mny = cb[1], mxy = cb[1];
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const double x = cb[2 * c], y = cb[2 * c + 1];
        ok = ok && x == x && y == y;
        mnx = fmin(mnx, x);
        mxx = fmax(mxx, x);
        mny = fmin(mny, y);
        mxy = fmax(mxy, y);
    }
    if (!ok || !(mnx > -1e6 && mxx < 1e6 && mny > -1e6 && mxy < 1e6)) return s;
    const int x_lo = max(0, (int)floor(mnx) - 1), x_hi = min(sw - 1, (int)floor(mxx) + 2);
    const int y_lo = max(0, (int)floor(mny) - 1), y_hi = min(sh - 1, (int)floor(mxy) + 2);
    if (x_lo > x_hi || y_lo > y_hi) return s;
    const int rows = y_hi - y_lo + 1;
    const int dw = (3 * (x_hi - x_lo + 1) * (int)sizeof(T) + 6) / 4;
    if (rows > 4 * G::NR || dw > 64 * G::NC) return s;
    s.ok = 1;
    s.x_lo = x_lo;
    s.y_lo = y_lo;
    s.rows = rows;
    s.dw = dw;
    // the dword-aligned span of the first and last staged rows inside [lead, end of the image)
    const uint32_t lead = (uint32_t)((uintptr_t)f.src & 3u);
    const uint32_t total = f.sw * f.sh * 3u * (uint32_t)sizeof(T) + lead;
    const uint32_t first = (lead + ((uint32_t)y_lo * f.sw + (uint32_t)x_lo) * 3u * (uint32_t)sizeof(T)) & ~3u;
    const uint32_t last =
        ((lead + ((uint32_t)(y_lo + rows - 1) * f.sw + (uint32_t)x_lo) * 3u * (uint32_t)sizeof(T)) & ~3u) + 4u * dw;
    s.interior = first >= lead && last <= total;
    constexpr double m = 1e-6;
    const bool same_sign = (cd[0] > 2e-6 && cd[1] > 2e-6 && cd[2] > 2e-6 && cd[3] > 2e-6) ||
                           (cd[0] < -2e-6 && cd[1] < -2e-6 && cd[2] < -2e-6 && cd[3] < -2e-6);
    // (sources at most 2^24 wide: the staged loop forms x0 * 3 * sizeof(T) by a 24-bit multiply)
    s.inside = same_sign && mnx >= m && mny >= m && mxx <= (double)(sw - 1) - m && mxy <= (double)(sh - 1) - m &&
               sw <= (1 << 24);
    return s;
}

// Byte offset, within the dword-aligned view of `src`, of source row y's first staged byte.
template <typename T>
__device__ __forceinline__ uint32_t stage_row_byte(const FrameDesc& f, const StageBox& s, int y, uint32_t lead) {
    return lead + ((uint32_t)y * f.sw + (uint32_t)s.x_lo) * 3u * (uint32_t)sizeof(T);
}

// Loads this thread's dword slots of the staged box into registers.  Dwords that straddle the
// ends of the source buffer are assembled from in-bounds bytes.
template <typename T>
__device__ __forceinline__ void stage_load(const FrameDesc& f, const StageBox& s, int lane, int wave,
                                           uint32_t* pf) {
    using G = StageGeom<T>;
    const uintptr_t base = (uintptr_t)f.src;
    const uint32_t lead = (uint32_t)(base & 3u);
    const __attribute__((address_space(1))) uint32_t* dwp =
        (const __attribute__((address_space(1))) uint32_t*)(base - lead);
    const uint32_t total = f.sw * f.sh * 3u * (uint32_t)sizeof(T) + lead;  // end of the image in the view
    // block-uniform: plain dword loads.  16-bit frames only: the 8-bit kernel runs at 128 VGPRs, where
    // this second copy of the loads costs more in register spills than its checks save
    if (sizeof(T) == 2 && s.interior) {
        // a wave-uniform row pointer and this lane's fixed offset in the row: scalar row arithmetic,
        // no per-load address computation.  Slots outside the box are left unset: stage_store skips
        // them under the same conditions.
#pragma unroll
        for (int ir = 0; ir < G::NR; ir++) {
            const int row = wave + 4 * ir;  // wave-uniform
            if (row >= s.rows) continue;
            const __attribute__((address_space(1))) uint32_t* rp =
                dwp + (stage_row_byte<T>(f, s, s.y_lo + row, lead) >> 2);
#pragma unroll
            for (int ic = 0; ic < G::NC; ic++) {
                const int col = lane + 64 * ic;
                if (col < s.dw) pf[ir * G::NC + ic] = rp[col];
            }
        }
        return;
    }
#pragma unroll
    for (int ir = 0; ir < G::NR; ir++) {
        const int row = wave + 4 * ir;
        const uint32_t rb = row < s.rows ? (stage_row_byte<T>(f, s, s.y_lo + row, lead) & ~3u) : 0u;
#pragma unroll
        for (int ic = 0; ic < G::NC; ic++) {
            const int col = lane + 64 * ic;
            uint32_t v = 0;
            if (row < s.rows && col < s.dw) {
                const uint32_t byte = rb + 4u * (uint32_t)col;
                if (byte >= lead && byte + 4u <= total) {
                    v = dwp[byte >> 2];
                } else {
                    const __attribute__((address_space(1))) uint8_t* bp =
                        (const __attribute__((address_space(1))) uint8_t*)(base - lead);
                    for (int i = 0; i < 4; i++)
                        if (byte + i >= lead && byte + i < total) v |= (uint32_t)bp[byte + i] << (8 * i);
                }
            }
            pf[ir * G::NC + ic] = v;
        }
    }
}

template <typename T>
__device__ __forceinline__ void stage_store(const StageBox& s, int lane, int wave, const uint32_t* pf,
                                            uint32_t* stage) {
    using G = StageGeom<T>;
#pragma unroll
    for (int ir = 0; ir < G::NR; ir++) {
        const int row = wave + 4 * ir;
#pragma unroll
        for (int ic = 0; ic < G::NC; ic++) {
            const int col = lane + 64 * ic;
            if (row < s.rows && col < s.dw) stage[row * kStagePitch<T> + col] = pf[ir * G::NC + ic];
        }
    }
}

// Frame descriptor k, read through the constant address space: the host writes the table before the
// launch and no kernel writes it, so its uniform fields load as scalars (s_load, counted by
// lgkmcnt).  Through a global pointer they would be vector loads, and waiting for them (vmcnt)
// would also wait for every global store the wave has in flight.
__device__ __forceinline__ FrameDesc frame_desc(const FrameDesc* frames, int k) {
#if defined(__HIP_DEVICE_COMPILE__)  // the host pass of this file has no constant address space
    return ((const __attribute__((address_space(4))) FrameDesc*)frames)[k];
#else
    return frames[k];
#endif
}

// A frame's fast-path constants, copied out of its descriptor once per frame (the descriptor sits
// in global memory, where the kernel's own stores could alias it for the compiler).
struct FastFrame {
    double b1f, b4f, b7;         // the homography's column for the pano y coordinate (x and y rows times focal)
    double dcy;
    double sx0, sy0;             // source centre + centre shift
    int roi_top, legacy;
    uint32_t sw;
    double wlim, hlim;           // interior bounds: x < sw - 1, y < sh - 1
};

__device__ __forceinline__ FastFrame fast_frame(const FrameDesc& f) {
    const RemapParams& p = f.p;
    return FastFrame{p.focal * p.B[1], p.focal * p.B[4], p.B[7], p.dcy, p.scx + p.csx, p.scy + p.csy, p.roi_top,
                     p.legacy, f.sw, (double)f.sw - 1.0, (double)f.sh - 1.0};
}

// The interior arithmetic of samplePixelBilinear on taps t[12] (rows y0, y0+1 x columns x0, x0+1
// x rgb) with the reference's f64 weights; false within `band` of a rounding tie.
__device__ __forceinline__ bool bilinear_taps(const uint32_t t[12], double fx, double fy, double band, uint32_t o[3]) {
    const double omx = 1.0 - fx, omy = 1.0 - fy;
    const double w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
    const double sup = ((w00 + w10) + w01) + w11;
    const double inv = rcp_f64(sup);
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double v = (((double)t[c] * w00 + (double)t[3 + c] * w10) + (double)t[6 + c] * w01) + (double)t[9 + c] * w11;
        const double q = v * inv + 0.5;  // > 0: truncation is the floor
        const double d = __builtin_amdgcn_fract(q);
        ok = ok && d > band && d < 1.0 - band;
        o[c] = (uint32_t)q;
    }
    return ok;
}

// u8 interior taps in f32, the lerp form on exact integer differences: the value is within 1e-4 of
// the exact one (weight rounding 2 x 255 x 2^-25, f32 roundings of at most 2^-17 each at
// magnitudes below 256, and their propagation), and q = value + (0.5 - 2.5e-4) within 1.16e-4 of
// exact + 0.5 - 2.5e-4 (one more rounding).  Where fract(q) < 1 - 5e-4, exact + 0.5 lies at least
// 1.34e-4 from an integer on either side, so the exact value, and so the reference's f64 value,
// rounds to the truncation of q; elsewhere the pixel is redone in f64 by bilinear_taps, whose own
// band then sends true near-ties to the exact path.
__device__ __forceinline__ bool bilinear_taps_u8(const uint32_t t[12], double fx, double fy, double band,
                                                 uint32_t o[3]) {
    const float fx32 = (float)fx, fy32 = (float)fy;
    bool near = false;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float p00 = (float)t[c], p10 = (float)t[3 + c], p01 = (float)t[6 + c], p11 = (float)t[9 + c];
        const float top = fmaf(fx32, p10 - p00, p00);
        const float bot = fmaf(fx32, p11 - p01, p01);
        const float q = fmaf(fy32, bot - top, top) + (0.5f - 2.5e-4f);  // > 0: truncation is the floor
        near = near || __builtin_amdgcn_fractf(q) >= 1.0f - 5e-4f;
        o[c] = (uint32_t)q;
    }
    return near ? bilinear_taps(t, fx, fy, band, o) : true;
}

// u16 interior taps in f64 on exact integer differences, without the weights' sum and its
// reciprocal: value = p00 + fx d0 + fy (dd + fx e), d0 = p10 - p00, dd = p01 - p00,
// e = (p11 - p01) - d0, three fused products.  Against the reference's sum(p w) / sum(w)
// (weights rounded in f64, sum(w) = 1 to within a few ulp) this differs by a few ulp of 65535,
// about 1e-10; `band` (4e-6, set by the reciprocal coordinates) is far wider, so outside it both
// round the same way, and inside it the pixel goes to the exact path.  The tie test is one compare:
// q = value + (0.5 - bw) with bw = band + 2^-30 (more than the addition's rounding, at most 2^-37
// below 65536), and where fract(q) < 1 - 2 bw, value + 0.5 lies more than `band` from an integer
// and truncating q gives its floor, the rounded value.
__device__ __forceinline__ bool bilinear_taps_lerp64(const uint32_t t[12], double fx, double fy, double band,
                                                     uint32_t o[3]) {
    const double bw = band + 0x1p-30;
    const double shift = 0.5 - bw, lim = 1.0 - 2.0 * bw;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int p00 = (int)t[c], p10 = (int)t[3 + c], p01 = (int)t[6 + c], p11 = (int)t[9 + c];
        const int d0 = p10 - p00, dd = p01 - p00, e = (p11 - p01) - d0;
        const double top = fma(fx, (double)d0, (double)p00);
        const double in = fma(fx, (double)e, (double)dd);
        const double q = fma(fy, in, top) + shift;  // > 0: truncation is the floor
        ok = ok && __builtin_amdgcn_fract(q) < lim;
        o[c] = (uint32_t)q;
    }
    return ok;
}

// The constant term of a staged tap's byte offset for rows of a multiple of 4 bytes (remap_fast,
// ALIGNED): the byte of the box's first column within its staged dword, less the offsets of the
// box's first row and column.  Formed once per frame, and opaque to the compiler, which would
// otherwise fold its column term back into the per-pixel product and lose the 24-bit multiply.
template <typename T>
__device__ __forceinline__ uint32_t aligned_stage_k(const StageBox& s, uint32_t lead) {
    constexpr uint32_t C3 = 3u * (uint32_t)sizeof(T), PB = 4u * (uint32_t)kStagePitch<T>;
    uint32_t k = (((lead + (uint32_t)s.x_lo * C3) & 3u) - (uint32_t)s.x_lo * C3) - (uint32_t)s.y_lo * PB;
    asm volatile("" : "+v"(k));
    return k;
}

// Region pixel in row gy of the homography fast path: coordinate by one reciprocal
// (fast_coord), interior taps from the staged box (or from global memory when the frame is not
// staged).  False where the exact path has to decide.  STAGED is the block-uniform choice between
// the two tap sources: a loop that only reads the staged box holds no global loads, so no
// iteration waits on the vector-memory counter, which also counts the loop's own RGBA stores.
template <typename T, bool STAGED, bool INSIDE = false, bool ALIGNED = false>
__device__ __forceinline__ bool remap_fast(const FastFrame& ff, const ColHomog& ch, const StageBox& s,
                                           const uint8_t* stage, uint32_t lead, const T* __restrict__ src, int gy,
                                           double band, uint32_t o[3], double* seed = nullptr, uint32_t kst = 0) {
    // INSIDE (staged tiles whose whole footprint is in the source interior and the box, StageBox::
    // inside): the checks below hold for every pixel and are skipped
    // the reference's c + (f * (l / denom) + shift) with the focal length folded into the homography
    // rows and fused products: a few ulp apart, which the tie band covers
    const double pyd = (double)(ff.roi_top + gy) - ff.dcy;
    const double lx = fma(ff.b1f, pyd, ch.cx);
    const double ly = fma(ff.b4f, pyd, ch.cy);
    const double denom = -fma(ff.b7, pyd, ch.cz);
    if (!INSIDE && !(ff.legacy ? denom > 1e-6 : fabs(denom) > 1e-6)) return false;
    double inv;
    if constexpr (INSIDE) {
        // 1/denom by two Newton steps from the reciprocal of this column's previous row (`seed`; the
        // denominator is affine in the row and changes little between them): the seed's relative
        // error squares at each step, so from 2^-20 or less the result is as accurate as
        // rcp_f64's.  A wave where some lane's seed is further off (the frame's first row, strong
        // perspective) starts from the hardware estimate instead.
        double r = *seed;
        double e = fma(-denom, r, 1.0);
        if (__builtin_amdgcn_ballot_w64(!(fabs(e) <= 0x1p-20))) {
            asm volatile("" ::: "memory");  // a real branch: the estimate is not speculated
            r = __builtin_amdgcn_rcp(denom);
            e = fma(-denom, r, 1.0);
        }
        r = fma(r, e, r);
        e = fma(-denom, r, 1.0);
        inv = fma(r, e, r);
        *seed = inv;
    } else {
        inv = rcp_f64(denom);
    }
    const double x = fma(lx, inv, ff.sx0);
    const double y = fma(ly, inv, ff.sy0);
    if (!INSIDE && !(x >= 0.0 && y >= 0.0 && x < ff.wlim && y < ff.hlim)) return false;
    // x, y >= 0 here: truncation is the floor, and fract(x) = x - floor(x) exactly
    const int x0 = (int)x, y0 = (int)y;
    const double fx = __builtin_amdgcn_fract(x), fy = __builtin_amdgcn_fract(y);
    uint32_t t[12];
    if (STAGED) {
        constexpr uint32_t C3 = 3u * (uint32_t)sizeof(T), PB = 4u * (uint32_t)kStagePitch<T>;
        uint32_t a0, a1;
        if constexpr (INSIDE && ALIGNED) {
            // rows of a multiple of 4 bytes: every staged row starts at the same byte of its first
            // dword, so the offset is y0 * PB + x0 * C3 + kst, a constant of the box (aligned_stage_k;
            // modulo 2^32, x0 < 2^24 for the 24-bit multiply, StageBox::inside), and the next row's
            // is PB further
            a0 = __umul24((uint32_t)x0, C3) + (uint32_t)y0 * PB + kst;
            a1 = a0 + PB;
        } else {
            if (!INSIDE && (x0 < s.x_lo || y0 < s.y_lo || y0 + 1 >= s.y_lo + s.rows)) return false;
            const uint32_t xb = (uint32_t)(x0 - s.x_lo) * C3;
            if (!INSIDE && xb + 6u * (uint32_t)sizeof(T) + 3u > 4u * (uint32_t)s.dw) return false;
            const uint32_t rowb = ff.sw * C3;
            const uint32_t g0 = lead + ((uint32_t)y0 * ff.sw + (uint32_t)s.x_lo) * C3;
            const uint32_t r0 = (uint32_t)(y0 - s.y_lo);
            a0 = r0 * PB + (g0 & 3u) + xb;
            a1 = (r0 + 1u) * PB + ((g0 + rowb) & 3u) + xb;
        }
#pragma unroll
        for (int c = 0; c < 6; c++) {
            if (sizeof(T) == 1) {
                t[c] = stage[a0 + c];
                t[6 + c] = stage[a1 + c];
            } else {
                t[c] = *(const uint16_t*)&stage[a0 + 2 * c];
                t[6 + c] = *(const uint16_t*)&stage[a1 + 2 * c];
            }
        }
    } else {
        const __attribute__((address_space(1))) T* px = as_global(src);
        const uint32_t rs = ff.sw * 3u;
        const uint32_t b00 = (uint32_t)y0 * rs + (uint32_t)x0 * 3u;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            t[c] = px[b00 + c];
            t[6 + c] = px[b00 + rs + c];
        }
    }
    if (sizeof(T) == 1) return bilinear_taps_u8(t, fx, fy, band, o);
    return bilinear_taps_lerp64(t, fx, fy, band, o);
}

// Cross-lane moves within a wave (DPP wave_shr:1 / wave_shl:1): the value of lane - 1 / lane + 1,
// 0 at the edge lane (bound_ctrl: an out-of-range source reads 0, so no old value is needed).
// Every lane of the wave must execute them.  The window sums below only feed lanes b..63-b, whose
// windows never leave the wave, so the edge value is never used.
__device__ __forceinline__ int lane_prev(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int lane_next(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ double lane_prev(double v) {
    return __hiloint2double(lane_prev(__double2hiint(v)), lane_prev(__double2loint(v)));
}
__device__ __forceinline__ double lane_next(double v) {
    return __hiloint2double(lane_next(__double2hiint(v)), lane_next(__double2loint(v)));
}

// Window sum over lanes c-b .. c+b, added in lane order (the order of the reference's column
// sums, computeRange src/fuse/contrast.zig:218-281).
template <int B, typename V>
__device__ __forceinline__ V lane_window_sum_b(V x) {  // the 2B neighbours kept: 2B moves
    V m[B], p[B];
    m[0] = lane_prev(x);
    p[0] = lane_next(x);
#pragma unroll
    for (int i = 1; i < B; i++) {
        m[i] = lane_prev(m[i - 1]);
        p[i] = lane_next(p[i - 1]);
    }
    V s = m[B - 1];
#pragma unroll
    for (int i = B - 2; i >= 0; i--) s = s + m[i];
    s = s + x;
#pragma unroll
    for (int i = 0; i < B; i++) s = s + p[i];
    return s;
}

template <typename V>
__device__ __forceinline__ V lane_window_sum(V x, int b) {
    if (b == 2) return lane_window_sum_b<2>(x);
    if (b == 1) return lane_window_sum_b<1>(x);
    if (b == 3) return lane_window_sum_b<3>(x);
    // wider windows: forward sums x(c) + ... + x(c+2b), then moved back by b lanes
    V s = x, y = x;
    for (int i = 0; i < 2 * b; i++) {
        y = lane_next(y);
        s = s + y;
    }
    for (int i = 0; i < b; i++) s = lane_prev(s);
    return s;
}

// COL: hybrid or hardmask winners without a frame cache (the region colour in LDS, new winners
// written to the hard image at once); compiled apart so that the other launches carry none of it.
// PACK: the u8 hybrid with its frames cached, whose LDS is packed (alpha bytes, union support as
// row lane masks) to keep four blocks per CU; likewise compiled apart.
template <typename T, int RH, bool COL, bool PACK, bool LENS>
__device__ __forceinline__ void pass1_tile(const Pass1Args& a, const int bx, const int by) {
    using P = Pack<T>;
    using PT = typename P::type;
    using G = StageGeom<T>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = (int)a.window / 2;
    const int th = RH - 2 * b, tw = P1C - 2 * b;
    const int rpw = (th + 3) / 4;
    const bool acc = a.norm_sum != nullptr;
    const bool hyb = a.best != nullptr;
    // hybrid with the frames cached: the winner is kept as its frame slot, its colour read back
    // from the frame cache when the tile is written (the RGBA the region colour would have held)
    const bool hyb_slot = !COL && hyb;  // the launch picks COL whenever there are winners and no cache
    const Pass1Lds L = pass1_lds(RH, th, (int)sizeof(T), hyb, pass1_slot_winners(a), a.norm_sum != nullptr, PACK);
    float* gray = (float*)(smem + L.gray);            // [RH][64] gray, -1 where alpha == 0 or outside
    using AT = std::conditional_t<PACK, uint8_t, uint16_t>;
    AT* alph = (AT*)(smem + L.alpha);                 // [RH][64] alpha
    PT* colour = (PT*)(smem + L.colour);              // [RH][64] hybrid: packed rgb
    uint32_t* stage = (uint32_t*)(smem + L.stage);    // staged source box of the current frame
    float* tS = (float*)(smem + L.tS);                // [th][64] norm sum of the tile
    // union support of the tile: one 64-bit lane mask per tile row for 16-bit samples and the u8
    // hybrid (whose staging / winner arrays fill LDS: the mask keeps them to four blocks per CU),
    // else a byte per pixel
    const bool umask = acc && (PACK || sizeof(T) == 2);
    uint8_t* tUb = smem + L.tU;                       // [th][64]
    uint64_t* tUw = (uint64_t*)(smem + L.tU);         // [th]
    float* tB = (float*)(smem + L.tB);                // [th][64] hybrid: best weight
    uint16_t* tK = (uint16_t*)(smem + L.tK);          // [th][64] hybrid, frames cached: its slot
    __shared__ double corners[8];                     // wave 0: region corners in a frame's source
    __shared__ double cden[4];                        // the homography denominators there
    __shared__ StageBox sbox[2];                      // the staging boxes of frames k, k+1 (by k & 1)
    __shared__ int sFull[4];                          // per wave: every region pixel it remapped has full alpha

    const int W = (int)a.w, H = (int)a.h;
    const int X0 = bx * tw, Y0 = by * th;
    // the wave index made wave-uniform: row arithmetic on it (loop bounds, staging rows, tile rows)
    // then runs on the scalar unit
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const bool contrast_ok = W >= (int)a.window && H >= (int)a.window;
    const size_t fpx = (size_t)W * H;
    const int r0 = wave * rpw;                        // this wave's output rows [r0, r0 + rpw) of the tile
    const int r1 = r0 + rpw < th ? r0 + rpw : th;
    const int ox = X0 + lane - b;                     // output column of this lane
    const bool col_ok = lane >= b && lane < P1C - b && ox < W;
    const int nfull = (int)(a.window * a.window);
    const double nf_full = (double)nfull, nf_full1 = (double)(nfull - 1);
    const double r_full = 1.0 / nf_full, r_full1 = 1.0 / nf_full1;
    const double band = sizeof(T) == 1 ? 1e-8 : 4e-6;  // covers the reciprocal coordinates' error
    const int gx = X0 - b + lane;                     // region column of this lane
    const bool keep_rgba = a.rgba != nullptr, keep_w = a.weights != nullptr;

    // The staging box of the region in frame k's source, formed by wave 0 alone and published in
    // slot k & 1: threads 0-3 map the region's corners (NaN where not staged), the wave takes their
    // bounding box (stage_box), lane 0 stores it.
    auto stage_next = [&](int k) {
        if (wave != 0) return;
        // the corner's position is formed anew for every frame: held across the frame loop, it was
        // spilled, and its reload waited (vmcnt) for every store of wave 0 before the barrier
        int tid = lane;
        asm volatile("" : "+v"(tid));
        if (tid < 4) {
            double x = __builtin_nan(""), y = __builtin_nan(""), d = 0.0;
            if (k < a.k1) {
                const FrameDesc f = frame_desc(a.frames, k);
                if (f.kind == 0 && f.p.mode == REMAP_HOMOGRAPHY) {
                    const int cx = X0 - b + ((tid & 1) ? P1C - 1 : 0);
                    const int cy = Y0 - b + ((tid & 2) ? RH - 1 : 0);
                    double sx, sy;
                    if (fast_coord(f.p, col_homog(f.p, cx), cy, &sx, &sy, &d)) {
                        x = sx;
                        y = sy;
                    }
                }
            }
            corners[2 * tid] = x;
            corners[2 * tid + 1] = y;
            cden[tid] = d;
        }
        __builtin_amdgcn_wave_barrier();  // one wave: its LDS accesses take effect in order
        const StageBox bx = k < a.k1 ? stage_box<T>(corners, cden, frame_desc(a.frames, k))
                                     : StageBox{0, 0, 0, 0, 0, 0, 0};
        if (lane == 0) sbox[k & 1] = bx;
        __builtin_amdgcn_wave_barrier();  // every lane has read the corners before the next frame's
    };

#pragma unroll 1
    for (int r = r0; r < r1; r++) {
        const int y = Y0 + r;
        float sv = 0.0f, bv = 0.0f;
        uint32_t uv = 0;
        if (col_ok && y < H) {
            const size_t i = (size_t)y * W + ox;
            if (acc) {
                sv = a.norm_sum[i];
                uv = a.union_sup[i];
            }
            if (hyb) bv = a.best[i];
        }
        if (acc) {  // the single-pass methods have no norm sum: tS and tU take no LDS
            tS[r * P1C + lane] = sv;
            if (umask) {
                const uint64_t um = __ballot(uv != 0);  // every lane of the wave is here
                if (lane == 0) tUw[r] = um;
            } else {
                tUb[r * P1C + lane] = (uint8_t)uv;
            }
        }
        if (hyb) tB[r * P1C + lane] = bv;
        if (hyb_slot) tK[r * P1C + lane] = 0xffffu;   // no winner in this launch yet
    }

    // stage the first frame
    uint32_t pf[G::NR * G::NC];
    stage_next(a.k0);
    stage_next(a.k0 + 1);
    lds_barrier();
    // frames are staged where the region lies inside the stack and its footprint inside the source
    // (StageBox::inside): the staged remap loop then needs no per-pixel checks.  The other tiles
    // (a border of the stack, a source edge) gather their taps from global memory.
    const bool region_in = X0 - b >= 0 && Y0 - b >= 0 && X0 - b + P1C <= W && Y0 - b + RH <= H;
    const FrameDesc f0 = frame_desc(a.frames, a.k0);
    StageBox box = sbox[a.k0 & 1];
    box.ok = box.ok && box.inside && region_in;
    if (box.ok) {
        stage_load<T>(f0, box, lane, wave, pf);
        stage_store<T>(box, lane, wave, pf, stage);
    }
    lds_barrier();

    for (int k = a.k0; k < a.k1; k++) {
        // the next frame's source box (its corners were published by the previous barrier): its
        // loads are issued before this frame's stores and are in flight through the remap, so
        // waiting for them (before the contrast) does not wait for the contrast's stores
        const FrameDesc fn = frame_desc(a.frames, k + 1 < a.k1 ? k + 1 : k);
        StageBox nbox = sbox[(k + 1) & 1];  // none past the last frame
        nbox.ok = nbox.ok && nbox.inside && region_in;
        if (nbox.ok) stage_load<T>(fn, nbox, lane, wave, pf);
        const FrameDesc f = frame_desc(a.frames, k);
        const size_t slot = (size_t)(k - a.slot_base);
        T* rgba = keep_rgba ? (T*)a.rgba + slot * fpx * 4 : nullptr;
        float* wout = keep_w ? a.weights + slot * fpx : nullptr;
        const bool fast = f.kind == 0 && f.p.mode == REMAP_HOMOGRAPHY;
        const uint32_t lead = (uint32_t)((uintptr_t)f.src & 3u);
        const FastFrame ff = fast_frame(f);
        const T* fsrc = (const T*)f.src;
        const FrameDesc* fptr = a.frames + k;
        const ColHomog chm = fast ? col_homog_focal(f.p, gx) : ColHomog{0.0, 0.0, 0.0};
        // remap the region (remapRigidImage); gray (fillAverageFromLoaded) where alpha > 0.  One
        // loop per tap source (staged box or global gather), box.ok being the same for the block.
        // every region pixel inside the stack and at full alpha (a wave vote after the remap): the
        // contrast then needs no support tests and every window holds window^2 samples
        bool lane_full = true;
        // region pixel (r, gy) of this lane as sample o: gray, alpha, hybrid colour, RGBA cache.  IN:
        // the region lies inside the stack and o came from the fast path (alpha full)
        auto emit = [&](int r, int gy, const uint32_t o[4], auto in_c) {
            constexpr bool IN = decltype(in_c)::value;
            float g = -1.0f;
            if (IN || o[3] > 0) {
                const float sc = (float)((o[0] + o[1]) + o[2]);
                g = div_markstein(sc, SampleMax<T>::gray3, 1.0f / SampleMax<T>::gray3);
            }
            const PT c = P::make(o[0], o[1], o[2], o[3]);
            if (keep_rgba && r >= b && r < RH - b && lane >= b && lane < P1C - b && (IN || (gx < W && gy < H))) {
                // an output pixel of the tile; 32-bit byte offsets (frames hold fewer than 2^28 pixels)
                const uint32_t i = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
                if (sizeof(T) == 1)
                    *(uint32_t*)((char*)rgba + 4u * i) = *(const uint32_t*)&c;
                else
                    *(uint2*)((char*)rgba + 8u * i) = *(const uint2*)&c;
            }
            gray[r * P1C + lane] = g;
            alph[r * P1C + lane] = (AT)o[3];
            if (COL) colour[r * P1C + lane] = c;
            lane_full = lane_full && o[3] == SampleMax<T>::i;
        };
        // the fast path where it decides; the rest (pixels within the rounding-tie band, frames the
        // fast path does not cover) is marked in `redo` and remapped by the exact path after the
        // loop, the one call site, which keeps the call out of the loops
        uint32_t redo = 0;
        if (box.ok) {  // staged: no per-pixel checks
            // ALIGNED: source rows of a multiple of 4 bytes, whose staged rows share one alignment
            auto staged_rows = [&](auto aligned_c) {
                constexpr bool AL = decltype(aligned_c)::value;
                const uint32_t kst = AL ? aligned_stage_k<T>(box, lead) : 0u;
                double seed = 0.0;  // the previous row's reciprocal denominator (none yet)
#pragma unroll 1
                for (int r = wave, i = 0; r < RH; r += 4, i++) {
                    const int gy = Y0 - b + r;
                    uint32_t o[4] = {0, 0, 0, SampleMax<T>::i};
                    const bool ok = remap_fast<T, true, true, AL>(ff, chm, box, (const uint8_t*)stage, lead, fsrc, gy,
                                                                  band, o, &seed, kst);
                    // near-ties are rare: the mask is updated behind a wave vote
                    if (__builtin_amdgcn_ballot_w64(!ok)) {
                        asm volatile("" ::: "memory");
                        if (!ok) redo |= 1u << i;
                    }
                    emit(r, gy, o, std::true_type{});
                }
            };
            if (((ff.sw * 3u * (uint32_t)sizeof(T)) & 3u) == 0)
                staged_rows(std::true_type{});
            else
                staged_rows(std::false_type{});
        } else {
#pragma unroll 1
            for (int r = wave, i = 0; r < RH; r += 4, i++) {
                const int gy = Y0 - b + r;
                uint32_t o[4] = {0, 0, 0, 0};
                if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
                    if (fast && remap_fast<T, false>(ff, chm, box, (const uint8_t*)stage, lead, fsrc, gy, band, o))
                        o[3] = SampleMax<T>::i;
                    else
                        redo |= 1u << i;
                }
                emit(r, gy, o, std::false_type{});
            }
        }
#pragma unroll 1
        for (int r = wave, i = 0; redo != 0; r += 4, i++) {
            if (redo & (1u << i)) {
                redo &= ~(1u << i);
                const int gy = Y0 - b + r;
                const uint64_t e = frame_rgba_exact<T, LENS>(*fptr, gx, gy);
                const uint32_t o[4] = {(uint32_t)e & 0xffffu, (uint32_t)(e >> 16) & 0xffffu,
                                       (uint32_t)(e >> 32) & 0xffffu, (uint32_t)(e >> 48)};
                emit(r, gy, o, std::false_type{});
            }
        }
        {
            const bool wf = __all(lane_full ? 1 : 0);
            if (lane == 0) sFull[wave] = wf ? 1 : 0;
        }
        stage_next(k + 2);  // its slot held frame k's box, read by every wave before the last barrier
        lds_barrier();
        const bool full = contrast_ok && sFull[0] && sFull[1] && sFull[2] && sFull[3];
        // every wave is done with this frame's staged box: the next one goes in
        box = nbox;
        if (box.ok) stage_store<T>(box, lane, wave, pf, stage);
        // contrast of this wave's output rows: window-column sums slide down the column in
        // registers (lane = column), the window's columns are added across lanes
        // (computeRange src/fuse/contrast.zig:172-282: f64 sums of x and of the f32 product x*x).
        // Compiled for the default window 5 (BW = 2) and for any window (BW = 0: b at run time),
        // each with and without the fully supported region (FULL: every sample counts, so the
        // count is window^2 and alpha is full; the sums take the same operations either way).
        auto contrast_rows = [&](auto bw, auto full_c) {
            constexpr int BW = decltype(bw)::value;
            constexpr bool FULL = decltype(full_c)::value;
            const int bb = BW > 0 ? BW : b;
            auto wsum = [&](auto x) {
                if constexpr (BW > 0)
                    return lane_window_sum_b<BW>(x);
                else
                    return lane_window_sum(x, bb);
            };
            double vs = 0.0, vq = 0.0;
            int vn = 0;
#pragma unroll 1
            for (int r = r0; r < r1; r++) {
                if (contrast_ok) {
                    if (r == r0) {
                        for (int dy = 0; dy <= 2 * bb; dy++) {
                            const float v = gray[(r0 + dy) * P1C + lane];
                            if (FULL || v >= 0.0f) {
                                vs += (double)v;
                                vq += (double)(v * v);
                                vn++;
                            }
                        }
                    } else {
                        const float vi = gray[(r + 2 * bb) * P1C + lane];
                        const float vo = gray[(r - 1) * P1C + lane];
                        if (FULL || vi >= 0.0f) {
                            vs += (double)vi;
                            vq += (double)(vi * vi);
                            vn++;
                        }
                        if (FULL || vo >= 0.0f) {
                            vs -= (double)vo;
                            vq -= (double)(vo * vo);
                            vn--;
                        }
                    }
                }
                double sw = 0.0, qw = 0.0;
                int n = FULL ? nfull : 0;
                if (contrast_ok) {  // every lane takes part in the cross-lane sums
                    sw = wsum(vs);
                    qw = wsum(vq);
                    if (!FULL) n = wsum(vn);
                }
                const int y = Y0 + r;
                if (umask) {  // accumulateBinarySupportMax as one lane-mask OR per row
                    const uint64_t um = __ballot(col_ok && y < H && (FULL || alph[(r + bb) * P1C + lane] > 0));
                    if (lane == 0) atomicOr((unsigned long long*)&tUw[r], (unsigned long long)um);
                }
                if (!col_ok || y >= H) continue;
                const size_t i = (size_t)y * W + ox;
                float stdw = 0.0f;
                if (contrast_ok && ox >= bb && ox < W - bb && y >= bb && y < H - bb &&
                    (FULL || (gray[(r + bb) * P1C + lane] >= 0.0f && n > 1))) {  // sampleStdDev :284-290
                    double m, var;
                    if (FULL || n == nfull) {
                        m = div_markstein(sw * sw, nf_full, r_full);
                        var = div_markstein(qw - m, nf_full1, r_full1);
                    } else {
                        const double nf = (double)n, den = (double)(n - 1);
                        m = (sw * sw) / nf;
                        var = (qw - m) / den;
                    }
                    // a positive variance of 8/16-bit gray windows is far above 2^-767 (the
                    // smallest is about 2^-90); a wave holding one below takes the library's sqrt
                    if (__builtin_amdgcn_ballot_w64(var > 0.0 && var < 0x1p-767)) {
                        asm volatile("" ::: "memory");  // a real branch: never speculated
                        stdw = var > 0.0 ? (float)sqrt(var) : 0.0f;
                    } else {
                        stdw = var > 0.0 ? (float)sqrt_normal(var) : 0.0f;
                    }
                }
                const uint32_t alpha = FULL ? SampleMax<T>::i : (uint32_t)alph[(r + bb) * P1C + lane];
                const float af = FULL || alpha == SampleMax<T>::i
                                     ? 1.0f
                                     : div_markstein((float)alpha, SampleMax<T>::f, 1.0f / SampleMax<T>::f);
                const float wv = stdw * af;  // applySupportInto
                if (keep_w) *(float*)((char*)wout + 4u * ((uint32_t)y * (uint32_t)W + (uint32_t)ox)) = wv;
                const int t = r * P1C + lane;
                if (acc) {
                    tS[t] = tS[t] + wv;
                    if (!umask && alpha > 0) tUb[t] = 1;
                }
                if (hyb && wv > tB[t]) {  // updateRange: strictly greater wins, first frame on ties
                    tB[t] = wv;
                    if (!COL) {
                        tK[t] = (uint16_t)slot;
                    } else {  // only this thread writes this pixel of the hard image
                        const PT c = colour[(r + bb) * P1C + lane];
                        T* hd = (T*)a.hard + i * 3;
                        hd[0] = (T)P::ch(c, 0);
                        hd[1] = (T)P::ch(c, 1);
                        hd[2] = (T)P::ch(c, 2);
                    }
                }
            }
        };
        if (b == 2) {
            if (full)
                contrast_rows(std::integral_constant<int, 2>{}, std::true_type{});
            else
                contrast_rows(std::integral_constant<int, 2>{}, std::false_type{});
        } else {
            if (full)
                contrast_rows(std::integral_constant<int, 0>{}, std::true_type{});
            else
                contrast_rows(std::integral_constant<int, 0>{}, std::false_type{});
        }
        lds_barrier();
    }
    // the winners' RGBA were stored to the frame cache by whichever wave remapped their row: a
    // workgroup barrier with a global-memory fence makes them visible to the reading thread
    if (hyb_slot) __syncthreads();

#pragma unroll 1
    for (int r = r0; r < r1; r++) {
        const int y = Y0 + r;
        if (!col_ok || y >= H) continue;
        const size_t i = (size_t)y * W + ox;
        const int t = r * P1C + lane;
        if (acc) {
            a.norm_sum[i] = tS[t];
            a.union_sup[i] = umask ? (uint8_t)((tUw[r] >> lane) & 1u) : tUb[t];
        }
        if (hyb) a.best[i] = tB[t];
        if (hyb_slot) {
            const uint32_t k = tK[t];
            if (k != 0xffffu) {  // the winner's RGBA, stored by this block's remap of that frame
                const T* px = (const T*)a.rgba + ((size_t)k * fpx + i) * 4;
                T* hd = (T*)a.hard + i * 3;
                hd[0] = px[0];
                hd[1] = px[1];
                hd[2] = px[2];
            }
        }
    }
}

// Tile of this workgroup in a 2D grid, XCD-aware.  Workgroups are dealt to the 8 XCDs round-robin
// (linear workgroup id i on XCD i % 8), each XCD with its own L2.  Every XCD takes a contiguous run
// of tiles in an order that walks the grid in strips of FG_TUNE_XCD_STRIP tile rows, column by
// column within a strip (1: row-major): the tiles one XCD runs at the same time are neighbours,
// whose shared halo columns (and the cache lines a tile's row ends only partly use) come into that
// L2 once.  A bijection of the grid for any strip height.
__device__ __forceinline__ void xcd_tile(int* bx, int* by) {
    const int nx = (int)gridDim.x, ny = (int)gridDim.y, n = nx * ny;
    const int lin = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    const int q = n >> 3, r = n & 7, x = lin & 7, sl = lin >> 3;
    const int t = x < r ? x * (q + 1) + sl : r * (q + 1) + (x - r) * q + sl;
    constexpr int kStrip = FG_TUNE_XCD_STRIP > 1 ? FG_TUNE_XCD_STRIP : 1;
    const int st = t / (kStrip * nx), w = t - st * (kStrip * nx);
    const int h = min(kStrip, ny - st * kStrip);
    *bx = w / h;
    *by = st * kStrip + w % h;
}

template <typename T, int RH, bool COL, bool PACK, bool LENS>
__global__ void __launch_bounds__(256, 4) k_pass1(Pass1Args a) {
    int bx, by;
    xcd_tile(&bx, &by);
    pass1_tile<T, RH, COL, PACK, LENS>(a, bx, by);
}

// Pass 1 of many stacks in one persistent launch: the blocks pull work items (a stack and one of its
// tiles, every frame of that stack) from a queue in the order of `items` (longest first) until it
// runs dry; every block reaches the end of the queue and returns.
template <typename T, int RH, bool COL, bool PACK, bool LENS>
__global__ void __launch_bounds__(256, sizeof(T) == 1 ? 4 : 3)
    k_pass1_batch(const Pass1Args* __restrict__ stacks, const BatchItem* __restrict__ items, int n_items,
                  int* __restrict__ queue) {
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        // the same item for the whole block: made wave-uniform so that its arguments load as scalars
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();  // every thread has the item before thread 0 takes the next one
        if (it >= n_items) return;
        const int job = __builtin_amdgcn_readfirstlane((int)items[it].job);
        const int bx = __builtin_amdgcn_readfirstlane((int)items[it].bx);
        const int by = __builtin_amdgcn_readfirstlane((int)items[it].by);
        const Pass1Args a = stacks[job];  // one scalar copy per item, as kernel arguments would be
        pass1_tile<T, RH, COL, PACK, LENS>(a, bx, by);
        __syncthreads();  // the next item reuses the block's LDS
    }
}

template <typename T, int RH, bool COL, bool PACK, bool LENS>
static hipError_t launch_pass1_kl(const Pass1Args& a, hipStream_t s) {
    const size_t lds = pass1_lds_bytes((int)a.window, (int)sizeof(T), a.best != nullptr, pass1_slot_winners(a),
                                       a.norm_sum != nullptr, PACK);
    const int b = (int)a.window / 2, th = RH - 2 * b, tw = P1C - 2 * b;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_pass1<T, RH, COL, PACK, LENS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    dim3 grid((a.w + tw - 1) / tw, (a.h + th - 1) / th);
    hipLaunchKernelGGL((k_pass1<T, RH, COL, PACK, LENS>), grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

template <typename T, int RH, bool COL, bool PACK>
static hipError_t launch_pass1_k(const Pass1Args& a, hipStream_t s) {
    return a.lens ? launch_pass1_kl<T, RH, COL, PACK, true>(a, s) : launch_pass1_kl<T, RH, COL, PACK, false>(a, s);
}

// Hybrid winners of frames [k0, k1) from their cached weights (updateRange, src/fuse/blend.zig:
// 237-266): per pixel the strictly greater weight wins, so the first frame wins ties; `best` carries
// the running maximum across launches, and a new winner's colour comes from the frame cache.  The
// weights are the ones pass 1 stored, so the winners are those pass 1 would keep; kept apart, the
// pass-1 launch carries no winner state in LDS and runs four blocks per CU.
template <typename T>
__global__ void __launch_bounds__(256) k_hybrid_winners(Pass1Args a) {
    const size_t npx = (size_t)a.w * a.h;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= npx) return;
    const float* wp = a.weights + (size_t)(a.k0 - a.slot_base) * npx + i;
    const int n = a.k1 - a.k0;
    float b = a.best[i];
    int win = -1;
    int k = 0;
    for (; k + 8 <= n; k += 8) {  // eight loads in flight
        float w[8];
#pragma unroll
        for (int j = 0; j < 8; j++) w[j] = wp[(size_t)(k + j) * npx];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (w[j] > b) {
                b = w[j];
                win = k + j;
            }
        }
    }
    for (; k < n; k++) {
        const float w = wp[(size_t)k * npx];
        if (w > b) {
            b = w;
            win = k;
        }
    }
    if (win >= 0) {
        a.best[i] = b;
        const T* px = (const T*)a.rgba + ((size_t)(a.k0 - a.slot_base + win) * npx + i) * 4;
        T* hd = (T*)a.hard + i * 3;
        hd[0] = px[0];
        hd[1] = px[1];
        hd[2] = px[2];
    }
}

template <typename T, int RH>
static hipError_t launch_pass1_t(const Pass1Args& a, hipStream_t s) {
    const bool slots = pass1_slot_winners(a);
    if (a.best != nullptr && !slots) return launch_pass1_k<T, RH, true, false>(a, s);
    if constexpr (sizeof(T) == 2) {
        if (a.best != nullptr && a.norm_sum != nullptr && a.weights != nullptr) {
            // cached 16-bit hybrid: the plain pass 1 (four blocks per CU, where the winners' LDS
            // left three), then the winners from the weights it stored.  The 8-bit hybrid keeps
            // them in the packed pass-1 layout, which is faster there.
            Pass1Args p = a;
            p.best = nullptr;
            p.hard = nullptr;
            hipError_t e = launch_pass1_k<T, RH, false, false>(p, s);
            if (e != hipSuccess) return e;
            const size_t npx = (size_t)a.w * a.h;
            hipLaunchKernelGGL(k_hybrid_winners<T>, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, s, a);
            return hipGetLastError();
        }
    } else {
        if (a.best != nullptr && a.norm_sum != nullptr) return launch_pass1_k<T, RH, false, true>(a, s);
    }
    return launch_pass1_k<T, RH, false, false>(a, s);
}

void pass1_grid(uint32_t w, uint32_t h, uint32_t window, uint32_t* gx, uint32_t* gy) {
    const int b = (int)window / 2, rh = pass1_rows((int)window);
    const uint32_t tw = (uint32_t)(P1C - 2 * b), th = (uint32_t)(rh - 2 * b);
    *gx = (w + tw - 1) / tw;
    *gy = (h + th - 1) / th;
}

size_t pass1_batch_lds(int sb, uint32_t window) { return pass1_lds_bytes((int)window, sb, false, true, true, false); }

template <typename T, int RH>
static int pass1_batch_occupancy(uint32_t window) {
    int per_cu = 0;
    const size_t lds = pass1_batch_lds((int)sizeof(T), window);
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_pass1_batch<T, RH, false, false, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_pass1_batch<T, RH, false, false, false>,
                                                     256, lds) != hipSuccess)
        per_cu = 1;
    return per_cu > 0 ? per_cu : 1;
}

int pass1_resident_blocks(int sb, int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    const int per_cu = sb == 1 ? pass1_batch_occupancy<uint8_t, 32>(5) : pass1_batch_occupancy<uint16_t, 32>(5);
    return per_cu * cus;
}

template <typename T, int RH, bool LENS>
static hipError_t launch_pass1_batch_t(uint32_t window, const Pass1Args* stacks, const BatchItem* items, int n_items,
                                       int* queue, int blocks, hipStream_t s) {
    const size_t lds = pass1_batch_lds((int)sizeof(T), window);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_pass1_batch<T, RH, false, false, LENS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_pass1_batch<T, RH, false, false, LENS>), dim3((unsigned)blocks), dim3(256), lds, s, stacks,
                       items, n_items, queue);
    return hipGetLastError();
}

template <typename T, int RH>
static hipError_t launch_pass1_batch_l(bool lens, uint32_t window, const Pass1Args* stacks, const BatchItem* items,
                                       int n_items, int* queue, int blocks, hipStream_t s) {
    return lens ? launch_pass1_batch_t<T, RH, true>(window, stacks, items, n_items, queue, blocks, s)
                : launch_pass1_batch_t<T, RH, false>(window, stacks, items, n_items, queue, blocks, s);
}

hipError_t launch_pass1_batch(int sb, uint32_t window, bool lens, const Pass1Args* stacks, const BatchItem* items,
                              int n_items, int* queue, int blocks, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (window < 3 || window > 31 || (window & 1) == 0 || blocks <= 0) return hipErrorInvalidValue;
    const bool big = pass1_rows((int)window) == 64;
    if (sb == 1)
        return big ? launch_pass1_batch_l<uint8_t, 64>(lens, window, stacks, items, n_items, queue, blocks, s)
                   : launch_pass1_batch_l<uint8_t, 32>(lens, window, stacks, items, n_items, queue, blocks, s);
    return big ? launch_pass1_batch_l<uint16_t, 64>(lens, window, stacks, items, n_items, queue, blocks, s)
               : launch_pass1_batch_l<uint16_t, 32>(lens, window, stacks, items, n_items, queue, blocks, s);
}

hipError_t launch_pass1(const Pass1Args& a, hipStream_t s) {
    if (a.k1 <= a.k0) return hipSuccess;
    if (a.window < 3 || a.window > 31 || (a.window & 1) == 0) return hipErrorInvalidValue;
    const bool big = pass1_rows((int)a.window) == 64;
    if (a.sample_bytes == 1) return big ? launch_pass1_t<uint8_t, 64>(a, s) : launch_pass1_t<uint8_t, 32>(a, s);
    return big ? launch_pass1_t<uint16_t, 64>(a, s) : launch_pass1_t<uint16_t, 32>(a, s);
}

// ---------------------------------------------------------------------------
// pass 2
// ---------------------------------------------------------------------------
// Level data in LDS and in the per-frame level buffers is one float4 per pixel:
//   x = red,   or -(red + 1) when the pixel has no image support
//   y, z = green, blue
//   w = mask,  or -1 when the pixel has no mask support (level 0: no frame has alpha > 0,
//              where the mask is exactly 1/N since every weight there is 0)
// Unsupported pixels above level 0 have rgb 0, so -(0 + 1) = -1 decodes to 0.

__device__ __forceinline__ float dec_r(float r) { return r < 0.0f ? -r - 1.0f : r; }

template <int TW, int TH>
struct Tile {
    static constexpr int RW = TW + 8, RH = TH + 7;          // level-l region: tile + 4 left/top, + 3 right/bottom
                                                            // (rows padded to an even length for 8-byte LDS reads)
    static constexpr int NW = TW / 2 + 2, NH = TH / 2 + 2;  // level-(l+1) region: tile/2 + 1 on each side
    static constexpr int PPT = TW * TH / 256;
    static constexpr int STEP = 256 / TW;
};

// Divisor of expandRgbExactNoWrap (src/fuse/pyramid.zig:1167-1354) for output (2i+a, 2j+b) of a
// source level sw x sh: the in-bounds kernel weight, except where the reference divides by the
// constants of the wider case (one-pixel-wide or one-pixel-tall source levels).
__device__ __forceinline__ float nowrap_div(int i, int j, bool ev_x, bool ev_y, int sw, int sh, int y) {
    const float hw = ev_x ? (float)((i > 0 ? 1 : 0) + 6 + (i + 1 < sw ? 1 : 0)) : (float)(4 + (i + 1 < sw ? 4 : 0));
    const float vw = ev_y ? (float)((j > 0 ? 1 : 0) + 6 + (j + 1 < sh ? 1 : 0)) : (float)(1 + (j + 1 < sh ? 1 : 0));
    float div = hw * vw;
    if (sw == 1 && sh > 1 && j <= sh - 2 && ev_x) div = ev_y ? (j == 0 ? 49.0f : 56.0f) : 14.0f;
    if (sh == 1 && sw > 1 && y == 0 && i == sw - 1) div = ev_x ? 49.0f : 28.0f;
    return div;
}

// Running sums of one support-aware reduce output (reduceRgbWithSupportRowsSequential /
// reduceScalarWithSupportRowsSequential, src/fuse/pyramid.zig:997-1099).
struct RedAcc {
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, ws = 0.0f, mv = 0.0f, mws = 0.0f;
};

__device__ __forceinline__ float kw_of(int r) { return r == 2 ? 6.0f : ((r == 1 || r == 3) ? 4.0f : 1.0f); }

// Adds the five taps of one LDS row (columns base..base+4, kx order) to the outputs it feeds, with
// vertical kernel weights wya / wyb.  A tap without support adds value * 0 (= +-0) and weight 0,
// which leaves every sum unchanged, exactly as the reference's `continue`.
template <bool DO_A, bool DO_B>
__device__ __forceinline__ void reduce_row(const float* rR, const float* rG, const float* rB, const float* rM, int base,
                                           float wya, float wyb, RedAcc& a, RedAcc& b) {
    const float KW[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    float tr[5], tg[5], tb[5], tm[5];
    // 8-byte reads throughout (the fifth tap's pair ends inside the row: base + 5 < RW), which
    // keeps the stride-2 access pattern free of bank conflicts
    {
        const float2 u0 = *(const float2*)&rR[base], u1 = *(const float2*)&rR[base + 2], u2 = *(const float2*)&rR[base + 4];
        tr[0] = u0.x; tr[1] = u0.y; tr[2] = u1.x; tr[3] = u1.y; tr[4] = u2.x;
    }
    {
        const float2 u0 = *(const float2*)&rG[base], u1 = *(const float2*)&rG[base + 2], u2 = *(const float2*)&rG[base + 4];
        tg[0] = u0.x; tg[1] = u0.y; tg[2] = u1.x; tg[3] = u1.y; tg[4] = u2.x;
    }
    {
        const float2 u0 = *(const float2*)&rB[base], u1 = *(const float2*)&rB[base + 2], u2 = *(const float2*)&rB[base + 4];
        tb[0] = u0.x; tb[1] = u0.y; tb[2] = u1.x; tb[3] = u1.y; tb[4] = u2.x;
    }
    {
        const float2 u0 = *(const float2*)&rM[base], u1 = *(const float2*)&rM[base + 2], u2 = *(const float2*)&rM[base + 4];
        tm[0] = u0.x; tm[1] = u0.y; tm[2] = u1.x; tm[3] = u1.y; tm[4] = u2.x;
    }
#pragma unroll
    for (int kx = 0; kx < 5; kx++) {
        const bool gs = tr[kx] >= 0.0f, ms = tm[kx] >= 0.0f;
        if (DO_A) {
            const float w = wya * KW[kx];
            const float gw = gs ? w : 0.0f, mw = ms ? w : 0.0f;
            a.s0 += tr[kx] * gw;
            a.s1 += tg[kx] * gw;
            a.s2 += tb[kx] * gw;
            a.ws += gw;
            a.mv += tm[kx] * mw;
            a.mws += mw;
        }
        if (DO_B) {
            const float w = wyb * KW[kx];
            const float gw = gs ? w : 0.0f, mw = ms ? w : 0.0f;
            b.s0 += tr[kx] * gw;
            b.s1 += tg[kx] * gw;
            b.s2 += tb[kx] * gw;
            b.ws += gw;
            b.mv += tm[kx] * mw;
            b.mws += mw;
        }
    }
}

// One tap of the reduce on a fully supported window: sum + value * W as the reference rounds it.
// For power-of-two W the product is exact, so the fused form rounds once, exactly like
// RN(sum + RN(value * W)); the other weights (6, 24, 36) keep the separate product.
template <int W>
__device__ __forceinline__ float red_tap(float sum, float v) {
    if constexpr ((W & (W - 1)) == 0)
        return fmaf(v, (float)W, sum);
    else
        return sum + v * (float)W;
}

// Adds one LDS row (columns base..base+4, kx order) with vertical weight KY to the four sums
// (r, g, b, mask) of a window whose 25 taps all have image and mask support.
template <int KY>
__device__ __forceinline__ void reduce_row_full(const float* rR, const float* rG, const float* rB, const float* rM,
                                                int base, float s[4]) {
    const float* ch[4] = {rR, rG, rB, rM};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float2 u0 = *(const float2*)&ch[c][base], u1 = *(const float2*)&ch[c][base + 2];
        const float u4 = (*(const float2*)&ch[c][base + 4]).x;
        float a = s[c];
        a = red_tap<KY * 1>(a, u0.x);
        a = red_tap<KY * 4>(a, u0.y);
        a = red_tap<KY * 6>(a, u1.x);
        a = red_tap<KY * 4>(a, u1.y);
        a = red_tap<KY * 1>(a, u4);
        s[c] = a;
    }
}

// Level 0 with every tap supported: the rgb samples are integers below 2^16 and the kernel weights
// sum to 256, so every partial sum of a reduce output is an integer below 2^24, exact in f32, and
// the sum equals the reference's sequential one in any order.  hsum5: the horizontal (1,4,6,4,1)
// sum of one LDS row (columns base..base+4); the vertical sums of these give the outputs.
__device__ __forceinline__ float hsum5_int(const float* c, int base) {
    const float2 u0 = *(const float2*)&c[base], u1 = *(const float2*)&c[base + 2];
    const float u4 = (*(const float2*)&c[base + 4]).x;
    return fmaf(6.0f, u1.x, fmaf(4.0f, u0.y + u1.y, u0.x + u4));
}
__device__ __forceinline__ float vsum5_int(float h0, float h1, float h2, float h3, float h4) {
    return fmaf(6.0f, h2, fmaf(4.0f, h1 + h3, h0 + h4));
}

// One row of reduce_row_full for a single channel (the mask's sum keeps the reference's order).
template <int KY>
__device__ __forceinline__ float reduce_row_full1(const float* c, int base, float a) {
    const float2 u0 = *(const float2*)&c[base], u1 = *(const float2*)&c[base + 2];
    const float u4 = (*(const float2*)&c[base + 4]).x;
    a = red_tap<KY * 1>(a, u0.x);
    a = red_tap<KY * 4>(a, u0.y);
    a = red_tap<KY * 6>(a, u1.x);
    a = red_tap<KY * 4>(a, u1.y);
    return red_tap<KY * 1>(a, u4);
}

// w / s correctly rounded from y = RN(1/s) (Markstein correction); the quotient of a tiny
// numerator, where the correction's residual could leave the normal range, is divided directly.
__device__ __forceinline__ float div_by_recip(float w, float s, float y) {
    if (w != 0.0f && fabsf(w) < 1e-30f) return w / s;
    const float q = w * y;
    const float r = fmaf(-q, s, w);
    return fmaf(r, y, q);
}

template <typename T, bool L0, int TW, int TH, bool SPLIT>
__device__ __forceinline__ void pass2_tile(const Pass2Args& a, const LevelArgs& g, const int bx, const int by,
                                           const int bz) {
    using TL = Tile<TW, TH>;
    using PT = typename Pack<T>::type;
    constexpr int RW = TL::RW, RH = TL::RH, NW = TL::NW, NH = TL::NH;
    constexpr int NREG = RH * RW;
    constexpr int NPF = (NREG + 255) / 256;  // region pixels per thread
    // level-l region, planar: red (support flag in the sign), green, blue, mask (flag in the sign)
    __shared__ __align__(16) float rR[NREG];
    __shared__ __align__(16) float rG[NREG];
    __shared__ __align__(16) float rB[NREG];
    __shared__ __align__(16) float rM[NREG];
    // norm sum and RN(1 / norm sum) where some frame has support and the sum exceeds 1e-12; elsewhere
    // -1 and the pixel's mask value instead (1/N with support, -1 without)
    __shared__ float2 sSY[L0 ? NREG : 1];
    __shared__ float nR[NH * NW], nG[NH * NW], nB[NH * NW];   // level-(l+1) gaussian (values)
    __shared__ float hR[NH * TW], hG[NH * TW], hB[NH * TW];   // its horizontal expansion per tile column
    __shared__ int sPart[2];                   // per frame parity: some region pixel lacks support

    const int X0 = bx * TW, Y0 = by * TH;
    const int tx = threadIdx.x % TW, ty = threadIdx.x / TW;
    const int n = g.w * g.h;
    float* accp = a.acc + g.acc_off;
    const float m_unsup = L0 ? 1.0f / (float)a.frame_count : 0.0f;
    const int kb = SPLIT ? a.k0 + bz : a.k0;
    const int ke = SPLIT ? kb + 1 : a.k1;
    const bool clamp_regime = (g.w > g.h ? g.w : g.h) > 512;  // exact_expand_max_dim, src/fuse/pyramid.zig:8
    const int ox1 = X0 / 2 - 1, oy1 = Y0 / 2 - 1;
    // a tile whose expand taps never reach a level edge: the clamp-to-edge expand needs no clamping
    const bool inner = clamp_regime && !g.top && X0 >= 4 && Y0 >= 4 && 2 * (X0 / 2 + TW / 2 + 2) <= 2 * g.nw - 2 &&
                       2 * (Y0 / 2 + TH / 2 + 2) <= 2 * g.nh - 2 && X0 + TW <= g.w && Y0 + TH <= g.h;

    // a launch that starts at frame 0 starts its sums from 0.0f, as the accumulator would hold
    float ar[TL::PPT], ag[TL::PPT], ab[TL::PPT];
#pragma unroll
    for (int j = 0; j < TL::PPT; j++) {
        ar[j] = ag[j] = ab[j] = 0.0f;
        const int x = X0 + tx, y = Y0 + ty + j * TL::STEP;
        if (!SPLIT && a.k0 > 0 && x < g.w && y < g.h) {
            const int i = y * g.w + x;
            ar[j] = accp[i];
            ag[j] = accp[i + n];
            ab[j] = accp[i + 2 * n];
        }
    }
    if (threadIdx.x < 2) sPart[threadIdx.x] = 0;
    if (L0) {
#pragma unroll 1
        for (int idx = threadIdx.x; idx < NREG; idx += 256) {
            const int gx = X0 - 4 + idx % RW, gy = Y0 - 4 + idx / RW;
            float v = -1.0f, y = -1.0f;
            if (gx >= 0 && gy >= 0 && gx < g.w && gy < g.h) {
                const int i = gy * g.w + gx;
                if (a.union_sup[i]) {
                    v = a.norm_sum[i];
                    y = 1.0f / v;
                    if (!(v > 1e-12f)) {
                        v = -1.0f;
                        y = m_unsup;
                    }
                }
            }
            sSY[idx] = make_float2(v, y);
        }
    }

    // level index of this thread's region pixels (-1 outside the level), the same for every frame
    int gidx[NPF];
#pragma unroll
    for (int j = 0; j < NPF; j++) {
        const int idx = threadIdx.x + 256 * j;
        const int gx = X0 - 4 + idx % RW, gy = Y0 - 4 + idx / RW;
        gidx[j] = idx < NREG && gx >= 0 && gy >= 0 && gx < g.w && gy < g.h ? gy * g.w + gx : -1;
    }
    // the next frame's level-l region, in flight while the current frame is reduced and expanded
    PT pf_px[L0 ? NPF : 1];
    float pf_w[L0 ? NPF : 1];
    float4 pf_v[L0 ? 1 : NPF];
    // Loads without branches: a pixel outside the level loads element 0 of the same frame (its value
    // is never used), and the address is the frame's base (wave-uniform) plus a 32-bit byte offset
    // (levels hold fewer than 2^28 pixels).
    auto prefetch = [&](int k) {
        const size_t slot = (size_t)(k - a.slot_base);
        const char* rg = L0 ? (const char*)((const PT*)a.rgba + slot * (size_t)n) : nullptr;
        const char* wts = L0 ? (const char*)(a.weights + slot * (size_t)n) : nullptr;
        const char* lv = L0 ? nullptr : (const char*)(a.lev + slot * a.lev_stride + g.lev_off);
#pragma unroll
        for (int j = 0; j < NPF; j++) {
            const uint32_t e = gidx[j] >= 0 ? (uint32_t)gidx[j] : 0u;
            if (L0) {
                pf_px[j] = *(const PT*)(rg + e * (uint32_t)sizeof(PT));
                pf_w[j] = *(const float*)(wts + e * 4u);
            } else {
                pf_v[j] = *(const float4*)(lv + e * 16u);
            }
        }
    };
    prefetch(kb);
    lds_barrier();

    for (int k = kb; k < ke; k++) {
        const size_t slot = (size_t)(k - a.slot_base);
        // A: level-l region of frame k from the prefetch registers
        {
            bool full = true;
#pragma unroll
            for (int j = 0; j < NPF; j++) {
                const int idx = threadIdx.x + 256 * j;
                if (idx >= NREG) continue;
                float4 v = make_float4(-1.0f, 0.0f, 0.0f, -1.0f);
                if (gidx[j] >= 0) {
                    if (L0) {
                        // fillRgbBase + fillBinarySupport (src/fuse/pyramid.zig:502-503) and
                        // normalizeWeightsInto (:303-316) with the union support (:202)
                        const PT q = pf_px[j];
                        const float r = (float)Pack<T>::ch(q, 0);
                        v.x = Pack<T>::ch(q, 3) > 0 ? r : -(r + 1.0f);
                        v.y = (float)Pack<T>::ch(q, 1);
                        v.z = (float)Pack<T>::ch(q, 2);
                        // w / S by the Markstein correction from RN(1/S) (div_by_recip), formed
                        // without branches; where the norm sum is marked (-1), the mask value held
                        // in place of its reciprocal (1/N where S <= 1e-12, -1 without support)
                        const float2 sy = sSY[idx];
                        const float sv = sy.x, w = pf_w[j];
                        const float qd = w * sy.y;
                        float mw = fmaf(fmaf(-qd, sv, w), sy.y, qd);
                        // tiny numerators divide directly, behind a wave-uniform branch: weights
                        // are stds of 8/16-bit gray values, far above 1e-30 where nonzero
                        const bool tiny = w != 0.0f && fabsf(w) < 1e-30f;
                        if (__builtin_amdgcn_ballot_w64(tiny)) {
                            asm volatile("" ::: "memory");  // a real branch: the division is not speculated
                            if (tiny) mw = w / sv;
                        }
                        v.w = sv > 0.0f ? mw : sy.y;
                    } else {
                        v = pf_v[j];
                    }
                }
                full = full && v.x >= 0.0f && v.w >= 0.0f;
                rR[idx] = v.x;
                rG[idx] = v.y;
                rB[idx] = v.z;
                rM[idx] = v.w;
            }
            if (!full) sPart[k & 1] = 1;
            if (threadIdx.x == 0) sPart[(k + 1) & 1] = 0;
        }
        lds_barrier();
        if (k + 1 < ke) prefetch(k + 1);
        if (!g.top) {
            // B: reduceRgbWithSupport / reduceScalarWithSupport (src/fuse/pyramid.zig:997-1099),
            // two vertically adjacent outputs per thread.
            const bool full = sPart[k & 1] == 0;
            constexpr int NP = NW * (NH / 2);
            float4* lvn = a.lev + slot * a.lev_stride + g.nlev_off;
            for (int p = threadIdx.x; p < NP; p += 256) {
                const int oi = p % NW, oj = 2 * (p / NW);
                const int base0 = (2 * oj) * RW + 2 * oi;
                float a0, a1, a2, aw, am, amw, b0, b1, b2, bw, bm, bmw;
                if (L0 && full) {
                    // every tap in bounds and supported: weight sums are exactly 256; rgb in the
                    // separable integer form (exact), the mask in the reference's order
                    float ca[3], cb[3];
                    const float* rgb[3] = {rR, rG, rB};
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float h[7];
#pragma unroll
                        for (int rr = 0; rr < 7; rr++) h[rr] = hsum5_int(rgb[c], base0 + rr * RW);
                        ca[c] = vsum5_int(h[0], h[1], h[2], h[3], h[4]);
                        cb[c] = vsum5_int(h[2], h[3], h[4], h[5], h[6]);
                    }
                    float ma = 0.0f, mb = 0.0f;
                    ma = reduce_row_full1<1>(rM, base0, ma);
                    ma = reduce_row_full1<4>(rM, base0 + RW, ma);
                    ma = reduce_row_full1<6>(rM, base0 + 2 * RW, ma);
                    mb = reduce_row_full1<1>(rM, base0 + 2 * RW, mb);
                    ma = reduce_row_full1<4>(rM, base0 + 3 * RW, ma);
                    mb = reduce_row_full1<4>(rM, base0 + 3 * RW, mb);
                    ma = reduce_row_full1<1>(rM, base0 + 4 * RW, ma);
                    mb = reduce_row_full1<6>(rM, base0 + 4 * RW, mb);
                    mb = reduce_row_full1<4>(rM, base0 + 5 * RW, mb);
                    mb = reduce_row_full1<1>(rM, base0 + 6 * RW, mb);
                    a0 = ca[0]; a1 = ca[1]; a2 = ca[2]; am = ma; aw = amw = 256.0f;
                    b0 = cb[0]; b1 = cb[1]; b2 = cb[2]; bm = mb; bw = bmw = 256.0f;
                } else if (full) {
                    // every tap in bounds and supported: weight sums are exactly 256
                    float sa[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    reduce_row_full<1>(rR, rG, rB, rM, base0, sa);
                    reduce_row_full<4>(rR, rG, rB, rM, base0 + RW, sa);
                    reduce_row_full<6>(rR, rG, rB, rM, base0 + 2 * RW, sa);
                    reduce_row_full<1>(rR, rG, rB, rM, base0 + 2 * RW, sb);
                    reduce_row_full<4>(rR, rG, rB, rM, base0 + 3 * RW, sa);
                    reduce_row_full<4>(rR, rG, rB, rM, base0 + 3 * RW, sb);
                    reduce_row_full<1>(rR, rG, rB, rM, base0 + 4 * RW, sa);
                    reduce_row_full<6>(rR, rG, rB, rM, base0 + 4 * RW, sb);
                    reduce_row_full<4>(rR, rG, rB, rM, base0 + 5 * RW, sb);
                    reduce_row_full<1>(rR, rG, rB, rM, base0 + 6 * RW, sb);
                    a0 = sa[0]; a1 = sa[1]; a2 = sa[2]; am = sa[3]; aw = amw = 256.0f;
                    b0 = sb[0]; b1 = sb[1]; b2 = sb[2]; bm = sb[3]; bw = bmw = 256.0f;
                } else {
                    // A tap without support adds value * 0 (= +-0) and weight 0, which leaves every
                    // sum unchanged, exactly as the reference's `continue`.
                    RedAcc ra, rb;
                    // rows 0-1 feed the upper output only, rows 2-4 both, rows 5-6 the lower only
#pragma unroll 1
                    for (int rr = 0; rr < 2; rr++)
                        reduce_row<true, false>(rR, rG, rB, rM, base0 + rr * RW, kw_of(rr), 0.0f, ra, rb);
#pragma unroll 1
                    for (int rr = 2; rr < 5; rr++)
                        reduce_row<true, true>(rR, rG, rB, rM, base0 + rr * RW, kw_of(rr), kw_of(rr - 2), ra, rb);
#pragma unroll 1
                    for (int rr = 5; rr < 7; rr++)
                        reduce_row<false, true>(rR, rG, rB, rM, base0 + rr * RW, 0.0f, kw_of(rr - 2), ra, rb);
                    a0 = ra.s0; a1 = ra.s1; a2 = ra.s2; aw = ra.ws; am = ra.mv; amw = ra.mws;
                    b0 = rb.s0; b1 = rb.s1; b2 = rb.s2; bw = rb.ws; bm = rb.mv; bmw = rb.mws;
                }
                const int gi = ox1 + oi;
#pragma unroll
                for (int h2 = 0; h2 < 2; h2++) {
                    const int o = (oj + h2) * NW + oi;
                    const int gj = oy1 + oj + h2;
                    const float s0 = h2 ? b0 : a0, s1 = h2 ? b1 : a1, s2 = h2 ? b2 : a2, ws = h2 ? bw : aw;
                    const float mv = h2 ? bm : am, mws = h2 ? bmw : amw;
                    float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
                    if (gi >= 0 && gj >= 0 && gi < g.nw && gj < g.nh) {
                        float4 outv = make_float4(-1.0f, 0.0f, 0.0f, 0.0f);
                        if (full) {  // division by 256 is exact
                            e0 = s0 * (1.0f / 256.0f);
                            e1 = s1 * (1.0f / 256.0f);
                            e2 = s2 * (1.0f / 256.0f);
                            outv = make_float4(e0, e1, e2, mv * (1.0f / 256.0f));
                        } else {
                            if (ws > 0.0f) {
                                e0 = s0 / ws;
                                e1 = s1 / ws;
                                e2 = s2 / ws;
                                outv = make_float4(e0, e1, e2, 0.0f);
                            }
                            outv.w = mws > 0.0f ? mv / mws : -1.0f;
                        }
                        if (oi >= 1 && oi <= TW / 2 && oj + h2 >= 1 && oj + h2 <= TH / 2)
                            lvn[gj * g.nw + gi] = outv;
                    }
                    nR[o] = e0;
                    nG[o] = e1;
                    nB[o] = e2;
                }
            }
            lds_barrier();
            // C: horizontal pass of EXPAND for every level-(l+1) row of the region and tile column
            // (horizontalExpandRgbRow :1748-1775 above 512 px; the running sums of
            // expandRgbExactNoWrap :1197-1214 at or below 512 px, zero outside the level)
#pragma unroll 1
            for (int idx = threadIdx.x; idx < NH * TW; idx += 256) {
                const int r = idx / TW, x = X0 + idx % TW;
                const int gj = oy1 + r;
                float h0 = 0.0f, h1 = 0.0f, h2 = 0.0f;
                if (inner) {  // horizontalExpandRgbRow without the clamps it cannot need here
                    // even and odd columns interleave in a wave: one form for both, without
                    // branches.  Even: ((a + 6b) + c) / 8; odd: ((b + c) + -0) / 2, where adding -0
                    // returns the sum unchanged (its sign of zero included)
                    const int c1 = r * NW - ox1 + (x >> 1);
                    const bool ev = (x & 1) == 0;
                    const float sc = ev ? 0.125f : 0.5f;
                    auto hx = [&](const float* nv) {
                        const float a = nv[c1 - 1], b = nv[c1], c = nv[c1 + 1];
                        const float u = ev ? a : b, v = ev ? 6.0f * b : c, w = ev ? c : -0.0f;
                        return ((u + v) + w) * sc;
                    };
                    h0 = hx(nR);
                    h1 = hx(nG);
                    h2 = hx(nB);
                } else if (x < g.w && gj >= 0 && gj < g.nh) {
                    const int i = x >> 1;
                    const int rb = r * NW - ox1;
                    if (clamp_regime) {
                        if ((x & 1) == 0) {
                            const int c0 = rb + clamp_coord(i - 1, g.nw), c1 = rb + clamp_coord(i, g.nw),
                                      c2 = rb + clamp_coord(i + 1, g.nw);
                            h0 = ((nR[c0] + 6.0f * nR[c1]) + nR[c2]) * 0.125f;
                            h1 = ((nG[c0] + 6.0f * nG[c1]) + nG[c2]) * 0.125f;
                            h2 = ((nB[c0] + 6.0f * nB[c1]) + nB[c2]) * 0.125f;
                        } else {
                            const int c0 = rb + clamp_coord(i, g.nw), c1 = rb + clamp_coord(i + 1, g.nw);
                            h0 = (nR[c0] + nR[c1]) * 0.5f;
                            h1 = (nG[c0] + nG[c1]) * 0.5f;
                            h2 = (nB[c0] + nB[c1]) * 0.5f;
                        }
                    } else {
                        const bool l_in = i - 1 >= 0, r_in = i + 1 < g.nw;
                        const int c1 = rb + i;
                        if ((x & 1) == 0) {
                            const float z0 = l_in ? nR[c1 - 1] : 0.0f, z2 = r_in ? nR[c1 + 1] : 0.0f;
                            const float y0 = l_in ? nG[c1 - 1] : 0.0f, y2 = r_in ? nG[c1 + 1] : 0.0f;
                            const float w0 = l_in ? nB[c1 - 1] : 0.0f, w2 = r_in ? nB[c1 + 1] : 0.0f;
                            h0 = (z0 + 6.0f * nR[c1]) + z2;
                            h1 = (y0 + 6.0f * nG[c1]) + y2;
                            h2 = (w0 + 6.0f * nB[c1]) + w2;
                        } else {
                            const float z2 = r_in ? nR[c1 + 1] : 0.0f;
                            const float y2 = r_in ? nG[c1 + 1] : 0.0f;
                            const float w2 = r_in ? nB[c1 + 1] : 0.0f;
                            h0 = 4.0f * (nR[c1] + z2);
                            h1 = 4.0f * (nG[c1] + y2);
                            h2 = 4.0f * (nB[c1] + w2);
                        }
                    }
                }
                hR[idx] = h0;
                hG[idx] = h1;
                hB[idx] = h2;
            }
            lds_barrier();
        }
        // D: vertical pass of EXPAND, laplacian, accumulate
        // acc[l] += (G[l] - EXPAND(G[l+1])) * M[l]  (accumulateLapRowsSequential :1506-1531),
        // top level acc += G * M (accumulateGaussianRowsSequential :1533-1554)
#pragma unroll 1
        for (int j = 0; j < TL::PPT; j++) {
            const int x = X0 + tx, y = Y0 + ty + j * TL::STEP;
            if (x >= g.w || y >= g.h) continue;
            const int ci = (y - Y0 + 4) * RW + (x - X0 + 4);
            f3 lap = mk3(dec_r(rR[ci]), rG[ci], rB[ci]);
            const float mraw = rM[ci];
            const float m = mraw < 0.0f ? m_unsup : mraw;
            if (!g.top) {
                const int jj = y >> 1;
                const int xx = x - X0;
                f3 e;
                if (inner) {
                    const int q1 = (jj - oy1) * TW + xx;
                    if ((y & 1) == 0) {
                        const int q0 = q1 - TW, q2 = q1 + TW;
                        e = mk3(((hR[q0] + 6.0f * hR[q1]) + hR[q2]) * 0.125f,
                                ((hG[q0] + 6.0f * hG[q1]) + hG[q2]) * 0.125f,
                                ((hB[q0] + 6.0f * hB[q1]) + hB[q2]) * 0.125f);
                    } else {
                        e = mk3((hR[q1] + hR[q1]) * 0.5f, (hG[q1] + hG[q1]) * 0.5f, (hB[q1] + hB[q1]) * 0.5f);
                    }
                } else if (clamp_regime) {
                    if ((y & 1) == 0) {
                        const int q0 = (clamp_coord(jj - 1, g.nh) - oy1) * TW + xx;
                        const int q1 = (clamp_coord(jj, g.nh) - oy1) * TW + xx;
                        const int q2 = (clamp_coord(jj + 1, g.nh) - oy1) * TW + xx;
                        e = mk3(((hR[q0] + 6.0f * hR[q1]) + hR[q2]) * 0.125f,
                                ((hG[q0] + 6.0f * hG[q1]) + hG[q2]) * 0.125f,
                                ((hB[q0] + 6.0f * hB[q1]) + hB[q2]) * 0.125f);
                    } else {
                        const int q = (clamp_coord(jj, g.nh) - oy1) * TW + xx;
                        e = mk3((hR[q] + hR[q]) * 0.5f, (hG[q] + hG[q]) * 0.5f, (hB[q] + hB[q]) * 0.5f);
                    }
                } else {
                    const int q1 = (jj - oy1) * TW + xx;
                    f3 num;
                    if ((y & 1) == 0) {
                        const int q0 = q1 - TW, q2 = q1 + TW;
                        num = mk3((hR[q0] + 6.0f * hR[q1]) + hR[q2], (hG[q0] + 6.0f * hG[q1]) + hG[q2],
                                  (hB[q0] + 6.0f * hB[q1]) + hB[q2]);
                    } else {
                        const int q2 = q1 + TW;
                        num = mk3(hR[q1] + hR[q2], hG[q1] + hG[q2], hB[q1] + hB[q2]);
                    }
                    e = div3(num, nowrap_div(x >> 1, jj, (x & 1) == 0, (y & 1) == 0, g.nw, g.nh, y));
                }
                lap = sub3(lap, e);
            }
            if (SPLIT) {
                const int i = y * g.w + x;
                float* cp = a.contrib + (size_t)(k - a.k0) * 3 * (size_t)n;
                cp[i] = lap.x * m;
                cp[i + n] = lap.y * m;
                cp[i + 2 * n] = lap.z * m;
            } else {
                ar[j] = ar[j] + lap.x * m;
                ag[j] = ag[j] + lap.y * m;
                ab[j] = ab[j] + lap.z * m;
            }
        }
        lds_barrier();
    }

    if (!SPLIT) {
#pragma unroll
        for (int j = 0; j < TL::PPT; j++) {
            const int x = X0 + tx, y = Y0 + ty + j * TL::STEP;
            if (x >= g.w || y >= g.h) continue;
            const int i = y * g.w + x;
            accp[i] = ar[j];
            accp[i + n] = ag[j];
            accp[i + 2 * n] = ab[j];
        }
    }
}

template <typename T, bool L0, int TW, int TH, bool SPLIT>
__global__ void __launch_bounds__(256) k_pass2(Pass2Args a, LevelArgs g) {
    if constexpr (SPLIT) {
        pass2_tile<T, L0, TW, TH, SPLIT>(a, g, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
    } else {
        int bx, by;
        xcd_tile(&bx, &by);
        pass2_tile<T, L0, TW, TH, SPLIT>(a, g, bx, by, 0);
    }
}

// One pyramid level of many stacks in one launch: block b runs item b (a stack's pass-2 arguments at
// this level and one of its tiles) over every frame of that stack.
template <typename T, bool L0, int TW, int TH>
__global__ void __launch_bounds__(256) k_pass2_batch(const Pass2Job* __restrict__ jobs,
                                                     const BatchItem* __restrict__ items) {
    const int job = __builtin_amdgcn_readfirstlane((int)items[blockIdx.x].job);
    const int bx = __builtin_amdgcn_readfirstlane((int)items[blockIdx.x].bx);
    const int by = __builtin_amdgcn_readfirstlane((int)items[blockIdx.x].by);
    const Pass2Job j = jobs[job];
    pass2_tile<T, L0, TW, TH, false>(j.a, j.g, bx, by, 0);
}

// acc += contribution of each frame, in frame order (from 0.0f when the frames start at frame 0)
__global__ void __launch_bounds__(256) k_ordered_sum(float* __restrict__ acc, const float* __restrict__ contrib,
                                                     int nk, size_t n3, int fresh) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n3) return;
    float v = fresh ? 0.0f : acc[i];
    for (int k = 0; k < nk; k++) v = v + contrib[(size_t)k * n3 + i];
    acc[i] = v;
}

Tuning& tuning() {
    static Tuning t{FG_TUNE_PASS2_MIN_TILES, FG_TUNE_BATCH_WIDE_LEVEL, FG_TUNE_BATCH_PASS1_BLOCKS_PER_CU,
                    FG_TUNE_BATCH_GROUP_MIB, FG_TUNE_BATCH_STREAMS,
                    FG_TUNE_BATCH_LARGE_KPIX};
    return t;
}

static int tiles(int w, int h, int tw, int th) { return ((w + tw - 1) / tw) * ((h + th - 1) / th); }

size_t pass2_contrib_floats(const LevelTable* lt, uint32_t frames) {
    size_t mx = 0;
    for (int l = 0; l < lt->n; l++) {
        const int w = (int)lt->lv[l].w, h = (int)lt->lv[l].h;
        const int mt = tuning().pass2_min_tiles;
        if (tiles(w, h, 64, 16) >= mt || tiles(w, h, 16, 16) >= mt) continue;
        const size_t need = (size_t)3 * w * h * frames;
        mx = need > mx ? need : mx;
    }
    return mx;
}

template <typename T, bool L0>
static hipError_t launch_level(const Pass2Args& a, const LevelArgs& g, int nk, hipStream_t s) {
    const int mt = tuning().pass2_min_tiles;
    if (tiles(g.w, g.h, 64, 16) >= mt) {
        dim3 grid((g.w + 63) / 64, (g.h + 15) / 16);
        hipLaunchKernelGGL((k_pass2<T, L0, 64, 16, false>), grid, dim3(256), 0, s, a, g);
    } else if (tiles(g.w, g.h, 16, 16) >= mt || nk == 1) {
        dim3 grid((g.w + 15) / 16, (g.h + 15) / 16);
        hipLaunchKernelGGL((k_pass2<T, L0, 16, 16, false>), grid, dim3(256), 0, s, a, g);
    } else {
        const size_t n3 = (size_t)3 * g.w * g.h;
        if (n3 * (size_t)nk > a.contrib_floats) return hipErrorInvalidValue;
        dim3 grid((g.w + 15) / 16, (g.h + 15) / 16, nk);
        hipLaunchKernelGGL((k_pass2<T, L0, 16, 16, true>), grid, dim3(256), 0, s, a, g);
        hipLaunchKernelGGL(k_ordered_sum, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, a.acc + g.acc_off,
                           (const float*)a.contrib, nk, n3, (int)(a.k0 == 0));
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t pass2_t(const Pass2Args& a, hipStream_t s) {
    const LevelTable& lt = *a.lt;
    const int L = lt.n;
    const int nk = a.k1 - a.k0;
    if (nk <= 0) return hipSuccess;
    const int lend = a.level_end > 0 && a.level_end < L ? a.level_end : L;
    for (int l = a.level_begin; l < lend; l++) {
        const LevelArgs g = level_args(lt, l);
        hipError_t e = l == 0 ? launch_level<T, true>(a, g, nk, s) : launch_level<T, false>(a, g, nk, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_pass2(const Pass2Args& a, hipStream_t s) {
    return a.sample_bytes == 1 ? pass2_t<uint8_t>(a, s) : pass2_t<uint16_t>(a, s);
}

template <typename T, bool L0>
static hipError_t launch_pass2_batch_t(int tw, const Pass2Job* jobs, const BatchItem* items, int n_items,
                                       hipStream_t s) {
    if (tw == 64)
        hipLaunchKernelGGL((k_pass2_batch<T, L0, 64, 16>), dim3((unsigned)n_items), dim3(256), 0, s, jobs, items);
    else
        hipLaunchKernelGGL((k_pass2_batch<T, L0, 16, 16>), dim3((unsigned)n_items), dim3(256), 0, s, jobs, items);
    return hipGetLastError();
}

hipError_t launch_pass2_batch(int sb, bool level0, int tw, int th, const Pass2Job* jobs, const BatchItem* items,
                              int n_items, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (!((tw == 64 && th == 16) || (tw == 16 && th == 16))) return hipErrorInvalidValue;
    if (sb == 1)
        return level0 ? launch_pass2_batch_t<uint8_t, true>(tw, jobs, items, n_items, s)
                      : launch_pass2_batch_t<uint8_t, false>(tw, jobs, items, n_items, s);
    return level0 ? launch_pass2_batch_t<uint16_t, true>(tw, jobs, items, n_items, s)
                  : launch_pass2_batch_t<uint16_t, false>(tw, jobs, items, n_items, s);
}

// level-0 gaussian + mask of one cached frame in the level-buffer encoding (debug readback)
template <typename T>
__global__ void k_level0_encode(Pass2Args a, size_t n, int slot, float4* dst) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const T* q = (const T*)a.rgba + ((size_t)slot * n + i) * 4;
    const float r = (float)q[0];
    float4 v;
    v.x = q[3] > 0 ? r : -(r + 1.0f);
    v.y = (float)q[1];
    v.z = (float)q[2];
    const float def = 1.0f / (float)a.frame_count;
    v.w = a.union_sup[i] ? normalized_mask(a.weights[(size_t)slot * n + i], a.norm_sum[i], def) : -1.0f;
    dst[i] = v;
}

hipError_t launch_level0_encode(const Pass2Args& a, int slot, float4* dst, hipStream_t s) {
    const size_t n = (size_t)a.lt->lv[0].w * a.lt->lv[0].h;
    dim3 grid((unsigned)((n + 255) / 256));
    if (a.sample_bytes == 1)
        hipLaunchKernelGGL(k_level0_encode<uint8_t>, grid, dim3(256), 0, s, a, n, slot, dst);
    else
        hipLaunchKernelGGL(k_level0_encode<uint16_t>, grid, dim3(256), 0, s, a, n, slot, dst);
    return hipGetLastError();
}

}  // namespace fg
