// Shared declarations of the HIP kernels (fg_kernels.hip) and the C-ABI host layer (fg_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fg {

// How an output (panorama) pixel maps to a source pixel.
enum RemapMode : int32_t {
    REMAP_HOMOGRAPHY = 0,   // basic rectilinear: the legacy chain, evaluated as its homography
    REMAP_LENS = 1,         // the lens row walk: homography (+ Newton undistort with radial terms)
    REMAP_TRANSLATION = 2,  // translation-plane path
};

// Per-frame constants of the inverse transform, prepared on the host from the pose.
// local = B * (x_pano - dcx, y_pano - dcy, neg_d);  denom = -local.z
// radial = focal * local.xy / denom + center_shift;   source = src_center + undistort(radial)
struct RemapParams {
    double B[9];
    double neg_d;
    double dcx, dcy;
    double focal;
    double csx, csy;
    double scx, scy;
    double ra, rb, rc, inv_norm;
    double pn[3];
    double tx, ty, tz;
    int32_t mode;
    int32_t legacy;   // basic rectilinear: only rays in front of the camera land in the image
    int32_t radial;   // lens mode: radial terms present
    int32_t roi_left, roi_top;
};

// Level l of a pyramid: size and the offset (in pixels) of its first pixel in a buffer that
// stores all levels back to back.
struct Level {
    uint32_t w, h;
    size_t off;
};

constexpr int kMaxLevels = 40;

struct LevelTable {
    int n;
    Level lv[kMaxLevels];
    size_t total;  // pixels over all levels
};

void make_levels(uint32_t w, uint32_t h, uint32_t max_levels, LevelTable* t);
void make_remap_params(const double pose[15], uint32_t w, uint32_t h, int32_t roi_left, int32_t roi_top,
                       RemapParams* p);

// ---- launchers (all asynchronous on `s`) ----
hipError_t launch_remap(const void* src, int sample_bytes, int channels, uint32_t sw, uint32_t sh,
                        const RemapParams& p, uint32_t ow, uint32_t oh, void* dst, hipStream_t s);
// One frame of a stack as pass 1 sees it.
struct FrameDesc {
    const void* src;   // device pixels
    uint32_t sw, sh;   // source size
    int32_t kind;      // 0: rgb source remapped under p into the stack's ROI; 1: aligned frame
    int32_t extra;     // aligned frame carries an alpha channel
    RemapParams p;
};
inline bool frames_use_lens(const FrameDesc* f, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (f[i].kind == 0 && f[i].p.mode == REMAP_LENS) return true;
    return false;
}

// Pass 1 over frames [k0, k1): remap (or load) each frame into the RGBA cache, contrast
// weight * alpha into the weight cache, and norm sum / union support / hybrid winners.
struct Pass1Args {
    const FrameDesc* frames;  // device array indexed by frame
    int k0, k1;
    int slot_base;            // frame index held in cache slot 0
    int sample_bytes;
    uint32_t w, h, window;
    void* rgba;               // cache slot j at rgba + j * w*h*4 samples (nullptr: not kept)
    float* weights;           // cache slot j at weights + j * w*h (nullptr: not kept)
    float* norm_sum;          // nullptr: weights only (recompute in an uncached second pass)
    uint8_t* union_sup;
    float* best;              // hybrid only, else nullptr
    void* hard;
    int lens;                 // some frame of [k0, k1) is remapped by the lens row walk (REMAP_LENS)
};
hipError_t launch_pass1(const Pass1Args& a, hipStream_t s);

// Pass 2 over frames [k0, k1) for every level: mask and gaussian pyramids, laplacian
// accumulation into the planar accumulators.
struct Pass2Args {
    int sample_bytes;
    const LevelTable* lt;
    int k0, k1;
    int slot_base;            // frame held in slot 0 of the rgba / weight / level caches
    const void* rgba;
    const float* weights;
    const float* norm_sum;
    const uint8_t* union_sup;
    uint32_t frame_count;     // N of normalizeWeightsInto's 1/N fallback
    float4* lev;              // per-frame levels >= 1 (rgb + mask, support flags in signs)
    size_t lev_stride;        // float4s per frame slot
    float* acc;
    float* contrib;           // scratch for the ordered sum of small levels
    size_t contrib_floats;
    int level_begin, level_end;   // levels [level_begin, level_end) of the table (end 0 = all)
};
hipError_t launch_pass2(const Pass2Args& a, hipStream_t s);

// Level l of a stack's table as the pass-2 kernels read it.
struct LevelArgs {
    int w, h;          // level l
    int nw, nh;        // level l + 1 (unused at the top level)
    int top;           // l == L - 1: no reduce, no expand
    size_t lev_off;    // level l within a frame's level buffer (levels >= 1)
    size_t nlev_off;   // level l + 1 within a frame's level buffer
    size_t acc_off;    // level l within the planar accumulators
};
inline LevelArgs level_args(const LevelTable& lt, int l) {
    const int L = lt.n;
    const size_t base1 = L > 1 ? lt.lv[1].off : 0;
    LevelArgs g;
    g.w = (int)lt.lv[l].w;
    g.h = (int)lt.lv[l].h;
    g.top = l == L - 1;
    g.nw = g.top ? 0 : (int)lt.lv[l + 1].w;
    g.nh = g.top ? 0 : (int)lt.lv[l + 1].h;
    g.lev_off = l >= 1 ? lt.lv[l].off - base1 : 0;
    g.nlev_off = g.top ? 0 : lt.lv[l + 1].off - base1;
    g.acc_off = 3 * lt.lv[l].off;
    return g;
}

// ---- launch shapes (defaults in fg_tuning.h, set by the autotuner) ----
struct Tuning {
    int pass2_min_tiles;
    int batch_wide_level;
    int batch_pass1_blocks_per_cu;
    int batch_group_mib;
    int batch_streams;
    int batch_large_kpix;
};
Tuning& tuning();

// ---- many stacks in one schedule (fg_batch.cpp) ----
// One unit of work of a batched launch: a job (a stack, or a stack at one level) and a tile of it.
struct BatchItem {
    uint32_t job;
    uint16_t bx, by;
};
// A stack at one pyramid level: what k_pass2 takes as its two arguments.
struct Pass2Job {
    Pass2Args a;
    LevelArgs g;
};
// The tiles of pass 1 over a w x h stack (the grid of a plain launch) and its dynamic LDS bytes.
void pass1_grid(uint32_t w, uint32_t h, uint32_t window, uint32_t* gx, uint32_t* gy);
size_t pass1_batch_lds(int sample_bytes, uint32_t window);
// Blocks of a persistent pass-1 launch that are resident at once on the device.
int pass1_resident_blocks(int sample_bytes, int device);
// Pass 1 (pyramid-contrast, frames cached) of every stack in `stacks` over the tiles in `items`,
// pulled by `blocks` persistent blocks through the counter `queue` (zero before the launch).
// `lens`: some frame of the launch is remapped by the lens row walk.
hipError_t launch_pass1_batch(int sample_bytes, uint32_t window, bool lens, const Pass1Args* stacks,
                              const BatchItem* items, int n_items, int* queue, int blocks, hipStream_t s);
// One pyramid level of many stacks: item b on block b, tiles of tw x th (64x16 or 16x16).
hipError_t launch_pass2_batch(int sample_bytes, bool level0, int tw, int th, const Pass2Job* jobs,
                              const BatchItem* items, int n_items, hipStream_t s);
// Tiles of the pass-2 kernel of shape tw x th over a w x h level.
inline void pass2_tiles(int w, int h, int tw, int th, int* gx, int* gy) {
    *gx = (w + tw - 1) / tw;
    *gy = (h + th - 1) / th;
}
// The collapse of many stacks: parent level += EXPAND(child level), or, with `out`, the last step
// fused with the quantize.  Item tiles are 64 x 4 pixels of the parent level.
struct CollapseJob {
    float* parent;
    const float* child;
    int pw, ph, cw, ch;
    void* out;        // final step: the stack's quantized output, else nullptr
};
hipError_t launch_collapse_batch(const CollapseJob* jobs, const BatchItem* items, int n_items, bool final_step,
                                 int sample_bytes, hipStream_t s);
inline void collapse_tiles(int w, int h, int* gx, int* gy) {
    *gx = (w + 63) / 64;
    *gy = (h + 3) / 4;
}
// Zeroes many word ranges in one launch: item (range, chunk) clears words [chunk * 4096, +4096).
struct ZeroJob {
    uint32_t* p;
    size_t words;
};
hipError_t launch_zero_batch(const ZeroJob* jobs, const BatchItem* items, int n_items, hipStream_t s);

// What the batch engine reads of a cached stack: its pass arguments (Pass1Args::frames left for the
// batch's own frame table), level table, output staging and the sums it clears per run.
struct StackView {
    Pass1Args p1;
    Pass2Args p2;
    const LevelTable* lt;
    void* out_dev;
    float* S;
    uint8_t* U;
    size_t px;
    int sb;
};
constexpr size_t kZeroChunkWords = 4096;
size_t pass2_contrib_floats(const LevelTable* lt, uint32_t frames);
// level-0 gaussian + mask of slot j, encoded as the level buffers are
hipError_t launch_level0_encode(const Pass2Args& a, int slot, float4* dst, hipStream_t s);

hipError_t launch_quantize(const float* acc0, uint32_t w, uint32_t h, int sample_bytes, void* out, hipStream_t s);
// collapse with its last step fused with the quantize (the collapsed level 0 is not kept)
hipError_t launch_collapse_quantize(float* acc, const LevelTable* lt, int sample_bytes, void* out, hipStream_t s);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s);

// softmask-contrast over frames [k0, k1) of the weight / rgba cache slots (slot 0 holds frame
// slot_base): blur each weight map, add weight and weight * rgb into wsum / psum (planar rgb)
struct SoftArgs {
    int sample_bytes;
    int k0, k1, slot_base;
    uint32_t w, h;
    const void* rgba;
    const float* weights;
    float* wsum;
    float* psum;
};
hipError_t launch_soft(const SoftArgs& a, hipStream_t s);
hipError_t launch_soft_finalize(const float* wsum, const float* psum, uint32_t w, uint32_t h, int sample_bytes,
                                void* out, hipStream_t s);

// hybrid merge (mergeHybridImages): Laplacian pyramids of soft and hard, mix, collapse, quantize.
// gs and gh hold gaussian levels 1..n-1 only (level l at 3 * (off_l - off_1) floats; level 0 is
// read from the images); out_pyr is laid out like the accumulator.
hipError_t launch_hybrid_merge(const void* soft, const void* hard, int sample_bytes, const LevelTable* lt,
                               uint32_t hard_start, uint32_t hard_count, float mix, float* gs, float* gh,
                               float* out_pyr, void* out, hipStream_t s);

// planar acc level -> interleaved host-order buffer (debug)
hipError_t launch_planar_to_interleaved(const float* planar, size_t n, float* out, hipStream_t s);

}  // namespace fg
