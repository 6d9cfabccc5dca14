/*
 * focus_oracle.h — CPU restatement of the focus-stacker remap + pyramid-contrast
 * path, used ONLY to check results (tests/, __graft_entry__.smoke(), and the
 * cpu_baseline leg of bench.py).  Nothing in the product library links it.
 *
 * Every function follows a reference file:line (cited in focus_oracle.c).  The
 * library is built twice from the same source:
 *   liboracle.so    OR_REAL = float  : the reference's own f32/f64 mix
 *   liboracle64.so  OR_REAL = double : every f32 storage point promoted to f64,
 *                                      the "mathematically ideal" restatement
 * Remap coordinates/weights and the contrast window sums are f64 in both, as
 * they are in the reference.
 *
 * Layouts follow the reference: interleaved row-major pixels, stride =
 * width * channels; f32 RGB levels interleaved (RGBRGB...).
 */
#ifndef FOCUS_ORACLE_H
#define FOCUS_ORACLE_H

#include <stdint.h>

#ifndef OR_REAL
#define OR_REAL float
#endif
typedef OR_REAL or_real;

#ifdef __cplusplus
extern "C" {
#endif

/* The 15 f64 fields of ImagePose in declaration order (src/optimize.zig:32-48). */
typedef struct {
    double yaw, pitch, roll, hfov_delta;
    double trans_x, trans_y, trans_z;
    double translation_plane_yaw, translation_plane_pitch;
    double radial_a, radial_b, radial_c;
    double center_shift_x, center_shift_y;
    double base_hfov_degrees;
} or_pose;

/* InverseTransformCache (src/optimize.zig:3527-3553). */
typedef struct {
    int basic_rectilinear, has_translation;
    double src_center_x, src_center_y, dest_center_x, dest_center_y;
    double pano_distance, half_turn, resize_scale, dest_focal, image_focal;
    double pitch_roll[3][3], world_to_local[3][3];
    double yaw_turn, yaw;
    double radial_a, radial_b, radial_c, center_shift_x, center_shift_y, radial_inv_norm;
    double trans_x, trans_y, trans_z;
    double plane_normal[3];
} or_inv_cache;

/* Threads of the parallel loops (OpenMP; default: OMP_NUM_THREADS or the cores available).
 * The results are the same bits for every thread count. */
void or_set_threads(int n);
int or_get_threads(void);

/* ---- pose math (a1-a3, a6) ---- */
int or_pose_is_basic_rectilinear(const or_pose* pose);
int or_pose_has_translation(const or_pose* pose);
void or_init_inverse_cache(const or_pose* pose, uint32_t width, uint32_t height, or_inv_cache* out);
void or_inverse_transform_cached(const or_inv_cache* c, double x, double y, double* sx, double* sy);
void or_inverse_legacy_chain(const or_inv_cache* c, double x, double y, double* sx, double* sy);
void or_transform_point(const or_pose* pose, double x, double y, uint32_t w, uint32_t h, double* ox, double* oy);
/* returns 1 and fills roi[4] = {left, top, right, bottom} if the overlap is non-empty, else 0 */
int or_common_overlap_roi(int n, const uint8_t* active, const uint32_t* widths, const uint32_t* heights,
                          const or_pose* poses, int32_t* roi);

/* ---- remap (a4, a5) ----
 * src: width*height*channels samples (channels 1 or 3, no extra channel).
 * roi: {left, top, right, bottom} or NULL for the full frame.
 * dst: roi_w*roi_h*(channels+1) samples; last channel = support alpha. */
void or_remap_u8(const uint8_t* src, uint32_t width, uint32_t height, int channels, const or_pose* pose,
                 const int32_t* roi, uint8_t* dst);
void or_remap_u16(const uint16_t* src, uint32_t width, uint32_t height, int channels, const or_pose* pose,
                  const int32_t* roi, uint16_t* dst);

/* ---- fusion building blocks (a7-a10, a12-a16, a18) ----
 * "img" is an interleaved image of `color` (=3) colour channels plus `extra`
 * (0 or 1) alpha channel; sample_bytes 1 (u8) or 2 (u16). */
void or_gray(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra, or_real* gray);
void or_binary_support(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra, or_real* out);
void or_union_support(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra, or_real* out);
void or_apply_support(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra, or_real* weights);
void or_contrast(const or_real* gray, const or_real* support, uint32_t w, uint32_t h, uint32_t window, or_real* weights);
double or_sample_std(double sum, double sum_sqr, uint32_t n);
void or_normalize(const or_real* weights, const or_real* norm, uint32_t count, uint32_t total_images, or_real* out);

uint32_t or_level_count(uint32_t w, uint32_t h);
uint32_t or_next_level_size(uint32_t v);
void or_reduce_scalar_support(uint32_t sw, uint32_t sh, const or_real* src, const or_real* support,
                              uint32_t dw, uint32_t dh, or_real* dst, or_real* dst_support);
void or_reduce_rgb_support(uint32_t sw, uint32_t sh, const or_real* src, const or_real* support,
                           uint32_t dw, uint32_t dh, or_real* dst, or_real* dst_support);
void or_reduce_rgb_clamped(uint32_t sw, uint32_t sh, const or_real* src, uint32_t dw, uint32_t dh, or_real* dst);
void or_expand_rgb(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh, const or_real* src, or_real* dst);
void or_fill_rgb_base(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra, or_real* dst);

/* Whole-stack fusion of already-remapped (or pre-aligned) frames. */
enum { OR_METHOD_PYRAMID = 0, OR_METHOD_HYBRID = 1, OR_METHOD_HARDMASK = 2, OR_METHOD_SOFTMASK = 3 };
typedef struct {
    int method;               /* OR_METHOD_* */
    uint32_t window;          /* contrast window, odd >= 3 (reference default 5) */
    float sharpness;          /* hybrid mix in [0,1] (reference default 0.35) */
    uint32_t max_levels;      /* 0 = full depth down to 1x1 (reference behaviour) */
} or_fuse_opts;

/* Optional debug outputs (any pointer may be NULL).
 *   norm_sum, union_support : w*h
 *   weights                 : n*w*h   (contrast * alpha, i.e. after applySupport)
 *   acc_levels              : sum over levels of lw*lh*3 (accumulator before collapse)
 *   mask_levels, gauss_levels: levels of frame `frame_index` (sum lw*lh and lw*lh*3)
 *   soft, hard              : w*h*3 samples (hybrid only)                       */
typedef struct {
    or_real* norm_sum;
    or_real* union_support;
    or_real* weights;
    or_real* acc_levels;
    or_real* mask_levels;
    or_real* gauss_levels;
    int frame_index;
    void* soft;
    void* hard;
} or_fuse_debug;

int or_fuse_frames(int n, const void* const* frames, uint32_t w, uint32_t h, int sample_bytes, int extra,
                   const or_fuse_opts* opts, void* out, or_fuse_debug* dbg);

/* Remap every source under its pose into `roi`, then fuse (the focus_stack_zig path). */
int or_stack(int n, const void* const* sources, uint32_t src_w, uint32_t src_h, int sample_bytes,
             const or_pose* poses, const int32_t* roi, const or_fuse_opts* opts, void* out, or_fuse_debug* dbg);

/* Single-pass pieces: blurFiveTapInto (src/fuse/masks.zig:86-207), in == out allowed, and the
 * soft blend accumulate / finalize (src/fuse/blend.zig:51-186). */
void or_blur_five_tap(uint32_t w, uint32_t h, const or_real* in, or_real* scratch, or_real* out);
void or_soft_accumulate(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra,
                        const or_real* weights, or_real* weight_sums, or_real* pixel_sums);
void or_soft_finalize(const or_real* weight_sums, const or_real* pixel_sums, uint32_t w, uint32_t h, int sample_bytes,
                      void* out);

/* Hybrid pieces (a18). */
void or_update_winners(const void* img, int sample_bytes, uint32_t w, uint32_t h, int color, int extra,
                       const or_real* raw_weights, or_real* best, void* out);
void or_merge_hybrid(const void* soft, const void* hard, int sample_bytes, uint32_t w, uint32_t h,
                     uint32_t level_start, uint32_t level_count, float mix, uint32_t max_levels, void* out);
void or_collapse(uint32_t w, uint32_t h, uint32_t levels, const or_real* acc_levels, int sample_bytes, void* out);

#ifdef __cplusplus
}
#endif
#endif
