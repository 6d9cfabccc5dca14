/*
 * focus_oracle.c — CPU restatement of the reference's remap + pyramid-contrast
 * path.  TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg, to check results; the product library never
 * links or calls it.
 *
 * Each function names the reference lines (asingleoat/focus-stacker, paths
 * relative to the repository root) whose arithmetic it restates.  Operation
 * order, f32 rounding points and branch structure follow the reference so that
 * the f32 build (OR_REAL=float) reproduces the reference bit for bit wherever
 * the reference itself is deterministic.  Build with -ffp-contract=off (the
 * reference's Zig float mode is strict: no FMA contraction).
 *
 * Parity status of each part (what pins it) is documented in tests/ and in
 * the design notes; the parts with no external pin are marked "parity
 * unpinned" there.
 */
#include "focus_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef or_real R;
static const double PI = 3.14159265358979323846;

/* Threading.  Loops whose iterations are independent (one output pixel or row each) run in
 * parallel; every iteration performs exactly the operations of the serial loop, so the result is
 * the same bits for any thread count.  The reference's own order-dependent sums (the contrast's
 * sliding column sums, the norm-sum and accumulator frame order, the no-wrap expand's running
 * sums) keep their serial order (or_contrast splits its slide into a per-column phase and a
 * per-row phase, each in the reference's order). */
#define OR_PAR_MIN 16384
#define OR_STR(x) #x
#define OR_PRAGMA(x) _Pragma(OR_STR(x))
#define PAR_FOR(n) OR_PRAGMA(omp parallel for schedule(static) if ((size_t)(n) >= OR_PAR_MIN))

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* pose math — src/optimize.zig                                               */
/* ------------------------------------------------------------------------- */

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* focalLengthPixels, src/optimize.zig:3705-3708 */
static double focal_length_px(uint32_t width, double hfov_degrees) {
    double h = clampd(hfov_degrees, 1e-3, 179.0);
    return ((double)width * 0.5) / tan(h * PI / 360.0);
}
/* effectiveHfovDegrees, src/optimize.zig:3721-3723 */
static double effective_hfov(const or_pose* p) { return clampd(p->base_hfov_degrees + p->hfov_delta, 1e-3, 179.0); }

/* hasTranslation src/optimize.zig:3962-3964 */
int or_pose_has_translation(const or_pose* p) {
    return fabs(p->trans_x) > 1e-12 || fabs(p->trans_y) > 1e-12 || fabs(p->trans_z) > 1e-12;
}
/* hasBasicRectilinearPose src/optimize.zig:3074-3081 */
int or_pose_is_basic_rectilinear(const or_pose* p) {
    return !or_pose_has_translation(p) && fabs(p->radial_a) <= 1e-12 && fabs(p->radial_b) <= 1e-12 &&
           fabs(p->radial_c) <= 1e-12 && fabs(p->center_shift_x) <= 1e-12 && fabs(p->center_shift_y) <= 1e-12;
}

/* matrixMatrixMul src/optimize.zig:3951-3959 */
static void mat_mul(double a[3][3], double b[3][3], double out[3][3]) {
    double r[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) r[i][k] = a[i][0] * b[0][k] + a[i][1] * b[1][k] + a[i][2] * b[2][k];
    memcpy(out, r, sizeof r);
}
/* setMatrix src/optimize.zig:3901-3921 */
static void set_matrix(double a, double b, double c, int cl, double m[3][3]) {
    double mx[3][3] = {{1, 0, 0}, {0, cos(a), sin(a)}, {0, -sin(a), cos(a)}};
    double my[3][3] = {{cos(b), 0, -sin(b)}, {0, 1, 0}, {sin(b), 0, cos(b)}};
    double mz[3][3] = {{cos(c), sin(c), 0}, {-sin(c), cos(c), 0}, {0, 0, 1}};
    double dummy[3][3];
    if (cl) mat_mul(mz, mx, dummy); else mat_mul(mx, mz, dummy);
    mat_mul(dummy, my, m);
}
/* yawMatrix src/optimize.zig:3891-3899 */
static void yaw_matrix(double yaw, double m[3][3]) {
    double c = cos(yaw), s = sin(yaw);
    double t[3][3] = {{c, 0, -s}, {0, 1, 0}, {s, 0, c}};
    memcpy(m, t, sizeof t);
}
/* radialNormalizer src/optimize.zig:4150-4156 */
static double radial_normalizer(uint32_t w, uint32_t h) {
    double hw = ((double)w - 1.0) * 0.5, hh = ((double)h - 1.0) * 0.5;
    double d = hw * hw + hh * hh;
    return d >= 1e-12 ? 1.0 / d : 1.0;
}
/* cartErect src/optimize.zig:4045-4053 ; translationPlaneNormal :3966-3968 */
static void plane_normal(const or_pose* p, double n[3]) {
    double phi = p->translation_plane_yaw / 1.0;
    double tz = PI * 0.5 - (-p->translation_plane_pitch / 1.0);
    n[0] = sin(tz) * sin(phi);
    n[1] = cos(tz);
    n[2] = sin(tz) * -cos(phi);
}

/* initInverseTransformCache src/optimize.zig:3555-3586 */
void or_init_inverse_cache(const or_pose* p, uint32_t w, uint32_t h, or_inv_cache* c) {
    memset(c, 0, sizeof *c);
    double cx = ((double)w - 1.0) * 0.5, cy = ((double)h - 1.0) * 0.5;
    double pano = focal_length_px(w, p->base_hfov_degrees);
    double img_f = focal_length_px(w, effective_hfov(p));
    c->basic_rectilinear = or_pose_is_basic_rectilinear(p);
    c->has_translation = or_pose_has_translation(p);
    c->src_center_x = cx; c->src_center_y = cy;
    c->dest_center_x = cx; c->dest_center_y = cy;
    c->pano_distance = pano;
    c->half_turn = pano * PI;
    c->resize_scale = img_f / pano;
    c->dest_focal = pano;
    c->image_focal = img_f;
    set_matrix(-p->pitch, 0.0, p->roll, 1, c->pitch_roll);
    double ym[3][3];
    yaw_matrix(-p->yaw, ym);
    mat_mul(c->pitch_roll, ym, c->world_to_local);
    c->yaw_turn = -p->yaw * pano;
    c->yaw = p->yaw;
    c->radial_a = p->radial_a; c->radial_b = p->radial_b; c->radial_c = p->radial_c;
    c->center_shift_x = p->center_shift_x; c->center_shift_y = p->center_shift_y;
    c->radial_inv_norm = radial_normalizer(w, h);
    c->trans_x = p->trans_x; c->trans_y = p->trans_y; c->trans_z = p->trans_z;
    plane_normal(p, c->plane_normal);
}

/* erectRectPoint src/optimize.zig:3337-3345 */
static void erect_rect(double* x, double* y, double d) {
    double px = *x, py = *y;
    *x = d * atan2(px, d);
    *y = d * atan2(py, sqrt(d * d + px * px));
}
/* rotateErectPoint src/optimize.zig:3347-3362 */
static void rotate_erect(double* x, double half_turn, double turn) {
    double v = *x + turn, full = 2.0 * half_turn;
    if (v < -half_turn) { v += full; while (v < -half_turn) v += full; }
    else if (v > half_turn) { v -= full; while (v > half_turn) v -= full; }
    *x = v;
}
/* sphereTpErectPoint src/optimize.zig:3247-3270 */
static void sphere_tp_erect(double* x, double* y, double d) {
    double phi = *x / d;
    double theta = -(*y / d) + PI * 0.5;
    if (theta < 0.0) { theta = -theta; phi += PI; }
    if (theta > PI) { theta = PI - (theta - PI); phi += PI; }
    double s = sin(theta), v0 = s * sin(phi), v1 = cos(theta);
    double r = sqrt(v1 * v1 + v0 * v0);
    double scale = r == 0.0 ? 0.0 : d * atan2(r, s * cos(phi)) / r;
    *x = scale * v0; *y = scale * v1;
}
/* perspSpherePoint src/optimize.zig:3272-3292 (m applied transposed) */
static void persp_sphere(double* x, double* y, double m[3][3], double d) {
    double px = *x, py = *y;
    double r = sqrt(px * px + py * py);
    double theta = r / d;
    double st = sin(theta);
    double rs = r == 0.0 ? 0.0 : st / r;
    double vx = rs * px, vy = rs * py, vz = cos(theta);
    double tx = m[0][0] * vx + m[1][0] * vy + m[2][0] * vz;
    double ty = m[0][1] * vx + m[1][1] * vy + m[2][1] * vz;
    double tz = m[0][2] * vx + m[1][2] * vy + m[2][2] * vz;
    double r2 = sqrt(tx * tx + ty * ty);
    double scale = r2 == 0.0 ? 0.0 : d * atan2(r2, tz) / r2;
    *x = scale * tx; *y = scale * ty;
}
/* rectSphereTpPoint src/optimize.zig:3309-3324 */
static void rect_sphere_tp(double* x, double* y, double d) {
    double r = sqrt(*x * *x + *y * *y);
    double theta = r / d;
    double rho = theta >= PI * 0.5 ? 1.6e16 : (theta == 0.0 ? 1.0 : tan(theta) / theta);
    *x = rho * *x; *y = rho * *y;
}

/* basicRectilinearInverseTransformPointCachedLegacy src/optimize.zig:3618-3634 */
void or_inverse_legacy_chain(const or_inv_cache* c, double ox, double oy, double* sx, double* sy) {
    double x = ox - c->dest_center_x, y = oy - c->dest_center_y;
    double pr[3][3];
    memcpy(pr, c->pitch_roll, sizeof pr);
    erect_rect(&x, &y, c->pano_distance);
    rotate_erect(&x, c->half_turn, c->yaw_turn);
    sphere_tp_erect(&x, &y, c->pano_distance);
    persp_sphere(&x, &y, pr, c->pano_distance);
    rect_sphere_tp(&x, &y, c->pano_distance);
    x *= c->resize_scale; y *= c->resize_scale;
    *sx = c->src_center_x + x; *sy = c->src_center_y + y;
}

/* invertRadialDistortionCached src/optimize.zig:4204-4235 (returns the offset) */
static void invert_radial_cached(const or_inv_cache* c, double dx, double dy, double* ux, double* uy) {
    double q = sqrt(dx * dx + dy * dy);
    if (q < 1e-12) { *ux = dx; *uy = dy; return; }
    double radius = q;
    for (int it = 0; it < 6; it++) {
        double nr2 = radius * radius * c->radial_inv_norm;
        double nr4 = nr2 * nr2;
        double factor = 1.0 + c->radial_a * nr2 + c->radial_b * nr4 + c->radial_c * nr4 * nr2;
        double fd = c->radial_a + 2.0 * c->radial_b * nr2 + 3.0 * c->radial_c * nr4;
        double f = radius * factor - q;
        if (fabs(f) <= 1e-9) break;
        double df = factor + 2.0 * radius * radius * c->radial_inv_norm * fd;
        if (fabs(df) < 1e-12) break;
        radius -= f / df;
        radius = radius > 0.0 ? radius : 0.0;
    }
    double scale = radius / q;
    *ux = dx * scale; *uy = dy * scale;
}

/* imagePlaneFromWorldRayCached src/optimize.zig:3859-3881 */
static void image_plane_cached(const or_inv_cache* c, const double r[3], double* x, double* y) {
    const double (*m)[3] = c->world_to_local;
    double lx = m[0][0] * r[0] + m[0][1] * r[1] + m[0][2] * r[2];
    double ly = m[1][0] * r[0] + m[1][1] * r[1] + m[1][2] * r[2];
    double lz = m[2][0] * r[0] + m[2][1] * r[1] + m[2][2] * r[2];
    double denom = -lz;
    if (fabs(denom) < 1e-12) { *x = -1.0; *y = -1.0; return; }
    *x = c->image_focal * (lx / denom) + c->center_shift_x;
    *y = c->image_focal * (ly / denom) + c->center_shift_y;
}

/* inverseTransformPointCached src/optimize.zig:3593-3616 (+ translation path
 * rayFromTranslationPlaneCachedUnnormalized :4009-4025, linePlaneIntersection :4027-4043) */
void or_inverse_transform_cached(const or_inv_cache* c, double ox, double oy, double* sx, double* sy) {
    if (c->basic_rectilinear) { or_inverse_legacy_chain(c, ox, oy, sx, sy); return; }
    double ray[3] = {ox - c->dest_center_x, oy - c->dest_center_y, -c->dest_focal};
    if (c->has_translation) {
        const double* n = c->plane_normal;
        double num = 1.0 - (n[0] * 0.0 + n[1] * 0.0 + n[2] * 0.0);
        double den = n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2];
        if (fabs(den) < 1e-12) { *sx = -1.0; *sy = -1.0; return; }
        double u = num / den;
        double hit[3] = {0.0 + u * ray[0], 0.0 + u * ray[1], 0.0 + u * ray[2]};
        ray[0] = hit[0] - c->trans_x; ray[1] = hit[1] - c->trans_y; ray[2] = hit[2] - c->trans_z;
    }
    double rx, ry, ux, uy;
    image_plane_cached(c, ray, &rx, &ry);
    invert_radial_cached(c, rx, ry, &ux, &uy);
    *sx = c->src_center_x + ux; *sy = c->src_center_y + uy;
}

/* ---- forward transform (computeCommonOverlapRoi only) ---- */

/* sphereTpRectPoint src/optimize.zig:3238-3245 */
static void sphere_tp_rect(double* x, double* y, double d) {
    double r = sqrt(*x * *x + *y * *y) / d;
    double th = r == 0.0 ? 1.0 : atan(r) / r;
    *x = th * *x; *y = th * *y;
}
/* erectSphereTpPoint src/optimize.zig:3294-3307 */
static void erect_sphere_tp(double* x, double* y, double d) {
    double r = sqrt(*x * *x + *y * *y);
    double th = r / d;
    double s = th == 0.0 ? 1.0 / d : sin(th) / r;
    double v1 = s * *x, v0 = cos(th);
    double py = *y;
    *x = d * atan2(v1, v0);
    *y = d * atan((s * py) / sqrt(v0 * v0 + v1 * v1));
}
/* rectErectPoint src/optimize.zig:3326-3335 ; returns 0 when invalid */
static int rect_erect(double* x, double* y, double d) {
    double phi = *x / d;
    if (phi < -PI * 0.5 || phi > PI * 0.5) return 0;
    double theta = -(*y / d) + PI * 0.5;
    *x = d * tan(phi);
    *y = d / (tan(theta) * cos(phi));
    return 1;
}
static void normalize3(double v[3]) {
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n < 1e-12) { v[0] = 0; v[1] = 0; v[2] = 1; return; }
    v[0] = v[0] / n; v[1] = v[1] / n; v[2] = v[2] / n;
}

/* transformPoint src/optimize.zig:3506-3525 (+ basicRectilinearTransformPoint :3662-3683,
 * sourceWorldRay :3781-3793, applyRadialDistortion :4164-4172, projectViaTranslationPlane
 * :3970-3978, projectRectilinear :4073-4080) */
void or_transform_point(const or_pose* p, double x, double y, uint32_t w, uint32_t h, double* ox, double* oy) {
    double cx = ((double)w - 1.0) * 0.5, cy = ((double)h - 1.0) * 0.5;
    double pano = focal_length_px(w, p->base_hfov_degrees);
    if (or_pose_is_basic_rectilinear(p)) {
        double img_f = focal_length_px(w, effective_hfov(p));
        double rs = pano / img_f;
        double px = (x - cx) * rs, py = (y - cy) * rs;
        double m[3][3];
        sphere_tp_rect(&px, &py, pano);
        set_matrix(p->pitch, 0.0, p->roll, 1, m);
        persp_sphere(&px, &py, m, pano);
        erect_sphere_tp(&px, &py, pano);
        rotate_erect(&px, pano * PI, p->yaw * pano);
        if (!rect_erect(&px, &py, pano)) { *ox = -1.0; *oy = -1.0; return; }
        *ox = cx + px; *oy = cy + py;
        return;
    }
    double dx = (x - cx) - p->center_shift_x, dy = (y - cy) - p->center_shift_y;
    double nr2 = (dx * dx + dy * dy) * radial_normalizer(w, h);
    double r4 = nr2 * nr2, r6 = r4 * nr2;
    double factor = 1.0 + p->radial_a * nr2 + p->radial_b * r4 + p->radial_c * r6;
    double v[3] = {dx * factor, dy * factor, -focal_length_px(w, effective_hfov(p))};
    normalize3(v);
    double m[3][3];
    set_matrix(-p->pitch, 0.0, p->roll, 1, m);
    double t[3] = {m[0][0] * v[0] + m[1][0] * v[1] + m[2][0] * v[2],
                   m[0][1] * v[0] + m[1][1] * v[1] + m[2][1] * v[2],
                   m[0][2] * v[0] + m[1][2] * v[1] + m[2][2] * v[2]};
    double c = cos(p->yaw), s = sin(p->yaw);
    double ray[3] = {c * t[0] - s * t[2], t[1], s * t[0] + c * t[2]};
    if (or_pose_has_translation(p)) {
        double n[3];
        plane_normal(p, n);
        double cam[3] = {p->trans_x, p->trans_y, p->trans_z};
        double tgt[3] = {cam[0] + ray[0], cam[1] + ray[1], cam[2] + ray[2]};
        double dir[3] = {tgt[0] - cam[0], tgt[1] - cam[1], tgt[2] - cam[2]};
        double num = 1.0 - (n[0] * cam[0] + n[1] * cam[1] + n[2] * cam[2]);
        double den = n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2];
        if (fabs(den) < 1e-12) { ray[0] = 0.0; ray[1] = 0.0; ray[2] = 1.0; }
        else {
            double u = num / den;
            ray[0] = cam[0] + u * dir[0]; ray[1] = cam[1] + u * dir[1]; ray[2] = cam[2] + u * dir[2];
        }
    }
    double denom = -ray[2];
    if (denom <= 1e-12) { *ox = -1.0; *oy = -1.0; return; }
    *ox = cx + pano * (ray[0] / denom);
    *oy = cy + pano * (ray[1] / denom);
}

/* computeCommonOverlapRoi src/remap.zig:585-633 with clipPolygonToConvex :651-687,
 * isInsideEdge :689-691, lineIntersection :693-708 */
typedef struct { double x, y; } v2;
static int inside_edge(v2 p, v2 a, v2 b) { return ((b.x - a.x) * (p.y - a.y) - (b.y - a.y) * (p.x - a.x)) >= -1e-6; }
static v2 line_x(v2 p1, v2 p2, v2 q1, v2 q2) {
    double a1 = p2.y - p1.y, b1 = p1.x - p2.x, c1 = a1 * p1.x + b1 * p1.y;
    double a2 = q2.y - q1.y, b2 = q1.x - q2.x, c2 = a2 * q1.x + b2 * q1.y;
    double det = a1 * b2 - a2 * b1;
    if (fabs(det) < 1e-9) return p2;
    v2 r = {(b2 * c1 - b1 * c2) / det, (a1 * c2 - a2 * c1) / det};
    return r;
}
int or_common_overlap_roi(int n, const uint8_t* active, const uint32_t* widths, const uint32_t* heights,
                          const or_pose* poses, int32_t* roi) {
    size_t cap = 8 + 4 * (size_t)(n > 0 ? n : 1);
    v2* poly = malloc(sizeof(v2) * cap);
    v2* next = malloc(sizeof(v2) * cap);
    int np = 0, first = 1, found = 0;
    for (int i = 0; i < n; i++) {
        if (active && !active[i]) continue;
        uint32_t w = widths[i], h = heights[i];
        v2 q[4];
        double xs[4] = {0, (double)(w - 1), (double)(w - 1), 0}, ys[4] = {0, 0, (double)(h - 1), (double)(h - 1)};
        for (int k = 0; k < 4; k++) or_transform_point(&poses[i], xs[k], ys[k], w, h, &q[k].x, &q[k].y);
        if (first) { memcpy(poly, q, sizeof q); np = 4; first = 0; continue; }
        for (int e = 0; e < 4; e++) {
            v2 a = q[e], b = q[(e + 1) % 4];
            int nn = 0;
            if (np == 0) break;
            v2 prev = poly[np - 1];
            int prev_in = inside_edge(prev, a, b);
            for (int k = 0; k < np; k++) {
                v2 cur = poly[k];
                int cur_in = inside_edge(cur, a, b);
                if (cur_in != prev_in && (size_t)nn < cap) next[nn++] = line_x(prev, cur, a, b);
                if (cur_in && (size_t)nn < cap) next[nn++] = cur;
                prev = cur; prev_in = cur_in;
            }
            memcpy(poly, next, sizeof(v2) * nn);
            np = nn;
        }
        if (np == 0) goto out;
    }
    if (np == 0) goto out;
    double mnx = poly[0].x, mxx = poly[0].x, mny = poly[0].y, mxy = poly[0].y;
    for (int k = 1; k < np; k++) {
        mnx = fmin(mnx, poly[k].x); mxx = fmax(mxx, poly[k].x);
        mny = fmin(mny, poly[k].y); mxy = fmax(mxy, poly[k].y);
    }
    roi[0] = (int32_t)ceil(mnx); roi[1] = (int32_t)ceil(mny);
    roi[2] = (int32_t)floor(mxx) + 1; roi[3] = (int32_t)floor(mxy) + 1;
    found = !(roi[2] <= roi[0] || roi[3] <= roi[1]);
out:
    free(poly); free(next);
    return found;
}

/* ------------------------------------------------------------------------- */
/* remap — src/remap.zig                                                      */
/* ------------------------------------------------------------------------- */

/* samplePixelBilinearU8/U16 src/remap.zig:710-753, 841-884 ; interior :788-839, 886-937 ;
 * supportFromWeights :939-946 ; bilinearChannel*Soft :948-994.  The first guard
 * stands in for the reference's i32 conversion of far-away coordinates: every
 * finite coordinate outside [-1, W) x [-1, H) has zero in-bounds weight there. */
#define DEF_SAMPLE(NAME, T, MAXV)                                                                       \
    static void NAME(T* dst, const T* px, uint32_t w, uint32_t h, int ch, double x, double y) {        \
        if (!(x >= -1.0 && x < (double)w && y >= -1.0 && y < (double)h)) {                              \
            for (int c = 0; c <= ch; c++) dst[c] = 0;                                                   \
            return;                                                                                     \
        }                                                                                               \
        int32_t x0 = (int32_t)floor(x), y0 = (int32_t)floor(y), x1 = x0 + 1, y1 = y0 + 1;             \
        double fx = x - (double)x0, fy = y - (double)y0;                                                \
        double omx = 1.0 - fx, omy = 1.0 - fy;                                                          \
        double w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;                          \
        if (x0 >= 0 && y0 >= 0 && x1 < (int32_t)w && y1 < (int32_t)h) {                                \
            size_t rs = (size_t)w * ch;                                                                 \
            size_t b00 = (size_t)y0 * rs + (size_t)x0 * ch, b10 = b00 + ch, b01 = b00 + rs, b11 = b01 + ch; \
            double sup = w00 + w10 + w01 + w11;                                                         \
            for (int c = 0; c < ch; c++) {                                                              \
                double v = (double)px[b00 + c] * w00 + (double)px[b10 + c] * w10 +                      \
                           (double)px[b01 + c] * w01 + (double)px[b11 + c] * w11;                       \
                dst[c] = (T)round(v / sup);                                                             \
            }                                                                                           \
            dst[ch] = MAXV;                                                                             \
            return;                                                                                     \
        }                                                                                               \
        int i00 = x0 >= 0 && y0 >= 0 && x0 < (int32_t)w && y0 < (int32_t)h;                            \
        int i10 = x1 >= 0 && y0 >= 0 && x1 < (int32_t)w && y0 < (int32_t)h;                            \
        int i01 = x0 >= 0 && y1 >= 0 && x0 < (int32_t)w && y1 < (int32_t)h;                            \
        int i11 = x1 >= 0 && y1 >= 0 && x1 < (int32_t)w && y1 < (int32_t)h;                            \
        double sup = 0.0;                                                                               \
        if (i00) sup += w00;                                                                            \
        if (i10) sup += w10;                                                                            \
        if (i01) sup += w01;                                                                            \
        if (i11) sup += w11;                                                                            \
        if (sup <= 0.0) {                                                                               \
            for (int c = 0; c <= ch; c++) dst[c] = 0;                                                   \
            return;                                                                                     \
        }                                                                                               \
        for (int c = 0; c < ch; c++) {                                                                  \
            double v = 0.0;                                                                             \
            if (i00) v += (double)px[((size_t)y0 * w + x0) * ch + c] * w00;                             \
            if (i10) v += (double)px[((size_t)y0 * w + x1) * ch + c] * w10;                             \
            if (i01) v += (double)px[((size_t)y1 * w + x0) * ch + c] * w01;                             \
            if (i11) v += (double)px[((size_t)y1 * w + x1) * ch + c] * w11;                             \
            dst[c] = (T)round(v / sup);                                                                 \
        }                                                                                               \
        dst[ch] = (T)round(sup * (double)MAXV);                                                         \
    }
DEF_SAMPLE(sample_u8, uint8_t, 255)
DEF_SAMPLE(sample_u16, uint16_t, 65535)

/* undistortSourcePoint src/remap.zig:755-786 */
static void undistort_source(const or_inv_cache* c, double dx, double dy, double* sx, double* sy) {
    double ux, uy;
    double q = sqrt(dx * dx + dy * dy);
    if (q < 1e-12) { *sx = c->src_center_x + dx; *sy = c->src_center_y + dy; return; }
    invert_radial_cached(c, dx, dy, &ux, &uy);
    *sx = c->src_center_x + ux; *sy = c->src_center_y + uy;
}

/* remapRigidImage/remapU8/remapU16 src/remap.zig:166-336 with the row kernels
 * remapU8Rows :343-375 and remapU8RowsNoTranslationLens :377-441 (U16 twins :443-541),
 * run single-threaded (jobs = 1). */
#define DEF_REMAP(NAME, T, SAMPLE)                                                                      \
    void NAME(const T* src, uint32_t w, uint32_t h, int ch, const or_pose* pose, const int32_t* roi_in, T* dst) { \
        int32_t roi[4] = {0, 0, (int32_t)w, (int32_t)h};                                                \
        if (roi_in) memcpy(roi, roi_in, sizeof roi);                                                    \
        uint32_t ow = (uint32_t)(roi[2] - roi[0]), oh = (uint32_t)(roi[3] - roi[1]);                    \
        int dch = ch + 1;                                                                               \
        or_inv_cache c;                                                                                 \
        or_init_inverse_cache(pose, w, h, &c);                                                          \
        double rl = (double)roi[0], rt = (double)roi[1];                                                \
        if (!c.basic_rectilinear && !c.has_translation) {                                               \
            double stx = c.world_to_local[0][0], sty = c.world_to_local[1][0], stz = c.world_to_local[2][0]; \
            double cz = -c.dest_focal, bx = rl - c.dest_center_x;                                       \
            int radial = c.radial_a != 0.0 || c.radial_b != 0.0 || c.radial_c != 0.0;                  \
            PAR_FOR((size_t)ow * oh)                                                                    \
            for (uint32_t y = 0; y < oh; y++) {                                                         \
                double py = (rt + (double)y) - c.dest_center_y;                                         \
                double lx = c.world_to_local[0][0] * bx + c.world_to_local[0][1] * py + c.world_to_local[0][2] * cz; \
                double ly = c.world_to_local[1][0] * bx + c.world_to_local[1][1] * py + c.world_to_local[1][2] * cz; \
                double lz = c.world_to_local[2][0] * bx + c.world_to_local[2][1] * py + c.world_to_local[2][2] * cz; \
                T* d = dst + (size_t)y * ow * dch;                                                      \
                for (uint32_t x = 0; x < ow; x++, d += dch) {                                           \
                    double denom = -lz;                                                                 \
                    if (fabs(denom) < 1e-12) {                                                          \
                        for (int k = 0; k < dch; k++) d[k] = 0;                                         \
                    } else {                                                                            \
                        double rx = c.image_focal * (lx / denom) + c.center_shift_x;                    \
                        double ry = c.image_focal * (ly / denom) + c.center_shift_y;                    \
                        double sx, sy;                                                                  \
                        if (radial) undistort_source(&c, rx, ry, &sx, &sy);                             \
                        else { sx = c.src_center_x + rx; sy = c.src_center_y + ry; }                    \
                        SAMPLE(d, src, w, h, ch, sx, sy);                                               \
                    }                                                                                   \
                    lx += stx; ly += sty; lz += stz;                                                    \
                }                                                                                       \
            }                                                                                           \
            return;                                                                                     \
        }                                                                                               \
        PAR_FOR((size_t)ow * oh)                                                                        \
        for (uint32_t y = 0; y < oh; y++) {                                                             \
            double wy = rt + (double)y, wx = rl;                                                        \
            T* d = dst + (size_t)y * ow * dch;                                                          \
            for (uint32_t x = 0; x < ow; x++, d += dch) {                                               \
                double sx, sy;                                                                          \
                or_inverse_transform_cached(&c, wx, wy, &sx, &sy);                                      \
                SAMPLE(d, src, w, h, ch, sx, sy);                                                       \
                wx += 1.0;                                                                              \
            }                                                                                           \
        }                                                                                               \
    }
DEF_REMAP(or_remap_u8, uint8_t, sample_u8)
DEF_REMAP(or_remap_u16, uint16_t, sample_u16)

/* ------------------------------------------------------------------------- */
/* fusion building blocks — src/fuse/                                    */
/* ------------------------------------------------------------------------- */

static inline uint32_t sample_at(const void* img, int sb, size_t i) {
    return sb == 1 ? ((const uint8_t*)img)[i] : ((const uint16_t*)img)[i];
}

/* fillAverageU8/U16 src/fuse/grayscale.zig:21-57 */
void or_gray(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, R* gray) {
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++) {
        if (color == 1) {
            gray[i] = sb == 1 ? (R)sample_at(img, sb, i * st) / (R)255.0 : (R)sample_at(img, sb, i * st) / (R)65535.0;
        } else {
            uint32_t s = sample_at(img, sb, i * st) + sample_at(img, sb, i * st + 1) + sample_at(img, sb, i * st + 2);
            gray[i] = sb == 1 ? (R)s / (R)(3.0 * 255.0) : (R)s / (R)(3.0 * 65535.0);
        }
    }
}
/* fillBinarySupport src/fuse/masks.zig:32-57 */
void or_binary_support(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, R* out) {
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++) out[i] = extra == 0 ? (R)1 : (sample_at(img, sb, i * st + color) > 0 ? (R)1 : (R)0);
}
/* accumulateBinarySupportMax src/fuse/masks.zig:59-84 */
void or_union_support(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, R* out) {
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++)
        if (extra == 0 || sample_at(img, sb, i * st + color) > 0) out[i] = (R)1;
}
/* applySupportInto src/fuse/masks.zig:6-30 */
void or_apply_support(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, R* weights) {
    if (extra == 0) return;
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++) {
        R s = (R)sample_at(img, sb, i * st + color) / (sb == 1 ? (R)255.0 : (R)65535.0);
        weights[i] *= s;
    }
}

/* sampleStdDev src/fuse/contrast.zig:284-290 */
double or_sample_std(double sum, double sum_sqr, uint32_t n) {
    if (n <= 1) return 0.0;
    double denom = (double)(n - 1), nf = (double)n;
    double var = (sum_sqr - ((sum * sum) / nf)) / denom;
    if (!(var > 0.0)) var = 0.0;
    return sqrt(var);
}

/* computeLocalContrastWeightsWithWorkspace src/fuse/contrast.zig:96-154 with computeRange
 * :172-282 at jobs = 1.  The square of each sample is formed in the gray type (f32 in the
 * reference: `value * value` of two f32) and then widened into the f64 running sums. */
void or_contrast(const R* px, const R* support, uint32_t w, uint32_t h, uint32_t window, R* weights) {
    size_t count = (size_t)w * h;
    memset(weights, 0, count * sizeof(R));
    if (window < 3 || w < window || h < window) return;
    uint32_t border = window / 2;
    uint32_t first_row = border, last_row = h - border;
    if (first_row >= last_row) return;
    /* Phase 1, per column (independent columns): the column sums of each output row, formed by
     * the reference's slide (initial window top to bottom, then per row: remove the row leaving,
     * add the row entering).  Phase 2, per row (independent rows): the horizontal slide over those
     * sums.  Every value takes the operations of the single-threaded slide in its order. */
    size_t rows = last_row - first_row;
    double* sums = malloc(sizeof(double) * rows * w);
    double* sqr = malloc(sizeof(double) * rows * w);
    uint32_t* cnt = malloc(sizeof(uint32_t) * rows * w);
    PAR_FOR((size_t)w * h)
    for (uint32_t c = 0; c < w; c++) {
        double cs = 0, cq = 0;
        uint32_t cn = 0;
        for (uint32_t yy = first_row - border; yy <= first_row + border; yy++) {
            size_t idx = (size_t)yy * w + c;
            if (support[idx] <= (R)0) continue;
            R v = px[idx];
            cs += (double)v;
            cq += (double)(R)(v * v);
            cn += 1;
        }
        for (uint32_t row = first_row; row < last_row; row++) {
            size_t o = (size_t)(row - first_row) * w + c;
            sums[o] = cs; sqr[o] = cq; cnt[o] = cn;
            if (row + 1 >= last_row) break;
            size_t rb = (size_t)(row - border) * w + c, ab = (size_t)(row + border + 1) * w + c;
            if (support[rb] > (R)0) {
                R v = px[rb];
                cs -= (double)v; cq -= (double)(R)(v * v); cn -= 1;
            }
            if (support[ab] > (R)0) {
                R v = px[ab];
                cs += (double)v; cq += (double)(R)(v * v); cn += 1;
            }
        }
    }
    PAR_FOR((size_t)w * h)
    for (uint32_t row = first_row; row < last_row; row++) {
        const double* rs = sums + (size_t)(row - first_row) * w;
        const double* rq = sqr + (size_t)(row - first_row) * w;
        const uint32_t* rn = cnt + (size_t)(row - first_row) * w;
        double sum = 0, sum_sqr = 0;
        uint32_t n = 0;
        for (uint32_t c = 0; c < window; c++) { sum += rs[c]; sum_sqr += rq[c]; n += rn[c]; }
        uint32_t x = border;
        size_t out = (size_t)row * w + border;
        for (;;) {
            weights[out] = support[out] > (R)0 ? (R)or_sample_std(sum, sum_sqr, n) : (R)0;
            if (x == w - border - 1) break;
            uint32_t nx = x + border + 1, old = x - border;
            sum += rs[nx] - rs[old];
            sum_sqr += rq[nx] - rq[old];
            n += rn[nx] - rn[old];
            x++; out++;
        }
    }
    free(sums); free(sqr); free(cnt);
}

/* normalizeWeightsInto src/fuse/pyramid.zig:303-316 */
void or_normalize(const R* wgt, const R* norm, uint32_t count, uint32_t total, R* out) {
    R def = (R)1.0 / (R)total;
    PAR_FOR(count)
    for (uint32_t i = 0; i < count; i++) out[i] = norm[i] > (R)1e-12 ? wgt[i] / norm[i] : def;
}

/* computeLevelCount / nextLevelSize src/fuse/pyramid.zig:163-176 */
uint32_t or_next_level_size(uint32_t v) { return v > 1 ? (v + 1) / 2 : 1; }
uint32_t or_level_count(uint32_t w, uint32_t h) {
    uint32_t levels = 1;
    while (w > 1 || h > 1) { w = or_next_level_size(w); h = or_next_level_size(h); levels++; }
    return levels;
}

static const R KW[5] = {1, 4, 6, 4, 1}; /* kernelWeight src/fuse/pyramid.zig:1480-1487 */

/* reduceScalarWithSupportRowsSequential src/fuse/pyramid.zig:997-1042 */
void or_reduce_scalar_support(uint32_t sw, uint32_t sh, const R* src, const R* sup, uint32_t dw, uint32_t dh,
                              R* dst, R* dsup) {
    PAR_FOR((size_t)dw * dh)
    for (uint32_t dy = 0; dy < dh; dy++)
        for (uint32_t dx = 0; dx < dw; dx++) {
            int32_t cy = (int32_t)dy * 2, cx = (int32_t)dx * 2;
            R vs = 0, ws = 0;
            for (int ky = 0; ky < 5; ky++) {
                int32_t sy = cy + ky - 2;
                if (sy < 0 || sy >= (int32_t)sh) continue;
                R wy = KW[ky];
                for (int kx = 0; kx < 5; kx++) {
                    int32_t sx = cx + kx - 2;
                    if (sx < 0 || sx >= (int32_t)sw) continue;
                    size_t i = (size_t)sy * sw + sx;
                    R s = sup[i];
                    if (s <= (R)0) continue;
                    R wt = wy * KW[kx] * s;
                    vs += src[i] * wt;
                    ws += wt;
                }
            }
            size_t o = (size_t)dy * dw + dx;
            if (ws > (R)0) { dst[o] = vs / ws; dsup[o] = 1; }
            else { dst[o] = 0; dsup[o] = 0; }
        }
}

/* reduceRgbWithSupportRowsSequential src/fuse/pyramid.zig:1044-1099 */
void or_reduce_rgb_support(uint32_t sw, uint32_t sh, const R* src, const R* sup, uint32_t dw, uint32_t dh,
                           R* dst, R* dsup) {
    PAR_FOR((size_t)dw * dh)
    for (uint32_t dy = 0; dy < dh; dy++)
        for (uint32_t dx = 0; dx < dw; dx++) {
            int32_t cy = (int32_t)dy * 2, cx = (int32_t)dx * 2;
            R s0 = 0, s1 = 0, s2 = 0, ws = 0;
            for (int ky = 0; ky < 5; ky++) {
                int32_t sy = cy + ky - 2;
                if (sy < 0 || sy >= (int32_t)sh) continue;
                R wy = KW[ky];
                for (int kx = 0; kx < 5; kx++) {
                    int32_t sx = cx + kx - 2;
                    if (sx < 0 || sx >= (int32_t)sw) continue;
                    size_t i = (size_t)sy * sw + sx;
                    R s = sup[i];
                    if (s <= (R)0) continue;
                    R wt = wy * KW[kx] * s;
                    s0 += src[i * 3 + 0] * wt;
                    s1 += src[i * 3 + 1] * wt;
                    s2 += src[i * 3 + 2] * wt;
                    ws += wt;
                }
            }
            size_t o = (size_t)dy * dw + dx;
            if (ws > (R)0) {
                dst[o * 3 + 0] = s0 / ws; dst[o * 3 + 1] = s1 / ws; dst[o * 3 + 2] = s2 / ws; dsup[o] = 1;
            } else {
                dst[o * 3 + 0] = 0; dst[o * 3 + 1] = 0; dst[o * 3 + 2] = 0; dsup[o] = 0;
            }
        }
}

/* clampCoord src/fuse/pyramid.zig:1837-1842 */
static uint32_t clamp_coord(int32_t c, uint32_t limit) {
    if (c <= 0) return 0;
    int32_t mx = (int32_t)limit - 1;
    return c >= mx ? (uint32_t)mx : (uint32_t)c;
}

/* reduceRgbRowsSequential src/fuse/pyramid.zig:952-995 with horizontalReduceRgbRow :1726-1746 */
void or_reduce_rgb_clamped(uint32_t sw, uint32_t sh, const R* src, uint32_t dw, uint32_t dh, R* dst) {
    OR_PRAGMA(omp parallel if ((size_t)dw * dh >= OR_PAR_MIN))
    {
    R* rows = malloc(sizeof(R) * 5 * (size_t)dw * 3);
    OR_PRAGMA(omp for schedule(static))
    for (uint32_t dy = 0; dy < dh; dy++) {
        for (int ky = 0; ky < 5; ky++) {
            uint32_t y = clamp_coord((int32_t)dy * 2 + ky - 2, sh);
            const R* s = src + (size_t)y * sw * 3;
            R* r = rows + (size_t)ky * dw * 3;
            for (uint32_t dx = 0; dx < dw; dx++) {
                int32_t sx = (int32_t)dx * 2;
                uint32_t x0 = clamp_coord(sx - 2, sw), x1 = clamp_coord(sx - 1, sw), x2 = clamp_coord(sx, sw),
                         x3 = clamp_coord(sx + 1, sw), x4 = clamp_coord(sx + 2, sw);
                for (int c = 0; c < 3; c++)
                    r[dx * 3 + c] = (s[x0 * 3 + c] + (R)4.0 * s[x1 * 3 + c] + (R)6.0 * s[x2 * 3 + c] +
                                     (R)4.0 * s[x3 * 3 + c] + s[x4 * 3 + c]) * ((R)1.0 / (R)16.0);
            }
        }
        for (uint32_t dx = 0; dx < dw; dx++)
            for (int c = 0; c < 3; c++) {
                size_t k = (size_t)dx * 3 + c, L = (size_t)dw * 3;
                dst[((size_t)dy * dw + dx) * 3 + c] =
                    (rows[k] + (R)4.0 * rows[L + k] + (R)6.0 * rows[2 * L + k] + (R)4.0 * rows[3 * L + k] +
                     rows[4 * L + k]) * ((R)1.0 / (R)16.0);
            }
    }
    free(rows);
    }
}

/* horizontalExpandRgbRow src/fuse/pyramid.zig:1748-1775 */
static void h_expand_row(uint32_t sw, const R* srow, uint32_t dw, R* out) {
    for (uint32_t dx = 0; dx < dw; dx++) {
        int32_t sx = (int32_t)(dx / 2);
        if ((dx & 1) == 0) {
            uint32_t x0 = clamp_coord(sx - 1, sw), x1 = clamp_coord(sx, sw), x2 = clamp_coord(sx + 1, sw);
            for (int c = 0; c < 3; c++)
                out[dx * 3 + c] = (srow[x0 * 3 + c] + (R)6.0 * srow[x1 * 3 + c] + srow[x2 * 3 + c]) * ((R)1.0 / (R)8.0);
        } else {
            uint32_t x0 = clamp_coord(sx, sw), x1 = clamp_coord(sx + 1, sw);
            for (int c = 0; c < 3; c++) out[dx * 3 + c] = (srow[x0 * 3 + c] + srow[x1 * 3 + c]) * (R)0.5;
        }
    }
}

/* expandRgbRowsSequential src/fuse/pyramid.zig:1101-1149.  For odd output rows the
 * reference fetches y0 = y1 = clamp(dy/2) (:1127-1128), so the row is the horizontal
 * expansion of source row dy/2 alone, not the average of rows dy/2 and dy/2+1. */
static void expand_clamp(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh, const R* src, R* dst) {
    OR_PRAGMA(omp parallel if ((size_t)dw * dh >= OR_PAR_MIN))
    {
    R* r0 = malloc(sizeof(R) * (size_t)dw * 3);
    R* r1 = malloc(sizeof(R) * (size_t)dw * 3);
    R* r2 = malloc(sizeof(R) * (size_t)dw * 3);
    OR_PRAGMA(omp for schedule(static))
    for (uint32_t dy = 0; dy < dh; dy++) {
        int even = (dy & 1) == 0;
        int32_t by = (int32_t)(dy / 2);
        uint32_t y0 = even ? clamp_coord(by - 1, sh) : clamp_coord(by, sh);
        uint32_t y1 = clamp_coord(by, sh);
        uint32_t y2 = clamp_coord(by + 1, sh);
        h_expand_row(sw, src + (size_t)y0 * sw * 3, dw, r0);
        h_expand_row(sw, src + (size_t)y1 * sw * 3, dw, r1);
        if (even) h_expand_row(sw, src + (size_t)y2 * sw * 3, dw, r2);
        R* d = dst + (size_t)dy * dw * 3;
        for (size_t k = 0; k < (size_t)dw * 3; k++)
            d[k] = even ? (r0[k] + (R)6.0 * r1[k] + r2[k]) * ((R)1.0 / (R)8.0) : (r0[k] + r1[k]) * (R)0.5;
    }
    free(r0); free(r1); free(r2);
    }
}

/* expandRgbExactNoWrap src/fuse/pyramid.zig:1167-1354, restated with the same running
 * column sums (sc0a/sc0b current source row, sc1a/sc1b previous) and the same divisor
 * constants, including the ones that do not equal the in-bounds kernel weight when the
 * source level is one pixel wide or one pixel tall. */
typedef struct { R v[3]; } v3;
static inline v3 v3z(void) { v3 r = {{0, 0, 0}}; return r; }
static inline v3 v3ld(const R* s, size_t i) { v3 r = {{s[i * 3], s[i * 3 + 1], s[i * 3 + 2]}}; return r; }
static inline v3 v3add(v3 a, v3 b) { v3 r; for (int c = 0; c < 3; c++) r.v[c] = a.v[c] + b.v[c]; return r; }
static inline v3 v3mul(R k, v3 a) { v3 r; for (int c = 0; c < 3; c++) r.v[c] = k * a.v[c]; return r; }
static inline v3 v3div(v3 a, R k) { v3 r; for (int c = 0; c < 3; c++) r.v[c] = a.v[c] / k; return r; }
static inline void wr(R* dst, uint32_t dw, uint32_t x, uint32_t y, v3 v) {
    size_t b = ((size_t)y * dw + x) * 3;
    dst[b] = v.v[0]; dst[b + 1] = v.v[1]; dst[b + 2] = v.v[2];
}
static void expand_nowrap(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh, const R* src, R* dst) {
    memset(dst, 0, sizeof(R) * (size_t)dw * dh * 3);
    if (sw == 0 || sh == 0) return;
    size_t n = (size_t)sw + 1;
    v3* sc0a = calloc(n, sizeof(v3)); v3* sc0b = calloc(n, sizeof(v3));
    v3* sc1a = calloc(n, sizeof(v3)); v3* sc1b = calloc(n, sizeof(v3));
    int dwe = (dw & 1) == 0, dhe = (dh & 1) == 0;
    const R six = 6.0, four = 4.0;
    {
        v3 sr0 = v3ld(src, 0), sr1 = v3z();
        for (size_t x = 1; x < sw; x++) {
            v3 cur = v3ld(src, x);
            sc0a[x] = v3add(v3add(sr1, v3mul(six, sr0)), cur);
            sc0b[x] = v3mul(four, v3add(sr0, cur));
            sc1a[x] = v3z(); sc1b[x] = v3z();
            sr1 = sr0; sr0 = cur;
        }
        sc0a[sw] = v3add(sr1, v3mul(six, sr0));
        sc0b[sw] = v3mul(four, sr0);
        sc1a[sw] = v3z(); sc1b[sw] = v3z();
    }
    if (sh == 1) {
        if (sw == 1) {
            wr(dst, dw, 0, 0, v3div(v3add(sc1a[1], v3mul(six, sc0a[1])), (R)36));
            if (dwe) wr(dst, dw, 1, 0, v3div(v3add(sc1b[1], v3mul(six, sc0b[1])), (R)24));
            if (dhe) {
                wr(dst, dw, 0, 1, v3div(sc0a[1], (R)6));
                if (dwe) wr(dst, dw, 1, 1, v3div(sc0b[1], (R)4));
            }
            goto done;
        }
        uint32_t dx = 0;
        for (size_t x = 1; x < sw; x++) {
            v3 o00 = v3div(v3add(sc1a[x], v3mul(six, sc0a[x])), x == 1 ? (R)42 : (R)48);
            v3 o10 = v3div(v3add(sc1b[x], v3mul(six, sc0b[x])), (R)48);
            wr(dst, dw, dx, 0, o00); dx++;
            wr(dst, dw, dx, 0, o10); dx++;
            if (dhe) {
                wr(dst, dw, dx - 2, 1, v3div(sc0a[x], x == 1 ? (R)7 : (R)8));
                wr(dst, dw, dx - 1, 1, v3div(sc0b[x], (R)8));
            }
        }
        wr(dst, dw, dx, 0, v3div(v3add(sc1a[sw], v3mul(six, sc0a[sw])), (R)49));
        if (dwe) wr(dst, dw, dx + 1, 0, v3div(v3add(sc1b[sw], v3mul(six, sc0b[sw])), (R)28));
        if (dhe) {
            wr(dst, dw, dx, 1, v3div(sc0a[sw], (R)7));
            if (dwe) wr(dst, dw, dx + 1, 1, v3div(sc0b[sw], (R)4));
        }
        goto done;
    }
    for (size_t y = 1; y < sh; y++) {
        v3 sr0 = v3ld(src, y * sw), sr1 = v3z();
        uint32_t dx = 0;
        uint32_t top = (uint32_t)((y - 1) * 2), bot = top + 1;
        for (size_t x = 1; x < sw; x++) {
            v3 cur = v3ld(src, y * sw + x);
            v3 o00 = v3add(sc1a[x], v3mul(six, sc0a[x]));
            v3 o10 = v3add(sc1b[x], v3mul(six, sc0b[x]));
            v3 o01 = sc0a[x], o11 = sc0b[x];
            sc1a[x] = sc0a[x]; sc1b[x] = sc0b[x];
            sc0a[x] = v3add(v3add(sr1, v3mul(six, sr0)), cur);
            sc0b[x] = v3mul(four, v3add(sr0, cur));
            sr1 = sr0; sr0 = cur;
            o00 = v3add(o00, sc0a[x]); o10 = v3add(o10, sc0b[x]);
            o01 = v3add(o01, sc0a[x]); o11 = v3add(o11, sc0b[x]);
            int left = x == 1;
            R s00 = y == 1 ? (left ? 49 : 56) : (left ? 56 : 64);
            R s10 = y == 1 ? 56 : 64;
            R s01 = left ? 14 : 16;
            wr(dst, dw, dx, top, v3div(o00, s00)); dx++;
            wr(dst, dw, dx, top, v3div(o10, s10)); dx++;
            wr(dst, dw, dx - 2, bot, v3div(o01, s01));
            wr(dst, dw, dx - 1, bot, v3div(o11, (R)16));
        }
        v3 o00 = v3add(sc1a[sw], v3mul(six, sc0a[sw]));
        v3 o01 = sc0a[sw];
        v3 o10 = v3add(sc1b[sw], v3mul(six, sc0b[sw]));
        v3 o11 = sc0b[sw];
        sc1a[sw] = sc0a[sw]; sc1b[sw] = sc0b[sw];
        sc0a[sw] = v3add(sr1, v3mul(six, sr0));
        sc0b[sw] = v3mul(four, sr0);
        o00 = v3add(o00, sc0a[sw]); o01 = v3add(o01, sc0a[sw]);
        wr(dst, dw, dx, top, v3div(o00, y == 1 ? (R)49 : (R)56));
        wr(dst, dw, dx, bot, v3div(o01, (R)14));
        if (dwe) {
            o10 = v3add(o10, sc0b[sw]); o11 = v3add(o11, sc0b[sw]);
            wr(dst, dw, dx + 1, top, v3div(o10, y == 1 ? (R)28 : (R)32));
            wr(dst, dw, dx + 1, bot, v3div(o11, (R)8));
        }
    }
    {
        uint32_t top = (sh - 1) * 2, bot = top + 1;
        if (sw == 1) {
            wr(dst, dw, 0, top, v3div(v3add(sc1a[1], v3mul(six, sc0a[1])), (R)42));
            if (dwe) wr(dst, dw, 1, top, v3div(v3add(sc1b[1], v3mul(six, sc0b[1])), (R)28));
            if (dhe) {
                wr(dst, dw, 0, bot, v3div(sc0a[1], (R)6));
                if (dwe) wr(dst, dw, 1, bot, v3div(sc0b[1], (R)4));
            }
            goto done;
        }
        uint32_t dx = 0;
        for (size_t x = 1; x < sw; x++) {
            R s00 = x == 1 ? 49 : 56, s01 = x == 1 ? 7 : 8;
            wr(dst, dw, dx, top, v3div(v3add(sc1a[x], v3mul(six, sc0a[x])), s00)); dx++;
            wr(dst, dw, dx, top, v3div(v3add(sc1b[x], v3mul(six, sc0b[x])), (R)56)); dx++;
            if (dhe) {
                wr(dst, dw, dx - 2, bot, v3div(sc0a[x], s01));
                wr(dst, dw, dx - 1, bot, v3div(sc0b[x], (R)8));
            }
        }
        wr(dst, dw, dx, top, v3div(v3add(sc1a[sw], v3mul(six, sc0a[sw])), (R)49));
        if (dwe) wr(dst, dw, dx + 1, top, v3div(v3add(sc1b[sw], v3mul(six, sc0b[sw])), (R)28));
        if (dhe) {
            wr(dst, dw, dx, bot, v3div(sc0a[sw], (R)7));
            if (dwe) wr(dst, dw, dx + 1, bot, v3div(sc0b[sw], (R)4));
        }
    }
done:
    free(sc0a); free(sc0b); free(sc1a); free(sc1b);
}

/* expandRgb src/fuse/pyramid.zig:849-909 (regime switch at exact_expand_max_dim = 512, :8) */
void or_expand_rgb(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh, const R* src, R* dst) {
    if ((dw > dh ? dw : dh) <= 512) expand_nowrap(dw, dh, sw, sh, src, dst);
    else expand_clamp(dw, dh, sw, sh, src, dst);
}

/* fillRgbBaseRows src/fuse/pyramid.zig:606-639 */
void or_fill_rgb_base(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, R* dst) {
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) dst[i * 3 + c] = (R)sample_at(img, sb, i * st + c);
}

/* collapseToImageWithJobsAndDebug src/fuse/pyramid.zig:223-269 (levels given concatenated) */
static void level_dims(uint32_t w, uint32_t h, uint32_t levels, uint32_t* lw, uint32_t* lh, size_t* off) {
    size_t o = 0;
    for (uint32_t l = 0; l < levels; l++) {
        lw[l] = w; lh[l] = h; off[l] = o;
        o += (size_t)w * h;
        w = or_next_level_size(w); h = or_next_level_size(h);
    }
    off[levels] = o;
}
static void quantize(const R* base, size_t count, int sb, void* out) {
    PAR_FOR(count)
    for (size_t i = 0; i < count; i++) {
        R v = base[i] + (R)0.5;
        R hi = sb == 1 ? (R)255.0 : (R)65535.0;
        v = v < (R)0 ? (R)0 : (v > hi ? hi : v);
        if (sb == 1) ((uint8_t*)out)[i] = (uint8_t)v;
        else ((uint16_t*)out)[i] = (uint16_t)v;
    }
}
void or_collapse(uint32_t w, uint32_t h, uint32_t levels, const R* acc, int sb, void* out) {
    uint32_t lw[40], lh[40];
    size_t off[41];
    level_dims(w, h, levels, lw, lh, off);
    R* c = malloc(sizeof(R) * off[levels] * 3);
    memcpy(c, acc, sizeof(R) * off[levels] * 3);
    R* tmp = malloc(sizeof(R) * (size_t)w * h * 3);
    for (uint32_t l = levels - 1; l >= 1; l--) {
        R* parent = c + off[l - 1] * 3;
        or_expand_rgb(lw[l - 1], lh[l - 1], lw[l], lh[l], c + off[l] * 3, tmp);
        size_t pn = (size_t)lw[l - 1] * lh[l - 1] * 3;
        PAR_FOR(pn)
        for (size_t k = 0; k < pn; k++) parent[k] += tmp[k];
    }
    quantize(c, (size_t)w * h * 3, sb, out);
    free(c); free(tmp);
}

/* buildImageLaplacianPyramid src/fuse/pyramid.zig:418-487 + mergeHybridImages :271-301 */
void or_merge_hybrid(const void* soft, const void* hard, int sb, uint32_t w, uint32_t h, uint32_t start,
                     uint32_t count, float mix, uint32_t max_levels, void* out) {
    uint32_t levels = or_level_count(w, h);
    if (max_levels && max_levels < levels) levels = max_levels;
    uint32_t lw[40], lh[40];
    size_t off[41];
    level_dims(w, h, levels, lw, lh, off);
    R* lap[2];
    const void* imgs[2] = {soft, hard};
    R* g = malloc(sizeof(R) * off[levels] * 3);
    R* tmp = malloc(sizeof(R) * (size_t)w * h * 3);
    for (int k = 0; k < 2; k++) {
        lap[k] = malloc(sizeof(R) * off[levels] * 3);
        or_fill_rgb_base(imgs[k], sb, w, h, 3, 0, g);
        for (uint32_t l = 1; l < levels; l++)
            or_reduce_rgb_clamped(lw[l - 1], lh[l - 1], g + off[l - 1] * 3, lw[l], lh[l], g + off[l] * 3);
        for (uint32_t l = 0; l + 1 < levels; l++) {
            or_expand_rgb(lw[l], lh[l], lw[l + 1], lh[l + 1], g + off[l + 1] * 3, tmp);
            size_t ln = (size_t)lw[l] * lh[l] * 3;
            R* lk = lap[k];
            PAR_FOR(ln)
            for (size_t i = 0; i < ln; i++) lk[off[l] * 3 + i] = g[off[l] * 3 + i] - tmp[i];
        }
        memcpy(lap[k] + off[levels - 1] * 3, g + off[levels - 1] * 3, sizeof(R) * (size_t)lw[levels - 1] * lh[levels - 1] * 3);
    }
    uint32_t hs = start < levels ? start : levels;
    uint32_t he = hs + count < levels ? hs + count : levels;
    R m = (R)mix;
    for (uint32_t l = 0; l < levels; l++) {
        R* l0 = lap[0];
        const R* l1 = lap[1];
        PAR_FOR((off[l + 1] - off[l]) * 3)
        for (size_t i = off[l] * 3; i < off[l + 1] * 3; i++)
            l0[i] = (l >= hs && l < he) ? l0[i] + (l1[i] - l0[i]) * m : l0[i];
    }
    or_collapse(w, h, levels, lap[0], sb, out);
    free(lap[0]); free(lap[1]); free(g); free(tmp);
}

/* updateRange src/fuse/blend.zig:237-288 with effectiveWeightU8/U16 :290-314.
 * raw_weights are the contrast weights before applySupport. */
void or_update_winners(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, const R* raw,
                       R* best, void* out) {
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++) {
        R wt = raw[i];
        if (extra) wt = wt * ((R)sample_at(img, sb, i * st + color) / (sb == 1 ? (R)255.0 : (R)65535.0));
        if (wt <= best[i]) continue;
        best[i] = wt;
        for (int c = 0; c < 3; c++) {
            uint32_t v = sample_at(img, sb, i * st + c);
            if (sb == 1) ((uint8_t*)out)[i * 3 + c] = (uint8_t)v;
            else ((uint16_t*)out)[i * 3 + c] = (uint16_t)v;
        }
    }
}

/* blurFiveTapInto src/fuse/masks.zig:86-105 with horizontalPassRange :151-177 and
 * verticalPassRange :179-207: clamped (1,4,6,4,1)/16, summed left to right, rows then columns */
void or_blur_five_tap(uint32_t w, uint32_t h, const R* in, R* scratch, R* out) {
    PAR_FOR((size_t)w * h)
    for (size_t y = 0; y < h; y++) {
        const R* r = in + y * w;
        for (size_t x = 0; x < w; x++) {
            size_t xm2 = x >= 2 ? x - 2 : 0, xm1 = x >= 1 ? x - 1 : 0;
            size_t xp1 = x + 1 < w ? x + 1 : w - 1, xp2 = x + 2 < w ? x + 2 : w - 1;
            scratch[y * w + x] = ((((r[xm2] + (R)4.0 * r[xm1]) + (R)6.0 * r[x]) + (R)4.0 * r[xp1]) + r[xp2]) / (R)16.0;
        }
    }
    PAR_FOR((size_t)w * h)
    for (size_t y = 0; y < h; y++) {
        size_t ym2 = y >= 2 ? y - 2 : 0, ym1 = y >= 1 ? y - 1 : 0;
        size_t yp1 = y + 1 < h ? y + 1 : h - 1, yp2 = y + 2 < h ? y + 2 : h - 1;
        for (size_t x = 0; x < w; x++)
            out[y * w + x] = ((((scratch[ym2 * w + x] + (R)4.0 * scratch[ym1 * w + x]) + (R)6.0 * scratch[y * w + x]) +
                               (R)4.0 * scratch[yp1 * w + x]) +
                              scratch[yp2 * w + x]) /
                             (R)16.0;
    }
}

/* accumulateRange src/fuse/blend.zig:97-148: pixels with weight > 0 add weight and weight*rgb */
void or_soft_accumulate(const void* img, int sb, uint32_t w, uint32_t h, int color, int extra, const R* wts,
                        R* ws, R* ps) {
    size_t n = (size_t)w * h, st = (size_t)(color + extra);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++) {
        R wt = wts[i];
        if (wt <= 0) continue;
        ws[i] += wt;
        for (int c = 0; c < 3; c++) ps[i * 3 + c] += wt * (R)sample_at(img, sb, i * st + c);
    }
}

/* finalizeSoft src/fuse/blend.zig:150-186 */
void or_soft_finalize(const R* ws, const R* ps, uint32_t w, uint32_t h, int sb, void* out) {
    size_t n = (size_t)w * h;
    R hi = sb == 1 ? (R)255.0 : (R)65535.0;
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++) {
        R wt = ws[i];
        for (int c = 0; c < 3; c++) {
            uint32_t q = 0;
            if (wt > (R)1e-12) {
                R inv = (R)1.0 / wt;
                R v = ps[i * 3 + c] * inv + (R)0.5;
                v = v < (R)0 ? (R)0 : (v > hi ? hi : v);
                q = (uint32_t)v;
            }
            if (sb == 1) ((uint8_t*)out)[i * 3 + c] = (uint8_t)q;
            else ((uint16_t*)out)[i * 3 + c] = (uint16_t)q;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* stack drivers — src/stack/pipeline.zig:361-605 and src/fuse/pipeline.zig:254-445   */
/* ------------------------------------------------------------------------- */

/* computeWeightMapForRemapped src/stack/pipeline.zig:653-671 */
static void weight_map(const void* img, int sb, uint32_t w, uint32_t h, int extra, uint32_t window, R* gray, R* sup,
                       R* out) {
    or_gray(img, sb, w, h, 3, extra, gray);
    or_binary_support(img, sb, w, h, 3, extra, sup);
    or_contrast(gray, sup, w, h, window, out);
}

/* accumulateImageWithWorkspace src/fuse/pyramid.zig:190-204 (mask pyramid :384-416,
 * image pyramid + accumulation :489-542, accumulateLap/Gaussian :1506-1554) */
static void accumulate_frame(const void* img, int sb, uint32_t w, uint32_t h, int extra, const R* mask0,
                             const R* union_sup, uint32_t levels, R* acc, R* mlev, R* msup, R* glev, R* gsup,
                             R* tmp) {
    uint32_t lw[40], lh[40];
    size_t off[41];
    level_dims(w, h, levels, lw, lh, off);
    memcpy(mlev, mask0, sizeof(R) * (size_t)w * h);
    memcpy(msup, union_sup, sizeof(R) * (size_t)w * h);
    for (uint32_t l = 1; l < levels; l++)
        or_reduce_scalar_support(lw[l - 1], lh[l - 1], mlev + off[l - 1], msup + off[l - 1], lw[l], lh[l],
                                 mlev + off[l], msup + off[l]);
    or_fill_rgb_base(img, sb, w, h, 3, extra, glev);
    or_binary_support(img, sb, w, h, 3, extra, gsup);
    for (uint32_t l = 1; l < levels; l++)
        or_reduce_rgb_support(lw[l - 1], lh[l - 1], glev + off[l - 1] * 3, gsup + off[l - 1], lw[l], lh[l],
                              glev + off[l] * 3, gsup + off[l]);
    for (uint32_t l = 0; l + 1 < levels; l++) {
        or_expand_rgb(lw[l], lh[l], lw[l + 1], lh[l + 1], glev + off[l + 1] * 3, tmp);
        const R* cur = glev + off[l] * 3;
        const R* m = mlev + off[l];
        R* d = acc + off[l] * 3;
        PAR_FOR((size_t)lw[l] * lh[l])
        for (size_t p = 0; p < (size_t)lw[l] * lh[l]; p++) {
            R wt = m[p];
            R l0 = cur[p * 3] - tmp[p * 3], l1 = cur[p * 3 + 1] - tmp[p * 3 + 1], l2 = cur[p * 3 + 2] - tmp[p * 3 + 2];
            d[p * 3] += l0 * wt; d[p * 3 + 1] += l1 * wt; d[p * 3 + 2] += l2 * wt;
        }
    }
    {
        uint32_t l = levels - 1;
        const R* g = glev + off[l] * 3;
        const R* m = mlev + off[l];
        R* d = acc + off[l] * 3;
        for (size_t p = 0; p < (size_t)lw[l] * lh[l]; p++)
            for (int c = 0; c < 3; c++) d[p * 3 + c] += g[p * 3 + c] * m[p];
    }
}

/* runRemappedSinglePass src/stack/pipeline.zig:245-312 / runSinglePass src/fuse/pipeline.zig:189-252:
 * hardmask = winners of the contrast weight times alpha; softmask = support-scaled weights,
 * blurred, blended, finalised (:135-158) */
static int single_pass(int n, const void* const* frames, uint32_t w, uint32_t h, int sb, int extra,
                       const or_fuse_opts* o, void* out, or_fuse_debug* dbg) {
    size_t count = (size_t)w * h;
    R* gray = malloc(sizeof(R) * count);
    R* sup = malloc(sizeof(R) * count);
    R* wk = malloc(sizeof(R) * count);
    R* scratch = malloc(sizeof(R) * count);
    if (o->method == OR_METHOD_HARDMASK) {
        R* best = malloc(sizeof(R) * count);
        for (size_t i = 0; i < count; i++) best[i] = -INFINITY;
        for (int k = 0; k < n; k++) {
            weight_map(frames[k], sb, w, h, extra, o->window, gray, sup, wk);
            or_update_winners(frames[k], sb, w, h, 3, extra, wk, best, out);
        }
        free(best);
    } else {
        R* ws = calloc(count, sizeof(R));
        R* ps = calloc(count * 3, sizeof(R));
        for (int k = 0; k < n; k++) {
            weight_map(frames[k], sb, w, h, extra, o->window, gray, sup, wk);
            or_apply_support(frames[k], sb, w, h, 3, extra, wk);
            or_blur_five_tap(w, h, wk, scratch, wk);
            if (dbg && dbg->weights) memcpy(dbg->weights + count * k, wk, sizeof(R) * count);
            or_soft_accumulate(frames[k], sb, w, h, 3, extra, wk, ws, ps);
        }
        if (dbg && dbg->norm_sum) memcpy(dbg->norm_sum, ws, sizeof(R) * count);
        or_soft_finalize(ws, ps, w, h, sb, out);
        free(ws); free(ps);
    }
    free(gray); free(sup); free(wk); free(scratch);
    return 0;
}

int or_fuse_frames(int n, const void* const* frames, uint32_t w, uint32_t h, int sb, int extra, const or_fuse_opts* o,
                   void* out, or_fuse_debug* dbg) {
    if (n < 1 || w == 0 || h == 0) return -1;
    if (o->method == OR_METHOD_HARDMASK || o->method == OR_METHOD_SOFTMASK)
        return single_pass(n, frames, w, h, sb, extra, o, out, dbg);
    size_t count = (size_t)w * h;
    uint32_t levels = or_level_count(w, h);
    if (o->max_levels && o->max_levels < levels) levels = o->max_levels;
    uint32_t lw[40], lh[40];
    size_t off[41];
    level_dims(w, h, levels, lw, lh, off);
    R* gray = malloc(sizeof(R) * count);
    R* sup = malloc(sizeof(R) * count);
    R* S = calloc(count, sizeof(R));
    R* U = calloc(count, sizeof(R));
    R* wts = malloc(sizeof(R) * count * n);
    /* pass 1: src/stack/pipeline.zig:412-477 */
    for (int k = 0; k < n; k++) {
        R* wk = wts + count * k;
        weight_map(frames[k], sb, w, h, extra, o->window, gray, sup, wk);
        or_apply_support(frames[k], sb, w, h, 3, extra, wk);
        or_union_support(frames[k], sb, w, h, 3, extra, U);
        PAR_FOR(count)
        for (size_t i = 0; i < count; i++) S[i] += wk[i];
    }
    /* pass 2: src/stack/pipeline.zig:526-597 */
    R* acc = calloc(off[levels] * 3, sizeof(R));
    R* mlev = malloc(sizeof(R) * off[levels]);
    R* msup = malloc(sizeof(R) * off[levels]);
    R* glev = malloc(sizeof(R) * off[levels] * 3);
    R* gsup = malloc(sizeof(R) * off[levels]);
    R* tmp = malloc(sizeof(R) * count * 3);
    for (int k = 0; k < n; k++) {
        or_normalize(wts + count * k, S, (uint32_t)count, (uint32_t)n, gray);
        accumulate_frame(frames[k], sb, w, h, extra, gray, U, levels, acc, mlev, msup, glev, gsup, tmp);
        if (dbg && dbg->frame_index == k) {
            if (dbg->mask_levels) memcpy(dbg->mask_levels, mlev, sizeof(R) * off[levels]);
            if (dbg->gauss_levels) memcpy(dbg->gauss_levels, glev, sizeof(R) * off[levels] * 3);
        }
    }
    if (dbg) {
        if (dbg->norm_sum) memcpy(dbg->norm_sum, S, sizeof(R) * count);
        if (dbg->union_support) memcpy(dbg->union_support, U, sizeof(R) * count);
        if (dbg->weights) memcpy(dbg->weights, wts, sizeof(R) * count * n);
        if (dbg->acc_levels) memcpy(dbg->acc_levels, acc, sizeof(R) * off[levels] * 3);
    }
    size_t ob = count * 3 * (size_t)sb;
    void* soft = malloc(ob);
    or_collapse(w, h, levels, acc, sb, soft);
    if (o->method == OR_METHOD_HYBRID) {
        /* src/stack/pipeline.zig:177-236: soft = plain pyramid result, hard = winners, merge */
        void* hard = malloc(ob);
        R* best = malloc(sizeof(R) * count);
        for (size_t i = 0; i < count; i++) best[i] = -INFINITY;
        for (int k = 0; k < n; k++) {
            weight_map(frames[k], sb, w, h, extra, o->window, gray, sup, tmp);
            or_update_winners(frames[k], sb, w, h, 3, extra, tmp, best, hard);
        }
        if (dbg && dbg->soft) memcpy(dbg->soft, soft, ob);
        if (dbg && dbg->hard) memcpy(dbg->hard, hard, ob);
        or_merge_hybrid(soft, hard, sb, w, h, 2, 4, o->sharpness, o->max_levels, out);
        free(hard); free(best);
    } else {
        memcpy(out, soft, ob);
    }
    free(soft);
    free(gray); free(sup); free(S); free(U); free(wts);
    free(acc); free(mlev); free(msup); free(glev); free(gsup); free(tmp);
    return 0;
}

int or_stack(int n, const void* const* sources, uint32_t sw, uint32_t sh, int sb, const or_pose* poses,
             const int32_t* roi, const or_fuse_opts* o, void* out, or_fuse_debug* dbg) {
    int32_t r[4] = {0, 0, (int32_t)sw, (int32_t)sh};
    if (n < 1) return -1;
    if (roi) memcpy(r, roi, sizeof r);
    uint32_t w = (uint32_t)(r[2] - r[0]), h = (uint32_t)(r[3] - r[1]);
    void** rem = calloc((size_t)n, sizeof(void*));
    for (int k = 0; k < n; k++) {
        rem[k] = malloc((size_t)w * h * 4 * sb);
        if (sb == 1) or_remap_u8(sources[k], sw, sh, 3, &poses[k], r, rem[k]);
        else or_remap_u16(sources[k], sw, sh, 3, &poses[k], r, rem[k]);
    }
    int rc = or_fuse_frames(n, (const void* const*)rem, w, h, sb, 1, o, out, dbg);
    for (int k = 0; k < n; k++) free(rem[k]);
    free(rem);
    return rc;
}
