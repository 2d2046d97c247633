/*
 * CPU oracle: PatchMatch multi-view stereo.  TEST INFRASTRUCTURE ONLY (see lr3d_oracle.h).
 *
 * Restates, line by line, the WGSL module `PATCHMATCH_WGSL`
 * (frontend/src/app/vision/mvs/patchmatch.wgsl.ts, cited as PM:<line>), the packing helpers
 * (frontend/src/app/vision/mvs/depth-packing.wgsl.ts, PK:<line>) and the host schedule of
 * DenseStereo (frontend/src/app/vision/mvs/dense-stereo.ts, DS:<line>), under the arithmetic
 * definition stated in lr3d_oracle.h.  Built with -ffp-contract=off.
 */
#include "lr3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* PM:38-45 */
#define SAMPLES 36
#define INVALID 2.0f
#define GEOMETRIC_CAP 3.0f
#define FILTER_COST 7u
#define FILTER_PIXELS 1.0f
#define FILTER_DEPTH 0.01f
/* PM:39-40 and PM:154 rescaled to u8 units: 1 / (2 (0.06*255)^2), 1 / (2 * 5^2), 1e-5 * 255^2 */
#define COLOR_K ((float)(1.0 / (2.0 * (0.06 * 255.0) * (0.06 * 255.0))))
#define SPACE_K 0.02f
#define DEGENERATE_DEN ((float)(1e-5 * 255.0 * 255.0))

static int g_threads = 0;

void orc_set_threads(int32_t n) { g_threads = n; }
int32_t orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* RNG                                                                                        */
/* ------------------------------------------------------------------------------------------ */

/* PM:56-60: PCG-RXS-M-XS hash (Jarzynski & Olano, JCGT 2020). */
uint32_t orc_pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

/* DS:332-334: Math.imul(ref + 1, 0x9e3779b1) + iteration*97 + phase*13 + flags*7, then >>> 0. */
uint32_t orc_host_seed(int32_t ref, uint32_t iteration, uint32_t phase, uint32_t flags) {
    return (uint32_t)(ref + 1) * 0x9e3779b1u + iteration * 97u + phase * 13u + flags * 7u;
}

/* PM:251-253 */
uint32_t orc_pixel_rng(uint32_t seed, uint32_t iteration, uint32_t phase, uint32_t index) {
    return orc_pcg(index ^ orc_pcg(seed + iteration * 7919u + phase * 104729u));
}

/* PM:62-65 */
float orc_random(uint32_t *rng) {
    *rng = orc_pcg(*rng);
    return (float)(*rng >> 8u) / 16777216.0f;
}

/* ------------------------------------------------------------------------------------------ */
/* Small vector helpers (the arithmetic definition of lr3d_oracle.h)                         */
/* ------------------------------------------------------------------------------------------ */

static inline float dot3f(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline void normalize3f(const float v[3], float out[3]) {
    float inv = 1.0f / sqrtf(dot3f(v, v));
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
}

static inline float expf_d(float x) { return (float)exp((double)x); }
static inline float cosf_d(float x) { return (float)cos((double)x); }
static inline float sinf_d(float x) { return (float)sin((double)x); }
static inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

/* WGSL round(): nearest, ties to even (the default rounding mode of rintf). */
static inline int32_t round_even(float v) { return (int32_t)rintf(v); }

/* x_cam = R x + t, each row fmaf(R2, x2, fmaf(R1, x1, fmaf(R0, x0, t))) */
static inline void rigid3f(const float R[9], const float x[3], const float t[3], float out[3]) {
    for (int i = 0; i < 3; i++)
        out[i] = fmaf(R[i * 3 + 2], x[2], fmaf(R[i * 3 + 1], x[1], fmaf(R[i * 3], x[0], t[i])));
}

/* R^T e, each column fmaf(R2j, e2, fmaf(R1j, e1, R0j * e0)) */
static inline void rotate_t3f(const float R[9], const float e[3], float out[3]) {
    for (int j = 0; j < 3; j++)
        out[j] = fmaf(R[6 + j], e[2], fmaf(R[3 + j], e[1], R[j] * e[0]));
}

/* PM:80-83 */
static inline void ray_of(const float k[4], float px, float py, float out[3]) {
    out[0] = (px - k[2]) / k[0];
    out[1] = (py - k[3]) / k[1];
    out[2] = 1.0f;
}

/* ------------------------------------------------------------------------------------------ */
/* f16 and packing                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* IEEE binary32 -> binary16, round to nearest even (subnormals, overflow to inf, NaN kept). */
uint16_t orc_f32_to_f16(float v) {
    uint32_t x;
    memcpy(&x, &v, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u | (man >> 13) : 0u));
    int32_t e = (int32_t)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        uint32_t m = man | 0x800000u; /* implicit bit */
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t mid = 1u << (shift - 1u);
        if (rem > mid || (rem == mid && (half & 1u))) half++;
        return (uint16_t)(sign | half);
    }
    uint32_t h = ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++; /* may carry into the exponent */
    return (uint16_t)(sign | h);
}

float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 31u;
    uint32_t man = (uint32_t)h & 1023u;
    uint32_t x;
    if (exp == 0) {
        float f = ldexpf((float)man, -24);
        return sign ? -f : f;
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* PK:3-5 */
float orc_unpack_depth(uint32_t packed) { return orc_f16_to_f32((uint16_t)(packed & 0xffffu)); }

/* PK:7-15 */
void orc_unpack_normal(uint32_t packed, float out[3]) {
    float ex = ((float)((packed >> 16u) & 63u) / 63.0f) * 2.0f - 1.0f;
    float ey = ((float)((packed >> 22u) & 63u) / 63.0f) * 2.0f - 1.0f;
    float n[3] = {ex, ey, (1.0f - fabsf(ex)) - fabsf(ey)};
    if (n[2] < 0.0f) {
        float sx = n[0] >= 0.0f ? 1.0f : -1.0f;
        float sy = n[1] >= 0.0f ? 1.0f : -1.0f;
        float nx = (1.0f - fabsf(n[1])) * sx;
        float ny = (1.0f - fabsf(n[0])) * sy;
        n[0] = nx;
        n[1] = ny;
    }
    normalize3f(n, out);
}

/* PK:17-19 */
uint32_t orc_unpack_cost(uint32_t packed) { return packed >> 28u; }

/* A whole packed map: depth (f16 -> double, 0 where the sample is 0 or not positive), the normal
 * (orc_unpack_normal, widened to double) and the cost code, for the FP64 TSDF adjoint. */
void orc_decode_map(const uint32_t *packed, int64_t n, double *depth, double *normal, uint8_t *cost) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const uint32_t m = packed[i];
        depth[i] = 0.0;
        normal[i * 3] = normal[i * 3 + 1] = normal[i * 3 + 2] = 0.0;
        cost[i] = (uint8_t)(m >> 28u);
        if (m == 0u) continue;
        const float d = orc_unpack_depth(m);
        depth[i] = d > 0.0f ? (double)d : 0.0;
        float nn[3];
        orc_unpack_normal(m, nn);
        for (int k = 0; k < 3; k++) normal[i * 3 + k] = (double)nn[k];
    }
}

/* PK:21-33 */
uint32_t orc_pack_sample(float depth, const float normal[3], float cost) {
    if (!(depth > 0.0f) || cost >= 1.0f) return 0u;
    float l1 = (fabsf(normal[0]) + fabsf(normal[1])) + fabsf(normal[2]);
    float ex = normal[0] / l1;
    float ey = normal[1] / l1;
    if (normal[2] < 0.0f) {
        float sx = ex >= 0.0f ? 1.0f : -1.0f;
        float sy = ey >= 0.0f ? 1.0f : -1.0f;
        float nx = (1.0f - fabsf(ey)) * sx;
        float ny = (1.0f - fabsf(ex)) * sy;
        ex = nx;
        ey = ny;
    }
    uint32_t qx = (uint32_t)(clampf(ex * 0.5f + 0.5f, 0.0f, 1.0f) * 63.0f + 0.5f);
    uint32_t qy = (uint32_t)(clampf(ey * 0.5f + 0.5f, 0.0f, 1.0f) * 63.0f + 0.5f);
    uint32_t c = (uint32_t)(cost * 15.0f);
    if (c > 15u) c = 15u;
    return (uint32_t)orc_f32_to_f16(depth) | (qx << 16u) | (qy << 22u) | (c << 28u);
}

/* ------------------------------------------------------------------------------------------ */
/* Host parameters                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* CAM:73-77 relativePose(a, b) with CAM's multiply3 / apply3 (math/mat3.ts:9-29), in double */
static void relative_pose(const orc_view *a, const orc_view *b, double R[9], double t[3]) {
    double aT[9] = {a->R[0], a->R[3], a->R[6], a->R[1], a->R[4], a->R[7], a->R[2], a->R[5], a->R[8]};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            R[r * 3 + c] = b->R[r * 3] * aT[c] + b->R[r * 3 + 1] * aT[3 + c] + b->R[r * 3 + 2] * aT[6 + c];
    for (int r = 0; r < 3; r++) {
        double rt = R[r * 3] * a->t[0] + R[r * 3 + 1] * a->t[1] + R[r * 3 + 2] * a->t[2];
        t[r] = b->t[r] - rt;
    }
}

/* DS:314-355 */
void orc_make_params(const orc_view *views, int32_t n_views, int32_t ref, const int32_t *sources,
                     int32_t n_src, double min_depth, double max_depth, uint32_t iteration,
                     uint32_t phase, double step, uint32_t flags, uint32_t needed, int32_t topk,
                     orc_params *out) {
    (void)n_views;
    memset(out, 0, sizeof(*out));
    const orc_view *rv = &views[ref];
    out->width = rv->width;
    out->height = rv->height;
    out->ref = ref;
    out->k[0] = (float)rv->fx;
    out->k[1] = (float)rv->fy;
    out->k[2] = (float)rv->cx;
    out->k[3] = (float)rv->cy;
    out->inv_min = (float)(1.0 / min_depth);
    out->inv_max = (float)(1.0 / max_depth);
    out->step = (float)step;
    out->geo_weight = 0.2f; /* DS:46 GEOMETRIC_WEIGHT */
    out->iteration = iteration;
    out->phase = phase;
    out->seed = orc_host_seed(ref, iteration, phase, flags);
    out->flags = flags;
    out->needed = needed;
    if (n_src > ORC_MAX_SOURCES) n_src = ORC_MAX_SOURCES; /* DS:186 */
    out->n_src = n_src;
    out->topk = topk;
    out->ref_off = orc_map_offset(views, ref);
    for (int s = 0; s < n_src; s++) {
        const orc_view *sv = &views[sources[s]];
        double R[9], t[3];
        relative_pose(rv, sv, R, t);
        orc_src *o = &out->src[s];
        o->k[0] = (float)sv->fx;
        o->k[1] = (float)sv->fy;
        o->k[2] = (float)sv->cx;
        o->k[3] = (float)sv->cy;
        for (int i = 0; i < 9; i++) o->R[i] = (float)R[i];
        for (int i = 0; i < 3; i++) o->t[i] = (float)t[i];
        o->layer = sources[s];
        o->width = sv->width;
        o->height = sv->height;
        o->map_off = orc_map_offset(views, sources[s]);
        /* H = Ks (R + t n^T / rho) Kr^-1 = G + (e / rho) (Kr^-T n)^T: G = (Ks R) Kr^-1, e = Ks t */
        double A[9];
        for (int j = 0; j < 3; j++) {
            A[j] = sv->fx * R[j] + sv->cx * R[6 + j];
            A[3 + j] = sv->fy * R[3 + j] + sv->cy * R[6 + j];
            A[6 + j] = R[6 + j];
        }
        for (int i = 0; i < 3; i++) {
            o->G[i * 3 + 0] = (float)(A[i * 3 + 0] / rv->fx);
            o->G[i * 3 + 1] = (float)(A[i * 3 + 1] / rv->fy);
            o->G[i * 3 + 2] = (float)((A[i * 3 + 2] - A[i * 3 + 0] * rv->cx / rv->fx) - A[i * 3 + 1] * rv->cy / rv->fy);
        }
        o->e[0] = (float)(sv->fx * t[0] + sv->cx * t[2]);
        o->e[1] = (float)(sv->fy * t[1] + sv->cy * t[2]);
        o->e[2] = (float)t[2];
    }
}

int64_t orc_map_offset(const orc_view *views, int32_t layer) {
    int64_t off = 0;
    for (int32_t l = 0; l < layer; l++) off += (int64_t)views[l].width * (int64_t)views[l].height;
    return off;
}

/* ------------------------------------------------------------------------------------------ */
/* Cost                                                                                       */
/* ------------------------------------------------------------------------------------------ */

static inline int offset_x(int i) { return (i % 6) * 2 - 5; } /* PM:76-78 */
static inline int offset_y(int i) { return (i / 6) * 2 - 5; }

static inline float pixel_u8(const orc_images *img, int32_t layer, int32_t x, int32_t y) {
    int32_t w = img->widths[layer], h = img->heights[layer];
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (x > w - 1) x = w - 1;
    if (y > h - 1) y = h - 1;
    return (float)img->pixels[layer][(size_t)y * (size_t)w + (size_t)x];
}

/* PM:85-110 (ref_pixel + prepare_window) */
void orc_window(const orc_params *p, const orc_images *img, int32_t x, int32_t y, float *values,
                float *weights, float *stats) {
    float center = pixel_u8(img, p->ref, x, y);
    float sw = 0.0f, sr = 0.0f, srr = 0.0f;
    for (int i = 0; i < SAMPLES; i++) {
        int ox = offset_x(i), oy = offset_y(i);
        float v = pixel_u8(img, p->ref, x + ox, y + oy);
        float dc = v - center;
        float ds = (float)(ox * ox + oy * oy);
        float w = expf_d(-(dc * dc) * COLOR_K - ds * SPACE_K);
        values[i] = v;
        weights[i] = w;
        float wv = w * v;
        sw = sw + w;
        sr = sr + wv;
        srr = srr + wv * v;
    }
    float mean = sr / sw;
    stats[0] = sw;
    stats[1] = mean;
    stats[2] = fmaxf(srr / sw - mean * mean, 0.0f);
}

/* PM:112-128.  H = Ks (R + (t/rho) n^T) Kr^-1 = G + (e / rho) m^T with m = Kr^-T n =
 * (nx / fx, ny / fy, nz - nx cx / fx - ny cy / fy), evaluated in the fixed order below. */
void orc_homography(const orc_params *p, int32_t s, const float n[3], float rho, float H[9]) {
    const orc_src *v = &p->src[s];
    float ikx = 1.0f / p->k[0], iky = 1.0f / p->k[1];
    float kx2 = -p->k[2] / p->k[0], ky2 = -p->k[3] / p->k[1];
    float irho = 1.0f / rho;
    float m[3] = {n[0] * ikx, n[1] * iky, (n[0] * kx2 + n[1] * ky2) + n[2]};
    for (int i = 0; i < 3; i++) {
        float f = v->e[i] * irho;
        for (int j = 0; j < 3; j++) H[i * 3 + j] = fmaf(f, m[j], v->G[i * 3 + j]);
    }
}

/* Bilinear sample with clamp-to-edge at texel coordinates (u, v) in [0, w-1] x [0, h-1]
 * (the reference samples (uv + 0.5) / size with a linear, clamp-to-edge sampler, DS:141). */
static inline float bilinear(const uint8_t *im, int32_t w, int32_t h, float u, float v) {
    float x0 = floorf(u), y0 = floorf(v);
    float fx = u - x0, fy = v - y0;
    int32_t ix = (int32_t)x0, iy = (int32_t)y0;
    int32_t ix1 = ix + 1 < w ? ix + 1 : w - 1;
    int32_t iy1 = iy + 1 < h ? iy + 1 : h - 1;
    float a = (float)im[(size_t)iy * w + ix], b = (float)im[(size_t)iy * w + ix1];
    float c = (float)im[(size_t)iy1 * w + ix], d = (float)im[(size_t)iy1 * w + ix1];
    float top = fmaf(fx, b - a, a);
    float bot = fmaf(fx, d - c, c);
    return fmaf(fy, bot - top, top);
}

/* PM:130-159 */
float orc_view_cost(const orc_params *p, const orc_images *img, int32_t s, const float H[9],
                    int32_t x, int32_t y, const float *values, const float *weights,
                    const float *stats) {
    const orc_src *sv = &p->src[s];
    const uint8_t *im = img->pixels[sv->layer];
    int32_t w = sv->width, h = sv->height;
    float wmax = (float)w - 1.0f, hmax = (float)h - 1.0f;
    float ss = 0.0f, sss = 0.0f, srs = 0.0f;
    for (int i = 0; i < SAMPLES; i++) {
        float X = (float)(x + offset_x(i)), Y = (float)(y + offset_y(i));
        float qx = fmaf(H[0], X, fmaf(H[1], Y, H[2]));
        float qy = fmaf(H[3], X, fmaf(H[4], Y, H[5]));
        float qz = fmaf(H[6], X, fmaf(H[7], Y, H[8]));
        if (!(qz > 1e-6f)) return INVALID;
        float rz = 1.0f / qz;
        float u = qx * rz, v = qy * rz;
        if (!(u >= 0.0f && v >= 0.0f && u <= wmax && v <= hmax)) return INVALID;
        float val = bilinear(im, w, h, u, v);
        float wv = weights[i] * val;
        ss = ss + wv;
        sss = fmaf(wv, val, sss);
        srs = fmaf(wv, values[i], srs);
    }
    /* PM:151-158 with the sums scaled by 1 / sw and the correlation by 1 / den (IEEE reciprocals;
     * a product with a correctly rounded reciprocal is within WGSL's 2.5 ulp for division) */
    float isw = 1.0f / stats[0];
    float mean = ss * isw;
    float var_s = fmaxf(sss * isw - mean * mean, 0.0f);
    float den = sqrtf(stats[2] * var_s);
    if (den < DEGENERATE_DEN) return 1.0f;
    float cov = srs * isw - stats[1] * mean;
    return clampf(1.0f - cov * (1.0f / den), 0.0f, 2.0f);
}

/* PM:161-190: forward-backward reprojection through source s's packed depth map.
 * out = (pixel distance, back-projected depth, 0, valid) */
static void reproject(const orc_params *p, const uint32_t *depths, int32_t s, float px, float py,
                      const float x[3], float out[4]) {
    const orc_src *v = &p->src[s];
    out[0] = out[1] = out[2] = out[3] = -1.0f;
    float xs[3];
    rigid3f(v->R, x, v->t, xs);
    if (!(xs[2] > 1e-6f)) return;
    int32_t pxs = round_even((v->k[0] * xs[0]) / xs[2] + v->k[2]);
    int32_t pys = round_even((v->k[1] * xs[1]) / xs[2] + v->k[3]);
    if (pxs < 0 || pys < 0 || pxs >= v->width || pys >= v->height) return;
    uint32_t packed = depths[(size_t)v->map_off + (size_t)pys * v->width + pxs];
    float ds = orc_unpack_depth(packed);
    if (!(ds > 0.0f) || orc_unpack_cost(packed) > FILTER_COST) return;
    float back_s[3] = {(((float)pxs - v->k[2]) / v->k[0]) * ds, (((float)pys - v->k[3]) / v->k[1]) * ds,
                       ds};
    float e[3] = {back_s[0] - v->t[0], back_s[1] - v->t[1], back_s[2] - v->t[2]};
    float back[3];
    rotate_t3f(v->R, e, back);
    if (!(back[2] > 1e-6f)) return;
    float qx = (p->k[0] * back[0]) / back[2] + p->k[2];
    float qy = (p->k[1] * back[1]) / back[2] + p->k[3];
    float dx = qx - px, dy = qy - py;
    out[0] = sqrtf(dx * dx + dy * dy);
    out[1] = back[2];
    out[2] = 0.0f;
    out[3] = 1.0f;
}

void orc_reproject(const orc_params *p, const uint32_t *depths, int32_t s, float px, float py,
                   const float x[3], float out[4]) {
    reproject(p, depths, s, px, py, x, out);
}

static inline int32_t topk_of(const orc_params *p) {
    int32_t n = p->n_src;
    if (p->topk > 0) return p->topk < n ? p->topk : n;
    int32_t k = n - 1 < 3 ? n - 1 : 3; /* PM:218 */
    return k < 1 ? 1 : k;
}

/* PM:192-233 */
float orc_hypothesis_cost(const orc_params *p, const orc_images *img, const uint32_t *depths,
                          int32_t x, int32_t y, const float *values, const float *weights,
                          const float *stats, const float n[3], float d) {
    float inv = 1.0f / d;
    if (!(d > 0.0f) || inv > p->inv_min || inv < p->inv_max) return INVALID;
    float px = (float)x, py = (float)y;
    float ray[3];
    ray_of(p->k, px, py, ray);
    if (!(dot3f(n, ray) < -1e-3f)) return INVALID;
    float X[3] = {ray[0] * d, ray[1] * d, ray[2] * d};
    float rho = dot3f(n, X);
    int32_t count = p->n_src;
    if (count == 0) return INVALID;
    int geometric = (p->flags & 1u) != 0u;
    float costs[ORC_MAX_SOURCES];
    for (int s = 0; s < count; s++) {
        float H[9];
        orc_homography(p, s, n, rho, H);
        float c = orc_view_cost(p, img, s, H, x, y, values, weights, stats);
        if (geometric) {
            float r[4];
            reproject(p, depths, s, px, py, X, r);
            float e = r[3] > 0.0f ? fminf(r[0], GEOMETRIC_CAP) : GEOMETRIC_CAP;
            c = c + p->geo_weight * e;
        }
        costs[s] = c;
    }
    int32_t k = topk_of(p);
    float total = 0.0f;
    for (int pick = 0; pick < k; pick++) {
        int best = pick;
        for (int j = pick + 1; j < count; j++)
            if (costs[j] < costs[best]) best = j;
        float tmp = costs[pick];
        costs[pick] = costs[best];
        costs[best] = tmp;
        total = total + costs[pick];
    }
    return total / (float)k;
}

/* PM:235-245 */
static void random_normal(uint32_t *rng, const float ray[3], float out[3]) {
    float z = orc_random(rng) * 2.0f - 1.0f;
    float a = orc_random(rng) * 6.2831853f;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float n[3] = {r * cosf_d(a), r * sinf_d(a), z};
    float view[3];
    normalize3f(ray, view);
    if (dot3f(n, view) > 0.0f) {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
    float m[3] = {n[0] - view[0] * 0.5f, n[1] - view[1] * 0.5f, n[2] - view[2] * 0.5f};
    normalize3f(m, out);
}

/* PM:247-249; WGSL mix(a, b, t) = a * (1 - t) + b * t */
static float random_depth(uint32_t *rng, const orc_params *p) {
    float t = orc_random(rng);
    return 1.0f / (p->inv_max * (1.0f - t) + p->inv_min * t);
}

/* ------------------------------------------------------------------------------------------ */
/* Kernels                                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* PM:255-289 */
void orc_init(const orc_params *p, const orc_images *img, const float *prior,
              const uint32_t *depths, float *hyp, float *cost) {
    int32_t W = p->width, Hh = p->height;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int32_t y = 0; y < Hh; y++) {
        for (int32_t x = 0; x < W; x++) {
            uint32_t index = (uint32_t)y * (uint32_t)W + (uint32_t)x;
            uint32_t rng = orc_pixel_rng(p->seed, p->iteration, p->phase, index);
            float values[SAMPLES], weights[SAMPLES], stats[3];
            orc_window(p, img, x, y, values, weights, stats);
            float ray[3];
            ray_of(p->k, (float)x, (float)y, ray);
            float n[3];
            random_normal(&rng, ray, n);
            float d = random_depth(&rng, p);
            if ((p->flags & 1u) != 0u) {
                uint32_t packed = depths[(size_t)p->ref_off + index];
                if (orc_unpack_depth(packed) > 0.0f) {
                    d = orc_unpack_depth(packed);
                    orc_unpack_normal(packed, n);
                }
            }
            float c = orc_hypothesis_cost(p, img, depths, x, y, values, weights, stats, n, d);
            if ((p->flags & 1u) == 0u && prior != NULL && prior[index] > 0.0f) {
                float pn[3];
                normalize3f(ray, pn);
                pn[0] = -pn[0];
                pn[1] = -pn[1];
                pn[2] = -pn[2];
                float pc = orc_hypothesis_cost(p, img, depths, x, y, values, weights, stats, pn,
                                               prior[index]);
                if (pc < c) {
                    c = pc;
                    n[0] = pn[0];
                    n[1] = pn[1];
                    n[2] = pn[2];
                    d = prior[index];
                }
            }
            hyp[index * 4 + 0] = n[0];
            hyp[index * 4 + 1] = n[1];
            hyp[index * 4 + 2] = n[2];
            hyp[index * 4 + 3] = d;
            cost[index] = c;
        }
    }
}

/* PM:296-333 */
void orc_init_coarse(const orc_params *p, const orc_images *img, const float ck[4],
                     int32_t coarse_w, int32_t coarse_h, const float *coarse_hyp, float *hyp,
                     float *cost) {
    int32_t W = p->width, Hh = p->height;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int32_t y = 0; y < Hh; y++) {
        for (int32_t x = 0; x < W; x++) {
            uint32_t index = (uint32_t)y * (uint32_t)W + (uint32_t)x;
            uint32_t rng = orc_pixel_rng(p->seed, p->iteration, p->phase, index);
            float values[SAMPLES], weights[SAMPLES], stats[3];
            orc_window(p, img, x, y, values, weights, stats);
            float px = (float)x, py = (float)y;
            float ray[3];
            ray_of(p->k, px, py, ray);
            float best[4];
            random_normal(&rng, ray, best);
            best[3] = random_depth(&rng, p);
            float best_cost =
                orc_hypothesis_cost(p, img, NULL, x, y, values, weights, stats, best, best[3]);
            float qx = ((px + 0.5f) * (float)coarse_w) / (float)W - 0.5f;
            float qy = ((py + 0.5f) * (float)coarse_h) / (float)Hh - 0.5f;
            int32_t bx = (int32_t)floorf(qx), by = (int32_t)floorf(qy);
            for (int corner = 0; corner < 4; corner++) {
                int32_t cx = bx + (corner & 1), cy = by + (corner >> 1);
                if (cx < 0) cx = 0;
                if (cy < 0) cy = 0;
                if (cx > coarse_w - 1) cx = coarse_w - 1;
                if (cy > coarse_h - 1) cy = coarse_h - 1;
                const float *h = &coarse_hyp[((size_t)cy * coarse_w + cx) * 4];
                if (!(h[3] > 0.0f)) continue;
                float point[3] = {(((float)cx - ck[2]) / ck[0]) * h[3],
                                  (((float)cy - ck[3]) / ck[1]) * h[3], h[3]};
                float den = dot3f(h, ray);
                if (fabsf(den) < 1e-6f) continue;
                float d = dot3f(h, point) / den;
                float cand = orc_hypothesis_cost(p, img, NULL, x, y, values, weights, stats, h, d);
                if (cand < best_cost) {
                    best_cost = cand;
                    best[0] = h[0];
                    best[1] = h[1];
                    best[2] = h[2];
                    best[3] = d;
                }
            }
            memcpy(&hyp[index * 4], best, sizeof(best));
            cost[index] = best_cost;
        }
    }
}

/* PM:335-338 */
static inline void near_offset(int i, int *ox, int *oy) {
    int m = (i + 1) / 2;
    *ox = (i & 1) == 1 ? -m : m;
    *oy = -(m + 1);
}

/* PM:340-347 */
static inline void orient(int ox, int oy, int direction, int *rx, int *ry) {
    switch (direction) {
    case 1: *rx = ox; *ry = -oy; break;
    case 2: *rx = oy; *ry = ox; break;
    case 3: *rx = -oy; *ry = ox; break;
    default: *rx = ox; *ry = oy; break;
    }
}

/* PM:349-423 */
void orc_propagate(const orc_params *p, const orc_images *img, const uint32_t *depths, float *hyp,
                   float *cost) {
    int32_t W = p->width, Hh = p->height;
    int32_t cols = (W + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int32_t y = 0; y < Hh; y++) {
        for (int32_t g = 0; g < cols; g++) {
            int32_t x = g * 2 + (int32_t)(((uint32_t)y + p->phase) & 1u);
            if (x >= W) continue;
            uint32_t index = (uint32_t)y * (uint32_t)W + (uint32_t)x;
            float px = (float)x, py = (float)y;
            uint32_t rng = orc_pixel_rng(p->seed, p->iteration, p->phase, index);
            float values[SAMPLES], weights[SAMPLES], stats[3];
            orc_window(p, img, x, y, values, weights, stats);
            float ray[3];
            ray_of(p->k, px, py, ray);
            float best[4];
            memcpy(best, &hyp[index * 4], sizeof(best));
            float best_cost = cost[index];

            for (int region = 0; region < 8; region++) {
                int count = region < 4 ? 7 : 11;
                int32_t chosen = -1;
                float chosen_cost = 1e9f;
                for (int i = 0; i < count; i++) {
                    int ox = 0, oy = -3 - 2 * i;
                    if (region < 4) near_offset(i, &ox, &oy);
                    int rx, ry;
                    orient(ox, oy, region % 4, &rx, &ry);
                    int32_t qx = x + rx, qy = y + ry;
                    if (qx < 0 || qy < 0 || qx >= W || qy >= Hh) continue;
                    int32_t qi = qy * W + qx;
                    if (cost[qi] < chosen_cost) {
                        chosen_cost = cost[qi];
                        chosen = qi;
                    }
                }
                if (chosen < 0) continue;
                const float *h = &hyp[(size_t)chosen * 4];
                float qray[3];
                ray_of(p->k, (float)(chosen % W), (float)(chosen / W), qray);
                float den = dot3f(h, ray);
                if (fabsf(den) < 1e-6f) continue;
                float qp[3] = {qray[0] * h[3], qray[1] * h[3], qray[2] * h[3]};
                float d = dot3f(h, qp) / den;
                float c = orc_hypothesis_cost(p, img, depths, x, y, values, weights, stats, h, d);
                if (c < best_cost) {
                    best_cost = c;
                    best[0] = h[0];
                    best[1] = h[1];
                    best[2] = h[2];
                    best[3] = d;
                }
            }

            float step = p->step;
            float view[3];
            normalize3f(ray, view);
            float n_random[3];
            random_normal(&rng, ray, n_random);
            float d_random = random_depth(&rng, p);
            float d_perturbed =
                1.0f / ((1.0f / best[3]) * (1.0f + ((orc_random(&rng) * 2.0f - 1.0f) * 0.15f) * step));
            float r1 = orc_random(&rng), r2 = orc_random(&rng), r3 = orc_random(&rng);
            float np_raw[3] = {best[0] + ((r1 * 2.0f - 1.0f) * 0.6f) * step,
                               best[1] + ((r2 * 2.0f - 1.0f) * 0.6f) * step,
                               best[2] + ((r3 * 2.0f - 1.0f) * 0.6f) * step};
            float n_perturbed[3];
            normalize3f(np_raw, n_perturbed);
            if (!(dot3f(n_perturbed, view) < 0.0f)) memcpy(n_perturbed, best, 3 * sizeof(float));
            float cur_n[3] = {best[0], best[1], best[2]};
            float cur_d = best[3];
            const float *normals[6] = {n_random, n_perturbed, cur_n, n_random, cur_n, n_perturbed};
            float tries[6] = {d_random, d_perturbed, d_random, cur_d, d_perturbed, cur_d};
            for (int c = 0; c < 6; c++) {
                float cand = orc_hypothesis_cost(p, img, depths, x, y, values, weights, stats,
                                                 normals[c], tries[c]);
                if (cand < best_cost) {
                    best_cost = cand;
                    best[0] = normals[c][0];
                    best[1] = normals[c][1];
                    best[2] = normals[c][2];
                    best[3] = tries[c];
                }
            }
            memcpy(&hyp[index * 4], best, sizeof(best));
            cost[index] = best_cost;
        }
    }
}

/* PM:425-433 */
void orc_pack(const orc_params *p, const float *hyp, const float *cost, uint32_t *out) {
    size_t total = (size_t)p->width * (size_t)p->height;
    for (size_t i = 0; i < total; i++) out[i] = orc_pack_sample(hyp[i * 4 + 3], &hyp[i * 4], cost[i]);
}

/* PM:436-469 */
void orc_consistency(const orc_params *p, const uint32_t *depths, uint32_t *filtered) {
    int32_t W = p->width, Hh = p->height;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (int32_t y = 0; y < Hh; y++) {
        for (int32_t x = 0; x < W; x++) {
            size_t index = (size_t)y * W + x;
            uint32_t packed = depths[(size_t)p->ref_off + index];
            float d = orc_unpack_depth(packed);
            if (!(d > 0.0f) || orc_unpack_cost(packed) > FILTER_COST) {
                filtered[index] = 0u;
                continue;
            }
            float px = (float)x, py = (float)y;
            float ray[3];
            ray_of(p->k, px, py, ray);
            float X[3] = {ray[0] * d, ray[1] * d, ray[2] * d};
            uint32_t agree = 0;
            float sum = d;
            for (int s = 0; s < p->n_src; s++) {
                float r[4];
                reproject(p, depths, s, px, py, X, r);
                if (r[3] > 0.0f && r[0] < FILTER_PIXELS && fabsf(r[1] - d) < FILTER_DEPTH * d) {
                    agree++;
                    sum = sum + r[1];
                }
            }
            uint32_t needed = p->needed > 1u ? p->needed : 1u;
            if ((uint32_t)p->n_src < needed) needed = (uint32_t)p->n_src;
            if (agree < needed) {
                filtered[index] = 0u;
                continue;
            }
            float mean = sum / (float)(agree + 1u);
            filtered[index] = (packed & 0xffff0000u) | (uint32_t)orc_f32_to_f16(mean);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Host schedule                                                                              */
/* ------------------------------------------------------------------------------------------ */

/* DS:185-239 */
int32_t orc_solve(const orc_view *views, int32_t n_views, const orc_images *img, int32_t ref,
                  const int32_t *sources, int32_t n_src, double min_depth, double max_depth,
                  const float *prior, int32_t geometric, int32_t iterations, int32_t topk,
                  int32_t pack, uint32_t *depths, float *hyp, float *cost) {
    if (n_src > ORC_MAX_SOURCES) n_src = ORC_MAX_SOURCES;
    if (n_src == 0) return -1;
    uint32_t flags = geometric ? 1u : 0u;
    orc_params p;
    orc_make_params(views, n_views, ref, sources, n_src, min_depth, max_depth, 0, 0, 1.0, flags, 0,
                    topk, &p);
    orc_init(&p, img, prior, depths, hyp, cost);
    for (int32_t it = 0; it < iterations; it++) {
        double step = pow(0.5, (double)it);
        for (uint32_t phase = 0; phase < 2; phase++) {
            orc_make_params(views, n_views, ref, sources, n_src, min_depth, max_depth,
                            (uint32_t)it, phase, step, flags, 0, topk, &p);
            orc_propagate(&p, img, depths, hyp, cost);
        }
    }
    if (pack) {
        orc_make_params(views, n_views, ref, sources, n_src, min_depth, max_depth, 0, 0, 0.0, flags,
                        0, topk, &p);
        orc_pack(&p, hyp, cost, depths + p.ref_off);
    }
    return 0;
}

/* DS:250-263 */
void orc_filter(const orc_view *views, int32_t n_views, int32_t ref, const int32_t *checks,
                int32_t n_checks, int32_t needed, const uint32_t *depths, uint32_t *filtered) {
    orc_params p;
    if (n_checks > ORC_MAX_SOURCES) n_checks = ORC_MAX_SOURCES;
    orc_make_params(views, n_views, ref, checks, n_checks, 0.01, 100.0, 0, 0, 0.0, 0,
                    (uint32_t)needed, 0, &p);
    orc_consistency(&p, depths, filtered);
}
