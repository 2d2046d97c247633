/*
 * CPU oracle: dense-cloud fusion of packed depth maps.  TEST INFRASTRUCTURE ONLY (see
 * lr3d_oracle.h).
 *
 * Restates frontend/src/app/vision/fusion/point-accumulator.ts (PA:<line>: open-addressing hash
 * on the integer cell, Float32Array sums, growth at 60 % load, coarsening by re-binning the cell
 * means) and addDepthMap / fuseDepthMaps of frontend/src/app/vision/fusion/depth-fusion.ts
 * (DF:<line>), with unpackDepthSample (frontend/src/app/vision/mvs/depth-maps.ts:65-85) and
 * fromCamera / applyTransposed3 (geometry/camera.ts:57-59, math/mat3.ts:32-38), in double like the
 * JavaScript.  Vector length is sqrt(x^2 + y^2 + z^2) (the reference's Math.hypot is left to the
 * JavaScript engine: last-ulp parity unpinned).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lr3d_oracle.h"

#define EMPTY 0x7fffffff

typedef struct {
    int64_t capacity, used;
    int32_t *keys;
    float *sums;
    uint32_t *counts;
    double cell;
} accumulator;

/* PA:145-147 hash: (imul(x, 73856093) ^ imul(y, 19349663) ^ imul(z, 83492791)) >>> 0 */
static uint32_t hash3(int32_t x, int32_t y, int32_t z) {
    return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}

static int acc_init(accumulator *a, int64_t capacity, double cell) {
    a->capacity = capacity;
    a->used = 0;
    a->cell = cell;
    a->keys = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)capacity);
    a->sums = (float *)calloc((size_t)capacity * 9, sizeof(float));
    a->counts = (uint32_t *)calloc((size_t)capacity, sizeof(uint32_t));
    if (!a->keys || !a->sums || !a->counts) return -1;
    for (int64_t i = 0; i < capacity * 3; i++) a->keys[i] = EMPTY;
    return 0;
}

static void acc_free(accumulator *a) {
    free(a->keys);
    free(a->sums);
    free(a->counts);
}

static int64_t slot_of(accumulator *a, int32_t ix, int32_t iy, int32_t iz);

/* PA:103-131 rebuild: re-bin every cell by its mean position, merging sums in slot order */
static int rebuild(accumulator *a, int64_t capacity, double cell) {
    accumulator old = *a;
    if (acc_init(a, capacity, cell)) return -1;
    for (int64_t s = 0; s < old.capacity; s++) {
        if (old.keys[s * 3] == EMPTY) continue;
        uint32_t n = old.counts[s];
        int64_t slot = slot_of(a, (int32_t)floor((double)old.sums[s * 9] / n / cell),
                               (int32_t)floor((double)old.sums[s * 9 + 1] / n / cell),
                               (int32_t)floor((double)old.sums[s * 9 + 2] / n / cell));
        for (int i = 0; i < 9; i++) a->sums[slot * 9 + i] = (float)((double)a->sums[slot * 9 + i] + (double)old.sums[s * 9 + i]);
        a->counts[slot] += n;
    }
    acc_free(&old);
    return 0;
}

/* PA:66-85 */
static int64_t slot_of(accumulator *a, int32_t ix, int32_t iy, int32_t iz) {
    if (a->used * 10 > a->capacity * 6) rebuild(a, a->capacity * 2, a->cell);
    int64_t mask = a->capacity - 1;
    int64_t slot = hash3(ix, iy, iz) & mask;
    for (;;) {
        int64_t k = slot * 3;
        if (a->keys[k] == EMPTY) {
            a->keys[k] = ix;
            a->keys[k + 1] = iy;
            a->keys[k + 2] = iz;
            a->used++;
            return slot;
        }
        if (a->keys[k] == ix && a->keys[k + 1] == iy && a->keys[k + 2] == iz) return slot;
        slot = (slot + 1) & mask;
    }
}

/* PA:37-63 */
static void acc_add(accumulator *a, const double v[9]) {
    int64_t slot = slot_of(a, (int32_t)floor(v[0] / a->cell), (int32_t)floor(v[1] / a->cell),
                           (int32_t)floor(v[2] / a->cell));
    for (int i = 0; i < 9; i++) a->sums[slot * 9 + i] = (float)((double)a->sums[slot * 9 + i] + v[i]);
    a->counts[slot]++;
}

static double half_to_double(uint32_t h) {
    int exponent = (h >> 10) & 31, mantissa = h & 1023;
    double sign = (h & 0x8000) ? -1.0 : 1.0;
    if (exponent == 0) return sign * ldexp(1.0, -14) * (mantissa / 1024.0);
    if (exponent == 31) return NAN;
    return sign * ldexp(1.0, exponent - 15) * (1.0 + mantissa / 1024.0);
}

/* depth-maps.ts:9-41 upsampleDepth (double arithmetic, stored as float like the Float32Array) */
void orc_upsample_depth(const float *meters, int32_t gw, int32_t gh, int32_t width, int32_t height, float *out) {
    for (int32_t y = 0; y < height; y++) {
        const double gy = fmin(gh - 1, fmax(0.0, ((y + 0.5) / height) * gh - 0.5));
        const int32_t y0 = (int32_t)floor(gy);
        const int32_t y1 = y0 + 1 < gh - 1 ? y0 + 1 : gh - 1;
        const double wy = gy - y0;
        for (int32_t x = 0; x < width; x++) {
            const double gx = fmin(gw - 1, fmax(0.0, ((x + 0.5) / width) * gw - 0.5));
            const int32_t x0 = (int32_t)floor(gx);
            const int32_t x1 = x0 + 1 < gw - 1 ? x0 + 1 : gw - 1;
            const double wx = gx - x0;
            const double d[4] = {meters[y0 * gw + x0], meters[y0 * gw + x1], meters[y1 * gw + x0], meters[y1 * gw + x1]};
            const double w[4] = {(1 - wx) * (1 - wy), wx * (1 - wy), (1 - wx) * wy, wx * wy};
            double sum = 0, weight = 0;
            for (int k = 0; k < 4; k++) {
                if (d[k] > 0) {
                    sum += d[k] * w[k];
                    weight += w[k];
                }
            }
            out[(size_t)y * width + x] = weight > 0.5 ? (float)(sum / weight) : 0.0f;
        }
    }
}

/* depth-maps.ts:65-85 unpackDepthSample */
int32_t orc_unpack_depth_sample(uint32_t value, double out[4]) {
    if (value == 0) return 0;
    double d = (double)(float)half_to_double(value & 0xffff); /* halfTable is a Float32Array */
    if (!(d > 0)) return 0;
    double ex = (((value >> 16) & 63) / 63.0) * 2 - 1;
    double ey = (((value >> 22) & 63) / 63.0) * 2 - 1;
    double nx = ex, ny = ey;
    double nz = 1 - fabs(ex) - fabs(ey);
    if (nz < 0) {
        nx = (1 - fabs(ey)) * (ex >= 0 ? 1 : -1);
        ny = (1 - fabs(ex)) * (ey >= 0 ? 1 : -1);
    }
    double length = sqrt(nx * nx + ny * ny + nz * nz);
    if (length == 0) length = 1;
    out[0] = d;
    out[1] = nx / length;
    out[2] = ny / length;
    out[3] = nz / length;
    return 1;
}

/* DF:44-74 addDepthMap, then the coarsening loop of DF:36-37 */
int32_t orc_fuse_depth_maps(const orc_fusion_view *views, int32_t n_views, double cell, int64_t cell_budget,
                            int32_t min_support, orc_cloud *out, double *final_cell) {
    accumulator a;
    memset(out, 0, sizeof(*out));
    if (acc_init(&a, 1 << 16, cell)) return -1;
    double sample[4], v[9];
    for (int32_t l = 0; l < n_views; l++) {
        const orc_fusion_view *fv = &views[l];
        const int32_t W = fv->width, H = fv->height;
        double sx = (double)fv->thumb_width / W, sy = (double)fv->thumb_height / H;
        const double *R = fv->R, *t = fv->t;
        for (int32_t y = 0; y < H; y++) {
            for (int32_t x = 0; x < W; x++) {
                if (!orc_unpack_depth_sample(fv->packed[(size_t)y * W + x], sample)) continue;
                double d = sample[0];
                double c[3] = {((x - fv->cx) / fv->fx) * d - t[0], ((y - fv->cy) / fv->fy) * d - t[1], d - t[2]};
                v[0] = R[0] * c[0] + R[3] * c[1] + R[6] * c[2];
                v[1] = R[1] * c[0] + R[4] * c[1] + R[7] * c[2];
                v[2] = R[2] * c[0] + R[5] * c[1] + R[8] * c[2];
                v[3] = R[0] * sample[1] + R[3] * sample[2] + R[6] * sample[3];
                v[4] = R[1] * sample[1] + R[4] * sample[2] + R[7] * sample[3];
                v[5] = R[2] * sample[1] + R[5] * sample[2] + R[8] * sample[3];
                int32_t tx = (int32_t)floor((x + 0.5) * sx), ty = (int32_t)floor((y + 0.5) * sy);
                if (tx > fv->thumb_width - 1) tx = fv->thumb_width - 1;
                if (ty > fv->thumb_height - 1) ty = fv->thumb_height - 1;
                size_t i = ((size_t)ty * fv->thumb_width + tx) * 3;
                v[6] = fv->rgb[i];
                v[7] = fv->rgb[i + 1];
                v[8] = fv->rgb[i + 2];
                acc_add(&a, v);
            }
        }
        while (a.used > cell_budget)
            if (rebuild(&a, a.capacity, a.cell * 1.4)) return -1; /* PA:98-100 coarsen(1.4) */
    }
    /* PA:108-143 cloud(minSupport), slot order */
    int64_t total = 0;
    for (int64_t s = 0; s < a.capacity; s++)
        if (a.keys[s * 3] != EMPTY && a.counts[s] >= (uint32_t)min_support) total++;
    out->n = total;
    out->positions = (float *)malloc(sizeof(float) * 3 * (size_t)(total ? total : 1));
    out->normals = (float *)malloc(sizeof(float) * 3 * (size_t)(total ? total : 1));
    out->colors = (uint8_t *)malloc(3 * (size_t)(total ? total : 1));
    out->support = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(total ? total : 1));
    int64_t p = 0;
    for (int64_t s = 0; s < a.capacity; s++) {
        if (a.keys[s * 3] == EMPTY || a.counts[s] < (uint32_t)min_support) continue;
        uint32_t n = a.counts[s];
        const float *o = &a.sums[s * 9];
        for (int k = 0; k < 3; k++) out->positions[p * 3 + k] = (float)((double)o[k] / n);
        double length = sqrt((double)o[3] * o[3] + (double)o[4] * o[4] + (double)o[5] * o[5]);
        if (length == 0) length = 1;
        for (int k = 0; k < 3; k++) out->normals[p * 3 + k] = (float)((double)o[3 + k] / length);
        for (int k = 0; k < 3; k++) out->colors[p * 3 + k] = (uint8_t)(uint32_t)((double)o[6 + k] / n);
        out->support[p] = (uint16_t)(n < 65535 ? n : 65535);
        p++;
    }
    if (final_cell) *final_cell = a.cell;
    acc_free(&a);
    return 0;
}

/* The accumulator on its own (PA:17-143), for the reference's unit tests. */
void *orc_acc_create(double cell) {
    accumulator *a = (accumulator *)malloc(sizeof(accumulator));
    if (!a || acc_init(a, 1 << 16, cell)) return NULL;
    return a;
}
void orc_acc_add(void *h, const double v[9]) { acc_add((accumulator *)h, v); }
void orc_acc_coarsen(void *h, double factor) {
    accumulator *a = (accumulator *)h;
    rebuild(a, a->capacity, a->cell * factor);
}
int64_t orc_acc_size(void *h) { return ((accumulator *)h)->used; }
/* points with at least min_support samples, slot order: positions, colors (3 each), support */
int64_t orc_acc_cloud(void *h, int32_t min_support, float *positions, uint8_t *colors, uint16_t *support) {
    accumulator *a = (accumulator *)h;
    int64_t p = 0;
    for (int64_t s = 0; s < a->capacity; s++) {
        if (a->keys[s * 3] == EMPTY || a->counts[s] < (uint32_t)min_support) continue;
        uint32_t n = a->counts[s];
        if (positions)
            for (int k = 0; k < 3; k++) positions[p * 3 + k] = (float)((double)a->sums[s * 9 + k] / n);
        if (colors)
            for (int k = 0; k < 3; k++) colors[p * 3 + k] = (uint8_t)(uint32_t)((double)a->sums[s * 9 + 6 + k] / n);
        if (support) support[p] = (uint16_t)(n < 65535 ? n : 65535);
        p++;
    }
    return p;
}
void orc_acc_destroy(void *h) {
    if (!h) return;
    acc_free((accumulator *)h);
    free(h);
}

void orc_cloud_free(orc_cloud *c) {
    free(c->positions);
    free(c->normals);
    free(c->colors);
    free(c->support);
    memset(c, 0, sizeof(*c));
}
