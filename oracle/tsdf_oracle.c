/*
 * CPU oracle: TSDF depth-map integration.  TEST INFRASTRUCTURE ONLY (see lr3d_oracle.h).
 *
 * Restates frontend/src/app/vision/fusion/tsdf-volume.ts (TV:<line>): the brick layout of
 * BrickGrid (TV:13-65), the host culling visibleBricks (TV:220-234, in double precision like
 * the JavaScript), the uniform packing of TsdfVolume.integrate (TV:240-268) and the WGSL
 * `integrate` kernel (TV:103-154), under the arithmetic definition of lr3d_oracle.h.
 */
#include "lr3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BRICK 8
#define BRICK_VOXELS 512
#define MAX_WEIGHT 96.0f   /* TV:7 */
#define PRIOR_WEIGHT 0.12f /* TV:8 */

/* BrickGrid.build with every cell marked (TV:54-64): pool order = ascending grid index */
int32_t orc_dense_coords(const int32_t dims[3], int32_t *coords) {
    int32_t dx = dims[0], dy = dims[1], dz = dims[2];
    int32_t n = 0;
    for (int32_t i = 0; i < dx * dy * dz; i++) {
        coords[n * 4 + 0] = i % dx;
        coords[n * 4 + 1] = (i / dx) % dy;
        coords[n * 4 + 2] = i / (dx * dy);
        coords[n * 4 + 3] = 0;
        n++;
    }
    return n;
}

/* TV:220-234 with CAM:52-65 (toCamera/project) in double */
int32_t orc_visible_bricks(const orc_tsdf_grid *g, const orc_integration_view *v, uint8_t *flags) {
    double side = g->voxel * BRICK;
    double radius = g->voxel * BRICK * 0.87 + g->truncation;
    int32_t count = 0;
    for (int32_t b = 0; b < g->n_bricks; b++) {
        double c[3];
        for (int a = 0; a < 3; a++) c[a] = g->origin[a] + ((double)g->coords[b * 4 + a] + 0.5) * side;
        double p[3];
        for (int r = 0; r < 3; r++)
            p[r] = (v->R[r * 3] * c[0] + v->R[r * 3 + 1] * c[1] + v->R[r * 3 + 2] * c[2]) + v->t[r];
        double u = (v->fx * p[0]) / p[2] + v->cx;
        double w = (v->fy * p[1]) / p[2] + v->cy;
        double z = p[2];
        flags[b] = 0;
        if (z < 0.02 - radius || z > v->max_depth + radius) continue;
        double margin = (v->fx * radius) / fmax(0.02, z);
        if (u < -margin || w < -margin || u > v->width + margin || w > v->height + margin) continue;
        flags[b] = 1;
        count++;
    }
    return count;
}

static inline float dot3f(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* TV:240-268 (host) + TV:103-154 (kernel) */
int64_t orc_tsdf_integrate(const orc_tsdf_grid *g, const orc_integration_view *v,
                           const uint32_t *depth, const float *prior, float *tsdf, float *weight) {
    uint8_t *visible = (uint8_t *)malloc((size_t)g->n_bricks);
    if (!visible) return -1;
    int32_t n_visible = orc_visible_bricks(g, v, visible);
    int64_t updates = 0;
    if (n_visible == 0) {
        free(visible);
        return 0;
    }
    /* The 128-byte uniform (TV:244-253): every field as f32. */
    float k[4] = {(float)v->fx, (float)v->fy, (float)v->cx, (float)v->cy};
    float R[9];
    for (int i = 0; i < 9; i++) R[i] = (float)v->R[i];
    float t[3] = {(float)v->t[0], (float)v->t[1], (float)v->t[2]};
    float origin[3] = {(float)g->origin[0], (float)g->origin[1], (float)g->origin[2]};
    float voxel = (float)g->voxel;
    float mu = (float)g->truncation;
    float confidence = (float)v->confidence;
    int32_t W = v->width, H = v->height;

#pragma omp parallel for schedule(static) reduction(+ : updates) num_threads(orc_get_threads())
    for (int32_t b = 0; b < g->n_bricks; b++) {
        if (!visible[b]) continue;
        const int32_t *base = &g->coords[b * 4];
        for (int32_t local = 0; local < BRICK_VOXELS; local++) {
            int32_t vx = local % 8, vy = (local / 8) % 8, vz = local / 64;
            float world[3] = {
                origin[0] + ((float)(base[0] * 8 + vx) + 0.5f) * voxel,
                origin[1] + ((float)(base[1] * 8 + vy) + 0.5f) * voxel,
                origin[2] + ((float)(base[2] * 8 + vz) + 0.5f) * voxel,
            };
            float cam[3];
            for (int i = 0; i < 3; i++)
                cam[i] = fmaf(R[i * 3 + 2], world[2],
                              fmaf(R[i * 3 + 1], world[1], fmaf(R[i * 3], world[0], t[i])));
            if (!(cam[2] > 0.01f)) continue;
            int32_t px = (int32_t)rintf((k[0] * cam[0]) / cam[2] + k[2]);
            int32_t py = (int32_t)rintf((k[1] * cam[1]) / cam[2] + k[3]);
            if (px < 0 || py < 0 || px >= W || py >= H) continue;
            size_t i = (size_t)py * W + px;
            uint32_t packed = depth[i];
            float d = orc_unpack_depth(packed);
            float w;
            if (d > 0.0f) {
                float n[3];
                orc_unpack_normal(packed, n);
                float inv = 1.0f / sqrtf(dot3f(cam, cam));
                float cn[3] = {cam[0] * inv, cam[1] * inv, cam[2] * inv};
                w = fmaxf(0.2f, fabsf(dot3f(n, cn)));
                if (confidence > 0.0f) {
                    float c = (float)orc_unpack_cost(packed) / 8.0f;
                    float r = fminf(fmaxf(confidence / d, 0.5f), 2.0f);
                    /* `w *= max(..) * r * r` (TV:135): w * ((max * r) * r) */
                    w = w * ((fmaxf(0.15f, (1.0f - c) * (1.0f - c)) * r) * r);
                }
            } else {
                d = prior ? prior[i] : 0.0f;
                w = PRIOR_WEIGHT;
            }
            if (!(d > 0.0f)) continue;
            float sdf = d - cam[2];
            if (sdf < -mu) continue;
            float observed = fminf(1.0f, sdf / mu);
            size_t index = (size_t)b * BRICK_VOXELS + local;
            float w0 = weight[index];
            tsdf[index] = (tsdf[index] * w0 + observed * w) / (w0 + w);
            weight[index] = fminf(w0 + w, MAX_WEIGHT);
            updates++;
        }
    }
    free(visible);
    return updates;
}
