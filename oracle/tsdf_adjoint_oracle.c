/*
 * CPU oracle, FP64: TSDF integration of a sequence of views into a fresh volume (TV:103-154 per
 * view, in view order; brick culling TV:220-234) restricted to a list of voxels, the loss
 * L = sum_i (tsdf_i - target_i)^2 over those voxels, and its gradient with respect to the depth
 * used at every pixel (the map depth where the map has one, else the prior) and to every view's
 * world -> camera pose (R as 9 free entries, t).  TEST INFRASTRUCTURE ONLY (see lr3d_oracle.h).
 *
 * The reference has no gradients; this is the FP64 reference the HIP adjoint is checked against
 * (and which is itself checked against FP64 central finite differences).  Per voxel and view:
 *   cam = R world + t;  skipped unless cam.z > 0.01, the nearest pixel (round) is inside, the depth
 *   used d > 0 and sdf = d - cam.z >= -mu;
 *   o = min(1, sdf / mu);
 *   w = max(0.2, |n . cam/|cam||) [* max(0.15, (1 - c/8)^2) * clamp(conf / d, 0.5, 2)^2] (map), 0.12 (prior);
 *   tsdf <- (tsdf W + o w) / (W + w),  W <- min(W + w, 96),  starting from tsdf = 1, W = 0.
 * Reverse sweep with a = dL/dtsdf, b = dL/dW, D = W_prev + w, u = [D < 96]:
 *   dL/do = a w / D,  dL/dw = a (o - tsdf) / D + u b,
 *   a <- a W_prev / D,  b <- a (tsdf_prev - tsdf) / D + u b.
 * The nearest-pixel choice, the skips and the clamps carry no gradient (piecewise constant).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lr3d_oracle.h"

#define MAX_WEIGHT 96.0
#define PRIOR_WEIGHT 0.12

typedef struct {
    int applied;
    double o, w, D, tsdf_prev, W_prev, tsdf;
    double do_dd, dw_dd;      /* d o / d depth, d w / d depth */
    double do_dc[3], dw_dc[3]; /* d o / d cam, d w / d cam */
    size_t pixel;              /* index into the view's map */
    double kink;               /* distance of the projection from a rounding boundary (pixels) */
} step64;

/* One view's contribution to one voxel, with the partial derivatives of (o, w). */
static int view_step(const orc_tsdf_grid *g, const orc_integration_view *v, const orc64_depth_input *in,
                     const double world[3], step64 *s) {
    double cam[3];
    for (int i = 0; i < 3; i++) cam[i] = v->R[i * 3] * world[0] + v->R[i * 3 + 1] * world[1] + v->R[i * 3 + 2] * world[2] + v->t[i];
    const double mu = g->truncation;
    /* distance to the nearest decision boundary, recorded before any skip: the cam.z > 0.01 test
     * and the sdf >= -mu test (in units of mu) and the nearest-pixel rounding (pixels); a voxel this
     * close to one can take the other branch in lower precision */
    const double kz = fabs(cam[2] - 0.01) / mu;
    if (!(cam[2] > 0.01)) {
        s->kink = kz;
        return 0;
    }
    const double u = (v->fx * cam[0]) / cam[2] + v->cx, vv = (v->fy * cam[1]) / cam[2] + v->cy;
    s->kink = fmin(kz, fmin(fabs(fabs(u - floor(u)) - 0.5), fabs(fabs(vv - floor(vv)) - 0.5)));
    const double ru = rint(u), rv = rint(vv);
    if (ru < 0 || rv < 0 || ru >= v->width || rv >= v->height) return 0;
    const size_t i = (size_t)rv * v->width + (size_t)ru;
    double d = in->depth[i];
    const int from_map = d > 0.0;
    double kink = s->kink;
    memset(s, 0, sizeof(*s));
    s->kink = kink;
    if (!from_map) d = in->prior ? in->prior[i] : 0.0;
    if (!(d > 0.0)) return 0;
    const double sdf = d - cam[2];
    kink = fmin(kink, fabs(sdf + mu) / mu);
    s->kink = kink;
    if (sdf < -mu) return 0;
    s->pixel = i;
    if (sdf / mu < 1.0) {
        s->o = sdf / mu;
        s->do_dd = 1.0 / mu;
        s->do_dc[2] = -1.0 / mu;
    } else {
        s->o = 1.0;
    }
    if (from_map) {
        const double *n = &in->normal[i * 3];
        const double len = sqrt(cam[0] * cam[0] + cam[1] * cam[1] + cam[2] * cam[2]);
        const double c[3] = {cam[0] / len, cam[1] / len, cam[2] / len};
        const double nc = n[0] * c[0] + n[1] * c[1] + n[2] * c[2];
        double base = fabs(nc), dbase[3] = {0, 0, 0};
        if (base > 0.2) {
            const double sg = nc >= 0 ? 1.0 : -1.0;
            for (int k = 0; k < 3; k++) dbase[k] = sg * (n[k] - nc * c[k]) / len;
        } else {
            base = 0.2;
        }
        double f = 1.0, df_dd = 0.0;
        if (v->confidence > 0.0) {
            const double cc = (double)in->cost[i] / 8.0;
            const double q = fmax(0.15, (1.0 - cc) * (1.0 - cc));
            double r = v->confidence / d, dr = -v->confidence / (d * d);
            if (r < 0.5) {
                r = 0.5;
                dr = 0.0;
            } else if (r > 2.0) {
                r = 2.0;
                dr = 0.0;
            }
            f = q * r * r;
            df_dd = q * 2.0 * r * dr;
        }
        s->w = base * f;
        s->dw_dd = base * df_dd;
        for (int k = 0; k < 3; k++) s->dw_dc[k] = dbase[k] * f;
    } else {
        s->w = PRIOR_WEIGHT;
    }
    s->applied = 1;
    return 1;
}

int32_t orc64_tsdf_backward(const orc_tsdf_grid *g, const orc_integration_view *views,
                            const orc64_depth_input *inputs, int32_t n_views, const uint32_t *voxels,
                            const double *targets, int64_t n_voxels, double *tsdf_out,
                            double *grad_depth, double *grad_pose, double *kink_out, double *loss) {
    uint8_t *vis = (uint8_t *)malloc((size_t)n_views * g->n_bricks);
    step64 *steps = (step64 *)malloc(sizeof(step64) * (size_t)(n_views > 0 ? n_views : 1));
    size_t *offs = (size_t *)malloc(sizeof(size_t) * (size_t)(n_views + 1));
    if (!vis || !steps || !offs) {
        free(vis);
        free(steps);
        free(offs);
        return -1;
    }
    offs[0] = 0;
    for (int32_t v = 0; v < n_views; v++) {
        orc_visible_bricks(g, &views[v], vis + (size_t)v * g->n_bricks);
        offs[v + 1] = offs[v] + (size_t)views[v].width * views[v].height;
    }
    if (grad_depth) memset(grad_depth, 0, sizeof(double) * offs[n_views]);
    if (grad_pose) memset(grad_pose, 0, sizeof(double) * 12 * (size_t)n_views);
    double total = 0.0;
    for (int64_t q = 0; q < n_voxels; q++) {
        const uint32_t index = voxels[q];
        const int32_t b = (int32_t)(index / 512u), local = (int32_t)(index % 512u);
        const int32_t *c = &g->coords[b * 4];
        const double world[3] = {
            g->origin[0] + ((double)(c[0] * 8 + local % 8) + 0.5) * g->voxel,
            g->origin[1] + ((double)(c[1] * 8 + (local / 8) % 8) + 0.5) * g->voxel,
            g->origin[2] + ((double)(c[2] * 8 + local / 64) + 0.5) * g->voxel,
        };
        double ts = 1.0, W = 0.0, kink = 1e9;
        for (int32_t v = 0; v < n_views; v++) {
            step64 *s = &steps[v];
            s->applied = 0;
            if (!vis[(size_t)v * g->n_bricks + b]) continue;
            s->kink = 1e9;
            const int applied = view_step(g, &views[v], &inputs[v], world, s);
            kink = fmin(kink, s->kink);
            if (!applied) continue;
            s->tsdf_prev = ts;
            s->W_prev = W;
            s->D = W + s->w;
            ts = (ts * W + s->o * s->w) / s->D;
            W = fmin(s->D, MAX_WEIGHT);
            s->tsdf = ts;
        }
        const double r = ts - targets[q];
        total += r * r;
        if (tsdf_out) tsdf_out[q] = ts;
        if (kink_out) kink_out[q] = kink;
        double a = 2.0 * r, bw = 0.0;
        for (int32_t v = n_views - 1; v >= 0; v--) {
            const step64 *s = &steps[v];
            if (!s->applied) continue;
            const double uncapped = s->D < MAX_WEIGHT ? 1.0 : 0.0;
            const double g_o = a * s->w / s->D;
            const double g_w = a * (s->o - s->tsdf) / s->D + uncapped * bw;
            const double a_prev = a * s->W_prev / s->D;
            const double b_prev = a * (s->tsdf_prev - s->tsdf) / s->D + uncapped * bw;
            if (grad_depth) grad_depth[offs[v] + s->pixel] += g_o * s->do_dd + g_w * s->dw_dd;
            if (grad_pose) {
                double gc[3];
                for (int k = 0; k < 3; k++) gc[k] = g_o * s->do_dc[k] + g_w * s->dw_dc[k];
                double *gp = &grad_pose[(size_t)v * 12];
                for (int i = 0; i < 3; i++) {
                    for (int j = 0; j < 3; j++) gp[i * 3 + j] += gc[i] * world[j];
                    gp[9 + i] += gc[i];
                }
            }
            a = a_prev;
            bw = b_prev;
        }
    }
    if (loss) *loss = total;
    free(vis);
    free(steps);
    free(offs);
    return 0;
}

/* The loss alone, L = sum_i (tsdf_i - target_i)^2 (the forward of orc64_tsdf_backward), with the
 * voxels split over OpenMP threads; for central differences over large voxel lists.  vis: the
 * culling flags [view][brick] (orc_visible_bricks), or NULL to compute them.  terms (or NULL)
 * receives each voxel's (tsdf_i - target_i)^2; the total adds them in voxel order, so it does not
 * depend on the thread count. */
int32_t orc64_tsdf_loss(const orc_tsdf_grid *g, const orc_integration_view *views, const orc64_depth_input *inputs,
                        int32_t n_views, const uint32_t *voxels, const double *targets, int64_t n_voxels,
                        const uint8_t *vis, double *loss, double *terms) {
    uint8_t *own = NULL;
    double *own_terms = NULL;
    if (!terms) {
        own_terms = (double *)malloc((size_t)n_voxels * sizeof(double) + 1);
        if (!own_terms) return -1;
        terms = own_terms;
    }
    if (!vis) {
        own = (uint8_t *)malloc((size_t)n_views * g->n_bricks + 1);
        if (!own) {
            free(own_terms);
            return -1;
        }
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
        for (int32_t v = 0; v < n_views; v++) orc_visible_bricks(g, &views[v], own + (size_t)v * g->n_bricks);
        vis = own;
    }
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (int64_t q = 0; q < n_voxels; q++) {
        const uint32_t index = voxels[q];
        const int32_t b = (int32_t)(index / 512u), local = (int32_t)(index % 512u);
        const int32_t *c = &g->coords[b * 4];
        const double world[3] = {
            g->origin[0] + ((double)(c[0] * 8 + local % 8) + 0.5) * g->voxel,
            g->origin[1] + ((double)(c[1] * 8 + (local / 8) % 8) + 0.5) * g->voxel,
            g->origin[2] + ((double)(c[2] * 8 + local / 64) + 0.5) * g->voxel,
        };
        double ts = 1.0, W = 0.0;
        for (int32_t v = 0; v < n_views; v++) {
            if (!vis[(size_t)v * g->n_bricks + b]) continue;
            step64 s;
            s.kink = 1e9;
            if (!view_step(g, &views[v], &inputs[v], world, &s)) continue;
            const double D = W + s.w;
            ts = (ts * W + s.o * s.w) / D;
            W = fmin(D, MAX_WEIGHT);
        }
        const double r = ts - targets[q];
        terms[q] = r * r;
    }
    double total = 0.0;
    for (int64_t q = 0; q < n_voxels; q++) total += terms[q];
    *loss = total;
    free(own);
    free(own_terms);
    return 0;
}
