/*
 * CPU oracle, FP64: the aggregated PatchMatch cost of one plane hypothesis (PM:130-159, PM:192-233)
 * and its analytic gradient with respect to the plane (depth, normal) and to the relative pose
 * (R, t) of each source view.  TEST INFRASTRUCTURE ONLY (see lr3d_oracle.h).
 *
 * The reference has no gradients; this is the FP64 reference the HIP adjoint is checked against
 * (and which is itself checked against FP64 central finite differences).  The cost follows the
 * forward definition in double precision: bilateral window in u8 units, homography
 * H = Ks (R + t n^T / rho) Kr^-1 with rho = d (n . ray), bilinear sampling with clamp-to-edge,
 * INVALID = 2 for a view with a tap outside, 1 for a degenerate variance, clamp to [0, 2], mean of
 * the k smallest view costs.  Gradients are taken where the cost is differentiable (selected,
 * valid, non-degenerate, unclamped views); elsewhere a view contributes nothing.
 */
#include <math.h>
#include <string.h>

#include "lr3d_oracle.h"

#define SAMPLES 36

static double pix(const orc_images *img, int32_t layer, int32_t x, int32_t y) {
    int32_t w = img->widths[layer], h = img->heights[layer];
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    return (double)img->pixels[layer][(size_t)y * w + x];
}

static void window64(const orc_images *img, int32_t ref, int32_t x, int32_t y, double *v, double *w,
                     double *sw, double *mean, double *var) {
    double c = pix(img, ref, x, y);
    double s = 0, sr = 0, srr = 0;
    for (int i = 0; i < SAMPLES; i++) {
        int ox = (i % 6) * 2 - 5, oy = (i / 6) * 2 - 5;
        double val = pix(img, ref, x + ox, y + oy);
        double dc = (val - c) / 255.0;
        w[i] = exp(-dc * dc / (2 * 0.06 * 0.06) - (double)(ox * ox + oy * oy) / 50.0);
        v[i] = val;
        s += w[i];
        sr += w[i] * val;
        srr += w[i] * val * val;
    }
    *sw = s;
    *mean = sr / s;
    *var = fmax(srr / s - (*mean) * (*mean), 0.0);
}

/* Per-view cost with gradient d cost / d H (row-major 3x3). Returns the cost. */
static double view64(const orc64_params *p, const orc_images *img, int32_t s, const double H[9],
                     int32_t x, int32_t y, const double *rv, const double *rw, double sw, double rmean,
                     double rvar, double G[9], int *smooth, double *min_kink) {
    const orc64_src *sv = &p->src[s];
    const uint8_t *im = img->pixels[sv->layer];
    int32_t W = sv->width, Hh = sv->height;
    double u[SAMPLES], v[SAMPLES], qz[SAMPLES], val[SAMPLES], gu[SAMPLES], gv[SAMPLES];
    double ss = 0, sss = 0, srs = 0;
    *smooth = 0;
    memset(G, 0, 9 * sizeof(double));
    double kink = 1e9;
    for (int i = 0; i < SAMPLES; i++) {
        double X = x + (i % 6) * 2 - 5, Y = y + (i / 6) * 2 - 5;
        double q0 = H[0] * X + H[1] * Y + H[2];
        double q1 = H[3] * X + H[4] * Y + H[5];
        double q2 = H[6] * X + H[7] * Y + H[8];
        if (!(q2 > 1e-6)) return 2.0;
        u[i] = q0 / q2;
        v[i] = q1 / q2;
        qz[i] = q2;
        if (!(u[i] >= 0 && v[i] >= 0 && u[i] <= W - 1 && v[i] <= Hh - 1)) return 2.0;
        double x0 = floor(u[i]), y0 = floor(v[i]);
        double fx = u[i] - x0, fy = v[i] - y0;
        kink = fmin(kink, fmin(fmin(fx, 1 - fx), fmin(fy, 1 - fy)));
        int32_t ix = (int32_t)x0, iy = (int32_t)y0;
        int32_t ix1 = ix + 1 < W ? ix + 1 : W - 1, iy1 = iy + 1 < Hh ? iy + 1 : Hh - 1;
        double a = im[(size_t)iy * W + ix], b = im[(size_t)iy * W + ix1];
        double c = im[(size_t)iy1 * W + ix], e = im[(size_t)iy1 * W + ix1];
        double top = a + fx * (b - a), bot = c + fx * (e - c);
        val[i] = top + fy * (bot - top);
        gu[i] = (b - a) * (1 - fy) + (e - c) * fy;
        gv[i] = bot - top;
        ss += rw[i] * val[i];
        sss += rw[i] * val[i] * val[i];
        srs += rw[i] * val[i] * rv[i];
    }
    double mean = ss / sw;
    double var_s = fmax(sss / sw - mean * mean, 0.0);
    double den = sqrt(rvar * var_s);
    if (den < 1e-5 * 255.0 * 255.0) return 1.0;
    double cov = srs / sw - rmean * mean;
    double cost = 1.0 - cov / den;
    if (cost <= 0.0 || cost >= 2.0) return cost <= 0.0 ? 0.0 : 2.0;
    *smooth = 1;
    *min_kink = kink;
    for (int i = 0; i < SAMPLES; i++) {
        double X = x + (i % 6) * 2 - 5, Y = y + (i / 6) * 2 - 5;
        double dval = (rw[i] / sw) * (-(rv[i] - rmean) / den + cov * rvar * (val[i] - mean) / (den * den * den));
        double gx = dval * gu[i] / qz[i], gy = dval * gv[i] / qz[i];
        double gz = -(gx * u[i] + gy * v[i]);
        double xs[3] = {X, Y, 1.0};
        for (int k = 0; k < 3; k++) {
            G[k] += gx * xs[k];
            G[3 + k] += gy * xs[k];
            G[6 + k] += gz * xs[k];
        }
    }
    return cost;
}

/* d cost / d (depth, normal, R, t) of one view from d cost / d H. */
static void chain64(const orc64_params *p, int32_t s, const double n[3], double d, const double ray[3],
                    const double G[9], double scale, orc64_result *out) {
    const orc64_src *sv = &p->src[s];
    double rho = d * (n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2]);
    double Ks[9] = {sv->k[0], 0, sv->k[2], 0, sv->k[1], sv->k[3], 0, 0, 1};
    double Ki[9] = {1 / p->k[0], 0, -p->k[2] / p->k[0], 0, 1 / p->k[1], -p->k[3] / p->k[1], 0, 0, 1};
    /* G_M = Ks^T G Ki^T */
    double T[9], GM[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += Ks[k * 3 + i] * G[k * 3 + j];
            T[i * 3 + j] = acc;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += T[i * 3 + k] * Ki[j * 3 + k];
            GM[i * 3 + j] = acc;
        }
    /* M = R + t n^T / rho */
    double tGn = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tGn += sv->t[i] * GM[i * 3 + j] * n[j];
    double X[3] = {ray[0] * d, ray[1] * d, ray[2] * d};
    double nr = n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2];
    for (int k = 0; k < 3; k++) {
        double tg = 0;
        for (int i = 0; i < 3; i++) tg += sv->t[i] * GM[i * 3 + k];
        out->d_normal[k] += scale * (tg / rho - tGn / (rho * rho) * X[k]);
    }
    out->d_depth += scale * (-tGn * nr / (rho * rho));
    for (int i = 0; i < 9; i++) out->d_R[s][i] += scale * GM[i];
    for (int i = 0; i < 3; i++) {
        double g = 0;
        for (int j = 0; j < 3; j++) g += GM[i * 3 + j] * n[j];
        out->d_t[s][i] += scale * g / rho;
    }
}

void orc64_homography(const orc64_params *p, int32_t s, const double n[3], double rho, double H[9]) {
    const orc64_src *v = &p->src[s];
    double M[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i * 3 + j] = v->R[i * 3 + j] + v->t[i] * n[j] / rho;
    double Ks[9] = {v->k[0], 0, v->k[2], 0, v->k[1], v->k[3], 0, 0, 1};
    double Ki[9] = {1 / p->k[0], 0, -p->k[2] / p->k[0], 0, 1 / p->k[1], -p->k[3] / p->k[1], 0, 0, 1};
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += Ks[i * 3 + k] * M[k * 3 + j];
            T[i * 3 + j] = acc;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += T[i * 3 + k] * Ki[k * 3 + j];
            H[i * 3 + j] = acc;
        }
}

/* Aggregated cost of plane (n, d) at reference pixel (x, y) and its gradient. */
int32_t orc64_cost_grad(const orc64_params *p, const orc_images *img, int32_t ref, int32_t x, int32_t y,
                        const double n[3], double d, orc64_result *out) {
    memset(out, 0, sizeof(*out));
    out->min_kink = 1e9;
    double ray[3] = {(x - p->k[2]) / p->k[0], (y - p->k[3]) / p->k[1], 1.0};
    double inv = 1.0 / d;
    if (!(d > 0) || inv > p->inv_min || inv < p->inv_max) {
        out->cost = 2.0;
        return 0;
    }
    double nr = n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2];
    if (!(nr < -1e-3)) {
        out->cost = 2.0;
        return 0;
    }
    double rv[SAMPLES], rw[SAMPLES], sw, rmean, rvar;
    window64(img, ref, x, y, rv, rw, &sw, &rmean, &rvar);
    double rho = d * nr;
    double costs[ORC_MAX_SOURCES], G[ORC_MAX_SOURCES][9], kinks[ORC_MAX_SOURCES];
    int smooth[ORC_MAX_SOURCES];
    for (int s = 0; s < p->n_src; s++) {
        double H[9];
        orc64_homography(p, s, n, rho, H);
        kinks[s] = 1e9;
        costs[s] = view64(p, img, s, H, x, y, rv, rw, sw, rmean, rvar, G[s], &smooth[s], &kinks[s]);
        out->view_cost[s] = costs[s];
    }
    int32_t k = p->topk > 0 ? (p->topk < p->n_src ? p->topk : p->n_src)
                            : (p->n_src - 1 < 3 ? (p->n_src - 1 < 1 ? 1 : p->n_src - 1) : 3);
    /* the k smallest by (cost, index) */
    int used[ORC_MAX_SOURCES] = {0};
    double total = 0;
    for (int pick = 0; pick < k; pick++) {
        int best = -1;
        for (int s = 0; s < p->n_src; s++)
            if (!used[s] && (best < 0 || costs[s] < costs[best])) best = s;
        used[best] = 1;
        total += costs[best];
    }
    out->cost = total / k;
    for (int s = 0; s < p->n_src; s++) {
        out->selected[s] = used[s];
        if (!used[s] || !smooth[s]) continue;
        out->min_kink = fmin(out->min_kink, kinks[s]);
        chain64(p, s, n, d, ray, G[s], 1.0 / k, out);
    }
    return 0;
}
