/*
 * Seeded procedural scenes with analytic ground truth, for tests and benchmarks.
 *
 * The surface texture restates the reference's test-scene noise (frontend/tools/synthetic-scene.ts:
 * `hash3` :33-36 / :112-117, `lattice` :38-40 / :119-121, smoothstep-trilinear `noise` :42-52 /
 * :123-139) and its 3-octave albedo mix (:54-61); shading is the same fixed Lambertian light
 * (:63-65), world +y up, cameras built like `lookAt` (frontend/tools/scene-renderer.ts:11-20).
 * Frames are u8 gray (2x2 supersampled), ground-truth depth is the camera z of the pixel-centre
 * ray (like `depthGrid`, scene-renderer.ts:87-114), 0 where the ray hits nothing.
 *
 * Primitives: spheres, solid axis-aligned boxes, finite horizontal rectangles and a room (the
 * inside of an axis-aligned box).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define SG_MAX_PRIMS 64

enum { SG_SPHERE = 0, SG_BOX = 1, SG_RECT_Y = 2, SG_ROOM = 3 };

typedef struct {
    int32_t kind;
    double a[3]; /* sphere centre | box / room min | rectangle (x0, y, z0) */
    double b[3]; /* box / room max | rectangle (x1, -, z1) */
    double r;    /* sphere radius */
} sg_prim;

typedef struct {
    int32_t n_prims;
    sg_prim prims[SG_MAX_PRIMS];
    double noise_offset[3]; /* seeded shift of the noise domain */
    double freq;            /* noise frequency scale (1 = the reference's 70/210/18 per metre) */
    double sky;             /* u8 value of rays that hit nothing */
} sg_scene;

static uint32_t hash3(int32_t x, int32_t y, int32_t z) {
    uint32_t qx = (uint32_t)x * 1597334673u;
    uint32_t qy = (uint32_t)y * 3812015801u;
    uint32_t qz = (uint32_t)z * 2798796415u;
    return (qx ^ qy ^ qz) * 1597334673u;
}

static double lattice(int32_t x, int32_t y, int32_t z) { return (double)(hash3(x, y, z) >> 8) / 16777216.0; }

static double mixd(double a, double b, double t) { return a + (b - a) * t; }

static double noise(double x, double y, double z) {
    double ix = floor(x), iy = floor(y), iz = floor(z);
    double fx = x - ix, fy = y - iy, fz = z - iz;
    double wx = fx * fx * (3 - 2 * fx), wy = fy * fy * (3 - 2 * fy), wz = fz * fz * (3 - 2 * fz);
    int32_t cx = (int32_t)ix, cy = (int32_t)iy, cz = (int32_t)iz;
    double x00 = mixd(lattice(cx, cy, cz), lattice(cx + 1, cy, cz), wx);
    double x10 = mixd(lattice(cx, cy + 1, cz), lattice(cx + 1, cy + 1, cz), wx);
    double x01 = mixd(lattice(cx, cy, cz + 1), lattice(cx + 1, cy, cz + 1), wx);
    double x11 = mixd(lattice(cx, cy + 1, cz + 1), lattice(cx + 1, cy + 1, cz + 1), wx);
    return mixd(mixd(x00, x10, wy), mixd(x01, x11, wy), wz);
}

/* 3-octave albedo (synthetic-scene.ts:54-61) as gray in [0, 1] */
static double albedo(const sg_scene *s, const double p[3]) {
    double q[3] = {p[0] + s->noise_offset[0], p[1] + s->noise_offset[1], p[2] + s->noise_offset[2]};
    double f = s->freq;
    double a = noise(q[0] * 70 * f, q[1] * 70 * f, q[2] * 70 * f);
    double b = noise(q[0] * 210 * f + 11, q[1] * 210 * f + 11, q[2] * 210 * f + 11);
    double c = noise(q[0] * 18 * f + 5, q[1] * 18 * f + 5, q[2] * 18 * f + 5);
    double t = 0.55 * a + 0.3 * b + 0.25 * c - 0.05;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    return mixd(0.12, 0.95, t);
}

static double shade(const double n[3]) {
    const double l[3] = {0.4 / 1.1180339887498949, 1.0 / 1.1180339887498949, 0.3 / 1.1180339887498949};
    double d = n[0] * l[0] + n[1] * l[1] + n[2] * l[2];
    return 0.6 + 0.4 * (d > 0 ? d : 0);
}

/* Nearest hit along o + t*d (t > 1e-9).  Returns t (or -1) and the outward normal. */
static double trace(const sg_scene *s, const double o[3], const double d[3], double n[3]) {
    double best = -1;
    for (int32_t i = 0; i < s->n_prims; i++) {
        const sg_prim *p = &s->prims[i];
        if (p->kind == SG_SPHERE) {
            double oc[3] = {o[0] - p->a[0], o[1] - p->a[1], o[2] - p->a[2]};
            double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
            double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - p->r * p->r;
            double h = B * B - A * C;
            if (h <= 0) continue;
            double t = (-B - sqrt(h)) / A;
            if (t <= 1e-9) continue;
            if (best < 0 || t < best) {
                best = t;
                for (int k = 0; k < 3; k++) n[k] = (o[k] + d[k] * t - p->a[k]) / p->r;
            }
        } else if (p->kind == SG_RECT_Y) {
            if (d[1] == 0) continue;
            double t = (p->a[1] - o[1]) / d[1];
            if (t <= 1e-9) continue;
            double x = o[0] + d[0] * t, z = o[2] + d[2] * t;
            if (x < p->a[0] || x > p->b[0] || z < p->a[2] || z > p->b[2]) continue;
            if (best < 0 || t < best) {
                best = t;
                n[0] = 0;
                n[1] = d[1] < 0 ? 1 : -1;
                n[2] = 0;
            }
        } else if (p->kind == SG_BOX || p->kind == SG_ROOM) {
            double tmin = -INFINITY, tmax = INFINITY;
            int amin = 0, amax = 0;
            int miss = 0;
            for (int k = 0; k < 3; k++) {
                if (d[k] == 0) {
                    if (o[k] < p->a[k] || o[k] > p->b[k]) miss = 1;
                    continue;
                }
                double t0 = (p->a[k] - o[k]) / d[k], t1 = (p->b[k] - o[k]) / d[k];
                if (t0 > t1) {
                    double tmp = t0;
                    t0 = t1;
                    t1 = tmp;
                }
                if (t0 > tmin) {
                    tmin = t0;
                    amin = k;
                }
                if (t1 < tmax) {
                    tmax = t1;
                    amax = k;
                }
            }
            if (miss || tmin > tmax) continue;
            double t;
            int axis;
            double sign;
            if (p->kind == SG_BOX) {
                if (tmin <= 1e-9) continue;
                t = tmin;
                axis = amin;
                sign = d[axis] > 0 ? -1 : 1;
            } else {
                if (tmax <= 1e-9) continue;
                t = tmax; /* inside the room: the exit wall */
                axis = amax;
                sign = d[axis] > 0 ? -1 : 1;
            }
            if (best < 0 || t < best) {
                best = t;
                n[0] = n[1] = n[2] = 0;
                n[axis] = sign;
            }
        }
    }
    return best;
}

/* Render one camera.  pose: world -> camera (x_cam = R X + t).  gray: W*H u8, depth: W*H f32
 * (camera z, 0 = no surface); either output may be NULL. */
int32_t sg_render(const sg_scene *s, int32_t W, int32_t H, double fx, double fy, double cx,
                  double cy, const double R[9], const double t[3], uint8_t *gray, float *depth) {
    /* camera centre c = -R^T t */
    double c[3];
    for (int k = 0; k < 3; k++) c[k] = -(R[k] * t[0] + R[3 + k] * t[1] + R[6 + k] * t[2]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t y = 0; y < H; y++) {
        for (int32_t x = 0; x < W; x++) {
            double n[3];
            if (depth) {
                double rc[3] = {((double)x - cx) / fx, ((double)y - cy) / fy, 1.0};
                double d[3];
                for (int k = 0; k < 3; k++) d[k] = R[k] * rc[0] + R[3 + k] * rc[1] + R[6 + k] * rc[2];
                double hit = trace(s, c, d, n);
                /* the camera-frame ray has z = 1, so the ray parameter is the camera depth */
                depth[(size_t)y * W + x] = hit > 0 ? (float)hit : 0.0f;
            }
            if (gray) {
                double sum = 0;
                for (int sub = 0; sub < 4; sub++) {
                    double jx = (double)(sub & 1) * 0.5 - 0.25, jy = (double)(sub >> 1) * 0.5 - 0.25;
                    double rc[3] = {((double)x + jx - cx) / fx, ((double)y + jy - cy) / fy, 1.0};
                    double d[3];
                    for (int k = 0; k < 3; k++) d[k] = R[k] * rc[0] + R[3 + k] * rc[1] + R[6 + k] * rc[2];
                    double hit = trace(s, c, d, n);
                    if (hit > 0) {
                        double p[3] = {c[0] + d[0] * hit, c[1] + d[1] * hit, c[2] + d[2] * hit};
                        sum += albedo(s, p) * shade(n);
                    } else {
                        sum += s->sky / 255.0;
                    }
                }
                double v = floor(sum / 4.0 * 255.0 + 0.5);
                gray[(size_t)y * W + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
            }
        }
    }
    return 0;
}

/* Signed distance to the nearest surface (positive outside solids and inside the room's free
 * space), for n points (xyz). */
int32_t sg_signed_distance(const sg_scene *s, const double *points, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double *p = &points[i * 3];
        double best = INFINITY;
        for (int32_t k = 0; k < s->n_prims; k++) {
            const sg_prim *q = &s->prims[k];
            double dist;
            if (q->kind == SG_SPHERE) {
                double dx = p[0] - q->a[0], dy = p[1] - q->a[1], dz = p[2] - q->a[2];
                dist = sqrt(dx * dx + dy * dy + dz * dz) - q->r;
            } else if (q->kind == SG_RECT_Y) {
                double dx = fmax(fmax(q->a[0] - p[0], 0), p[0] - q->b[0]);
                double dz = fmax(fmax(q->a[2] - p[2], 0), p[2] - q->b[2]);
                double dy = p[1] - q->a[1];
                dist = sqrt(dx * dx + dz * dz + dy * dy);
                if (dx == 0 && dz == 0) dist = fabs(dy);
                if (dy < 0 && dx == 0 && dz == 0) dist = dy; /* below the slab top counts as inside */
            } else {
                double ctr[3], half[3], qv[3];
                for (int a = 0; a < 3; a++) {
                    ctr[a] = 0.5 * (q->a[a] + q->b[a]);
                    half[a] = 0.5 * (q->b[a] - q->a[a]);
                    qv[a] = fabs(p[a] - ctr[a]) - half[a];
                }
                double ox = fmax(qv[0], 0), oy = fmax(qv[1], 0), oz = fmax(qv[2], 0);
                double outside = sqrt(ox * ox + oy * oy + oz * oz);
                double inside = fmin(fmax(qv[0], fmax(qv[1], qv[2])), 0);
                dist = outside + inside;
                if (q->kind == SG_ROOM) dist = -dist;
            }
            if (fabs(dist) < fabs(best)) best = dist;
        }
        out[i] = best;
    }
    return 0;
}
