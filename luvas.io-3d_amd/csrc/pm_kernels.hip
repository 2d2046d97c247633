// PatchMatch multi-view stereo kernels for gfx950.
//
// Semantics follow the reference WGSL module (frontend/src/app/vision/mvs/patchmatch.wgsl.ts,
// cited PM:<line>) under a fixed FP32 arithmetic definition, so that results are reproducible
// bit for bit on the CPU: no contraction (built with -ffp-contract=off), fmaf only where
// written, IEEE division and sqrt, exp/sin/cos evaluated in double and rounded to float, an
// FP32 bilinear filter in u8 units with clamp-to-edge, f16 rounding to nearest even.
//
// Work mapping: one thread per pixel (per active colour pixel in `propagate`), blocks of
// tx * ty threads; every view of a batch owns a contiguous range of logical blocks, and the
// hardware block index is remapped so that consecutive logical blocks (neighbouring tiles of
// one view, which gather from the same source-image regions) land on the same XCD and share
// its L2.  The same kernels serve the per-view ("plain") path and the batched path: only the
// launch plan differs.
#include "lr3d_internal.h"

#include <math.h>

namespace lr3d {

namespace {

constexpr float kInvalid = 2.0f;        // PM:41
constexpr float kGeometricCap = 3.0f;   // PM:42
constexpr uint32_t kFilterCost = 7u;    // PM:43
constexpr float kFilterPixels = 1.0f;   // PM:44
constexpr float kFilterDepth = 0.01f;   // PM:45
// PM:39-40 and PM:154 in u8 units: 1 / (2 (0.06 * 255)^2), 1 / (2 * 5^2), 1e-5 * 255^2.
constexpr float kColorK = (float)(1.0 / (2.0 * (0.06 * 255.0) * (0.06 * 255.0)));
constexpr float kSpaceK = 0.02f;
constexpr float kDegenerateDen = (float)(1e-5 * 255.0 * 255.0);

// ---- RNG (PM:56-65, PM:251-253) --------------------------------------------------------------

__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ float next_random(uint32_t &rng) {
    rng = pcg(rng);
    return (float)(rng >> 8u) / 16777216.0f;
}

// ---- small helpers -------------------------------------------------------------------------

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

__device__ __forceinline__ void normalize3(float &x, float &y, float &z) {
    float inv = 1.0f / sqrtf(dot3(x, y, z, x, y, z));
    x = x * inv;
    y = y * inv;
    z = z * inv;
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

__device__ __forceinline__ uint32_t f32_to_f16_bits(float v) {
    _Float16 h = (_Float16)v;  // round to nearest even
    return (uint32_t)__builtin_bit_cast(uint16_t, h);
}

__device__ __forceinline__ float f16_bits_to_f32(uint32_t bits) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(bits & 0xffffu));
}

// PK:7-15
__device__ __forceinline__ void unpack_normal(uint32_t packed, float &nx, float &ny, float &nz) {
    float ex = ((float)((packed >> 16u) & 63u) / 63.0f) * 2.0f - 1.0f;
    float ey = ((float)((packed >> 22u) & 63u) / 63.0f) * 2.0f - 1.0f;
    nx = ex;
    ny = ey;
    nz = (1.0f - fabsf(ex)) - fabsf(ey);
    if (nz < 0.0f) {
        float sx = nx >= 0.0f ? 1.0f : -1.0f;
        float sy = ny >= 0.0f ? 1.0f : -1.0f;
        float ux = (1.0f - fabsf(ny)) * sx;
        float uy = (1.0f - fabsf(nx)) * sy;
        nx = ux;
        ny = uy;
    }
    normalize3(nx, ny, nz);
}

// PK:21-33
__device__ __forceinline__ uint32_t pack_sample(float depth, float nx, float ny, float nz, float cost) {
    if (!(depth > 0.0f) || cost >= 1.0f) return 0u;
    float l1 = (fabsf(nx) + fabsf(ny)) + fabsf(nz);
    float ex = nx / l1;
    float ey = ny / l1;
    if (nz < 0.0f) {
        float sx = ex >= 0.0f ? 1.0f : -1.0f;
        float sy = ey >= 0.0f ? 1.0f : -1.0f;
        float ux = (1.0f - fabsf(ey)) * sx;
        float uy = (1.0f - fabsf(ex)) * sy;
        ex = ux;
        ey = uy;
    }
    uint32_t qx = (uint32_t)(clampf(ex * 0.5f + 0.5f, 0.0f, 1.0f) * 63.0f + 0.5f);
    uint32_t qy = (uint32_t)(clampf(ey * 0.5f + 0.5f, 0.0f, 1.0f) * 63.0f + 0.5f);
    uint32_t c = (uint32_t)(cost * 15.0f);
    c = c > 15u ? 15u : c;
    return f32_to_f16_bits(depth) | (qx << 16u) | (qy << 22u) | (c << 28u);
}

// ---- work mapping ---------------------------------------------------------------------------

// Hardware block i runs on XCD i % 8; give each XCD a contiguous range of logical blocks.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    uint32_t xcd = i % kXcds, local = i / kXcds;
    uint32_t per = n / kXcds, rem = n % kXcds;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// View owning logical block b (uniform: block_start is read through the scalar cache).
__device__ __forceinline__ int32_t find_view(const uint32_t *__restrict__ block_start, int32_t n,
                                             uint32_t b) {
    int32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int32_t mid = (lo + hi) >> 1;
        if (block_start[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// ---- bilateral window (PM:76-110) -------------------------------------------------------------

struct Window {
    float v[kSamples];
    float w[kSamples];
    float sw, mean, var;
};

__device__ __forceinline__ float ref_pixel(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                           int32_t x, int32_t y) {
    x = min(max(x, 0), W - 1);
    y = min(max(y, 0), H - 1);
    return (float)(img[__umul24((uint32_t)y, (uint32_t)W) + (uint32_t)x] & 0xffu);
}

__device__ __forceinline__ void prepare_window(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                               int32_t x, int32_t y, Window &win) {
    float center = ref_pixel(img, W, H, x, y);
    float sw = 0.0f, sr = 0.0f, srr = 0.0f;
#pragma unroll
    for (int i = 0; i < kSamples; i++) {
        const int ox = (i % 6) * 2 - 5, oy = (i / 6) * 2 - 5;
        float v = ref_pixel(img, W, H, x + ox, y + oy);
        float dc = v - center;
        const float ds = (float)(ox * ox + oy * oy);
        float w = (float)exp((double)(-(dc * dc) * kColorK - ds * kSpaceK));
        win.v[i] = v;
        win.w[i] = w;
        float wv = w * v;
        sw = sw + w;
        sr = sr + wv;
        srr = srr + wv * v;
    }
    float mean = sr / sw;
    win.sw = sw;
    win.mean = mean;
    win.var = fmaxf(srr / sw - mean * mean, 0.0f);
}

// ---- plane homography (PM:112-128) -------------------------------------------------------------

struct Mat3 {
    float m[9];
};

__device__ __forceinline__ void homography(const PmDesc &d, const PmSrc &s, float nx, float ny,
                                           float nz, float rho, Mat3 &H) {
    const float n[3] = {nx, ny, nz};
    const float t[3] = {s.t[0] / rho, s.t[1] / rho, s.t[2] / rho};
    float M[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M[i * 3 + j] = s.R[i * 3 + j] + t[i] * n[j];
    float A[9];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        A[j] = s.k[0] * M[j] + s.k[2] * M[6 + j];
        A[3 + j] = s.k[1] * M[3 + j] + s.k[3] * M[6 + j];
        A[6 + j] = M[6 + j];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        H.m[i * 3 + 0] = A[i * 3 + 0] * d.ikx;
        H.m[i * 3 + 1] = A[i * 3 + 1] * d.iky;
        H.m[i * 3 + 2] = (A[i * 3 + 0] * d.kx2 + A[i * 3 + 1] * d.ky2) + A[i * 3 + 2];
    }
}

// ---- bilateral NCC of one source view (PM:130-159) ---------------------------------------------
//
// Tap i at offset (ox, oy) maps to q = H (X, Y, 1) with X = px + ox, Y = py + oy, evaluated as
// q = fma(H_0, X, fma(H_1, Y, H_2)) per row, so the Y part is shared by the six taps of a window
// row.  All 36 taps are evaluated; the view is INVALID if any tap is (the reference returns at
// the first such tap, with the same result).  Sample indices are clamped into the image so that
// invalid taps still read inside the source image.
//
// 1/q.z: v_rcp_f32 followed by one Newton step with FMA reproduces the IEEE quotient 1.0f / q.z
// bit for bit for every |q.z| in [2^-60, 2^61) (checked exhaustively on gfx950).  Valid taps
// have q.z > 1e-6 > 2^-60; a view whose window could reach |q.z| >= 2^60 takes the IEEE
// division for the whole wavefront.
__device__ __forceinline__ float rcp_exact(float z) {
    const float r = __builtin_amdgcn_rcpf(z);
    return fmaf(fmaf(-z, r, 1.0f), r, r);
}

template <bool FAST>
__device__ __forceinline__ void view_taps(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                          const Mat3 &Hm, float px, float py, const Window &win,
                                          float &ss, float &sss, float &srs, bool &ok) {
    const float wmax = (float)W - 1.0f, hmax = (float)H - 1.0f;
    const uint32_t stride = (uint32_t)W;
#pragma unroll
    for (int row = 0; row < 6; row++) {
        const float Y = py + (float)(row * 2 - 5);
        const float rx = fmaf(Hm.m[1], Y, Hm.m[2]);
        const float ry = fmaf(Hm.m[4], Y, Hm.m[5]);
        const float rw = fmaf(Hm.m[7], Y, Hm.m[8]);
#pragma unroll
        for (int col = 0; col < 6; col++) {
            const int i = row * 6 + col;
            const float X = px + (float)(col * 2 - 5);
            const float qx = fmaf(Hm.m[0], X, rx);
            const float qy = fmaf(Hm.m[3], X, ry);
            const float qz = fmaf(Hm.m[6], X, rw);
            const float rz = FAST ? rcp_exact(qz) : 1.0f / qz;
            const float u = qx * rz, v = qy * rz;
            // Clamp into the image (NaN-safe: maxnum returns the number); the tap is inside
            // exactly when the clamp leaves it unchanged (a NaN never compares equal).  Valid
            // taps have 0 <= u <= W-1, so trunc(u) = floor(u) and fract(u) = u - floor(u).
            const float uc = fminf(fmaxf(u, 0.0f), wmax), vc = fminf(fmaxf(v, 0.0f), hmax);
            ok = ok & (qz > 1e-6f) & (uc == u) & (vc == v);
            const int32_t ix = (int32_t)uc;
            const int32_t iy = (int32_t)vc;
            const float fx = __builtin_amdgcn_fractf(u), fy = __builtin_amdgcn_fractf(v);
            const uint32_t q = img[__umul24((uint32_t)iy, stride) + (uint32_t)ix];
            const float a = (float)(q & 0xffu), b = (float)((q >> 8) & 0xffu);
            const float c = (float)((q >> 16) & 0xffu), e = (float)(q >> 24);
            const float top = fmaf(fx, b - a, a);
            const float bot = fmaf(fx, e - c, c);
            const float val = fmaf(fy, bot - top, top);
            const float wv = win.w[i] * val;
            ss = ss + wv;
            sss = fmaf(wv, val, sss);
            srs = fmaf(wv, win.v[i], srs);
        }
    }
}

// Taps of a window known to lie inside the source image (see view_cost): no per-tap checks.
// The linear index is still clamped to the image, so no tap can read outside it.
__device__ __forceinline__ void view_taps_inside(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                                 const Mat3 &Hm, float px, float py,
                                                 const Window &win, float &ss, float &sss,
                                                 float &srs) {
    const uint32_t stride = (uint32_t)W;
    const uint32_t last = (uint32_t)W * (uint32_t)H - 1u;
#pragma unroll
    for (int row = 0; row < 6; row++) {
        const float Y = py + (float)(row * 2 - 5);
        const float rx = fmaf(Hm.m[1], Y, Hm.m[2]);
        const float ry = fmaf(Hm.m[4], Y, Hm.m[5]);
        const float rw = fmaf(Hm.m[7], Y, Hm.m[8]);
#pragma unroll
        for (int col = 0; col < 6; col++) {
            const int i = row * 6 + col;
            const float X = px + (float)(col * 2 - 5);
            const float qx = fmaf(Hm.m[0], X, rx);
            const float qy = fmaf(Hm.m[3], X, ry);
            const float qz = fmaf(Hm.m[6], X, rw);
            const float rz = rcp_exact(qz);
            const float u = qx * rz, v = qy * rz;
            const uint32_t ix = (uint32_t)u, iy = (uint32_t)v;
            const float fx = __builtin_amdgcn_fractf(u), fy = __builtin_amdgcn_fractf(v);
            const uint32_t q = img[min(__umul24(iy, stride) + ix, last)];
            const float a = (float)(q & 0xffu), b = (float)((q >> 8) & 0xffu);
            const float c = (float)((q >> 16) & 0xffu), e = (float)(q >> 24);
            const float top = fmaf(fx, b - a, a);
            const float bot = fmaf(fx, e - c, c);
            const float val = fmaf(fy, bot - top, top);
            const float wv = win.w[i] * val;
            ss = ss + wv;
            sss = fmaf(wv, val, sss);
            srs = fmaf(wv, win.v[i], srs);
        }
    }
}

// Window classes of one source view, decided from the four corner taps (offsets (+-5, +-5)):
//   kOutside  a corner tap is invalid, so the view is INVALID (the reference stops at the first
//             invalid tap; every tap, corners included, is evaluated exactly as in the full loop);
//   kInside   every corner lies inside the image by more than a rounding margin: q.z is affine in
//             (X, Y) and positive on the window, so the warped window is the convex hull of the
//             warped corners, and every tap is inside as computed;
//   kCareful  anything else (windows touching the border, huge |q.z|): per-tap checks.
enum { kOutside = 0, kInside = 1, kCareful = 2 };

__device__ __forceinline__ int classify_window(const PmSrc &s, const Mat3 &H, float px, float py) {
    const float wmax = (float)s.W - 1.0f, hmax = (float)s.H - 1.0f;
    const float ax = fabsf(px) + 5.0f, ay = fabsf(py) + 5.0f;
    // bounds on the magnitude of the terms of q.x, q.y, q.z over the window
    const float bx = fabsf(H.m[0]) * ax + fabsf(H.m[1]) * ay + fabsf(H.m[2]);
    const float by = fabsf(H.m[3]) * ax + fabsf(H.m[4]) * ay + fabsf(H.m[5]);
    const float bz = fabsf(H.m[6]) * ax + fabsf(H.m[7]) * ay + fabsf(H.m[8]);
    if (!(bz < 0x1p60f)) return kCareful;  // outside the range where rcp_exact == 1/q.z
    bool valid = true;
    float zmin = INFINITY, umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
#pragma unroll
    for (int corner = 0; corner < 4; corner++) {
        const float X = px + ((corner & 1) ? 5.0f : -5.0f);
        const float Y = py + ((corner & 2) ? 5.0f : -5.0f);
        const float qx = fmaf(H.m[0], X, fmaf(H.m[1], Y, H.m[2]));
        const float qy = fmaf(H.m[3], X, fmaf(H.m[4], Y, H.m[5]));
        const float qz = fmaf(H.m[6], X, fmaf(H.m[7], Y, H.m[8]));
        const float rz = rcp_exact(qz);
        const float u = qx * rz, v = qy * rz;
        valid = valid & (qz > 1e-6f) & (u >= 0.0f) & (v >= 0.0f) & (u <= wmax) & (v <= hmax);
        zmin = fminf(zmin, qz);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
    }
    if (!valid) return kOutside;
    // Each computed q component is within 2^-20 of its term bound of the exact value (two FMAs
    // round at most 2 ulp of the larger term), so exact q.z >= zmin - ez on the window.  A
    // computed u = q.x / q.z then lies within eu of the exact ratio; require corners inside
    // by 4 eu so that exact and computed interior taps stay inside.
    const float ez = bz * 0x1p-20f;
    const float zlow = zmin - 2.0f * ez;
    if (!(zlow > 2e-6f)) return kCareful;
    const float ilow = 1.0f / zlow;
    const float eu = (bx * 0x1p-20f + (wmax + 1.0f) * ez) * ilow + (wmax + 1.0f) * 0x1p-20f;
    const float ev = (by * 0x1p-20f + (hmax + 1.0f) * ez) * ilow + (hmax + 1.0f) * 0x1p-20f;
    const bool inside = umin >= 4.0f * eu && umax <= wmax - 4.0f * eu && vmin >= 4.0f * ev &&
                        vmax <= hmax - 4.0f * ev;
    return inside ? kInside : kCareful;
}

__device__ __forceinline__ float view_cost(const PmSrc &s, const Mat3 &H, float px, float py,
                                           const Window &win, const uint32_t *__restrict__ quad) {
    const uint32_t *__restrict__ img = quad + s.off;
    const int cls = classify_window(s, H, px, py);
    float ss = 0.0f, sss = 0.0f, srs = 0.0f;
    bool ok = cls != kOutside;
    if (__any(cls == kCareful)) {
        // the checked loop is exact for every class; lanes already known outside stay invalid
        const float bound = fabsf(H.m[6]) * (fabsf(px) + 5.0f) + fabsf(H.m[7]) * (fabsf(py) + 5.0f) +
                            fabsf(H.m[8]);
        bool tap_ok = true;
        if (__all(bound < 0x1p60f))
            view_taps<true>(img, s.W, s.H, H, px, py, win, ss, sss, srs, tap_ok);
        else
            view_taps<false>(img, s.W, s.H, H, px, py, win, ss, sss, srs, tap_ok);
        ok = ok & tap_ok;
    } else if (__any(cls == kInside)) {
        if (cls == kInside) view_taps_inside(img, s.W, s.H, H, px, py, win, ss, sss, srs);
    }
    if (!ok) return kInvalid;
    float mean = ss / win.sw;
    float var_s = fmaxf(sss / win.sw - mean * mean, 0.0f);
    float den = sqrtf(win.var * var_s);
    if (den < kDegenerateDen) return 1.0f;
    float cov = srs / win.sw - win.mean * mean;
    return clampf(1.0f - cov / den, 0.0f, 2.0f);
}

// PM:161-190: forward-backward reprojection through a source view's packed depth map.
// Returns the pixel distance in r0 and the back-projected depth in r1; false when invalid.
__device__ __forceinline__ bool reproject(const PmDesc &d, const PmSrc &s,
                                          const uint32_t *__restrict__ depth, float px, float py,
                                          float x0, float x1, float x2, float &r0, float &r1) {
    float xs0 = fmaf(s.R[2], x2, fmaf(s.R[1], x1, fmaf(s.R[0], x0, s.t[0])));
    float xs1 = fmaf(s.R[5], x2, fmaf(s.R[4], x1, fmaf(s.R[3], x0, s.t[1])));
    float xs2 = fmaf(s.R[8], x2, fmaf(s.R[7], x1, fmaf(s.R[6], x0, s.t[2])));
    if (!(xs2 > 1e-6f)) return false;
    int32_t pxs = (int32_t)rintf((s.k[0] * xs0) / xs2 + s.k[2]);
    int32_t pys = (int32_t)rintf((s.k[1] * xs1) / xs2 + s.k[3]);
    if (pxs < 0 || pys < 0 || pxs >= s.W || pys >= s.H) return false;
    uint32_t packed = depth[s.off + (uint32_t)pys * (uint32_t)s.W + (uint32_t)pxs];
    float ds = f16_bits_to_f32(packed);
    if (!(ds > 0.0f) || (packed >> 28u) > kFilterCost) return false;
    float b0 = (((float)pxs - s.k[2]) / s.k[0]) * ds;
    float b1 = (((float)pys - s.k[3]) / s.k[1]) * ds;
    float e0 = b0 - s.t[0], e1 = b1 - s.t[1], e2 = ds - s.t[2];
    float back0 = fmaf(s.R[6], e2, fmaf(s.R[3], e1, s.R[0] * e0));
    float back1 = fmaf(s.R[7], e2, fmaf(s.R[4], e1, s.R[1] * e0));
    float back2 = fmaf(s.R[8], e2, fmaf(s.R[5], e1, s.R[2] * e0));
    if (!(back2 > 1e-6f)) return false;
    float qx = (d.k[0] * back0) / back2 + d.k[2];
    float qy = (d.k[1] * back1) / back2 + d.k[3];
    float dx = qx - px, dy = qy - py;
    r0 = sqrtf(dx * dx + dy * dy);
    r1 = back2;
    return true;
}

// Sorted k smallest of the view costs.  The reference's partial selection sort (PM:218-231)
// adds the k smallest costs in ascending order; keeping them sorted in registers with a
// min/max insertion network gives the same values in the same order.
template <int KMAX>
struct TopK {
    float s[KMAX];
    __device__ __forceinline__ void reset() {
#pragma unroll
        for (int j = 0; j < KMAX; j++) s[j] = INFINITY;
    }
    __device__ __forceinline__ void insert(float c) {
#pragma unroll
        for (int j = KMAX - 1; j > 0; j--) s[j] = fminf(s[j], fmaxf(s[j - 1], c));
        s[0] = fminf(s[0], c);
    }
    __device__ __forceinline__ float mean(int32_t k) const {
        float total = 0.0f;
#pragma unroll
        for (int j = 0; j < KMAX; j++)
            if (j < k) total = total + s[j];
        return total / (float)k;
    }
};

// PM:192-233
template <int KMAX, bool GEOMETRIC>
__device__ __forceinline__ float hypothesis_cost(const PmDesc &d, const PmPools &pools, float px,
                                                 float py, float rx, float ry, const Window &win,
                                                 float nx, float ny, float nz, float depth) {
    float inv = 1.0f / depth;
    if (!(depth > 0.0f) || inv > d.inv_min || inv < d.inv_max) return kInvalid;
    if (!(dot3(nx, ny, nz, rx, ry, 1.0f) < -1e-3f)) return kInvalid;
    float X0 = rx * depth, X1 = ry * depth, X2 = 1.0f * depth;
    float rho = dot3(nx, ny, nz, X0, X1, X2);
    const int32_t count = d.n_src;
    if (count == 0) return kInvalid;
    TopK<KMAX> top;
    top.reset();
    for (int32_t s = 0; s < count; s++) {
        const PmSrc &src = d.src[s];
        Mat3 H;
        homography(d, src, nx, ny, nz, rho, H);
        float c = view_cost(src, H, px, py, win, pools.quad);
        if (GEOMETRIC && (d.flags & 1u) != 0u) {
            float r0 = 0.0f, r1 = 0.0f;
            bool valid = reproject(d, src, pools.depth, px, py, X0, X1, X2, r0, r1);
            float e = valid ? fminf(r0, kGeometricCap) : kGeometricCap;
            c = c + d.geo_weight * e;
        }
        top.insert(c);
    }
    return top.mean(d.topk);
}

// PM:235-245
__device__ __forceinline__ void random_normal(uint32_t &rng, float rx, float ry, float &nx,
                                              float &ny, float &nz) {
    float z = next_random(rng) * 2.0f - 1.0f;
    float a = next_random(rng) * 6.2831853f;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    nx = r * (float)cos((double)a);
    ny = r * (float)sin((double)a);
    nz = z;
    float vx = rx, vy = ry, vz = 1.0f;
    normalize3(vx, vy, vz);
    if (dot3(nx, ny, nz, vx, vy, vz) > 0.0f) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
    }
    nx = nx - vx * 0.5f;
    ny = ny - vy * 0.5f;
    nz = nz - vz * 0.5f;
    normalize3(nx, ny, nz);
}

// PM:247-249 with WGSL mix(a, b, t) = a * (1 - t) + b * t
__device__ __forceinline__ float random_depth(uint32_t &rng, const PmDesc &d) {
    float t = next_random(rng);
    return 1.0f / (d.inv_max * (1.0f - t) + d.inv_min * t);
}

__device__ __forceinline__ uint32_t pixel_rng(const PmDesc &d, uint32_t iteration, uint32_t phase,
                                              uint32_t index) {
    uint32_t seed = d.seed_base + iteration * 97u + phase * 13u;  // DS:379
    return pcg(index ^ pcg(seed + iteration * 7919u + phase * 104729u));  // PM:251-253
}

// ---- kernels -------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) build_quad_kernel(const uint8_t *__restrict__ gray, int32_t W,
                                                         int32_t H, uint32_t *__restrict__ quad) {
    int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    int32_t x1 = min(x + 1, W - 1), y1 = min(y + 1, H - 1);
    uint32_t a = gray[(size_t)y * W + x], b = gray[(size_t)y * W + x1];
    uint32_t c = gray[(size_t)y1 * W + x], e = gray[(size_t)y1 * W + x1];
    quad[(size_t)y * W + x] = a | (b << 8) | (c << 16) | (e << 24);
}

// Locate this thread's pixel: tile of (tx * cols_per_thread) x ty pixels.
struct PixelTask {
    int32_t view;
    int32_t gx, gy;  // thread column / row inside the view
};

__device__ __forceinline__ PixelTask locate(const PmPlan &plan, int32_t tile_w) {
    uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    PixelTask t;
    t.view = find_view(plan.block_start, plan.n, b);
    const PmDesc &d = plan.descs[t.view];
    uint32_t local = b - plan.block_start[t.view];
    int32_t tiles_x = (d.W + tile_w - 1) / tile_w;
    int32_t bx = (int32_t)(local % (uint32_t)tiles_x), by = (int32_t)(local / (uint32_t)tiles_x);
    t.gx = bx * (int32_t)blockDim.x + (int32_t)threadIdx.x;
    t.gy = by * (int32_t)blockDim.y + (int32_t)threadIdx.y;
    return t;
}

// PM:255-289 (init) and PM:296-333 (init_coarse): iteration 0, phase 0 (DS:238-247)
template <int KMAX, bool GEOMETRIC>
__global__ void __launch_bounds__(256) pm_init_kernel(PmPlan plan, PmPools pools) {
    PixelTask task = locate(plan, (int32_t)blockDim.x);
    const PmDesc &d = plan.descs[task.view];
    const int32_t x = task.gx, y = task.gy;
    if (x >= d.W || y >= d.H) return;
    const uint32_t index = (uint32_t)y * (uint32_t)d.W + (uint32_t)x;
    const float px = (float)x, py = (float)y;
    uint32_t rng = pixel_rng(d, 0u, 0u, index);
    Window win;
    prepare_window(pools.quad + d.off, d.W, d.H, x, y, win);
    const float rx = (px - d.k[2]) / d.k[0];
    const float ry = (py - d.k[3]) / d.k[1];
    float nx, ny, nz;
    random_normal(rng, rx, ry, nx, ny, nz);
    float depth = random_depth(rng, d);
    float c;
    if (d.from_coarse) {
        // PM:307-330
        c = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, nx, ny, nz, depth);
        const float qx = ((px + 0.5f) * (float)d.coarse_W) / (float)d.W - 0.5f;
        const float qy = ((py + 0.5f) * (float)d.coarse_H) / (float)d.H - 0.5f;
        const int32_t bx = (int32_t)floorf(qx), by = (int32_t)floorf(qy);
        for (int corner = 0; corner < 4; corner++) {
            int32_t cx = min(max(bx + (corner & 1), 0), d.coarse_W - 1);
            int32_t cy = min(max(by + (corner >> 1), 0), d.coarse_H - 1);
            float4 h = pools.coarse_hyp[d.coarse_off + (uint32_t)cy * (uint32_t)d.coarse_W + (uint32_t)cx];
            if (!(h.w > 0.0f)) continue;
            float p0 = (((float)cx - d.ck[2]) / d.ck[0]) * h.w;
            float p1 = (((float)cy - d.ck[3]) / d.ck[1]) * h.w;
            float p2 = h.w;
            float den = dot3(h.x, h.y, h.z, rx, ry, 1.0f);
            if (fabsf(den) < 1e-6f) continue;
            float cd = dot3(h.x, h.y, h.z, p0, p1, p2) / den;
            float cand = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, h.x, h.y,
                                                          h.z, cd);
            if (cand < c) {
                c = cand;
                nx = h.x;
                ny = h.y;
                nz = h.z;
                depth = cd;
            }
        }
    } else {
        if (GEOMETRIC && (d.flags & 1u) != 0u) {
            uint32_t packed = pools.depth[d.off + index];
            float pd = f16_bits_to_f32(packed);
            if (pd > 0.0f) {
                depth = pd;
                unpack_normal(packed, nx, ny, nz);
            }
        }
        c = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, nx, ny, nz, depth);
        if ((d.flags & 1u) == 0u && d.has_prior) {
            float pr = pools.prior[d.prior_off + index];
            if (pr > 0.0f) {
                float pnx = rx, pny = ry, pnz = 1.0f;
                normalize3(pnx, pny, pnz);
                pnx = -pnx;
                pny = -pny;
                pnz = -pnz;
                float pc = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, pnx, pny,
                                                            pnz, pr);
                if (pc < c) {
                    c = pc;
                    nx = pnx;
                    ny = pny;
                    nz = pnz;
                    depth = pr;
                }
            }
        }
    }
    pools.hyp[d.off + index] = make_float4(nx, ny, nz, depth);
    pools.cost[d.off + index] = c;
}

// PM:335-347: neighbour offset i of region r.
__device__ __forceinline__ void region_offset(int region, int i, int &ox, int &oy) {
    int bx, by;
    if (region < 4) {
        int m = (i + 1) / 2;  // near_offset
        bx = (i & 1) == 1 ? -m : m;
        by = -(m + 1);
    } else {
        bx = 0;
        by = -3 - 2 * i;
    }
    switch (region & 3) {  // orient
    case 1: ox = bx; oy = -by; break;
    case 2: ox = by; oy = bx; break;
    case 3: ox = -by; oy = bx; break;
    default: ox = bx; oy = by; break;
    }
}

// PM:349-423: one red-black phase.
#ifndef LR3D_PM_WAVES_PER_EU
#define LR3D_PM_WAVES_PER_EU 1
#endif
template <int KMAX, bool GEOMETRIC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LR3D_PM_WAVES_PER_EU, 8)))
pm_propagate_kernel(PmPlan plan, PmPools pools,
                                                           uint32_t iteration, uint32_t phase,
                                                           float step) {
    PixelTask task = locate(plan, 2 * (int32_t)blockDim.x);
    const PmDesc &d = plan.descs[task.view];
    if ((int32_t)iteration >= d.iterations) return;
    const int32_t y = task.gy;
    const int32_t x = task.gx * 2 + (int32_t)(((uint32_t)y + phase) & 1u);
    if (x >= d.W || y >= d.H) return;
    const int32_t W = d.W, H = d.H;
    const uint32_t index = (uint32_t)y * (uint32_t)W + (uint32_t)x;
    const float px = (float)x, py = (float)y;
    uint32_t rng = pixel_rng(d, iteration, phase, index);
    Window win;
    prepare_window(pools.quad + d.off, W, H, x, y, win);
    const float rx = (px - d.k[2]) / d.k[0];
    const float ry = (py - d.k[3]) / d.k[1];
    const float *__restrict__ cost = pools.cost + d.off;
    const float4 *__restrict__ hyp = pools.hyp + d.off;
    float4 best = hyp[index];
    float best_cost = cost[index];

    // Fourteen candidates in the reference's order, through one evaluation site: the plane of
    // the cheapest neighbour of each of the 8 regions (PM:365-399), then the 6 random /
    // perturbed planes (PM:401-420), drawn once the regions are done.
    float rnx = 0.0f, rny = 0.0f, rnz = 0.0f, pnx = 0.0f, pny = 0.0f, pnz = 0.0f;
    float cnx = 0.0f, cny = 0.0f, cnz = 0.0f, cdepth = 0.0f, d_random = 0.0f, d_perturbed = 0.0f;
#pragma unroll 1
    for (int c = 0; c < 14; c++) {
        float tx, ty, tz, td;
        if (c < 8) {
            const int region = c;
            const int count = region < 4 ? 7 : 11;
            int32_t chosen = -1;
            float chosen_cost = 1e9f;
            for (int i = 0; i < count; i++) {
                int ox, oy;
                region_offset(region, i, ox, oy);
                const int32_t qx = x + ox, qy = y + oy;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                const int32_t qi = qy * W + qx;
                const float qc = cost[qi];
                if (qc < chosen_cost) {
                    chosen_cost = qc;
                    chosen = qi;
                }
            }
            if (chosen < 0) continue;
            const float4 h = hyp[chosen];
            const float qrx = ((float)(chosen % W) - d.k[2]) / d.k[0];
            const float qry = ((float)(chosen / W) - d.k[3]) / d.k[1];
            const float den = dot3(h.x, h.y, h.z, rx, ry, 1.0f);
            if (fabsf(den) < 1e-6f) continue;
            tx = h.x;
            ty = h.y;
            tz = h.z;
            td = dot3(h.x, h.y, h.z, qrx * h.w, qry * h.w, 1.0f * h.w) / den;
        } else {
            if (c == 8) {
                float vx = rx, vy = ry, vz = 1.0f;
                normalize3(vx, vy, vz);
                random_normal(rng, rx, ry, rnx, rny, rnz);
                d_random = random_depth(rng, d);
                d_perturbed = 1.0f / ((1.0f / best.w) *
                                      (1.0f + ((next_random(rng) * 2.0f - 1.0f) * 0.15f) * step));
                const float r1 = next_random(rng), r2 = next_random(rng), r3 = next_random(rng);
                pnx = best.x + ((r1 * 2.0f - 1.0f) * 0.6f) * step;
                pny = best.y + ((r2 * 2.0f - 1.0f) * 0.6f) * step;
                pnz = best.z + ((r3 * 2.0f - 1.0f) * 0.6f) * step;
                normalize3(pnx, pny, pnz);
                if (!(dot3(pnx, pny, pnz, vx, vy, vz) < 0.0f)) {
                    pnx = best.x;
                    pny = best.y;
                    pnz = best.z;
                }
                cnx = best.x;
                cny = best.y;
                cnz = best.z;
                cdepth = best.w;
            }
            // (n_r,d_r) (n_p,d_p) (n,d_r) (n_r,d) (n,d_p) (n_p,d)
            const int k = c - 8;
            const bool nr = k == 0 || k == 3, np = k == 1 || k == 5;
            tx = nr ? rnx : (np ? pnx : cnx);
            ty = nr ? rny : (np ? pny : cny);
            tz = nr ? rnz : (np ? pnz : cnz);
            td = (k == 0 || k == 2) ? d_random : ((k == 1 || k == 4) ? d_perturbed : cdepth);
        }
        const float cand = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, tx, ty,
                                                            tz, td);
        if (cand < best_cost) {
            best_cost = cand;
            best = make_float4(tx, ty, tz, td);
        }
    }
    pools.hyp[d.off + index] = best;
    pools.cost[d.off + index] = best_cost;
}

// PM:425-433
__global__ void __launch_bounds__(256) pm_pack_kernel(PmPlan plan, PmPools pools) {
    PixelTask task = locate(plan, (int32_t)blockDim.x);
    const PmDesc &d = plan.descs[task.view];
    if (task.gx >= d.W || task.gy >= d.H) return;
    const uint32_t i = d.off + (uint32_t)task.gy * (uint32_t)d.W + (uint32_t)task.gx;
    const float4 h = pools.hyp[i];
    pools.depth[i] = pack_sample(h.w, h.x, h.y, h.z, pools.cost[i]);
}

// PM:436-469
__global__ void __launch_bounds__(256) pm_consistency_kernel(PmPlan plan, PmPools pools,
                                                             uint32_t needed) {
    PixelTask task = locate(plan, (int32_t)blockDim.x);
    const PmDesc &d = plan.descs[task.view];
    const int32_t x = task.gx, y = task.gy;
    if (x >= d.W || y >= d.H) return;
    const uint32_t index = d.off + (uint32_t)y * (uint32_t)d.W + (uint32_t)x;
    const uint32_t packed = pools.depth[index];
    const float depth = f16_bits_to_f32(packed);
    if (!(depth > 0.0f) || (packed >> 28u) > kFilterCost) {
        pools.filt[index] = 0u;
        return;
    }
    const float px = (float)x, py = (float)y;
    const float rx = (px - d.k[2]) / d.k[0];
    const float ry = (py - d.k[3]) / d.k[1];
    const float X0 = rx * depth, X1 = ry * depth, X2 = 1.0f * depth;
    uint32_t agree = 0;
    float sum = depth;
    for (int32_t s = 0; s < d.n_src; s++) {
        float r0 = 0.0f, r1 = 0.0f;
        if (reproject(d, d.src[s], pools.depth, px, py, X0, X1, X2, r0, r1) && r0 < kFilterPixels &&
            fabsf(r1 - depth) < kFilterDepth * depth) {
            agree++;
            sum = sum + r1;
        }
    }
    uint32_t need = needed > 1u ? needed : 1u;
    need = (uint32_t)d.n_src < need ? (uint32_t)d.n_src : need;
    if (agree < need) {
        pools.filt[index] = 0u;
        return;
    }
    const float mean = sum / (float)(agree + 1u);
    pools.filt[index] = (packed & 0xffff0000u) | f32_to_f16_bits(mean);
}

}  // namespace

// ---- launchers ------------------------------------------------------------------------------------

hipError_t pm_build_quad(hipStream_t s, const uint8_t *gray, int32_t W, int32_t H, uint32_t *quad) {
    dim3 block(64, 4);
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(build_quad_kernel, grid, block, 0, s, gray, W, H, quad);
    return hipGetLastError();
}

template <int KMAX>
static hipError_t launch_init_k(hipStream_t s, const PmPlan &plan, const PmPools &pools, bool geo) {
    dim3 block(plan.tx, plan.ty), grid(plan.blocks);
    if (geo)
        hipLaunchKernelGGL((pm_init_kernel<KMAX, true>), grid, block, 0, s, plan, pools);
    else
        hipLaunchKernelGGL((pm_init_kernel<KMAX, false>), grid, block, 0, s, plan, pools);
    return hipGetLastError();
}

template <int KMAX>
static hipError_t launch_propagate_k(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                                     bool geo, uint32_t it, uint32_t phase, float step) {
    dim3 block(plan.tx, plan.ty), grid(plan.blocks);
    if (geo)
        hipLaunchKernelGGL((pm_propagate_kernel<KMAX, true>), grid, block, 0, s, plan, pools, it,
                           phase, step);
    else
        hipLaunchKernelGGL((pm_propagate_kernel<KMAX, false>), grid, block, 0, s, plan, pools, it,
                           phase, step);
    return hipGetLastError();
}

hipError_t pm_launch_init(hipStream_t s, const PmPlan &plan, const PmPools &pools) {
    const bool geo = plan.geometric != 0;
    return plan.kmax <= 3 ? launch_init_k<3>(s, plan, pools, geo) : launch_init_k<8>(s, plan, pools, geo);
}

hipError_t pm_launch_propagate(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                               uint32_t iteration, uint32_t phase, float step) {
    const bool geo = plan.geometric != 0;
    return plan.kmax <= 3 ? launch_propagate_k<3>(s, plan, pools, geo, iteration, phase, step)
                          : launch_propagate_k<8>(s, plan, pools, geo, iteration, phase, step);
}

hipError_t pm_launch_pack(hipStream_t s, const PmPlan &plan, const PmPools &pools) {
    dim3 block(plan.tx, plan.ty), grid(plan.blocks);
    hipLaunchKernelGGL(pm_pack_kernel, grid, block, 0, s, plan, pools);
    return hipGetLastError();
}

hipError_t pm_launch_consistency(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                                 uint32_t needed) {
    dim3 block(plan.tx, plan.ty), grid(plan.blocks);
    hipLaunchKernelGGL(pm_consistency_kernel, grid, block, 0, s, plan, pools, needed);
    return hipGetLastError();
}

}  // namespace lr3d
