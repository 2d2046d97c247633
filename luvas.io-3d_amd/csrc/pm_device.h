// Device-side building blocks of the PatchMatch kernels (forward and adjoint): RNG, packing,
// bilateral window, plane homography, bilateral NCC, hypothesis cost.
//
// Semantics follow the reference WGSL module (frontend/src/app/vision/mvs/patchmatch.wgsl.ts,
// cited PM:<line>) under the FP32 arithmetic definition stated at the top of pm_kernels.hip.
// Included by one translation unit per kernel family; everything is internal to it.
#pragma once

#include "lr3d_internal.h"
#include "fp_exact.h"

#include <math.h>

// One 4-byte record of a structured buffer: the buffer load with the index register scaled by the
// descriptor's stride (the LLVM intrinsic behind it; clang has no builtin for it).
extern "C" __device__ uint32_t struct_buffer_load_u32(__amdgpu_buffer_rsrc_t rsrc, int32_t vindex, int32_t voffset,
                                           int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.i32");

namespace lr3d {
namespace {

constexpr float kInvalid = 2.0f;        // PM:41
constexpr float kGeometricCap = 3.0f;   // PM:42
constexpr uint32_t kFilterCost = 7u;    // PM:43
constexpr float kFilterPixels = 1.0f;   // PM:44
constexpr float kFilterDepth = 0.01f;   // PM:45
// PM:154 in u8 units: 1e-5 * 255^2 (the window's colour and space constants are in lr3d_internal.h).
constexpr float kDegenerateDen = (float)(1e-5 * 255.0 * 255.0);
// The smallest float x with sqrtf(x) >= kDegenerateDen (IEEE square root is monotone, so
// sqrtf(x) < kDegenerateDen exactly when x < kDegenerateSq).
constexpr float kDegenerateSq = 0x1.b0f90ep-2f;

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
    float isw;  // 1 / sw
};

// Texel (x, y) of a layer in the quad pool: column block x / 16, row y within it, column x % 16
// (see lr3d_internal.h).
__device__ __forceinline__ uint32_t quad_index(uint32_t x, uint32_t y, uint32_t qrows) {
    return ((__umul24(x >> 4, qrows) + y) << 4) | (x & 15u);
}

__device__ __forceinline__ float ref_pixel(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                           int32_t qrows, int32_t x, int32_t y) {
    x = min(max(x, 0), W - 1);
    y = min(max(y, 0), H - 1);
    return (float)(img[quad_index((uint32_t)x, (uint32_t)y, (uint32_t)qrows)] & 0xffu);
}

// wtab: the weight table (kWindowClasses x 256, see lr3d_internal.h): weight of tap i =
// wtab[class(i) * 256 + |dc|], the value exp(-(dc^2) kColorK - |o|^2 kSpaceK) rounded to f32.
__device__ __forceinline__ void prepare_window(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                               int32_t qrows, int32_t x, int32_t y,
                                               const float *__restrict__ wtab, Window &win) {
    float center = ref_pixel(img, W, H, qrows, x, y);
    float sw = 0.0f, sr = 0.0f, srr = 0.0f;
#pragma unroll
    for (int i = 0; i < kSamples; i++) {
        const int ox = (i % 6) * 2 - 5, oy = (i / 6) * 2 - 5;
        float v = ref_pixel(img, W, H, qrows, x + ox, y + oy);
        float dc = v - center;
        float w = wtab[window_class(ox, oy) * 256 + (int32_t)fabsf(dc)];
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
    win.isw = 1.0f / sw;
}

// ---- plane homography (PM:112-128) -------------------------------------------------------------
//
// H = Ks (R + t n^T / rho) Kr^-1 = G + (e / rho) m^T with G = Ks R Kr^-1 and e = Ks t per source
// (PmSrc, formed on the host) and m = Kr^-T n per hypothesis: the plane's part is formed once per
// hypothesis (PlaneTerms), each source view then costs 3 products and 9 FMAs.

struct Mat3 {
    float m[9];
};

struct PlaneTerms {
    float irho;        // 1 / rho
    float m0, m1, m2;  // Kr^-T n
};

__device__ __forceinline__ PlaneTerms plane_terms(const PmDesc &d, float nx, float ny, float nz, float rho) {
    PlaneTerms p;
    p.irho = 1.0f / rho;
    p.m0 = nx * d.ikx;
    p.m1 = ny * d.iky;
    p.m2 = (nx * d.kx2 + ny * d.ky2) + nz;
    return p;
}

__device__ __forceinline__ void homography(const PmSrc &s, const PlaneTerms &p, Mat3 &H) {
    const float m[3] = {p.m0, p.m1, p.m2};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float f = s.e[i] * p.irho;
#pragma unroll
        for (int j = 0; j < 3; j++) H.m[i * 3 + j] = fmaf(f, m[j], s.G[i * 3 + j]);
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
// 1/q.z: rcp_exact (fp_exact.h) is the IEEE quotient 1.0f / q.z for every |q.z| in
// [2^-60, 2^61).  Valid taps have q.z > 1e-6 > 2^-60; a view whose window could reach
// |q.z| >= 2^60 takes the IEEE division for the whole wavefront.

template <bool FAST>
__device__ __forceinline__ void view_taps(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                          int32_t qrows, const Mat3 &Hm, float px, float py,
                                          const Window &win, float &ss, float &sss, float &srs,
                                          bool &ok) {
    const float wmax = (float)W - 1.0f, hmax = (float)H - 1.0f;
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
            const uint32_t q = img[quad_index((uint32_t)ix, (uint32_t)iy, (uint32_t)qrows)];
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

// Texel (x, y) of a layer through a structured buffer resource of 4-byte records over it: record
// 16 ((x / 16) (qrows - 1) + y) + x = quad_index(x, y, qrows), formed with a shift, one 24-bit
// multiply-add and one shift-add into a single address register (an index and a byte offset in
// two registers would cost the texture addresser about twice the cycles per load, measured).  The
// hardware range-checks the record index against the layer's texel count, so a load can never
// leave the layer.
__device__ __forceinline__ uint32_t texel_load(__amdgpu_buffer_rsrc_t texels, uint32_t x, uint32_t y,
                                               uint32_t qrows_m1) {
    return struct_buffer_load_u32(texels, (int32_t)(((__umul24(x >> 4, qrows_m1) + y) << 4) + x), 0, 0, 0);
}

// Taps of a window known to lie inside the source image (see view_cost): no per-tap checks.
// Texels are read through texel_load (three address instructions per tap beyond the two
// conversions).  cu, cv: the corner taps (0, 5, 30, 35) as classify_window computed them, with
// the same expressions, so they are taken from there.
__device__ __forceinline__ void view_taps_inside(const uint32_t *__restrict__ img, int32_t W, int32_t H,
                                                 int32_t qrows, const Mat3 &Hm, float px, float py,
                                                 const Window &win, const float cu[4],
                                                 const float cv[4], float &ss, float &sss,
                                                 float &srs) {
    const __amdgpu_buffer_rsrc_t texels = __builtin_amdgcn_make_buffer_rsrc(
        (void *)img, (short)4, (int)quad_texels(W, H), 0x00020000);
    const uint32_t qrows_m1 = (uint32_t)(qrows - 1);
#pragma unroll
    for (int row = 0; row < 6; row++) {
        const float Y = py + (float)(row * 2 - 5);
        const float rx = fmaf(Hm.m[1], Y, Hm.m[2]);
        const float ry = fmaf(Hm.m[4], Y, Hm.m[5]);
        const float rw = fmaf(Hm.m[7], Y, Hm.m[8]);
#pragma unroll
        for (int col = 0; col < 6; col++) {
            const int i = row * 6 + col;
            const int corner = (row == 0 || row == 5) && (col == 0 || col == 5) ? (col / 5) + 2 * (row / 5) : -1;
            float u, v;
            if (corner >= 0) {
                u = cu[corner];
                v = cv[corner];
            } else {
                const float X = px + (float)(col * 2 - 5);
                const float qx = fmaf(Hm.m[0], X, rx);
                const float qy = fmaf(Hm.m[3], X, ry);
                const float qz = fmaf(Hm.m[6], X, rw);
                const float rz = rcp_exact(qz);
                u = qx * rz;
                v = qy * rz;
            }
            const uint32_t ix = (uint32_t)u, iy = (uint32_t)v;
            const float fx = __builtin_amdgcn_fractf(u), fy = __builtin_amdgcn_fractf(v);
            const uint32_t q = texel_load(texels, ix, iy, qrows_m1);
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
//   kCareful  anything else (windows touching the border, huge terms of q): per-tap checks.
enum { kOutside = 0, kInside = 1, kCareful = 2 };

__device__ __forceinline__ int classify_window(const PmSrc &s, const Mat3 &H, float px, float py,
                                               float cu[4], float cv[4]) {
    const float wmax = (float)s.W - 1.0f, hmax = (float)s.H - 1.0f;
    const float ax = fabsf(px) + 5.0f, ay = fabsf(py) + 5.0f;
    // bounds on the magnitude of the terms of q.x, q.y, q.z over the window (each within a few
    // units of 2^-24 of its exact value, far inside the 2^-20 slack the margins below carry)
    const float bx = fmaf(fabsf(H.m[0]), ax, fmaf(fabsf(H.m[1]), ay, fabsf(H.m[2])));
    const float by = fmaf(fabsf(H.m[3]), ax, fmaf(fabsf(H.m[4]), ay, fabsf(H.m[5])));
    const float bz = fmaf(fabsf(H.m[6]), ax, fmaf(fabsf(H.m[7]), ay, fabsf(H.m[8])));
    // Past this, rcp_exact may differ from 1 / q.z; below it every q component is finite, and so is
    // every u, v of a corner with q.z > 1e-6 (no NaN), so a corner is invalid exactly when the
    // minimum or maximum over the corners fails a test.
    if (!(bz < 0x1p60f) || !(bx < 0x1p60f) || !(by < 0x1p60f)) return kCareful;
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
        cu[corner] = u;
        cv[corner] = v;
        zmin = fminf(zmin, qz);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
    }
    // (u and v of a corner with q.z <= 1e-6 may be NaN; zmin decides those windows first)
    if (!(zmin > 1e-6f) || !(umin >= 0.0f) || !(vmin >= 0.0f) || !(umax <= wmax) || !(vmax <= hmax))
        return kOutside;
    // Each computed q component is within 2^-20 of its term bound of the exact value (two FMAs
    // round at most 2 ulp of the larger term), so exact q.z >= zmin - ez on the window.  A
    // computed u = q.x / q.z then lies within eu of the exact ratio; require corners inside
    // by 4 eu so that exact and computed interior taps stay inside.
    const float ez = bz * 0x1p-20f;
    const float zlow = fmaf(-2.0f, ez, zmin);
    if (!(zlow > 2e-6f)) return kCareful;
    // an upper bound of 1 / zlow: v_rcp_f32 is within 1 ulp, the factor covers it and its rounding
    const float ilow = __builtin_amdgcn_rcpf(zlow) * 1.000001f;
    // 4 eu and 4 ev, eu = (bx 2^-20 + W ez) / zlow + W 2^-20 (W = wmax + 1), likewise ev
    const float ez4 = bz * 0x1p-18f;
    const float w1 = wmax + 1.0f, h1 = hmax + 1.0f;
    const float eu4 = fmaf(fmaf(bx, 0x1p-18f, w1 * ez4), ilow, w1 * 0x1p-18f);
    const float ev4 = fmaf(fmaf(by, 0x1p-18f, h1 * ez4), ilow, h1 * 0x1p-18f);
    const bool inside = umin >= eu4 && umax <= wmax - eu4 && vmin >= ev4 && vmax <= hmax - ev4;
    return inside ? kInside : kCareful;
}

__device__ __forceinline__ float view_cost(const PmSrc &s, const Mat3 &H, float px, float py,
                                           const Window &win, const uint32_t *__restrict__ quad) {
    const uint32_t *__restrict__ img = quad + s.qoff;
    float cu[4], cv[4];
    const int cls = classify_window(s, H, px, py, cu, cv);
    float ss = 0.0f, sss = 0.0f, srs = 0.0f;
    bool ok = cls != kOutside;
    // the checked loop is exact for every class; lanes already known outside stay invalid
    if (__any(cls == kCareful)) {
        const float bound = fabsf(H.m[6]) * (fabsf(px) + 5.0f) + fabsf(H.m[7]) * (fabsf(py) + 5.0f) +
                            fabsf(H.m[8]);
        bool tap_ok = true;
        if (__all(bound < 0x1p60f))
            view_taps<true>(img, s.W, s.H, s.qrows, H, px, py, win, ss, sss, srs, tap_ok);
        else
            view_taps<false>(img, s.W, s.H, s.qrows, H, px, py, win, ss, sss, srs, tap_ok);
        ok = ok & tap_ok;
    } else if (__any(cls == kInside)) {
        if (cls == kInside) view_taps_inside(img, s.W, s.H, s.qrows, H, px, py, win, cu, cv, ss, sss, srs);
    }
    if (!ok) return kInvalid;
    // PM:151-158 with the window's sums scaled by 1 / sw and the correlation by 1 / den (IEEE
    // reciprocals; den >= kDegenerateDen lies in the range where rcp_exact is the IEEE reciprocal)
    float mean = ss * win.isw;
    float var_s = fmaxf(sss * win.isw - mean * mean, 0.0f);
    // den = sqrtf(var_r * var_s) and the test den < kDegenerateDen, with the test taken on the
    // square (finite and >= 0 here) so that the square root only sees x >= kDegenerateSq
    const float x = win.var * var_s;
    if (!(x >= kDegenerateSq)) return 1.0f;
    const float den = sqrt_normal(x);
    float cov = srs * win.isw - win.mean * mean;
    return clampf(1.0f - cov * rcp_exact(den), 0.0f, 2.0f);
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

// PM:192-233.  `bound` is the cost the caller will compare the result against with a strict
// `<` (the pixel's best cost so far; +inf when the result is kept whatever it is).  Once fewer
// than k views are left, the k smallest of (the costs so far, and 0 for every view left) bound
// the final mean from below: every view cost is >= 0, the i-th smallest of a set can only fall
// when members are lowered to 0, and rounded addition and division by k are monotone.  When
// that bound already reaches `bound`, the candidate cannot win and the remaining views are not
// evaluated; `bound` itself is returned instead of the mean, and the caller discards it as it
// would have discarded the mean.  A wavefront stops once all of its lanes have stopped.
// The views are taken last to first: the k smallest costs are summed in ascending order, so the
// order of evaluation changes no result, and source lists are ranked best first
// (mvs/view-selection.ts:46-48 sorts them by shared-point score), so leaving the best-matching
// views to the end raises the bound sooner.
template <int KMAX, bool GEOMETRIC>
__device__ __forceinline__ float hypothesis_cost(const PmDesc &d, const PmPools &pools, float px,
                                                 float py, float rx, float ry, const Window &win,
                                                 float nx, float ny, float nz, float depth,
                                                 float bound) {
    float inv = 1.0f / depth;
    if (!(depth > 0.0f) || inv > d.inv_min || inv < d.inv_max) return kInvalid;
    if (!(dot3(nx, ny, nz, rx, ry, 1.0f) < -1e-3f)) return kInvalid;
    float X0 = rx * depth, X1 = ry * depth, X2 = 1.0f * depth;
    float rho = dot3(nx, ny, nz, X0, X1, X2);
    const int32_t count = d.n_src;
    if (count == 0) return kInvalid;
    const PlaneTerms plane = plane_terms(d, nx, ny, nz, rho);
    TopK<KMAX> top;
    top.reset();
    const int32_t first_check = count - d.topk + 1;  // from here on fewer than k views are left
    for (int32_t i = 0; i < count; i++) {
        const int32_t s = count - 1 - i;  // last source first
        if (i >= first_check) {
            const int32_t known = i - first_check + 1;  // k minus the views left
            float total = 0.0f;
#pragma unroll
            for (int j = 0; j < KMAX; j++)
                if (j < known) total = total + top.s[j];
            // total >= k * bound exactly (the FMA rounds the exact difference once, keeping its
            // sign), so total / k, and the final mean above it, cannot fall below `bound`
            if (fmaf(-(float)d.topk, bound, total) >= 0.0f) return bound;
        }
        const PmSrc &src = d.src[s];
        Mat3 H;
        homography(src, plane, H);
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
    uint32_t seed = d.seed_base + iteration * 97u + phase * 13u;  // DS:333
    return pcg(index ^ pcg(seed + iteration * 7919u + phase * 104729u));  // PM:251-253
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

}  // namespace
}  // namespace lr3d
