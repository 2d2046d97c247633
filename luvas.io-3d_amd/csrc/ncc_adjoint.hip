// Adjoint of the aggregated PatchMatch cost (no counterpart in the reference, which has no
// gradients).
//
// For the current plane (n, d) of every pixel p of a reference view, C_p is the aggregated cost
// of PM:192-233 (mean of the k smallest bilateral-NCC view costs).  The kernel returns
// dL/dd_p, dL/dn_p and, reduced per block, dL/dR_s and dL/dt_s of the relative pose of every
// source view s, for L = sum_p weight_p C_p.  Windows and taps are recomputed (nothing from the
// forward pass is stored):
//   pass 1  every view's cost exactly as the forward kernels compute it (FP32), and the k
//           smallest by (cost, index);
//   pass 2  for each selected view whose cost is differentiable (all taps inside, variance above
//           the floor, cost strictly inside (0, 2)), one sweep over the 36 taps for the NCC sums
//           and dC/dH, chained through H = Ks (R + t n^T / rho) Kr^-1, rho = d (n . ray).
// Pass 2 runs in FP64 from the FP32 plane, the u8 images and the FP64 relative poses: in
// FP32 the taps q = H (X, Y, 1) at pixel coordinates in the hundreds carry ~1e-4 px of rounding,
// which moves single-pixel gradients by up to ~1 %.  The top-k selection, INVALID views and the
// clamp contribute no gradient (the cost is piecewise constant there).
//
// Optionally the same launch takes one gradient step on every plane and keeps it where the
// step's FP32 aggregated cost is lower than the current one (refinement never raises a pixel's
// cost for fixed poses).
#include "pm_device.h"

namespace lr3d {

namespace {

// The bilateral window of every lane, staged in LDS as [sample][lane] (bank-conflict free), so
// that the FP64 tap loops below can stay rolled without spilling the window to scratch.
struct WindowLds {
    float v[kSamples][64];
    float w[kSamples][64];
};

// Reference-window statistics in FP64 from the FP32 window (values and bilateral weights).  The
// forward pass forms them in FP32, where E[x^2] - E[x]^2 cancels.
struct RefStats {
    double sw, mean, var;
};

__device__ __forceinline__ RefStats ref_stats(const WindowLds &win, int lane) {
    double sw = 0.0, sr = 0.0;
#pragma unroll 6
    for (int i = 0; i < kSamples; i++) {
        sw += (double)win.w[i][lane];
        sr = fma((double)win.w[i][lane], (double)win.v[i][lane], sr);
    }
    const double mean = sr / sw;
    double var = 0.0;
#pragma unroll 6
    for (int i = 0; i < kSamples; i++) {
        const double dv = (double)win.v[i][lane] - mean;
        var = fma((double)win.w[i][lane] * dv, dv, var);
    }
    return RefStats{sw, mean, var / sw};
}

struct Mat3d {
    double m[9];
};

// Reciprocals of the reference intrinsics (Kr^-1 = [1/fx 0 -cx/fx; 0 1/fy -cy/fy; 0 0 1]).
struct KrInv {
    double ifx, ify, cxf, cyf;
};

// H = Ks (R + t n^T / rho) Kr^-1 in FP64, given 1 / rho.
__device__ __forceinline__ void homography64(const AdjGeom &g, int s, const double n[3], double inv_rho,
                                             const KrInv &ki, Mat3d &H) {
    const double nr[3] = {n[0] * inv_rho, n[1] * inv_rho, n[2] * inv_rho};
    double M[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M[i * 3 + j] = g.R[s][i * 3 + j] + g.t[s][i] * nr[j];
    const double *k = g.k[s];
    double A[9];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        A[j] = k[0] * M[j] + k[2] * M[6 + j];
        A[3 + j] = k[1] * M[3 + j] + k[3] * M[6 + j];
        A[6 + j] = M[6 + j];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        H.m[i * 3 + 0] = A[i * 3 + 0] * ki.ifx;
        H.m[i * 3 + 1] = A[i * 3 + 1] * ki.ify;
        H.m[i * 3 + 2] = A[i * 3 + 2] - A[i * 3 + 0] * ki.cxf - A[i * 3 + 1] * ki.cyf;
    }
}

// 1 / z to within an ulp or two of double precision: v_rcp_f64 and two Newton steps (the passes are
// compared with the FP64 reference to 1e-4, far above that; the IEEE quotient's scaling and
// special-case sequence is not needed for the normal, positive q.z of a differentiable view).
__device__ __forceinline__ double rcp64(double z) {
    double r = __builtin_amdgcn_rcp(z);
    r = fma(fma(-z, r, 1.0), r, r);
    return fma(fma(-z, r, 1.0), r, r);
}

// Position of a tap of a source view in FP64, clamped into the image so that its texel read stays
// inside: u, v, 1 / q_z, the bilinear fractions and the index of its 2 x 2 texel quad.
struct TapPos {
    double u, v, rz, fx, fy;
    uint32_t texel;
};

__device__ __forceinline__ TapPos tap_position(int32_t qrows, double wmax, double hmax, const Mat3d &H, double X,
                                               double Y) {
    const double qx = fma(H.m[0], X, fma(H.m[1], Y, H.m[2]));
    const double qy = fma(H.m[3], X, fma(H.m[4], Y, H.m[5]));
    const double qz = fma(H.m[6], X, fma(H.m[7], Y, H.m[8]));
    TapPos t;
    t.rz = rcp64(qz);
    t.u = fmin(fmax(qx * t.rz, 0.0), wmax);
    t.v = fmin(fmax(qy * t.rz, 0.0), hmax);
    const double x0 = floor(t.u), y0 = floor(t.v);
    t.fx = t.u - x0;
    t.fy = t.v - y0;
    t.texel = quad_index((uint32_t)x0, (uint32_t)y0, (uint32_t)qrows);
    return t;
}

// Bilinear sample of a texel quad (u8 texels a b / c e) and its texel-space gradient.
struct TapVal {
    double val, gu, gv;
};

// A byte as a double, exactly: the bits of 2^52 + b less 2^52 (integer and FP64 additions in place
// of conversions).
__device__ __forceinline__ double byte_double(uint32_t b) {
    return __hiloint2double(0x43300000, (int)b) - 4503599627370496.0;
}

__device__ __forceinline__ TapVal tap_sample(uint32_t q, double fx, double fy) {
    const double a = byte_double(q & 0xffu), b = byte_double((q >> 8) & 0xffu);
    const double c = byte_double((q >> 16) & 0xffu), e = byte_double(q >> 24);
    const double top = fma(fx, b - a, a), bot = fma(fx, e - c, c);
    TapVal t;
    t.val = fma(fy, bot - top, top);
    t.gu = (b - a) * (1.0 - fy) + (e - c) * fy;
    t.gv = bot - top;
    return t;
}

// acc += k (a X, a Y, a, b X, b Y, b, c X, c Y, c)
__device__ __forceinline__ void add_tap_terms(double acc[9], double k, double a, double b, double c,
                                              double X, double Y) {
    const double ka = k * a, kb = k * b, kc = k * c;
    acc[0] = fma(ka, X, acc[0]);
    acc[1] = fma(ka, Y, acc[1]);
    acc[2] += ka;
    acc[3] = fma(kb, X, acc[3]);
    acc[4] = fma(kb, Y, acc[4]);
    acc[5] += kb;
    acc[6] = fma(kc, X, acc[6]);
    acc[7] = fma(kc, Y, acc[7]);
    acc[8] += kc;
}

// dC_view/dH (row-major) of a differentiable view, in one sweep over the 36 taps.
//   dC/dval_i = (w_i / S) [B (val_i - mean_s) - (r_i - mean_r) A],  A = 1/den, B = cov var_r / den^3
//   du/dH_0j = x_j / q_z, dv/dH_1j = x_j / q_z, d(u, v)/dH_2j = -(u, v) x_j / q_z
// so dC/dH = sum_i dC/dval_i T_i with the tap term T_i = (g_u x, g_v x, -(g_u u + g_v v) x) / q_z
// (g the texel-space image gradient, x = (X, Y, 1)).  A, B and mean_s need the whole window's
// sums, so the sweep gathers P = sum w val T, Q = sum w T and R = sum w (r - mean_r) T beside
// the NCC sums, and dC/dH = (B (P - mean_s Q) - A R) / S.
__device__ __forceinline__ void view_grad_h(const PmSrc &s, const Mat3d &H, double px, double py,
                                            const WindowLds &win, int lane, const RefStats &rs,
                                            const uint32_t *__restrict__ quad, double G[9]) {
    const uint32_t *__restrict__ img = quad + s.qoff;
    const double wmax = (double)s.W - 1.0, hmax = (double)s.H - 1.0;
    double ss = 0.0, sss = 0.0, srs = 0.0;
    double P[9], Q[9], R[9];
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = Q[k] = R[k] = 0.0;
    const int32_t qrows = s.qrows;
    // rows of taps; within a row the column offsets are constants
#pragma unroll 1
    for (int row = 0; row < 6; row++) {
        const double Y = py + (double)(row * 2 - 5);
#pragma unroll
        for (int col = 0; col < 6; col++) {
            const int i = row * 6 + col;
            const double X = px + (double)(col * 2 - 5);
            const TapPos t = tap_position(qrows, wmax, hmax, H, X, Y);
            const TapVal tv = tap_sample(img[t.texel], t.fx, t.fy);
            const double w = (double)win.w[i][lane], r = (double)win.v[i][lane];
            const double wv = w * tv.val;
            ss += wv;
            sss = fma(wv, tv.val, sss);
            srs = fma(wv, r, srs);
            const double a = tv.gu * t.rz, b = tv.gv * t.rz;
            const double c = -(a * t.u + b * t.v);
            add_tap_terms(P, wv, a, b, c, X, Y);
            add_tap_terms(Q, w, a, b, c, X, Y);
            add_tap_terms(R, w * (r - rs.mean), a, b, c, X, Y);
        }
    }
    const double inv_sw = 1.0 / rs.sw;
    const double mean = ss * inv_sw;
    const double var_s = fmax(sss * inv_sw - mean * mean, 0.0);
    const double den = sqrt(rs.var * var_s);
    const double cov = srs * inv_sw - rs.mean * mean;
    const double A = den > 0.0 ? 1.0 / den : 0.0;
    const double B = cov * rs.var * A * A * A;
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = inv_sw * (B * (P[k] - mean * Q[k]) - A * R[k]);
}

// Number of set bits of a wavefront mask in the lanes below this one.
__device__ __forceinline__ uint32_t below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Sums of 16 values over the wavefront by recursive halving: at each step a lane keeps half of its
// values and adds its partner's copies of them (partners 32, 16, 8 and 4 lanes apart), then the
// last value is summed over the 4 lanes left (partners 2 and 1 apart).  Every lane returns the
// total of value (lane >> 2) & 15.  17 exchanges instead of 16 shuffle trees of 6; the pattern is
// fixed, so the sums are deterministic.  Every lane of the wavefront must call it.
__device__ __forceinline__ double wave_sums16(const double v[16], int lane) {
    double a8[8], a4[4], a2[2];
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
    for (int k = 0; k < 8; k++) a8[k] = (b5 ? v[8 + k] : v[k]) + __shfl_xor(b5 ? v[k] : v[8 + k], 32, 64);
#pragma unroll
    for (int k = 0; k < 4; k++) a4[k] = (b4 ? a8[4 + k] : a8[k]) + __shfl_xor(b4 ? a8[k] : a8[4 + k], 16, 64);
#pragma unroll
    for (int k = 0; k < 2; k++) a2[k] = (b3 ? a4[2 + k] : a4[k]) + __shfl_xor(b3 ? a4[k] : a4[2 + k], 8, 64);
    double a1 = (b2 ? a2[1] : a2[0]) + __shfl_xor(b2 ? a2[0] : a2[1], 4, 64);
    a1 += __shfl_xor(a1, 2, 64);
    return a1 + __shfl_xor(a1, 1, 64);
}

// Wavefront sum (fixed shuffle tree, deterministic), result valid in lane 0.  Every lane of the
// wavefront must call it.
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// One wavefront per block (8 x 8 threads): pose gradients are reduced per wavefront with a fixed
// shuffle tree and written as one partial per block (the host adds them in order).
__global__ void __launch_bounds__(64) ncc_backward_kernel(PmPlan plan, PmPools pools,
                                                          const AdjGeom *__restrict__ geom,
                                                          const float *__restrict__ weights,
                                                          float *__restrict__ out_cost,
                                                          float *__restrict__ out_gd,
                                                          float *__restrict__ out_gn,
                                                          double *__restrict__ pose_partials,
                                                          float step_depth, float step_normal,
                                                          int32_t apply) {
    PixelTask task = locate(plan, (int32_t)blockDim.x);
    const PmDesc &d = plan.descs[task.view];
    const AdjGeom &g = geom[task.view];
    const int32_t x = task.gx, y = task.gy;
    const bool inside = x < d.W && y < d.H;
    const uint32_t index = inside ? (uint32_t)y * (uint32_t)d.W + (uint32_t)x : 0u;
    const float px = (float)x, py = (float)y;
    const float4 h = inside ? pools.hyp[d.off + index] : make_float4(0.0f, 0.0f, -1.0f, 0.0f);
    const float weight = inside ? (weights ? weights[d.off + index] : 1.0f) : 0.0f;
    const float rx = (px - d.k[2]) / d.k[0], ry = (py - d.k[3]) / d.k[1];
    const float inv = 1.0f / h.w;
    const float nr = dot3(h.x, h.y, h.z, rx, ry, 1.0f);
    const bool valid = inside && h.w > 0.0f && !(inv > d.inv_min) && !(inv < d.inv_max) &&
                       nr < -1e-3f && d.n_src > 0;
    __shared__ WindowLds lds;
    const int lane = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    RefStats rs{1.0, 0.0, 0.0};
    float wsw = 1.0f, wmean = 0.0f, wvar = 0.0f;  // the FP32 window statistics of the forward pass
    float costs[kMaxSources];
    uint32_t selected = 0u;
    float cost = kInvalid;
#pragma unroll
    for (int s = 0; s < kMaxSources; s++) costs[s] = INFINITY;
    if (valid) {
        Window win;
        prepare_window(pools.quad + d.qoff, d.W, d.H, d.qrows, x, y, pools.wtab, win);
#pragma unroll
        for (int i = 0; i < kSamples; i++) {
            lds.v[i][lane] = win.v[i];
            lds.w[i][lane] = win.w[i];
        }
        wsw = win.sw;
        wmean = win.mean;
        wvar = win.var;
        rs = ref_stats(lds, lane);
        const float rho = dot3(h.x, h.y, h.z, rx * h.w, ry * h.w, 1.0f * h.w);
        const PlaneTerms plane = plane_terms(d, h.x, h.y, h.z, rho);
        // pass 1: view costs, exactly as the forward pass, and the k smallest by (cost, index)
#pragma unroll 1
        for (int s = 0; s < d.n_src; s++) {
            Mat3 H;
            homography(d.src[s], plane, H);
            const float c = view_cost(d.src[s], H, px, py, win, pools.quad);
#pragma unroll
            for (int q = 0; q < kMaxSources; q++)
                if (q == s) costs[q] = c;  // registers, not scratch
        }
        float total = 0.0f;
        for (int pick = 0; pick < d.topk; pick++) {
            int best = -1;
            float bc = INFINITY;
#pragma unroll
            for (int s = 0; s < kMaxSources; s++) {
                if (s < d.n_src && !(selected & (1u << s)) && (best < 0 || costs[s] < bc)) {
                    best = s;
                    bc = costs[s];
                }
            }
            selected |= 1u << best;
            total = total + bc;
        }
        cost = total / (float)d.topk;
    }
    const double scale = (double)weight / (double)d.topk;
    double dd = 0.0, dn0 = 0.0, dn1 = 0.0, dn2 = 0.0;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    double *partial = pose_partials ? pose_partials + (size_t)b * kAdjPartial : nullptr;
    const bool lead = (threadIdx.x | threadIdx.y) == 0;
    // Pass 2 runs over the wavefront's (pixel, view) items compacted across lanes: a pixel
    // differentiates at most k of its views, so a loop over the views with one lane per pixel
    // would leave most lanes idle.  Items are listed view-major (pixels ascending within a view):
    // view s owns list entries [first[s], first[s + 1]).  INVALID (2), degenerate (exactly 1) or
    // clamped (0, 2) views carry no gradient, the cost being constant there; a non-degenerate
    // view whose FP32 cost is exactly 1 is skipped as well.
    __shared__ uint8_t items[64 * kMaxSources];
    uint64_t owns[kMaxSources];
    uint32_t first[kMaxSources + 1];
    first[0] = 0u;
#pragma unroll
    for (int s = 0; s < kMaxSources; s++) {
        const float cs = costs[s];
        const bool smooth = s < d.n_src && valid && (selected & (1u << s)) && cs > 0.0f && cs < 2.0f && cs != 1.0f;
        owns[s] = __ballot(smooth);
        if (smooth) items[first[s] + below(owns[s])] = (uint8_t)lane;
        first[s + 1] = first[s] + (uint32_t)__popcll(owns[s]);
    }
    __syncthreads();
    const uint32_t n_items = first[kMaxSources];
    // divisions shared by every item of the block
    const double inv_topk = 1.0 / (double)d.topk;
    KrInv ki;
    ki.ifx = 1.0 / g.kr[0];
    ki.ify = 1.0 / g.kr[1];
    ki.cxf = g.kr[2] * ki.ifx;
    ki.cyf = g.kr[3] * ki.ify;
    for (uint32_t base = 0; base < n_items; base += 64) {
        // this lane's item, its view and pixel; the pixel's plane, weight and window statistics
        // come from the pixel's lane (every lane takes part in the exchanges)
        const uint32_t i = base + (uint32_t)lane;
        const bool act = i < n_items;
        const int32_t p = act ? (int32_t)items[i] : 0;
        int s = 0;
#pragma unroll
        for (int q = 1; q < kMaxSources; q++) s += i >= first[q] ? 1 : 0;
        const float hx = __shfl(h.x, p), hy = __shfl(h.y, p), hz = __shfl(h.z, p), hw = __shfl(h.w, p);
        const float wp = __shfl(weight, p);
        const int32_t xp = __shfl(x, p), yp = __shfl(y, p);
        const RefStats rsp{__shfl(rs.sw, p), __shfl(rs.mean, p), __shfl(rs.var, p)};
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        double gpose[12];
#pragma unroll
        for (int j = 0; j < 12; j++) gpose[j] = 0.0;
        if (act) {
            // FP64 plane: X = d ray, rho = n . X
            const double n[3] = {(double)hx, (double)hy, (double)hz};
            const double dep = (double)hw;
            const double rayd[3] = {((double)xp - g.kr[2]) * ki.ifx, ((double)yp - g.kr[3]) * ki.ify, 1.0};
            const double Xp[3] = {rayd[0] * dep, rayd[1] * dep, dep};
            const double nrd = n[0] * rayd[0] + n[1] * rayd[1] + n[2] * rayd[2];
            const double rhod = nrd * dep;
            const double inv_rho = 1.0 / rhod, inv_r2 = inv_rho * inv_rho;
            const double scp = (double)wp * inv_topk;
            const PmSrc &src = d.src[s];
            Mat3d H;
            homography64(g, s, n, inv_rho, ki, H);
            double G[9];
            view_grad_h(src, H, (double)xp, (double)yp, lds, p, rsp, pools.quad, G);
            // G_M = Ks^T G Kr^-T
            const double *k = g.k[s];
            double T[9];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                T[j] = k[0] * G[j];
                T[3 + j] = k[1] * G[3 + j];
                T[6 + j] = k[2] * G[j] + k[3] * G[3 + j] + G[6 + j];
            }
            double GM[9];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                GM[a * 3 + 0] = (T[a * 3 + 0] - T[a * 3 + 2] * g.kr[2]) * ki.ifx;
                GM[a * 3 + 1] = (T[a * 3 + 1] - T[a * 3 + 2] * g.kr[3]) * ki.ify;
                GM[a * 3 + 2] = T[a * 3 + 2];
            }
            const double *t = g.t[s];
            double gmn[3];
#pragma unroll
            for (int a = 0; a < 3; a++) gmn[a] = GM[a * 3] * n[0] + GM[a * 3 + 1] * n[1] + GM[a * 3 + 2] * n[2];
            const double tGn = t[0] * gmn[0] + t[1] * gmn[1] + t[2] * gmn[2];
            c[0] = -tGn * nrd * inv_r2;
            c[1] = (t[0] * GM[0] + t[1] * GM[3] + t[2] * GM[6]) * inv_rho - tGn * Xp[0] * inv_r2;
            c[2] = (t[0] * GM[1] + t[1] * GM[4] + t[2] * GM[7]) * inv_rho - tGn * Xp[1] * inv_r2;
            c[3] = (t[0] * GM[2] + t[1] * GM[5] + t[2] * GM[8]) * inv_rho - tGn * Xp[2] * inv_r2;
#pragma unroll
            for (int j = 0; j < 9; j++) gpose[j] = scp * GM[j];
#pragma unroll
            for (int j = 0; j < 3; j++) gpose[9 + j] = scp * gmn[j] * inv_rho;
        }
        // views with items in this batch of 64: s_lo .. s_hi (wave-uniform)
        const uint32_t last = min(base + 64u, n_items) - 1u;
        int s_lo = 0, s_hi = 0;
#pragma unroll
        for (int q = 1; q < kMaxSources; q++) {
            s_lo += base >= first[q] ? 1 : 0;
            s_hi += last >= first[q] ? 1 : 0;
        }
        // each pixel adds its items' terms in view order, as a loop over its views would
#pragma unroll
        for (int q = 0; q < kMaxSources; q++) {
            if (q < s_lo || q > s_hi) continue;
            const uint32_t at = first[q] + below(owns[q]);
            const bool mine = ((owns[q] >> lane) & 1u) && at >= base && at < base + 64u;
            const int32_t from = mine ? (int32_t)(at - base) : 0;
            const double t0 = __shfl(c[0], from), t1 = __shfl(c[1], from);
            const double t2 = __shfl(c[2], from), t3 = __shfl(c[3], from);
            if (mine) {
                dd += t0;
                dn0 += t1;
                dn1 += t2;
                dn2 += t3;
            }
        }
        // pose gradients: per view, the 12 sums over the batch's lanes (a fixed exchange pattern),
        // added in batch order to the block's partial (zeroed before the launch)
        if (partial) {
            for (int q = s_lo; q <= s_hi; q++) {
                double v[16];
#pragma unroll
                for (int j = 0; j < 16; j++) v[j] = j < 12 && act && s == q ? gpose[j] : 0.0;
                const double sum = wave_sums16(v, lane);
                const int j = (lane >> 2) & 15;
                if ((lane & 3) == 0 && j < 12) partial[2 + q * 12 + j] += sum;
            }
        }
    }
    const float gd = (float)(scale * dd);
    const float gn0 = (float)(scale * dn0), gn1 = (float)(scale * dn1), gn2 = (float)(scale * dn2);
    float after = cost;
    if (apply && valid) {
        // one gradient step on the plane (depth kept in range and within 0.5 % of the current
        // value, normal kept unit and facing the camera), kept if its aggregated cost is lower
        const float lo = 1.0f / d.inv_min, hi = 1.0f / d.inv_max;
        float step = -step_depth * gd;
        step = fminf(fmaxf(step, -0.005f * h.w), 0.005f * h.w);
        const float nd = fminf(fmaxf(h.w + step, lo), hi);
        const float gdot = gn0 * h.x + gn1 * h.y + gn2 * h.z;
        float mx = h.x - step_normal * (gn0 - gdot * h.x);
        float my = h.y - step_normal * (gn1 - gdot * h.y);
        float mz = h.z - step_normal * (gn2 - gdot * h.z);
        normalize3(mx, my, mz);
        if (!(dot3(mx, my, mz, rx, ry, 1.0f) < -1e-3f)) {
            mx = h.x;
            my = h.y;
            mz = h.z;
        }
        Window win;
#pragma unroll
        for (int i = 0; i < kSamples; i++) {
            win.v[i] = lds.v[i][lane];
            win.w[i] = lds.w[i][lane];
        }
        win.sw = wsw;
        win.mean = wmean;
        win.var = wvar;
        win.isw = 1.0f / wsw;
        const float c =
            hypothesis_cost<kMaxSources, false>(d, pools, px, py, rx, ry, win, mx, my, mz, nd, cost);
        if (c < cost) {
            after = c;
            pools.hyp[d.off + index] = make_float4(mx, my, mz, nd);
            pools.cost[d.off + index] = c;
        }
    }
    if (partial) {
        const double l0 = wave_sum(valid ? (double)cost * weight : 0.0);
        const double l1 = wave_sum(valid ? (double)after * weight : 0.0);
        if (lead) {
            partial[0] = l0;
            partial[1] = l1;
        }
    }
    if (!inside) return;
    if (out_cost) out_cost[d.off + index] = cost;
    if (out_gd) out_gd[d.off + index] = gd;
    if (out_gn) {
        out_gn[(size_t)(d.off + index) * 3 + 0] = gn0;
        out_gn[(size_t)(d.off + index) * 3 + 1] = gn1;
        out_gn[(size_t)(d.off + index) * 3 + 2] = gn2;
    }
}

}  // namespace

hipError_t ncc_launch_backward(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                               const AdjGeom *geom, const float *weights, float *cost, float *gd,
                               float *gn, double *pose_partials, float step_depth, float step_normal,
                               int32_t apply) {
    if (plan.tx * plan.ty != 64) return hipErrorInvalidConfiguration;
    dim3 block(plan.tx, plan.ty), grid(plan.blocks);
    hipLaunchKernelGGL(ncc_backward_kernel, grid, block, 0, s, plan, pools, geom, weights, cost, gd,
                       gn, pose_partials, step_depth, step_normal, apply);
    return hipGetLastError();
}

}  // namespace lr3d
