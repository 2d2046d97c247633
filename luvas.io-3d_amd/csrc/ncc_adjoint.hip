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
//           the floor, cost strictly inside (0, 2)), the NCC sums;
//   pass 3  dC/dH accumulated over the 36 taps and chained through
//           H = Ks (R + t n^T / rho) Kr^-1, rho = d (n . ray).
// Passes 2 and 3 run in FP64 from the FP32 plane, the u8 images and the FP64 relative poses: in
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

// H = Ks (R + t n^T / rho) Kr^-1 in FP64.
__device__ __forceinline__ void homography64(const AdjGeom &g, int s, const double n[3], double rho,
                                             Mat3d &H) {
    double M[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M[i * 3 + j] = g.R[s][i * 3 + j] + g.t[s][i] * n[j] / rho;
    const double *k = g.k[s];
    double A[9];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        A[j] = k[0] * M[j] + k[2] * M[6 + j];
        A[3 + j] = k[1] * M[3 + j] + k[3] * M[6 + j];
        A[6 + j] = M[6 + j];
    }
    const double *kr = g.kr;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        H.m[i * 3 + 0] = A[i * 3 + 0] / kr[0];
        H.m[i * 3 + 1] = A[i * 3 + 1] / kr[1];
        H.m[i * 3 + 2] = A[i * 3 + 2] - A[i * 3 + 0] * kr[2] / kr[0] - A[i * 3 + 1] * kr[3] / kr[1];
    }
}

// Tap of a source view in FP64: position (clamped into the image, so reads stay inside), the
// bilinear sample and its texel-space gradient.
struct Tap {
    double u, v, rz, val, gu, gv;
};

// 1 / z to within an ulp or two of double precision: v_rcp_f64 and two Newton steps (the passes are
// compared with the FP64 reference to 1e-4, far above that; the IEEE quotient's scaling and
// special-case sequence is not needed for the normal, positive q.z of a differentiable view).
__device__ __forceinline__ double rcp64(double z) {
    double r = __builtin_amdgcn_rcp(z);
    r = fma(fma(-z, r, 1.0), r, r);
    return fma(fma(-z, r, 1.0), r, r);
}

__device__ __forceinline__ Tap source_tap(const uint32_t *__restrict__ img, int32_t qrows, double wmax, double hmax,
                                          const Mat3d &H, double X, double Y) {
    const double qx = fma(H.m[0], X, fma(H.m[1], Y, H.m[2]));
    const double qy = fma(H.m[3], X, fma(H.m[4], Y, H.m[5]));
    const double qz = fma(H.m[6], X, fma(H.m[7], Y, H.m[8]));
    Tap t;
    t.rz = rcp64(qz);
    t.u = fmin(fmax(qx * t.rz, 0.0), wmax);
    t.v = fmin(fmax(qy * t.rz, 0.0), hmax);
    const double x0 = floor(t.u), y0 = floor(t.v);
    const double fx = t.u - x0, fy = t.v - y0;
    const uint32_t q = img[quad_index((uint32_t)x0, (uint32_t)y0, (uint32_t)qrows)];
    const double a = (double)(q & 0xffu), b = (double)((q >> 8) & 0xffu);
    const double c = (double)((q >> 16) & 0xffu), e = (double)(q >> 24);
    const double top = fma(fx, b - a, a), bot = fma(fx, e - c, c);
    t.val = fma(fy, bot - top, top);
    t.gu = (b - a) * (1.0 - fy) + (e - c) * fy;
    t.gv = bot - top;
    return t;
}

// dC_view/dH (row-major) of a differentiable view (passes 2 and 3).
//   dC/dval_i = (w_i / S) [B (val_i - mean_s) - (r_i - mean_r) A],  A = 1/den, B = cov var_r / den^3
//   du/dH_0j = x_j / q_z, dv/dH_1j = x_j / q_z, d(u, v)/dH_2j = -(u, v) x_j / q_z
__device__ __forceinline__ void view_grad_h(const PmSrc &s, const Mat3d &H, double px, double py,
                                            const WindowLds &win, int lane, const RefStats &rs,
                                            const uint32_t *__restrict__ quad, double G[9]) {
    const uint32_t *__restrict__ img = quad + s.qoff;
    const double wmax = (double)s.W - 1.0, hmax = (double)s.H - 1.0;
    double ss = 0.0, sss = 0.0, srs = 0.0;
#pragma unroll 1
    for (int i = 0; i < kSamples; i++) {
        const double X = px + (double)((i % 6) * 2 - 5), Y = py + (double)((i / 6) * 2 - 5);
        const Tap t = source_tap(img, s.qrows, wmax, hmax, H, X, Y);
        const double wv = (double)win.w[i][lane] * t.val;
        ss += wv;
        sss = fma(wv, t.val, sss);
        srs = fma(wv, (double)win.v[i][lane], srs);
    }
    const double inv_sw = 1.0 / rs.sw;
    const double mean = ss * inv_sw;
    const double var_s = fmax(sss * inv_sw - mean * mean, 0.0);
    const double den = sqrt(rs.var * var_s);
    const double cov = srs * inv_sw - rs.mean * mean;
    const double A = den > 0.0 ? 1.0 / den : 0.0;
    const double B = cov * rs.var * A * A * A;
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll 1
    for (int i = 0; i < kSamples; i++) {
        const double X = px + (double)((i % 6) * 2 - 5), Y = py + (double)((i / 6) * 2 - 5);
        const Tap t = source_tap(img, s.qrows, wmax, hmax, H, X, Y);
        const double dval = (double)win.w[i][lane] * inv_sw *
                            (B * (t.val - mean) - ((double)win.v[i][lane] - rs.mean) * A);
        const double gx = dval * t.gu * t.rz, gy = dval * t.gv * t.rz;
        const double gz = -(gx * t.u + gy * t.v);
        G[0] = fma(gx, X, G[0]);
        G[1] = fma(gx, Y, G[1]);
        G[2] += gx;
        G[3] = fma(gy, X, G[3]);
        G[4] = fma(gy, Y, G[4]);
        G[5] += gy;
        G[6] = fma(gz, X, G[6]);
        G[7] = fma(gz, Y, G[7]);
        G[8] += gz;
    }
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
    // FP64 plane: X = d ray, rho = n . X
    const double n[3] = {(double)h.x, (double)h.y, (double)h.z};
    const double dep = (double)h.w;
    const double rayd[3] = {((double)x - g.kr[2]) / g.kr[0], ((double)y - g.kr[3]) / g.kr[1], 1.0};
    const double Xp[3] = {rayd[0] * dep, rayd[1] * dep, dep};
    const double nrd = n[0] * rayd[0] + n[1] * rayd[1] + n[2] * rayd[2];
    const double rhod = nrd * dep, r2 = rhod * rhod;
    const double scale = (double)weight / (double)d.topk;
    double dd = 0.0, dn0 = 0.0, dn1 = 0.0, dn2 = 0.0;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    double *partial = pose_partials ? pose_partials + (size_t)b * kAdjPartial : nullptr;
    const bool lead = (threadIdx.x | threadIdx.y) == 0;
    // passes 2 and 3 per source view; the loop is uniform over the wavefront for the reduction
    for (int s = 0; s < d.n_src; s++) {
        double gpose[12];
#pragma unroll
        for (int j = 0; j < 12; j++) gpose[j] = 0.0;
        float cs = INFINITY;
#pragma unroll
        for (int q = 0; q < kMaxSources; q++)
            if (q == s) cs = costs[q];
        // INVALID (2), degenerate (exactly 1) or clamped (0, 2) views carry no gradient: the cost
        // is constant there.  A non-degenerate view whose FP32 cost is exactly 1 is skipped as well.
        const bool smooth = valid && (selected & (1u << s)) && cs > 0.0f && cs < 2.0f && cs != 1.0f;
        if (smooth) {
            const PmSrc &src = d.src[s];
            Mat3d H;
            homography64(g, s, n, rhod, H);
            double G[9];
            view_grad_h(src, H, (double)x, (double)y, lds, lane, rs, pools.quad, G);
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
            for (int i = 0; i < 3; i++) {
                GM[i * 3 + 0] = (T[i * 3 + 0] - T[i * 3 + 2] * g.kr[2]) / g.kr[0];
                GM[i * 3 + 1] = (T[i * 3 + 1] - T[i * 3 + 2] * g.kr[3]) / g.kr[1];
                GM[i * 3 + 2] = T[i * 3 + 2];
            }
            const double *t = g.t[s];
            double gmn[3];
#pragma unroll
            for (int i = 0; i < 3; i++) gmn[i] = GM[i * 3] * n[0] + GM[i * 3 + 1] * n[1] + GM[i * 3 + 2] * n[2];
            const double tGn = t[0] * gmn[0] + t[1] * gmn[1] + t[2] * gmn[2];
            dd += -tGn * nrd / r2;
            dn0 += (t[0] * GM[0] + t[1] * GM[3] + t[2] * GM[6]) / rhod - tGn * Xp[0] / r2;
            dn1 += (t[0] * GM[1] + t[1] * GM[4] + t[2] * GM[7]) / rhod - tGn * Xp[1] / r2;
            dn2 += (t[0] * GM[2] + t[1] * GM[5] + t[2] * GM[8]) / rhod - tGn * Xp[2] / r2;
#pragma unroll
            for (int j = 0; j < 9; j++) gpose[j] = scale * GM[j];
#pragma unroll
            for (int j = 0; j < 3; j++) gpose[9 + j] = scale * gmn[j] / rhod;
        }
        if (partial) {
#pragma unroll
            for (int j = 0; j < 12; j++) {
                const double sum = wave_sum(gpose[j]);
                if (lead) partial[2 + s * 12 + j] = sum;
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
