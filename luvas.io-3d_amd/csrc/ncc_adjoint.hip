// Adjoint of the aggregated PatchMatch cost (no counterpart in the reference, which has no
// gradients).
//
// For the current plane (n, d) of every pixel p of a reference view, C_p is the aggregated cost
// of PM:192-233 (mean of the k smallest bilateral-NCC view costs).  The kernel returns
// dL/dd_p, dL/dn_p and, reduced per block, dL/dR_s and dL/dt_s of the relative pose of every
// source view s, for L = sum_p weight_p C_p.  Windows and taps are recomputed (nothing from the
// forward pass is stored): pass 1 evaluates every view's cost exactly as the forward kernels do
// and selects the k smallest; for each selected view whose cost is differentiable (all taps inside,
// variance above the floor, cost strictly inside (0, 2)), pass 2 re-accumulates the NCC sums and
// pass 3 accumulates dC/dH over the 36 taps, which is chained through
// H = Ks (R + t n^T / rho) Kr^-1, rho = d (n . ray).  The top-k selection, INVALID views and the
// clamp contribute no gradient (the cost is piecewise constant there).
//
// Optionally the same launch applies one gradient step to the planes (refinement).
#include "pm_device.h"

namespace lr3d {

namespace {

// NCC sums of one source view in FP32 (the forward arithmetic) for the adjoint passes.
struct ViewSums {
    float ss, sss, srs;
};

__device__ __forceinline__ ViewSums view_sums(const PmSrc &s, const Mat3 &H, float px, float py,
                                              const Window &win, const uint32_t *__restrict__ quad) {
    const uint32_t *__restrict__ img = quad + s.off;
    const float wmax = (float)s.W - 1.0f, hmax = (float)s.H - 1.0f;
    ViewSums r{0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < kSamples; i++) {
        const float X = px + (float)((i % 6) * 2 - 5), Y = py + (float)((i / 6) * 2 - 5);
        const float qx = fmaf(H.m[0], X, fmaf(H.m[1], Y, H.m[2]));
        const float qy = fmaf(H.m[3], X, fmaf(H.m[4], Y, H.m[5]));
        const float qz = fmaf(H.m[6], X, fmaf(H.m[7], Y, H.m[8]));
        const float rz = 1.0f / qz;
        const float u = fminf(fmaxf(qx * rz, 0.0f), wmax), v = fminf(fmaxf(qy * rz, 0.0f), hmax);
        const float x0 = floorf(u), y0 = floorf(v);
        const float fx = u - x0, fy = v - y0;
        const uint32_t q = img[__umul24((uint32_t)y0, (uint32_t)s.W) + (uint32_t)x0];
        const float a = (float)(q & 0xffu), b = (float)((q >> 8) & 0xffu);
        const float c = (float)((q >> 16) & 0xffu), e = (float)(q >> 24);
        const float top = fmaf(fx, b - a, a), bot = fmaf(fx, e - c, c);
        const float val = fmaf(fy, bot - top, top);
        const float wv = win.w[i] * val;
        r.ss = r.ss + wv;
        r.sss = fmaf(wv, val, r.sss);
        r.srs = fmaf(wv, win.v[i], r.srs);
    }
    return r;
}

// dC_view/dH (row-major) of a differentiable view, given its sums.
__device__ __forceinline__ void view_grad_h(const PmSrc &s, const Mat3 &H, float px, float py,
                                            const Window &win, const uint32_t *__restrict__ quad,
                                            const ViewSums &sums, float G[9]) {
    const uint32_t *__restrict__ img = quad + s.off;
    const float wmax = (float)s.W - 1.0f, hmax = (float)s.H - 1.0f;
    const float inv_sw = 1.0f / win.sw;
    const float mean = sums.ss * inv_sw;
    const float var_s = fmaxf(sums.sss * inv_sw - mean * mean, 0.0f);
    const float den = sqrtf(win.var * var_s);
    const float cov = sums.srs * inv_sw - win.mean * mean;
    const float A = 1.0f / den;
    const float B = cov * win.var * A * A * A;
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0f;
#pragma unroll
    for (int i = 0; i < kSamples; i++) {
        const float X = px + (float)((i % 6) * 2 - 5), Y = py + (float)((i / 6) * 2 - 5);
        const float qx = fmaf(H.m[0], X, fmaf(H.m[1], Y, H.m[2]));
        const float qy = fmaf(H.m[3], X, fmaf(H.m[4], Y, H.m[5]));
        const float qz = fmaf(H.m[6], X, fmaf(H.m[7], Y, H.m[8]));
        const float rz = 1.0f / qz;
        const float u = fminf(fmaxf(qx * rz, 0.0f), wmax), v = fminf(fmaxf(qy * rz, 0.0f), hmax);
        const float x0 = floorf(u), y0 = floorf(v);
        const float fx = u - x0, fy = v - y0;
        const uint32_t q = img[__umul24((uint32_t)y0, (uint32_t)s.W) + (uint32_t)x0];
        const float a = (float)(q & 0xffu), b = (float)((q >> 8) & 0xffu);
        const float c = (float)((q >> 16) & 0xffu), e = (float)(q >> 24);
        const float top = fmaf(fx, b - a, a), bot = fmaf(fx, e - c, c);
        const float val = fmaf(fy, bot - top, top);
        // d cost / d val_i
        const float dval = win.w[i] * inv_sw * (B * (val - mean) - (win.v[i] - win.mean) * A);
        const float gu = (b - a) * (1.0f - fy) + (e - c) * fy;
        const float gv = bot - top;
        const float gx = dval * gu * rz, gy = dval * gv * rz;
        const float gz = -(gx * u + gy * v);
        G[0] = fmaf(gx, X, G[0]);
        G[1] = fmaf(gx, Y, G[1]);
        G[2] += gx;
        G[3] = fmaf(gy, X, G[3]);
        G[4] = fmaf(gy, Y, G[4]);
        G[5] += gy;
        G[6] = fmaf(gz, X, G[6]);
        G[7] = fmaf(gz, Y, G[7]);
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
                                                          const float *__restrict__ weights,
                                                          float *__restrict__ out_cost,
                                                          float *__restrict__ out_gd,
                                                          float *__restrict__ out_gn,
                                                          double *__restrict__ pose_partials,
                                                          float step_depth, float step_normal,
                                                          int32_t apply) {
    PixelTask task = locate(plan, (int32_t)blockDim.x);
    const PmDesc &d = plan.descs[task.view];
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
    Window win;
    float rho = 0.0f;
    float costs[kMaxSources];
    uint32_t selected = 0u;
    float cost = kInvalid;
#pragma unroll
    for (int s = 0; s < kMaxSources; s++) costs[s] = INFINITY;
    if (valid) {
        prepare_window(pools.quad + d.off, d.W, d.H, x, y, win);
        rho = dot3(h.x, h.y, h.z, rx * h.w, ry * h.w, 1.0f * h.w);
        // pass 1: view costs, exactly as the forward pass, and the k smallest by (cost, index)
#pragma unroll 1
        for (int s = 0; s < d.n_src; s++) {
            Mat3 H;
            homography(d, d.src[s], h.x, h.y, h.z, rho, H);
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
    const double X0 = (double)rx * h.w, X1 = (double)ry * h.w, X2 = h.w;
    const double n0 = h.x, n1 = h.y, n2 = h.z;
    double dd = 0.0, dn0 = 0.0, dn1 = 0.0, dn2 = 0.0;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    double *partial = pose_partials ? pose_partials + (size_t)b * (kMaxSources * 12 + 1) : nullptr;
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
        // INVALID, degenerate (exactly 1) or clamped views carry no gradient
        bool smooth = valid && (selected & (1u << s)) && cs > 0.0f && cs < 2.0f && cs != 1.0f;
        if (smooth) {
            const PmSrc &src = d.src[s];
            Mat3 H;
            homography(d, src, h.x, h.y, h.z, rho, H);
            const ViewSums sums = view_sums(src, H, px, py, win, pools.quad);
            const float mean = sums.ss / win.sw;
            const float var_s = fmaxf(sums.sss / win.sw - mean * mean, 0.0f);
            if (!(sqrtf(win.var * var_s) < kDegenerateDen)) {
                float G[9];
                view_grad_h(src, H, px, py, win, pools.quad, sums, G);
                // G_M = Ks^T G Kr^-T
                double T[9];
                for (int j = 0; j < 3; j++) {
                    T[j] = (double)src.k[0] * G[j];
                    T[3 + j] = (double)src.k[1] * G[3 + j];
                    T[6 + j] = (double)src.k[2] * G[j] + (double)src.k[3] * G[3 + j] + G[6 + j];
                }
                double GM[9];
                for (int i = 0; i < 3; i++) {
                    GM[i * 3 + 0] = T[i * 3 + 0] * d.ikx + T[i * 3 + 2] * d.kx2;
                    GM[i * 3 + 1] = T[i * 3 + 1] * d.iky + T[i * 3 + 2] * d.ky2;
                    GM[i * 3 + 2] = T[i * 3 + 2];
                }
                const double t0 = src.t[0], t1 = src.t[1], t2 = src.t[2];
                const double rhod = rho, r2 = rhod * rhod;
                double gmn[3];
                for (int i = 0; i < 3; i++) gmn[i] = GM[i * 3] * n0 + GM[i * 3 + 1] * n1 + GM[i * 3 + 2] * n2;
                const double tGn = t0 * gmn[0] + t1 * gmn[1] + t2 * gmn[2];
                dd += -tGn * (double)nr / r2;
                dn0 += (t0 * GM[0] + t1 * GM[3] + t2 * GM[6]) / rhod - tGn * X0 / r2;
                dn1 += (t0 * GM[1] + t1 * GM[4] + t2 * GM[7]) / rhod - tGn * X1 / r2;
                dn2 += (t0 * GM[2] + t1 * GM[5] + t2 * GM[8]) / rhod - tGn * X2 / r2;
#pragma unroll
                for (int j = 0; j < 9; j++) gpose[j] = scale * GM[j];
#pragma unroll
                for (int j = 0; j < 3; j++) gpose[9 + j] = scale * gmn[j] / rhod;
            }
        }
        if (partial) {
#pragma unroll
            for (int j = 0; j < 12; j++) {
                const double sum = wave_sum(gpose[j]);
                if (lead) partial[1 + s * 12 + j] = sum;
            }
        }
    }
    if (partial) {
        const double ls = wave_sum(valid ? (double)cost * weight : 0.0);
        if (lead) partial[0] = ls;
    }
    if (!inside) return;
    const float gd = (float)(scale * dd);
    const float gn0 = (float)(scale * dn0), gn1 = (float)(scale * dn1), gn2 = (float)(scale * dn2);
    if (out_cost) out_cost[d.off + index] = cost;
    if (out_gd) out_gd[d.off + index] = gd;
    if (out_gn) {
        out_gn[(size_t)(d.off + index) * 3 + 0] = gn0;
        out_gn[(size_t)(d.off + index) * 3 + 1] = gn1;
        out_gn[(size_t)(d.off + index) * 3 + 2] = gn2;
    }
    if (apply && valid) {
        // one gradient step on the plane, kept valid (depth range, facing the camera)
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
        pools.hyp[d.off + index] = make_float4(mx, my, mz, nd);
    }
}

}  // namespace

hipError_t ncc_launch_backward(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                               const float *weights, float *cost, float *gd, float *gn,
                               double *pose_partials, float step_depth, float step_normal,
                               int32_t apply) {
    if (plan.tx * plan.ty != 64) return hipErrorInvalidConfiguration;
    dim3 block(plan.tx, plan.ty), grid(plan.blocks);
    hipLaunchKernelGGL(ncc_backward_kernel, grid, block, 0, s, plan, pools, weights, cost, gd, gn,
                       pose_partials, step_depth, step_normal, apply);
    return hipGetLastError();
}

}  // namespace lr3d
