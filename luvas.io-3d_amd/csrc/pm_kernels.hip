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
#include "pm_device.h"

namespace lr3d {

namespace {

// ---- kernels -------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) build_quad_kernel(const uint8_t *__restrict__ gray, int32_t W,
                                                         int32_t H, int32_t qrows, uint32_t *__restrict__ quad) {
    int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    int32_t x1 = min(x + 1, W - 1), y1 = min(y + 1, H - 1);
    uint32_t a = gray[(size_t)y * W + x], b = gray[(size_t)y * W + x1];
    uint32_t c = gray[(size_t)y1 * W + x], e = gray[(size_t)y1 * W + x1];
    quad[quad_index((uint32_t)x, (uint32_t)y, (uint32_t)qrows)] = a | (b << 8) | (c << 16) | (e << 24);
}

// PM:255-289 (init) and PM:296-333 (init_coarse): iteration 0, phase 0 (DS:192-201)
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
    prepare_window(pools.quad + d.qoff, d.W, d.H, d.qrows, x, y, pools.wtab, win);
    const float rx = (px - d.k[2]) / d.k[0];
    const float ry = (py - d.k[3]) / d.k[1];
    float nx, ny, nz;
    random_normal(rng, rx, ry, nx, ny, nz);
    float depth = random_depth(rng, d);
    float c;
    if (d.from_coarse) {
        // PM:307-330
        c = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, nx, ny, nz, depth,
                                             INFINITY);
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
                                                          h.z, cd, c);
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
        c = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, nx, ny, nz, depth,
                                             INFINITY);
        if ((d.flags & 1u) == 0u && d.has_prior) {
            float pr = pools.prior[d.prior_off + index];
            if (pr > 0.0f) {
                float pnx = rx, pny = ry, pnz = 1.0f;
                normalize3(pnx, pny, pnz);
                pnx = -pnx;
                pny = -pny;
                pnz = -pnz;
                float pc = hypothesis_cost<KMAX, GEOMETRIC>(d, pools, px, py, rx, ry, win, pnx, pny,
                                                            pnz, pr, c);
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
template <int KMAX, bool GEOMETRIC>
__global__ void __launch_bounds__(256) pm_propagate_kernel(PmPlan plan, PmPools pools,
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
    prepare_window(pools.quad + d.qoff, W, H, d.qrows, x, y, pools.wtab, win);
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
                                                            tz, td, best_cost);
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

hipError_t pm_build_quad(hipStream_t s, const uint8_t *gray, int32_t W, int32_t H, int32_t qrows, uint32_t *quad) {
    if (qrows < H) return hipErrorInvalidValue;
    dim3 block(64, 4);
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(build_quad_kernel, grid, block, 0, s, gray, W, H, qrows, quad);
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
