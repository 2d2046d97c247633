// TSDF depth-map integration kernels for gfx950.
//
// Semantics follow the reference's `integrate` WGSL kernel and its host culling
// (frontend/src/app/vision/fusion/tsdf-volume.ts, cited TV:<line>) under the FP32 arithmetic
// definition of the PatchMatch kernels (no contraction, fmaf only where written, IEEE division
// and sqrt, ties-to-even rounding); the brick culling runs in double like the reference's
// JavaScript.
//
// Layout and mapping: the pool keeps the reference's brick layout (voxel (x, y, z) of brick b at
// b*512 + x + 8y + 64z).  One 64-lane wavefront owns one brick: lane = x + 8y, and each lane walks
// its z-column, so every z step is one contiguous 256-byte access per array.  A pass loads a
// brick's 8 tsdf/weight values per lane once, applies every visible view in order in registers,
// and stores the voxels that changed: per voxel the update sequence is exactly the reference's
// one-view-at-a-time sequence, so one pass over n views equals n single-view passes.
#include "lr3d_internal.h"
#include "fp_exact.h"

#include <math.h>

#include <algorithm>

namespace lr3d {

namespace {

#ifndef LR3D_TSDF_VIEW_GROUP
#define LR3D_TSDF_VIEW_GROUP 2
#endif

constexpr float kMaxWeight = 96.0f;   // TV:7
// Wavefronts of a list pass's grid (integrate_kernel, LIST): enough to fill 256 CUs at the pass's
// 5 waves per SIMD, with a margin.
constexpr uint32_t kIntegrateWaves = 8192;
constexpr float kPriorWeight = 0.12f; // TV:8

__device__ __forceinline__ float f16_bits_to_f32(uint32_t bits) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(bits & 0xffffu));
}

// PK:7-15 (same arithmetic as the PatchMatch kernels)
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
    float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
    nx = nx * inv;
    ny = ny * inv;
    nz = nz * inv;
}

// Largest depth pixel i of view V can give a voxel: the map's depth where positive, else the prior
// where the view has one (TV:122-126); 0 where no update can come from the pixel.
__device__ __forceinline__ float usable_depth(const TsdfView &V, const uint32_t *__restrict__ maps,
                                              const float *__restrict__ priors, uint32_t i) {
    float d = f16_bits_to_f32(maps[V.map_off + i]);
    if (!(d > 0.0f)) d = V.has_prior ? priors[V.prior_off + i] : 0.0f;
    return d > 0.0f ? d : 0.0f;
}

// Depth tiles: one 16x16 block per tile of kTsdfTile^2 pixels, grid (max tw, max th, views).
__global__ void __launch_bounds__(256) depth_tile_kernel(const TsdfView *__restrict__ views,
                                                         const uint32_t *__restrict__ maps,
                                                         const float *__restrict__ priors,
                                                         float *__restrict__ tiles) {
    const TsdfView &V = views[blockIdx.z];
    if ((int32_t)blockIdx.x >= V.tw || (int32_t)blockIdx.y >= V.th) return;  // whole block
    const int32_t x = (int32_t)(blockIdx.x * kTsdfTile + threadIdx.x);
    const int32_t y = (int32_t)(blockIdx.y * kTsdfTile + threadIdx.y);
    float d = 0.0f;
    if (x < V.W && y < V.H) d = usable_depth(V, maps, priors, (uint32_t)y * (uint32_t)V.W + (uint32_t)x);
    for (int off = 32; off > 0; off >>= 1) d = fmaxf(d, __shfl_xor(d, off, 64));
    __shared__ float part[4];
    const uint32_t t = threadIdx.y * kTsdfTile + threadIdx.x;
    if ((t & 63u) == 0) part[t >> 6] = d;
    __syncthreads();
    if (t == 0)
        tiles[V.tile_off + blockIdx.y * (uint32_t)V.tw + blockIdx.x] =
            fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
}

// Coarse tiles: one block per view, one thread per coarse tile (max of kTsdfTile2^2 tiles).
__global__ void __launch_bounds__(256) depth_tile2_kernel(const TsdfView *__restrict__ views,
                                                          float *__restrict__ tiles) {
    const TsdfView &V = views[blockIdx.x];
    for (int32_t j = (int32_t)threadIdx.x; j < V.tw2 * V.th2; j += (int32_t)blockDim.x) {
        const int32_t cx = j % V.tw2, cy = j / V.tw2;
        float m = 0.0f;
        for (int32_t ty = cy * kTsdfTile2; ty < min(V.th, (cy + 1) * kTsdfTile2); ty++)
            for (int32_t tx = cx * kTsdfTile2; tx < min(V.tw, (cx + 1) * kTsdfTile2); tx++)
                m = fmaxf(m, tiles[V.tile_off + (uint32_t)(ty * V.tw + tx)]);
        tiles[V.tile2_off + (uint32_t)j] = m;
    }
}

// Whether any voxel of brick c can take an update from view V.  False only when the brick's voxel
// centres all lie at camera depth >= 0.02 and either their projections miss the image or the
// largest usable depth over the tiles their pixels fall in is below the brick's nearest depth
// minus the truncation, so that every voxel meets sdf < -mu (TV:140).  Computed in double over the
// 8 corners of the box of voxel centres (the projection of the box lies in the hull of the
// projected corners), with margins of one pixel and 1e-4 m + 1e-6 z, far above the FP32 rounding of
// a voxel's projection and depth in integrate_kernel.
__device__ bool depth_reachable(const TsdfGridDev &g, const TsdfView &V, int4 c, const float *__restrict__ tiles) {
    double zmin = 1e300, umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        const double x = g.origin_d[0] + ((double)(c.x * 8) + ((k & 1) ? 7.5 : 0.5)) * g.voxel_d;
        const double y = g.origin_d[1] + ((double)(c.y * 8) + ((k & 2) ? 7.5 : 0.5)) * g.voxel_d;
        const double z = g.origin_d[2] + ((double)(c.z * 8) + ((k & 4) ? 7.5 : 0.5)) * g.voxel_d;
        const double p0 = (V.Rd[0] * x + V.Rd[1] * y + V.Rd[2] * z) + V.td[0];
        const double p1 = (V.Rd[3] * x + V.Rd[4] * y + V.Rd[5] * z) + V.td[1];
        const double p2 = (V.Rd[6] * x + V.Rd[7] * y + V.Rd[8] * z) + V.td[2];
        if (!(p2 >= 0.02)) return true;
        const double u = (V.kd[0] * p0) / p2 + V.kd[2];
        const double v = (V.kd[1] * p1) / p2 + V.kd[3];
        zmin = fmin(zmin, p2);
        umin = fmin(umin, u);
        umax = fmax(umax, u);
        vmin = fmin(vmin, v);
        vmax = fmax(vmax, v);
    }
    const double fx0 = fmax(0.0, floor(umin) - 1.0), fx1 = fmin((double)(V.W - 1), ceil(umax) + 1.0);
    const double fy0 = fmax(0.0, floor(vmin) - 1.0), fy1 = fmin((double)(V.H - 1), ceil(vmax) + 1.0);
    if (fx0 > fx1 || fy0 > fy1) return false;  // every projection is outside the image
    int32_t tx0 = (int32_t)fx0 / kTsdfTile, tx1 = (int32_t)fx1 / kTsdfTile;
    int32_t ty0 = (int32_t)fy0 / kTsdfTile, ty1 = (int32_t)fy1 / kTsdfTile;
    float dmax = 0.0f;
    if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) <= 64) {
        for (int32_t ty = ty0; ty <= ty1; ty++)
            for (int32_t tx = tx0; tx <= tx1; tx++) dmax = fmaxf(dmax, tiles[V.tile_off + (uint32_t)(ty * V.tw + tx)]);
    } else {
        tx0 /= kTsdfTile2;
        tx1 /= kTsdfTile2;
        ty0 /= kTsdfTile2;
        ty1 /= kTsdfTile2;
        if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) > 64) return true;
        for (int32_t ty = ty0; ty <= ty1; ty++)
            for (int32_t tx = tx0; tx <= tx1; tx++) dmax = fmaxf(dmax, tiles[V.tile2_off + (uint32_t)(ty * V.tw2 + tx)]);
    }
    if (!(dmax > 0.0f)) return false;
    return (double)dmax >= zmin - g.mu_d - (1e-4 + 1e-6 * zmin);
}

// TV:220-234, one lane per (brick, view).  A group of 2^lane_shift consecutive lanes holds one
// 32-view word of one brick (bit j of the word = view 32*word + j visible): 32 lanes when a pass
// has more than 32 views, else the number of views rounded up to a power of two, so that a pass
// of one view does not leave 31 lanes in 32 idle.  With depth tiles the reference's test is
// followed by depth_reachable, which drops only (brick, view) pairs that update no voxel.
__global__ void __launch_bounds__(256) visibility_kernel(TsdfGridDev g, const int4 *__restrict__ coords,
                                                         const TsdfView *__restrict__ views,
                                                         int32_t n_views, uint32_t *__restrict__ vis,
                                                         int32_t words, const float *__restrict__ tiles,
                                                         uint32_t *__restrict__ ctrs, uint32_t lane_shift) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0u) ctrs[0] = 0u;  // the brick list's length, counted by brick_list_kernel
    const uint32_t group = 1u << lane_shift;
    const uint32_t pair = gid >> lane_shift;
    const uint32_t brick = pair / (uint32_t)words;
    const int32_t word = (int32_t)(pair % (uint32_t)words);
    const int32_t v = word * 32 + (int32_t)(gid & (group - 1u));
    bool visible = false;
    if (brick < (uint32_t)g.n_bricks && v < n_views) {
        const int4 c = coords[brick];
        const double side = g.voxel_d * 8.0;
        const double radius = g.voxel_d * 8.0 * 0.87 + g.mu_d;
        const double cx = g.origin_d[0] + ((double)c.x + 0.5) * side;
        const double cy = g.origin_d[1] + ((double)c.y + 0.5) * side;
        const double cz = g.origin_d[2] + ((double)c.z + 0.5) * side;
        const TsdfView &V = views[v];
        const double p0 = (V.Rd[0] * cx + V.Rd[1] * cy + V.Rd[2] * cz) + V.td[0];
        const double p1 = (V.Rd[3] * cx + V.Rd[4] * cy + V.Rd[5] * cz) + V.td[1];
        const double p2 = (V.Rd[6] * cx + V.Rd[7] * cy + V.Rd[8] * cz) + V.td[2];
        const double u = (V.kd[0] * p0) / p2 + V.kd[2];
        const double w = (V.kd[1] * p1) / p2 + V.kd[3];
        const double z = p2;
        visible = !(z < 0.02 - radius || z > V.max_depth + radius);
        if (visible) {
            const double margin = (V.kd[0] * radius) / fmax(0.02, z);
            visible = !(u < -margin || w < -margin || u > (double)V.W + margin || w > (double)V.H + margin);
        }
        if (visible && tiles) visible = depth_reachable(g, V, c, tiles);
    }
    const uint64_t bits = __ballot(visible);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane & (group - 1u)) == 0u && brick < (uint32_t)g.n_bricks) {
        // the group starts at this lane: its bits are the ballot's bits lane .. lane + group - 1
        const uint32_t mask = group == 32u ? 0xffffffffu : (1u << group) - 1u;
        vis[(size_t)brick * words + word] = (uint32_t)(bits >> lane) & mask;
    }
}

// Depth samples of one view for a lane's 8 voxels: the pixel index (-1 where the voxel is skipped
// before the lookup: cam.z <= 0.01 or outside the image) and the packed map value there.
constexpr int kViewGroup = LR3D_TSDF_VIEW_GROUP;  // views gathered before their updates are applied

struct ViewGather {
    int32_t idx[8];
    uint32_t packed[8];
    float prior[8];  // the prior at the pixel (views with a prior), else 0
};

// Correctly rounded forms of the pass's divisions and square roots (fp_exact.h) are used when the
// host finds the pass inside their ranges (FAST; integrate_fast_ok): every view's camera
// coordinates of the grid's voxel centres below 2^40 in magnitude, |fx|, |fy| below 2^40, and mu in
// [2^-40, 2^40].  Then, with the same bits as the IEEE forms:
//   - the projection divides k*cam (below 2^80) by zc in [0.01, 2^40) with one shared reciprocal.
//     Only a quotient below 2^-70 in magnitude can have a residual that underflows and a last bit
//     that differs; it is then added to the principal point c: either |c| >= 2^-40 and both sums
//     round to c, or both sums are below 0.5 and round to pixel 0;
//   - sdf / mu: sdf = d - cam.z with d > 0 and cam.z > 0.01 is 0 or at least 2^-30 in magnitude
//     (both are multiples of 2^-30 when d >= 2^-7, and |sdf| > 0.002 otherwise), so quotients below
//     1 stay normal; quotients from 1 up, and the NaN an overflowing intermediate gives, all clamp to
//     1 in fminf(1, .);
//   - 1 / sqrt(|cam|^2) has |cam|^2 in [1e-4, 3 * 2^80];
//   - the packed normal's decode (two divisions by 63, a square root and a division) is looked up
//     in a 4096-entry table formed with the same IEEE arithmetic (normal_table_kernel).
// The running mean's division keeps the IEEE form: uploaded volumes may hold any values.

// TV:110-121 up to the map lookup.  cam = R world + t, row i = fmaf(R_i2, z, fmaf(R_i1, y, fmaf(R_i0, x, t_i))).
// Every lane loads (a skipped voxel reads pixel 0 and is marked idx = -1), so the loads of a view
// are straight-line code and the waits before their use count only the loads issued after them.
// PRIORS: the pass has views with a prior; the prior at the pixel is gathered for those views.
template <bool PRIORS, bool FAST>
__device__ __forceinline__ void gather_view(const TsdfView &V, const uint32_t *__restrict__ maps,
                                            const float *__restrict__ priors, float wx, float wy, const float wz[8],
                                            ViewGather &out) {
    const uint32_t *__restrict__ map = maps + V.map_off;
    const float c0 = fmaf(V.R[1], wy, fmaf(V.R[0], wx, V.t[0]));
    const float c1 = fmaf(V.R[4], wy, fmaf(V.R[3], wx, V.t[1]));
    const float c2 = fmaf(V.R[7], wy, fmaf(V.R[6], wx, V.t[2]));
#pragma unroll
    for (int z = 0; z < 8; z++) {
        const float cam0 = fmaf(V.R[2], wz[z], c0);
        const float cam1 = fmaf(V.R[5], wz[z], c1);
        const float cam2 = fmaf(V.R[8], wz[z], c2);
        const bool front = cam2 > 0.01f;
        const float zc = front ? cam2 : 1.0f;
        float qu, qv;
        if (FAST) {
            const float y = rcp_exact(zc);
            qu = div_rn(V.k[0] * cam0, zc, y);
            qv = div_rn(V.k[1] * cam1, zc, y);
        } else {
            qu = (V.k[0] * cam0) / zc;
            qv = (V.k[1] * cam1) / zc;
        }
        const int32_t px = (int32_t)rintf(qu + V.k[2]);
        const int32_t py = (int32_t)rintf(qv + V.k[3]);
        const bool in = front && px >= 0 && py >= 0 && px < V.W && py < V.H;
        const uint32_t i = in ? (uint32_t)py * (uint32_t)V.W + (uint32_t)px : 0u;
        out.idx[z] = in ? (int32_t)i : -1;
        out.packed[z] = map[i];
        out.prior[z] = 0.0f;
        if (PRIORS && V.has_prior) out.prior[z] = priors[V.prior_off + i];  // V is wave-uniform
    }
}

// TV:122-153 from the map lookup on: the update of a lane's 8 voxels by one view.
template <bool COUNT, bool FAST>
__device__ __forceinline__ void apply_view(const TsdfView &V, const ViewGather &gv, float mu, float inv_mu,
                                           const float4 *__restrict__ ntab, float wx, float wy,
                                           const float wz[8], float ts[8], float ws[8], uint32_t &changed,
                                           uint32_t &updates) {
    const float c0 = fmaf(V.R[1], wy, fmaf(V.R[0], wx, V.t[0]));
    const float c1 = fmaf(V.R[4], wy, fmaf(V.R[3], wx, V.t[1]));
    const float c2 = fmaf(V.R[7], wy, fmaf(V.R[6], wx, V.t[2]));
#pragma unroll
    for (int z = 0; z < 8; z++) {
        if (gv.idx[z] < 0) continue;
        const float cam0 = fmaf(V.R[2], wz[z], c0);
        const float cam1 = fmaf(V.R[5], wz[z], c1);
        const float cam2 = fmaf(V.R[8], wz[z], c2);
        const uint32_t packed = gv.packed[z];
        float d = f16_bits_to_f32(packed);
        const bool from_map = d > 0.0f;
        if (!from_map) d = V.has_prior ? gv.prior[z] : 0.0f;
        if (!(d > 0.0f)) continue;
        const float sdf = d - cam2;
        if (sdf < -mu) continue;
        // The weight only matters for an update, so it is evaluated after the skips.
        float w;
        if (from_map) {
            float nx, ny, nz, inv;
            const float ss = cam0 * cam0 + cam1 * cam1 + cam2 * cam2;
            if (FAST) {
                const float4 n = ntab[(packed >> 16u) & 4095u];
                nx = n.x;
                ny = n.y;
                nz = n.z;
                inv = rcp_exact(sqrt_normal(ss));
            } else {
                unpack_normal(packed, nx, ny, nz);
                inv = 1.0f / sqrtf(ss);
            }
            w = fmaxf(0.2f, fabsf(nx * (cam0 * inv) + ny * (cam1 * inv) + nz * (cam2 * inv)));
            if (V.confidence > 0.0f) {
                const float cc = (float)(packed >> 28u) / 8.0f;
                const float r = fminf(fmaxf(V.confidence / d, 0.5f), 2.0f);
                w = w * ((fmaxf(0.15f, (1.0f - cc) * (1.0f - cc)) * r) * r);
            }
        } else {
            w = kPriorWeight;
        }
        const float observed = fminf(1.0f, FAST ? div_rn(sdf, mu, inv_mu) : sdf / mu);
        const float w0 = ws[z];
        ts[z] = (ts[z] * w0 + observed * w) / (w0 + w);
        ws[z] = fminf(w0 + w, kMaxWeight);
        changed |= 1u << z;
        if (COUNT) updates++;
    }
}

// The decoded normal of every packed octahedral code: entry i = unpack_normal of the code with
// x = i % 64 and y = i / 64 (bits 16-27 of a packed sample), in the same IEEE arithmetic.
__global__ void __launch_bounds__(256) normal_table_kernel(float4 *__restrict__ tab) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4096u) return;
    float nx, ny, nz;
    unpack_normal(i << 16u, nx, ny, nz);
    tab[i] = make_float4(nx, ny, nz, 0.0f);
}

// Bricks a pass reads: the pool indices of the bricks with at least one kept view, written to
// list[0 .. ctrs[0]) (order within each run of 64 pool bricks kept; runs in any order).  ctrs[0]
// (the list length) is zeroed by the culling kernel before this one runs.
__global__ void __launch_bounds__(256) brick_list_kernel(const uint32_t *__restrict__ vis, int32_t words,
                                                         int32_t n_bricks, uint32_t *__restrict__ ctrs,
                                                         uint32_t *__restrict__ list) {
    const uint32_t brick = blockIdx.x * blockDim.x + threadIdx.x;
    bool kept = false;
    if (brick < (uint32_t)n_bricks) {
        const uint32_t *__restrict__ w = vis + (size_t)brick * words;
        for (int32_t k = 0; k < words; k++) kept = kept || w[k] != 0u;
    }
    const uint64_t mask = __ballot(kept);
    if (mask == 0u) return;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t base = 0u;
    if (lane == 0u) base = atomicAdd(ctrs, (uint32_t)__popcll(mask));
    base = __shfl(base, 0, 64);
    if (kept) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = brick;
}

// TV:103-154: one integration pass, one wavefront per brick at a time; lane = (x, y), each lane walks
// its z column, all visible views of the brick in order.  Two schedules:
//   LIST = false  one wavefront per pool brick (a brick without a kept view leaves at once).
//                 Hardware block i runs on XCD i % 8; blocks are dealt to the XCDs in chunks of
//                 `chunk` consecutive blocks, so neighbouring bricks run on one XCD and share the
//                 depth-map lines of their footprints in its L2, while every XCD still takes work
//                 from the whole pool (chunk 1: round robin; the blocks past the last whole round keep
//                 their index).  The hardware hands bricks to free wavefronts, which balances passes
//                 where bricks see very different numbers of views.
//   LIST = true   a fixed grid walks the list of bricks with a kept view (brick_list_kernel), for
//                 passes that reach few bricks (one view per pass, as the reference integrates
//                 keyframes): no wavefront per empty pool brick.  Block b (XCD b % 8) serves queue
//                 q = b % 8, which holds the list chunks q, q + 8, q + 16, ... of `chunk` entries;
//                 the queue's wavefronts take its entries in turn (static striding: no atomics, which
//                 on this memory are resolved beyond the L2).
// Which wavefront updates a brick changes no value: each brick is owned by one wavefront, which
// applies its views in order.  Visible views are taken kViewGroup at a time: the group's depth
// samples are all gathered before the first view's updates, so the later views' gathers are in
// flight while the earlier ones are applied (a brick's chain of views is not one gather latency per
// view); the updates are the same sequence as applying the views one at a time.  Groups of 2
// measured fastest (4 and 6 hold more registers).
template <bool COUNT, bool PRIORS, bool FAST, bool LIST>
__global__ void __launch_bounds__(256) integrate_kernel(TsdfGridDev g, const TsdfView *__restrict__ views,
                                                        int32_t n_views, const uint32_t *__restrict__ vis,
                                                        int32_t words, const int4 *__restrict__ coords,
                                                        const uint32_t *__restrict__ maps,
                                                        const float *__restrict__ priors,
                                                        const float4 *__restrict__ ntab,
                                                        float *__restrict__ tsdf, float *__restrict__ weight,
                                                        unsigned long long *__restrict__ counter,
                                                        const uint32_t *__restrict__ ctrs,
                                                        const uint32_t *__restrict__ list, uint32_t chunk) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6, waves_per_block = blockDim.x / 64u;
    // LIST: the queue, the list length and the queue's wavefront count; else the block's bricks
    const uint32_t q = blockIdx.x % 8u;
    const uint32_t count = LIST ? __builtin_amdgcn_readfirstlane(ctrs[0]) : 0u;
    const uint32_t stride = (gridDim.x / 8u) * waves_per_block;
    uint32_t block = blockIdx.x;
    if (!LIST) {
        const uint32_t per_round = 8u * chunk;
        if (chunk > 1u && block < gridDim.x / per_round * per_round) {
            const uint32_t round = block / per_round, r = block % per_round;
            block = round * per_round + (r % 8u) * chunk + r / 8u;
        }
    }
    uint32_t updates = 0, written = 0;       // per lane
    uint32_t views_applied = 0, bricks = 0;  // wave-uniform
    for (uint32_t k = (blockIdx.x / 8u) * waves_per_block + wave, first = 1u;; k += stride, first = 0u) {
        uint32_t brick;
        if (LIST) {
            const uint32_t pos = ((k / chunk) * 8u + q) * chunk + k % chunk;  // increasing in k
            if (pos >= count) break;
            brick = __builtin_amdgcn_readfirstlane(list[pos]);
        } else {
            brick = __builtin_amdgcn_readfirstlane(block * waves_per_block + wave);
            if (!first || brick >= (uint32_t)g.n_bricks) break;
        }
        // the brick, its visibility words and the views are wave-uniform (scalar loads, and the
        // prior branch below is decided per view, not per lane)
        const int4 c = coords[brick];
        const int32_t vx = (int32_t)(lane & 7u), vy = (int32_t)(lane >> 3);
        const float wx = g.origin[0] + ((float)(c.x * 8 + vx) + 0.5f) * g.voxel;
        const float wy = g.origin[1] + ((float)(c.y * 8 + vy) + 0.5f) * g.voxel;
        float wz[8];
#pragma unroll
        for (int z = 0; z < 8; z++) wz[z] = g.origin[2] + ((float)(c.z * 8 + z) + 0.5f) * g.voxel;
        const float mu = g.mu;
        const size_t base = (size_t)brick * 512u + lane;
        const uint32_t *__restrict__ bvis = vis + (size_t)brick * words;

        // visible views in order (bit j of word k = view 32k + j)
        int32_t word = 0;
        uint32_t bits = words > 0 ? __builtin_amdgcn_readfirstlane(bvis[0]) : 0u;
        auto next_view = [&]() -> int32_t {
            while (bits == 0u) {
                if (++word >= words) return -1;
                bits = __builtin_amdgcn_readfirstlane(bvis[word]);
            }
            const int32_t v = __builtin_amdgcn_readfirstlane(word * 32 + (int32_t)__builtin_ctz(bits));
            bits &= bits - 1u;
            return v;
        };
        int32_t v = next_view();
        if (v < 0) continue;  // no view reaches this brick: nothing read, nothing written
        bricks++;

        float ts[8], ws[8];
#pragma unroll
        for (int z = 0; z < 8; z++) {
            ts[z] = tsdf[base + 64u * z];
            ws[z] = weight[base + 64u * z];
        }
        uint32_t changed = 0;  // bit z: voxel z of this lane was updated
        while (v >= 0) {
            int32_t vs[kViewGroup];
            int32_t n = 1;
            vs[0] = v;
            for (; n < kViewGroup; n++) {
                vs[n] = next_view();
                if (vs[n] < 0) break;
            }
            views_applied += (uint32_t)n;
            ViewGather grp[kViewGroup];
#pragma unroll
            for (int j = 0; j < kViewGroup; j++)
                if (j < n) gather_view<PRIORS, FAST>(views[vs[j]], maps, priors, wx, wy, wz, grp[j]);
#pragma unroll
            for (int j = 0; j < kViewGroup; j++)
                if (j < n)
                    apply_view<COUNT, FAST>(views[vs[j]], grp[j], mu, g.inv_mu, ntab, wx, wy, wz, ts, ws, changed,
                                            updates);
            v = n == kViewGroup ? next_view() : -1;
        }
#pragma unroll
        for (int z = 0; z < 8; z++) {
            if (changed & (1u << z)) {
                tsdf[base + 64u * z] = ts[z];
                weight[base + 64u * z] = ws[z];
            }
        }
        if (COUNT) written += (uint32_t)__builtin_popcount(changed);
    }
    if (COUNT) {
        // wavefront sums, one atomic each: voxel updates, (voxel, view) pairs evaluated (512 per
        // visible view of a brick), bricks read, voxels written
        for (int offset = 32; offset > 0; offset >>= 1) {
            updates += __shfl_down(updates, offset, 64);
            written += __shfl_down(written, offset, 64);
        }
        if (lane == 0) {
            if (updates) atomicAdd(counter, (unsigned long long)updates);
            if (views_applied) atomicAdd(counter + 1, 512ull * views_applied);
            if (bricks) atomicAdd(counter + 2, (unsigned long long)bricks);
            if (written) atomicAdd(counter + 3, (unsigned long long)written);
        }
    }
}

// ---- adjoint (no counterpart in the reference) ------------------------------------------------
//
// L = sum_q (tsdf_q - target_q)^2 over listed voxels q, where tsdf_q is the value after applying
// views 0..n-1 in order to a fresh voxel (tsdf 1, weight 0) exactly as integrate_kernel does
// (same culling bits, same FP32 arithmetic).  One thread per listed voxel:
//   forward   replays every view, storing (tsdf, weight) before every 16th view (a checkpoint) in
//             `ckpt` ([segment][voxel]);
//   reverse   per segment, last first: replays the segment from its checkpoint with the state
//             before each view in LDS, then walks it backwards with a = dL/dtsdf, b = dL/dweight
//             (FP64):  D = W + w, u = [D < 96],
//               dL/do = a w / D,  dL/dw = a (o - tsdf) / D + u b,
//               a <- a W / D,     b <- a (tsdf_prev - tsdf) / D + u b.
// dL/do and dL/dw are chained to the depth used at the pixel (map depth, or the prior where the
// map has none) and to cam = R world + t.  Depth gradients are added per pixel and pose gradients
// per view (wavefront sums first) with FP64 atomics: the result does not depend on the order
// beyond FP64 rounding.  The nearest-pixel choice, the skips and the clamps carry no gradient.

constexpr int kSegment = 16;

struct StepOut {
    float o, w;
    uint32_t pixel;      // index into the view's map
    double do_dd, dw_dd; // partial derivatives of (o, w) with respect to the depth used
    double do_dz;        // d o / d cam.z
    double dw_dc[3];     // d w / d cam
    float cam[3];
};

// One view applied to one voxel (TV:110-153 in integrate_kernel's arithmetic); false if skipped.
template <bool DERIV>
__device__ __forceinline__ bool tsdf_step(const TsdfView &V, const uint32_t *__restrict__ maps,
                                          const float *__restrict__ priors, float mu, float wx, float wy,
                                          float wz, StepOut &out) {
    const float c0 = fmaf(V.R[1], wy, fmaf(V.R[0], wx, V.t[0]));
    const float c1 = fmaf(V.R[4], wy, fmaf(V.R[3], wx, V.t[1]));
    const float c2 = fmaf(V.R[7], wy, fmaf(V.R[6], wx, V.t[2]));
    const float cam0 = fmaf(V.R[2], wz, c0);
    const float cam1 = fmaf(V.R[5], wz, c1);
    const float cam2 = fmaf(V.R[8], wz, c2);
    if (!(cam2 > 0.01f)) return false;
    const int32_t px = (int32_t)rintf((V.k[0] * cam0) / cam2 + V.k[2]);
    const int32_t py = (int32_t)rintf((V.k[1] * cam1) / cam2 + V.k[3]);
    if (px < 0 || py < 0 || px >= V.W || py >= V.H) return false;
    const uint32_t i = (uint32_t)py * (uint32_t)V.W + (uint32_t)px;
    const uint32_t packed = maps[V.map_off + i];
    float d = f16_bits_to_f32(packed);
    const bool from_map = d > 0.0f;
    if (!from_map) d = V.has_prior ? priors[V.prior_off + i] : 0.0f;
    if (!(d > 0.0f)) return false;
    const float sdf = d - cam2;
    if (sdf < -mu) return false;
    float w;
    double dbase[3] = {0.0, 0.0, 0.0}, f = 1.0, df_dd = 0.0, base_d = 0.0;
    if (from_map) {
        float nx, ny, nz;
        unpack_normal(packed, nx, ny, nz);
        const float inv = 1.0f / sqrtf(cam0 * cam0 + cam1 * cam1 + cam2 * cam2);
        const float nc = nx * (cam0 * inv) + ny * (cam1 * inv) + nz * (cam2 * inv);
        w = fmaxf(0.2f, fabsf(nc));
        if (DERIV) {
            base_d = (double)w;
            if (fabsf(nc) > 0.2f) {
                const double len = sqrt((double)cam0 * cam0 + (double)cam1 * cam1 + (double)cam2 * cam2);
                const double cx = cam0 / len, cy = cam1 / len, cz = cam2 / len;
                const double ncd = nx * cx + ny * cy + nz * cz;
                const double sg = ncd >= 0.0 ? 1.0 : -1.0;
                dbase[0] = sg * (nx - ncd * cx) / len;
                dbase[1] = sg * (ny - ncd * cy) / len;
                dbase[2] = sg * (nz - ncd * cz) / len;
            }
        }
        if (V.confidence > 0.0f) {
            const float cc = (float)(packed >> 28u) / 8.0f;
            const float ratio = V.confidence / d;
            const float r = fminf(fmaxf(ratio, 0.5f), 2.0f);
            const float q = fmaxf(0.15f, (1.0f - cc) * (1.0f - cc));
            w = w * ((q * r) * r);
            if (DERIV) {
                f = (double)q * r * r;
                if (ratio > 0.5f && ratio < 2.0f)
                    df_dd = (double)q * 2.0 * r * (-(double)V.confidence / ((double)d * d));
            }
        }
    } else {
        w = kPriorWeight;
    }
    out.o = fminf(1.0f, sdf / mu);
    out.w = w;
    out.pixel = i;
    if (DERIV) {
        const bool open = sdf / mu < 1.0f;
        out.do_dd = open ? 1.0 / (double)mu : 0.0;
        out.do_dz = open ? -1.0 / (double)mu : 0.0;
        out.dw_dd = base_d * df_dd;
        for (int k = 0; k < 3; k++) out.dw_dc[k] = dbase[k] * f;
        out.cam[0] = cam0;
        out.cam[1] = cam1;
        out.cam[2] = cam2;
    }
    return true;
}

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(256) tsdf_backward_kernel(
    TsdfGridDev g, const TsdfView *__restrict__ views, int32_t n_views, const uint32_t *__restrict__ vis,
    int32_t words, const int4 *__restrict__ coords, const uint32_t *__restrict__ maps,
    const float *__restrict__ priors, const uint32_t *__restrict__ voxels, const float *__restrict__ targets,
    uint32_t n_voxels, const uint32_t *__restrict__ pixel_off, float2 *__restrict__ ckpt,
    float *__restrict__ tsdf_out, double *__restrict__ grad_depth, double *__restrict__ grad_pose,
    double *__restrict__ loss_partials) {
    __shared__ float2 seg_state[kSegment][256];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool active = q < n_voxels;
    const uint32_t index = active ? voxels[q] : 0u;
    const uint32_t brick = index >> 9, local = index & 511u;
    const int4 c = coords[brick];
    const float wx = g.origin[0] + ((float)(c.x * 8 + (int32_t)(local & 7u)) + 0.5f) * g.voxel;
    const float wy = g.origin[1] + ((float)(c.y * 8 + (int32_t)((local >> 3) & 7u)) + 0.5f) * g.voxel;
    const float wz = g.origin[2] + ((float)(c.z * 8 + (int32_t)(local >> 6)) + 0.5f) * g.voxel;
    const float mu = g.mu;
    const int32_t n_seg = (n_views + kSegment - 1) / kSegment;
    auto visible = [&](int32_t v) -> bool {
        return active && ((vis[(size_t)brick * words + (v >> 5)] >> (v & 31)) & 1u);
    };
    // forward with checkpoints
    float ts = 1.0f, ws = 0.0f;
    for (int32_t v = 0; v < n_views; v++) {
        if ((v % kSegment) == 0 && active) ckpt[(size_t)(v / kSegment) * n_voxels + q] = make_float2(ts, ws);
        StepOut st;
        if (visible(v) && tsdf_step<false>(views[v], maps, priors, mu, wx, wy, wz, st)) {
            const float w0 = ws;
            ts = (ts * w0 + st.o * st.w) / (w0 + st.w);
            ws = fminf(w0 + st.w, kMaxWeight);
        }
    }
    const double r = active ? (double)ts - (double)targets[q] : 0.0;
    if (active && tsdf_out) tsdf_out[q] = ts;
    const double l = wave_sum_d(r * r);
    if (lane == 0) loss_partials[blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6)] = l;
    // reverse
    double a = 2.0 * r, b = 0.0;
    for (int32_t sgm = n_seg - 1; sgm >= 0; sgm--) {
        const int32_t v0 = sgm * kSegment, v1 = min(n_views, v0 + kSegment);
        float2 state = active ? ckpt[(size_t)sgm * n_voxels + q] : make_float2(1.0f, 0.0f);
        for (int32_t v = v0; v < v1; v++) {
            seg_state[v - v0][threadIdx.x] = state;
            StepOut st;
            if (visible(v) && tsdf_step<false>(views[v], maps, priors, mu, wx, wy, wz, st)) {
                const float w0 = state.y;
                state.x = (state.x * w0 + st.o * st.w) / (w0 + st.w);
                state.y = fminf(w0 + st.w, kMaxWeight);
            }
        }
        for (int32_t v = v1 - 1; v >= v0; v--) {
            StepOut st;
            const bool applied = visible(v) && tsdf_step<true>(views[v], maps, priors, mu, wx, wy, wz, st);
            double gc[3] = {0.0, 0.0, 0.0};
            if (applied) {
                const float2 prev = seg_state[v - v0][threadIdx.x];
                const double W0 = prev.y, w = st.w, o = st.o, D = W0 + w;
                const double tsk = ((double)prev.x * W0 + o * w) / D;
                const double u = (prev.y + st.w) < kMaxWeight ? 1.0 : 0.0;
                const double g_o = a * w / D;
                const double g_w = a * (o - tsk) / D + u * b;
                const double a_prev = a * W0 / D;
                const double b_prev = a * ((double)prev.x - tsk) / D + u * b;
                const double gdep = g_o * st.do_dd + g_w * st.dw_dd;
                if (gdep != 0.0) atomicAdd(&grad_depth[pixel_off[v] + st.pixel], gdep);
                gc[0] = g_w * st.dw_dc[0];
                gc[1] = g_w * st.dw_dc[1];
                gc[2] = g_o * st.do_dz + g_w * st.dw_dc[2];
                a = a_prev;
                b = b_prev;
            }
            if (__any(applied)) {
                const double world[3] = {wx, wy, wz};
                double part[12];
#pragma unroll
                for (int i = 0; i < 3; i++) {
#pragma unroll
                    for (int j = 0; j < 3; j++) part[i * 3 + j] = gc[i] * world[j];
                    part[9 + i] = gc[i];
                }
#pragma unroll
                for (int j = 0; j < 12; j++) {
                    const double sum = wave_sum_d(part[j]);
                    if (lane == 0 && sum != 0.0) atomicAdd(&grad_pose[(size_t)v * 12 + j], sum);
                }
            }
        }
    }
}

__global__ void to_float_kernel(const double *__restrict__ in, float *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

}  // namespace

hipError_t tsdf_launch_depth_tiles(hipStream_t s, const TsdfView *views, int32_t n_views, int32_t max_tw,
                                   int32_t max_th, const uint32_t *maps, const float *priors, float *tiles) {
    if (n_views <= 0 || max_tw <= 0 || max_th <= 0) return hipSuccess;
    if (n_views > 65535) return hipErrorInvalidValue;
    dim3 block(kTsdfTile, kTsdfTile), grid((uint32_t)max_tw, (uint32_t)max_th, (uint32_t)n_views);
    hipLaunchKernelGGL(depth_tile_kernel, grid, block, 0, s, views, maps, priors, tiles);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(depth_tile2_kernel, dim3((uint32_t)n_views), dim3(256), 0, s, views, tiles);
    return hipGetLastError();
}

hipError_t tsdf_launch_visibility(hipStream_t s, const TsdfGridDev &g, const int4 *coords,
                                  const TsdfView *views, int32_t n_views, uint32_t *vis,
                                  int32_t words, const float *tiles, uint32_t *ctrs, uint32_t *list) {
    // lanes per (brick, word): 32, or for a one-word pass the views rounded up to a power of two
    uint32_t lane_shift = 5;
    if (words == 1) {
        lane_shift = 0;
        while ((1 << lane_shift) < n_views) lane_shift++;
    }
    const uint64_t threads = (uint64_t)g.n_bricks * (uint64_t)words << lane_shift;
    if (threads == 0) return hipSuccess;
    if (threads > 0xffffff00ull) return hipErrorInvalidValue;
    dim3 block(256), grid((uint32_t)((threads + 255u) / 256u));
    hipLaunchKernelGGL(visibility_kernel, grid, block, 0, s, g, coords, views, n_views, vis, words, tiles, ctrs,
                       lane_shift);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !list) return e;
    hipLaunchKernelGGL(brick_list_kernel, dim3(((uint32_t)g.n_bricks + 255u) / 256u), dim3(256), 0, s, vis, words,
                       g.n_bricks, ctrs, list);
    return hipGetLastError();
}

hipError_t tsdf_launch_normal_table(hipStream_t s, float4 *tab) {
    hipLaunchKernelGGL(normal_table_kernel, dim3(16), dim3(256), 0, s, tab);
    return hipGetLastError();
}

template <bool FAST, bool LIST>
static void launch_integrate(hipStream_t s, dim3 grid, dim3 block, const TsdfGridDev &g, const TsdfView *views,
                             int32_t n_views, const uint32_t *vis, int32_t words, const int4 *coords,
                             const uint32_t *maps, const float *priors, const float4 *ntab, float *tsdf,
                             float *weight, unsigned long long *counter, const uint32_t *ctrs, const uint32_t *list,
                             uint32_t chunk) {
    // priors != nullptr: some view of the pass has a prior (the caller passes nullptr otherwise)
    const bool c = counter != nullptr, p = priors != nullptr;
    auto k = c ? (p ? integrate_kernel<true, true, FAST, LIST> : integrate_kernel<true, false, FAST, LIST>)
               : (p ? integrate_kernel<false, true, FAST, LIST> : integrate_kernel<false, false, FAST, LIST>);
    hipLaunchKernelGGL(k, grid, block, 0, s, g, views, n_views, vis, words, coords, maps, priors, ntab, tsdf, weight,
                       counter, ctrs, list, chunk);
}

hipError_t tsdf_launch_integrate(hipStream_t s, const TsdfGridDev &g, const TsdfView *views,
                                 int32_t n_views, const uint32_t *vis, int32_t words,
                                 const int4 *coords, const uint32_t *maps, const float *priors,
                                 const float4 *ntab, bool fast, float *tsdf, float *weight,
                                 unsigned long long *counter, const uint32_t *ctrs, const uint32_t *list,
                                 int32_t bricks_per_block, int32_t chunk) {
    if (bricks_per_block != 1 && bricks_per_block != 2 && bricks_per_block != 4) return hipErrorInvalidValue;
    if (chunk < 1 || chunk > kMaxXcdChunk) return hipErrorInvalidValue;
    if (fast && !ntab) return hipErrorInvalidValue;
    if (list && !ctrs) return hipErrorInvalidValue;
    if (words <= 0) return hipSuccess;
    const uint32_t bpb = (uint32_t)bricks_per_block;
    dim3 block(64u * bpb), grid(((uint32_t)g.n_bricks + bpb - 1u) / bpb);
    if (list) {
        // a fixed grid of at most kIntegrateWaves wavefronts, in whole rounds of 8 blocks so that
        // every queue is served
        const uint32_t waves = std::min<uint32_t>(kIntegrateWaves, (uint32_t)g.n_bricks);
        grid = dim3(((waves + bpb - 1u) / bpb + 7u) / 8u * 8u);
        if (fast)
            launch_integrate<true, true>(s, grid, block, g, views, n_views, vis, words, coords, maps, priors, ntab,
                                         tsdf, weight, counter, ctrs, list, (uint32_t)chunk);
        else
            launch_integrate<false, true>(s, grid, block, g, views, n_views, vis, words, coords, maps, priors, ntab,
                                          tsdf, weight, counter, ctrs, list, (uint32_t)chunk);
    } else if (fast) {
        launch_integrate<true, false>(s, grid, block, g, views, n_views, vis, words, coords, maps, priors, ntab,
                                      tsdf, weight, counter, ctrs, list, (uint32_t)chunk);
    } else {
        launch_integrate<false, false>(s, grid, block, g, views, n_views, vis, words, coords, maps, priors, ntab,
                                       tsdf, weight, counter, ctrs, list, (uint32_t)chunk);
    }
    return hipGetLastError();
}

hipError_t tsdf_launch_backward(hipStream_t s, const TsdfGridDev &g, const TsdfView *views, int32_t n_views,
                                const uint32_t *vis, int32_t words, const int4 *coords, const uint32_t *maps,
                                const float *priors, const uint32_t *voxels, const float *targets,
                                uint32_t n_voxels, const uint32_t *pixel_off, float2 *ckpt, float *tsdf_out,
                                double *grad_depth, double *grad_pose, double *loss_partials) {
    dim3 block(256), grid((n_voxels + 255u) / 256u);
    hipLaunchKernelGGL(tsdf_backward_kernel, grid, block, 0, s, g, views, n_views, vis, words, coords, maps,
                       priors, voxels, targets, n_voxels, pixel_off, ckpt, tsdf_out, grad_depth, grad_pose,
                       loss_partials);
    return hipGetLastError();
}

hipError_t launch_to_float(hipStream_t s, const double *in, float *out, size_t n) {
    if (n == 0) return hipSuccess;
    dim3 block(256), grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(to_float_kernel, grid, block, 0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace lr3d
