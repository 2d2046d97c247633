// Dense-cloud fusion of packed depth maps on the GPU: addDepthMap / fuseDepthMaps
// (frontend/src/app/vision/fusion/depth-fusion.ts, DF:23-74) into the voxel-cell accumulator of
// frontend/src/app/vision/fusion/point-accumulator.ts (PA:17-143).
//
// The accumulator sums position, normal and colour per integer cell in Float32 (each += rounds
// to float) in the order the samples arrive (views in order, pixels in row-major order).  Here
// the valid samples of a chunk of views get their cell key in double (unpackDepthSample,
// fromCamera, floor(x/cell) as the JavaScript computes them), a stable radix sort groups them by
// cell keeping arrival order, and one thread per cell adds them in that order onto the cell's
// running sums, so every cell's sums equal the reference's.  The host ends a chunk where the
// reference re-bins its cells (a coarsening after a view, a table growth before a sample).  Cells
// are kept sorted by key rather than in hash-table order: the output order differs from the
// reference's slot order, and a re-binning (by mean position, PA:87-131) merges cells in key
// order where the reference merges in slot order.
#include "lr3d_internal.h"

#include <math.h>

namespace lr3d {

namespace {

constexpr int32_t kKeyBias = 1 << 20;
// Cells with more samples than this are added by a wavefront each (big_reduce_kernel).
constexpr uint32_t kBigCell = 64;
// Blocks of big_reduce_kernel (each strides over the list of big cells).
constexpr uint32_t kBigReduceBlocks = 2048;

__device__ __forceinline__ uint64_t pack_key(int64_t ix, int64_t iy, int64_t iz, uint32_t *overflow) {
    if (ix < -kKeyBias || ix >= kKeyBias || iy < -kKeyBias || iy >= kKeyBias || iz < -kKeyBias || iz >= kKeyBias) {
        atomicOr(overflow, 1u);
        return 0ull;
    }
    return ((uint64_t)(ix + kKeyBias) << 42) | ((uint64_t)(iy + kKeyBias) << 21) | (uint64_t)(iz + kKeyBias);
}

// depth-maps.ts:44-57 halfTable: exponent 31 decodes to NaN (never a valid depth)
__device__ __forceinline__ double half_to_double(uint32_t h) {
    const int exponent = (h >> 10) & 31, mantissa = h & 1023;
    const double sign = (h & 0x8000u) ? -1.0 : 1.0;
    if (exponent == 0) return sign * ldexp(1.0, -14) * (mantissa / 1024.0);
    if (exponent == 31) return __builtin_nan("");
    return sign * ldexp(1.0, exponent - 15) * (1.0 + mantissa / 1024.0);
}

// One sample of pixel i of view v (DF:52-71): position, normal (world), colour; false if empty.
__device__ __forceinline__ bool sample_of(const FusionView &v, const uint32_t *__restrict__ maps,
                                          const uint8_t *__restrict__ thumbs, uint32_t i, double out[9]) {
    const uint32_t value = maps[v.map_off + i];
    if (value == 0u) return false;
    const double d = (double)(float)half_to_double(value & 0xffffu);  // halfTable is a Float32Array
    if (!(d > 0.0)) return false;
    // depth-maps.ts:72-84
    const double ex = (((value >> 16) & 63u) / 63.0) * 2.0 - 1.0;
    const double ey = (((value >> 22) & 63u) / 63.0) * 2.0 - 1.0;
    double nx = ex, ny = ey;
    const double nz = 1.0 - fabs(ex) - fabs(ey);
    if (nz < 0.0) {
        nx = (1.0 - fabs(ey)) * (ex >= 0.0 ? 1.0 : -1.0);
        ny = (1.0 - fabs(ex)) * (ey >= 0.0 ? 1.0 : -1.0);
    }
    double length = sqrt(nx * nx + ny * ny + nz * nz);
    if (length == 0.0) length = 1.0;
    const double n[3] = {nx / length, ny / length, nz / length};
    const int32_t x = (int32_t)(i % (uint32_t)v.W), y = (int32_t)(i / (uint32_t)v.W);
    // fromCamera(pose, [((x - cx) / fx) d, ((y - cy) / fy) d, d]) = R^T (c - t)
    const double c[3] = {((x - v.k[2]) / v.k[0]) * d - v.t[0], ((y - v.k[3]) / v.k[1]) * d - v.t[1], d - v.t[2]};
    const double *R = v.R;
    out[0] = R[0] * c[0] + R[3] * c[1] + R[6] * c[2];
    out[1] = R[1] * c[0] + R[4] * c[1] + R[7] * c[2];
    out[2] = R[2] * c[0] + R[5] * c[1] + R[8] * c[2];
    out[3] = R[0] * n[0] + R[3] * n[1] + R[6] * n[2];
    out[4] = R[1] * n[0] + R[4] * n[1] + R[7] * n[2];
    out[5] = R[2] * n[0] + R[5] * n[1] + R[8] * n[2];
    const double sx = (double)v.tw / v.W, sy = (double)v.th / v.H;
    int32_t tx = (int32_t)floor((x + 0.5) * sx), ty = (int32_t)floor((y + 0.5) * sy);
    tx = min(tx, v.tw - 1);
    ty = min(ty, v.th - 1);
    const uint8_t *rgb = thumbs + v.thumb_off + ((size_t)ty * v.tw + tx) * 3;
    out[6] = rgb[0];
    out[7] = rgb[1];
    out[8] = rgb[2];
    return true;
}

// The chunk view holding sample number p: the last c with vstart[c] <= p (vstart[0] = 0).
__device__ __forceinline__ uint32_t view_of(const uint32_t *__restrict__ vstart, uint32_t nv, uint32_t p) {
    uint32_t lo = 0, hi = nv - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (vstart[mid] <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Over every pixel of a chunk of views at once (pixel p of the chunk is pixel p - vstart[c] of chunk
// view c): the valid samples (the chunk's first view from pixel `from` on), then (PA:37-48) the cell
// key floor(x / cell) of every valid sample with its sample number p.
__global__ void chunk_valid_kernel(const FusionView *__restrict__ views, const uint32_t *__restrict__ vstart,
                                   uint32_t nv, uint32_t from, const uint32_t *__restrict__ maps,
                                   uint32_t *__restrict__ flag) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= vstart[nv]) return;
    const uint32_t c = view_of(vstart, nv, p);
    const uint32_t i = p - vstart[c];
    const uint32_t value = (c > 0 || i >= from) ? maps[views[c].map_off + i] : 0u;
    const double d = value ? (double)(float)half_to_double(value & 0xffffu) : 0.0;
    flag[p] = d > 0.0 ? 1u : 0u;
}

__global__ void chunk_key_kernel(const FusionView *__restrict__ views, const uint32_t *__restrict__ vstart, uint32_t nv,
                                 const uint32_t *__restrict__ maps, const uint8_t *__restrict__ thumbs,
                                 const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, double cell,
                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ idx, uint32_t *__restrict__ overflow) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= vstart[nv] || !flag[p]) return;
    const uint32_t c = view_of(vstart, nv, p);
    double s[9];
    sample_of(views[c], maps, thumbs, p - vstart[c], s);
    const uint32_t at = pos[p];
    keys[at] = pack_key((int64_t)floor(s[0] / cell), (int64_t)floor(s[1] / cell), (int64_t)floor(s[2] / cell), overflow);
    idx[at] = p;
}

// New cells per chunk view of their first sample: a histogram per block in LDS, added to the
// global one bin by bin (integer atomics: the counts do not depend on the order of the additions).
__global__ void __launch_bounds__(256) first_view_kernel(const uint32_t *__restrict__ idx,
                                                         const uint32_t *__restrict__ start,
                                                         const uint32_t *__restrict__ is_new, uint32_t segments,
                                                         const uint32_t *__restrict__ vstart, uint32_t nv,
                                                         uint32_t *__restrict__ hist) {
    __shared__ uint32_t local[kFusionMaxChunkViews];
    for (uint32_t c = threadIdx.x; c < nv; c += blockDim.x) local[c] = 0u;
    __syncthreads();
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < segments && is_new[g]) atomicAdd(&local[view_of(vstart, nv, idx[start[g]])], 1u);
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < nv; c += blockDim.x)
        if (local[c]) atomicAdd(&hist[c], local[c]);
}

// Flags, over the pixels of chunk view c, the first samples of the new cells first seen there.
__global__ void new_cell_flag_kernel(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ start,
                                     const uint32_t *__restrict__ is_new, uint32_t segments, uint32_t lo, uint32_t hi,
                                     uint32_t *__restrict__ flag) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= segments || !is_new[g]) return;
    const uint32_t p = idx[start[g]];
    if (p >= lo && p < hi) flag[p - lo] = 1u;
}

// The pixel of the k-th flagged pixel (k from 0), given the exclusive scan of the flags.
__global__ void select_kernel(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ scan, uint32_t n,
                              uint32_t k, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i] && scan[i] == k) *out = i;
}

// Cells first seen at sample `limit` or later are not created in this pass.
__global__ void limit_kernel(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ start, uint32_t segments,
                             uint32_t limit, uint32_t *__restrict__ is_new) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < segments && idx[start[g]] >= limit) is_new[g] = 0u;
}

__global__ void head_kernel(const uint64_t *__restrict__ keys, uint32_t n, uint32_t *__restrict__ head) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void segment_start_kernel(const uint32_t *__restrict__ head, const uint32_t *__restrict__ seg, uint32_t n,
                                     uint32_t *__restrict__ start) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && head[i]) start[seg[i]] = i;
}

// existing cell of every segment (binary search in the sorted state keys), and whether it is new
__global__ void lookup_kernel(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ start, uint32_t segments,
                              const uint64_t *__restrict__ state_keys, uint32_t state_n, int32_t *__restrict__ exist,
                              uint32_t *__restrict__ is_new) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= segments) return;
    const uint64_t k = keys[start[g]];
    uint32_t lo = 0, hi = state_n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (state_keys[mid] < k) lo = mid + 1; else hi = mid;
    }
    const bool found = lo < state_n && state_keys[lo] == k;
    exist[g] = found ? (int32_t)lo : -1;
    is_new[g] = found ? 0u : 1u;
}

// PA:49-62: one thread per cell, the samples added in arrival order (views in order, pixels
// row-major: the stable sort kept it) onto the running sums; samples numbered `limit` or later
// belong to views after the chunk's end and are left out.
__global__ void reduce_kernel(const FusionView *__restrict__ views, const uint32_t *__restrict__ vstart, uint32_t nv,
                              uint32_t limit, const uint32_t *__restrict__ maps, const uint8_t *__restrict__ thumbs,
                              const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                              const uint32_t *__restrict__ start, uint32_t segments, uint32_t n,
                              const int32_t *__restrict__ exist, const uint32_t *__restrict__ is_new,
                              const uint32_t *__restrict__ new_pos, float *__restrict__ state_sums,
                              uint32_t *__restrict__ state_counts, uint64_t *__restrict__ new_keys,
                              float *__restrict__ new_sums, uint32_t *__restrict__ new_counts,
                              uint32_t *__restrict__ big, uint32_t *__restrict__ n_big) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= segments) return;
    const uint32_t a = start[g], b = g + 1 < segments ? start[g + 1] : n;
    const int32_t e = exist[g];
    if (e < 0 && !is_new[g]) return;  // a new cell first seen after the chunk's end
    if (b - a > kBigCell) {           // a wavefront takes it (big_reduce_kernel)
        big[atomicAdd(n_big, 1u)] = g;
        return;
    }
    float sums[9];
    uint32_t count = 0;
    for (int c = 0; c < 9; c++) sums[c] = e >= 0 ? state_sums[(size_t)e * 9 + c] : 0.0f;
    if (e >= 0) count = state_counts[e];
    for (uint32_t j = a; j < b; j++) {
        const uint32_t p = idx[j];
        if (p >= limit) break;
        const uint32_t c = view_of(vstart, nv, p);
        double s[9];
        sample_of(views[c], maps, thumbs, p - vstart[c], s);
        for (int k = 0; k < 9; k++) sums[k] = (float)((double)sums[k] + s[k]);
        count++;
    }
    float *dst = e >= 0 ? state_sums + (size_t)e * 9 : new_sums + (size_t)new_pos[g] * 9;
    for (int c = 0; c < 9; c++) dst[c] = sums[c];
    if (e >= 0) {
        state_counts[e] = count;
    } else {
        new_counts[new_pos[g]] = count;
        new_keys[new_pos[g]] = keys[a];
    }
}

// The cells with more than kBigCell samples, one wavefront (one 64-thread block) each, cells taken
// by striding over the list reduce_kernel made: 64 samples are decoded at once (one per lane, into
// LDS), then lanes 0-8 each add one of the 9 components of those samples in order, so every sum
// takes the same sequence of Float32 additions as a thread walking the cell would.
__global__ void __launch_bounds__(64) big_reduce_kernel(
    const FusionView *__restrict__ views, const uint32_t *__restrict__ vstart, uint32_t nv, uint32_t limit,
    const uint32_t *__restrict__ maps, const uint8_t *__restrict__ thumbs, const uint64_t *__restrict__ keys,
    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ start, uint32_t segments, uint32_t n,
    const int32_t *__restrict__ exist, const uint32_t *__restrict__ new_pos, float *__restrict__ state_sums,
    uint32_t *__restrict__ state_counts, uint64_t *__restrict__ new_keys, float *__restrict__ new_sums,
    uint32_t *__restrict__ new_counts, const uint32_t *__restrict__ big, const uint32_t *__restrict__ n_big) {
    __shared__ double vals[64][9];
    const uint32_t lane = threadIdx.x;
    const uint32_t cells = *n_big;
    for (uint32_t i = blockIdx.x; i < cells; i += gridDim.x) {
        const uint32_t g = big[i];
        const uint32_t a = start[g], b = g + 1 < segments ? start[g + 1] : n;
        const int32_t e = exist[g];
        float sum = 0.0f;
        uint32_t count = e >= 0 ? state_counts[e] : 0u;
        if (lane < 9 && e >= 0) sum = state_sums[(size_t)e * 9 + lane];
        for (uint32_t j0 = a; j0 < b; j0 += 64) {
            const uint32_t j = j0 + lane;
            const uint32_t p = j < b ? idx[j] : limit;
            const bool take = p < limit;  // samples below `limit` are a prefix of the cell's
            if (take) {
                const uint32_t c = view_of(vstart, nv, p);
                double sv[9];
                sample_of(views[c], maps, thumbs, p - vstart[c], sv);
                for (int k = 0; k < 9; k++) vals[lane][k] = sv[k];
            }
            const uint32_t m = (uint32_t)__popcll(__ballot(take));
            __syncthreads();
            if (lane < 9)
                for (uint32_t t = 0; t < m; t++) sum = (float)((double)sum + vals[t][lane]);
            count += m;
            __syncthreads();
            if (m < 64) break;
        }
        float *dst = e >= 0 ? state_sums + (size_t)e * 9 : new_sums + (size_t)new_pos[g] * 9;
        if (lane < 9) dst[lane] = sum;
        if (lane == 0) {
            if (e >= 0) {
                state_counts[e] = count;
            } else {
                new_counts[new_pos[g]] = count;
                new_keys[new_pos[g]] = keys[a];
            }
        }
    }
}

// gather a permutation of cells (sorted order) into fresh arrays
__global__ void gather_kernel(const uint32_t *__restrict__ perm, uint32_t n, const uint64_t *__restrict__ keys,
                              const float *__restrict__ sums, const uint32_t *__restrict__ counts,
                              uint64_t *__restrict__ out_keys, float *__restrict__ out_sums,
                              uint32_t *__restrict__ out_counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = perm[i];
    out_keys[i] = keys[p];
    for (int c = 0; c < 9; c++) out_sums[(size_t)i * 9 + c] = sums[(size_t)p * 9 + c];
    out_counts[i] = counts[p];
}

// PA:114-119: coarsened key of every cell from its mean position
__global__ void rekey_kernel(const float *__restrict__ sums, const uint32_t *__restrict__ counts, uint32_t n, double cell,
                             uint64_t *__restrict__ keys, uint32_t *__restrict__ idx, uint32_t *__restrict__ overflow) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c = counts[i];
    keys[i] = pack_key((int64_t)floor((double)sums[(size_t)i * 9] / c / cell),
                       (int64_t)floor((double)sums[(size_t)i * 9 + 1] / c / cell),
                       (int64_t)floor((double)sums[(size_t)i * 9 + 2] / c / cell), overflow);
    idx[i] = i;
}

// PA:120-121: merged cells, sums added in the sorted (old key) order
__global__ void merge_kernel(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                             const uint32_t *__restrict__ start, uint32_t segments, uint32_t n,
                             const float *__restrict__ sums, const uint32_t *__restrict__ counts,
                             uint64_t *__restrict__ out_keys, float *__restrict__ out_sums,
                             uint32_t *__restrict__ out_counts) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= segments) return;
    const uint32_t a = start[g], b = g + 1 < segments ? start[g + 1] : n;
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t count = 0;
    for (uint32_t j = a; j < b; j++) {
        const uint32_t p = idx[j];
        for (int c = 0; c < 9; c++) acc[c] = (float)((double)acc[c] + (double)sums[(size_t)p * 9 + c]);
        count += counts[p];
    }
    out_keys[g] = keys[a];
    for (int c = 0; c < 9; c++) out_sums[(size_t)g * 9 + c] = acc[c];
    out_counts[g] = count;
}

__global__ void support_kernel(const uint32_t *__restrict__ counts, uint32_t n, uint32_t min_support,
                               uint32_t *__restrict__ flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = counts[i] >= min_support ? 1u : 0u;
}

// PA:124-142 cloud(minSupport)
__global__ void cloud_kernel(const float *__restrict__ sums, const uint32_t *__restrict__ counts, uint32_t n,
                             const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                             float *__restrict__ positions, float *__restrict__ normals, uint8_t *__restrict__ colors,
                             uint16_t *__restrict__ support) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t p = pos[i];
    const double c = counts[i];
    const float *o = sums + (size_t)i * 9;
    for (int k = 0; k < 3; k++) positions[(size_t)p * 3 + k] = (float)((double)o[k] / c);
    double length = sqrt((double)o[3] * o[3] + (double)o[4] * o[4] + (double)o[5] * o[5]);
    if (length == 0.0) length = 1.0;
    for (int k = 0; k < 3; k++) normals[(size_t)p * 3 + k] = (float)((double)o[3 + k] / length);
    for (int k = 0; k < 3; k++) colors[(size_t)p * 3 + k] = (uint8_t)(uint32_t)((double)o[6 + k] / c);
    support[p] = (uint16_t)min(counts[i], 65535u);
}

inline dim3 grid_for(uint64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

// Launch and report the launch's own status: an earlier runtime call (a copy from pageable host
// memory, for one) can leave an error behind that hipGetLastError would otherwise report here.
#define LAUNCH(kernel, n, ...)                                                              \
    do {                                                                                    \
        (void)hipGetLastError();                                                            \
        if ((n) > 0) hipLaunchKernelGGL(kernel, grid_for(n), dim3(256), 0, s, __VA_ARGS__); \
        return hipGetLastError();                                                           \
    } while (0)

hipError_t fusion_launch_new_cell_flags(hipStream_t s, const uint32_t *idx, const uint32_t *start, const uint32_t *is_new,
                                        uint32_t segments, uint32_t lo, uint32_t hi, uint32_t *flag) {
    LAUNCH(new_cell_flag_kernel, segments, idx, start, is_new, segments, lo, hi, flag);
}
hipError_t fusion_launch_select(hipStream_t s, const uint32_t *flag, const uint32_t *scan, uint32_t n, uint32_t k,
                                uint32_t *out) {
    LAUNCH(select_kernel, n, flag, scan, n, k, out);
}
hipError_t fusion_launch_chunk_valid(hipStream_t s, const FusionView *views, const uint32_t *vstart, uint32_t nv,
                                     uint32_t pixels, uint32_t from, const uint32_t *maps, uint32_t *flag) {
    LAUNCH(chunk_valid_kernel, pixels, views, vstart, nv, from, maps, flag);
}
hipError_t fusion_launch_chunk_keys(hipStream_t s, const FusionView *views, const uint32_t *vstart, uint32_t nv,
                                    uint32_t pixels, const uint32_t *maps, const uint8_t *thumbs, const uint32_t *flag,
                                    const uint32_t *pos, double cell, uint64_t *keys, uint32_t *idx,
                                    uint32_t *overflow) {
    LAUNCH(chunk_key_kernel, pixels, views, vstart, nv, maps, thumbs, flag, pos, cell, keys, idx, overflow);
}
hipError_t fusion_launch_first_views(hipStream_t s, const uint32_t *idx, const uint32_t *start, const uint32_t *is_new,
                                     uint32_t segments, const uint32_t *vstart, uint32_t nv, uint32_t *hist) {
    LAUNCH(first_view_kernel, segments, idx, start, is_new, segments, vstart, nv, hist);
}
hipError_t fusion_launch_limit(hipStream_t s, const uint32_t *idx, const uint32_t *start, uint32_t segments,
                               uint32_t limit, uint32_t *is_new) {
    LAUNCH(limit_kernel, segments, idx, start, segments, limit, is_new);
}
hipError_t fusion_launch_heads(hipStream_t s, const uint64_t *keys, uint32_t n, uint32_t *head) {
    LAUNCH(head_kernel, n, keys, n, head);
}
hipError_t fusion_launch_starts(hipStream_t s, const uint32_t *head, const uint32_t *seg, uint32_t n, uint32_t *start) {
    LAUNCH(segment_start_kernel, n, head, seg, n, start);
}
hipError_t fusion_launch_lookup(hipStream_t s, const uint64_t *keys, const uint32_t *start, uint32_t segments,
                                const uint64_t *state_keys, uint32_t state_n, int32_t *exist, uint32_t *is_new) {
    LAUNCH(lookup_kernel, segments, keys, start, segments, state_keys, state_n, exist, is_new);
}
hipError_t fusion_launch_reduce(hipStream_t s, const FusionView *views, const uint32_t *vstart, uint32_t nv,
                                uint32_t limit, const uint32_t *maps, const uint8_t *thumbs, const uint64_t *keys,
                                const uint32_t *idx, const uint32_t *start, uint32_t segments, uint32_t n,
                                const int32_t *exist, const uint32_t *is_new, const uint32_t *new_pos,
                                float *state_sums, uint32_t *state_counts, uint64_t *new_keys, float *new_sums,
                                uint32_t *new_counts, uint32_t *big, uint32_t *n_big) {
    (void)hipGetLastError();
    if (segments == 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_kernel, grid_for(segments), dim3(256), 0, s, views, vstart, nv, limit, maps, thumbs, keys,
                       idx, start, segments, n, exist, is_new, new_pos, state_sums, state_counts, new_keys, new_sums,
                       new_counts, big, n_big);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(big_reduce_kernel, dim3(kBigReduceBlocks), dim3(64), 0, s, views, vstart, nv, limit, maps, thumbs,
                       keys, idx, start, segments, n, exist, new_pos, state_sums, state_counts, new_keys, new_sums,
                       new_counts, big, n_big);
    return hipGetLastError();
}
hipError_t fusion_launch_gather(hipStream_t s, const uint32_t *perm, uint32_t n, const uint64_t *keys, const float *sums,
                                const uint32_t *counts, uint64_t *out_keys, float *out_sums, uint32_t *out_counts) {
    LAUNCH(gather_kernel, n, perm, n, keys, sums, counts, out_keys, out_sums, out_counts);
}
hipError_t fusion_launch_rekey(hipStream_t s, const float *sums, const uint32_t *counts, uint32_t n, double cell,
                               uint64_t *keys, uint32_t *idx, uint32_t *overflow) {
    LAUNCH(rekey_kernel, n, sums, counts, n, cell, keys, idx, overflow);
}
hipError_t fusion_launch_merge(hipStream_t s, const uint64_t *keys, const uint32_t *idx, const uint32_t *start,
                               uint32_t segments, uint32_t n, const float *sums, const uint32_t *counts,
                               uint64_t *out_keys, float *out_sums, uint32_t *out_counts) {
    LAUNCH(merge_kernel, segments, keys, idx, start, segments, n, sums, counts, out_keys, out_sums, out_counts);
}
hipError_t fusion_launch_support(hipStream_t s, const uint32_t *counts, uint32_t n, uint32_t min_support, uint32_t *flag) {
    LAUNCH(support_kernel, n, counts, n, min_support, flag);
}
hipError_t fusion_launch_cloud(hipStream_t s, const float *sums, const uint32_t *counts, uint32_t n, const uint32_t *flag,
                               const uint32_t *pos, float *positions, float *normals, uint8_t *colors,
                               uint16_t *support) {
    LAUNCH(cloud_kernel, n, sums, counts, n, flag, pos, positions, normals, colors, support);
}

#undef LAUNCH

}  // namespace lr3d
