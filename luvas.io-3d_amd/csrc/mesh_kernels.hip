// Surface nets over the brick TSDF and the mesh post-processing of the reference, on the GPU.
//
// extractSurface (frontend/src/app/vision/fusion/surface-nets.ts, SN:47-166), computeNormals and
// removeSmallComponents (frontend/src/app/vision/fusion/mesh.ts, MS:9-90).  The reference walks
// the pool sequentially; here every ordering it produces comes from prefix sums, so the mesh is
// the same array for array:
//   vertices   one per active cell, numbered by an exclusive scan of the active flags in pool
//              order (SN:87-115); positions in double (JavaScript numbers), stored as float;
//   triangles  per pool voxel, per axis, the quads counted, scanned and written at their
//              sequential offsets (SN:117-156); the diagonal test uses the double positions;
//   normals    per vertex, the incident triangles visited in increasing triangle order, each
//              += rounding to float as the reference's Float32Array does (MS:9-32);
//   components union-find by hooking to the smaller label plus pointer jumping (labels are
//              order-independent), face counts per root, and the surviving vertices numbered by
//              first appearance in the kept index stream (MS:38-90), again by scans.
// Atomics are used on integers only, so results do not depend on scheduling.
#include "lr3d_internal.h"

#include <math.h>

namespace lr3d {

namespace {

__constant__ int8_t kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
__constant__ int8_t kEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                    {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// SN:53-58
__device__ __forceinline__ int64_t slot(const MeshVolume &v, int32_t x, int32_t y, int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= v.dims[0] * 8 || y >= v.dims[1] * 8 || z >= v.dims[2] * 8) return -1;
    const int32_t brick = v.lookup[((z >> 3) * v.dims[1] + (y >> 3)) * v.dims[0] + (x >> 3)];
    if (brick < 0) return -1;
    return (int64_t)brick * 512 + (x & 7) + ((y & 7) << 3) + ((z & 7) << 6);
}

__device__ __forceinline__ void cell_of(const MeshVolume &v, uint32_t s, int32_t &x, int32_t &y, int32_t &z) {
    const int4 c = v.coords[s >> 9];
    const uint32_t local = s & 511u;
    x = c.x * 8 + (int32_t)(local & 7u);
    y = c.y * 8 + (int32_t)((local >> 3) & 7u);
    z = c.z * 8 + (int32_t)(local >> 6);
}

// SN:76-84: all 8 corners present and observed (weight >= min), signs mixed.
__device__ __forceinline__ bool sample_cell(const MeshVolume &v, int32_t x, int32_t y, int32_t z, float values[8]) {
    int inside = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int64_t s = slot(v, x + kCorner[k][0], y + kCorner[k][1], z + kCorner[k][2]);
        if (s < 0 || (double)v.weight[s] < v.min_weight) return false;
        values[k] = v.tsdf[s];
        if (values[k] < 0.0f) inside++;
    }
    return inside > 0 && inside < 8;
}

// Pool voxel of thread t: the t-th voxel of the listed bricks, or of the pool without a list.
__device__ __forceinline__ uint32_t pool_voxel(const MeshVolume &v, uint32_t t) {
    return v.bricks ? v.bricks[t >> 9] * 512u + (t & 511u) : t;
}

// Bricks holding a voxel observed at the minimum weight: every cell and quad of the others is
// empty (a cell needs its own corner observed, a quad its own voxel), so the cell kernels skip them.
__global__ void __launch_bounds__(256) live_brick_kernel(MeshVolume v, uint32_t n_bricks, uint32_t *__restrict__ live) {
    const uint32_t b = blockIdx.x;
    if (b >= n_bricks) return;  // the whole block
    const size_t base = (size_t)b * 512u + threadIdx.x;
    const bool seen = (double)v.weight[base] >= v.min_weight || (double)v.weight[base + 256] >= v.min_weight;
    const int any = __syncthreads_or(seen ? 1 : 0);
    if (threadIdx.x == 0) live[b] = any ? 1u : 0u;
}

__global__ void brick_list_kernel(const uint32_t *__restrict__ live, const uint32_t *__restrict__ pos, uint32_t n,
                                  uint32_t *__restrict__ list) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n && live[b]) list[pos[b]] = b;
}

__global__ void __launch_bounds__(256) active_kernel(MeshVolume v, uint32_t cells, uint32_t *__restrict__ flag) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const uint32_t s = pool_voxel(v, t);
    int32_t x, y, z;
    cell_of(v, s, x, y, z);
    float values[8];
    flag[s] = sample_cell(v, x, y, z, values) ? 1u : 0u;
}

// SN:95-113, positions in double
__global__ void __launch_bounds__(256) vertex_kernel(MeshVolume v, uint32_t cells, const uint32_t *__restrict__ vidx,
                                                     int32_t *__restrict__ cell_vertex, double *__restrict__ pos_d,
                                                     float *__restrict__ pos_f) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const uint32_t s = pool_voxel(v, t);
    int32_t x, y, z;
    cell_of(v, s, x, y, z);
    float values[8];
    if (!sample_cell(v, x, y, z, values)) {
        cell_vertex[s] = -1;
        return;
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int n = 0;
    for (int e = 0; e < 12; e++) {
        const int a = kEdge[e][0], c = kEdge[e][1];
        const double va = values[a], vc = values[c];
        if ((va < 0.0) == (vc < 0.0)) continue;
        const double t = va / (va - vc);
        sx += (double)kCorner[a][0] + (double)(kCorner[c][0] - kCorner[a][0]) * t;
        sy += (double)kCorner[a][1] + (double)(kCorner[c][1] - kCorner[a][1]) * t;
        sz += (double)kCorner[a][2] + (double)(kCorner[c][2] - kCorner[a][2]) * t;
        n++;
    }
    const uint32_t i = vidx[s];
    cell_vertex[s] = (int32_t)i;
    const double p[3] = {v.origin[0] + ((double)x + 0.5 + sx / n) * v.voxel,
                         v.origin[1] + ((double)y + 0.5 + sy / n) * v.voxel,
                         v.origin[2] + ((double)z + 0.5 + sz / n) * v.voxel};
    for (int k = 0; k < 3; k++) {
        pos_d[(size_t)i * 3 + k] = p[k];
        pos_f[(size_t)i * 3 + k] = (float)p[k];
    }
}

// The quads of pool voxel s (SN:131-153): returns how many; writes their triangles if out != NULL.
__device__ __forceinline__ uint32_t quads_of(const MeshVolume &v, uint32_t s, const int32_t *__restrict__ cell_vertex,
                                             const double *__restrict__ pos, uint32_t *__restrict__ out) {
    if ((double)v.weight[s] < v.min_weight) return 0u;
    int32_t x, y, z;
    cell_of(v, s, x, y, z);
    const float v0 = v.tsdf[s];
    uint32_t count = 0;
    for (int a = 0; a < 3; a++) {
        const int ax = a == 0, ay = a == 1, az = a == 2;
        const int64_t s1 = slot(v, x + ax, y + ay, z + az);
        if (s1 < 0 || (double)v.weight[s1] < v.min_weight) continue;
        const float v1 = v.tsdf[s1];
        if ((v0 < 0.0f) == (v1 < 0.0f)) continue;
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        const int ux = b == 0, uy = b == 1, uz = b == 2;
        const int wx = c == 0, wy = c == 1, wz = c == 2;
        const int64_t sq[4] = {slot(v, x - ux - wx, y - uy - wy, z - uz - wz), slot(v, x - wx, y - wy, z - wz),
                               slot(v, x, y, z), slot(v, x - ux, y - uy, z - uz)};
        int32_t q[4];
        bool ok = true;
        for (int k = 0; k < 4; k++) {
            q[k] = sq[k] < 0 ? -1 : cell_vertex[sq[k]];
            ok &= q[k] >= 0;
        }
        if (!ok) continue;
        if (out) {
            uint32_t quad[4];
            if (v0 < 0.0f) {
                quad[0] = q[0]; quad[1] = q[1]; quad[2] = q[2]; quad[3] = q[3];
            } else {
                quad[0] = q[0]; quad[1] = q[3]; quad[2] = q[2]; quad[3] = q[1];
            }
            auto d2 = [&](uint32_t i, uint32_t j) {
                const double dx = pos[i * 3] - pos[j * 3], dy = pos[i * 3 + 1] - pos[j * 3 + 1];
                const double dz = pos[i * 3 + 2] - pos[j * 3 + 2];
                return dx * dx + dy * dy + dz * dz;
            };
            uint32_t *o = out + (size_t)count * 6;
            if (d2(quad[0], quad[2]) <= d2(quad[1], quad[3])) {
                o[0] = quad[0]; o[1] = quad[1]; o[2] = quad[2]; o[3] = quad[0]; o[4] = quad[2]; o[5] = quad[3];
            } else {
                o[0] = quad[0]; o[1] = quad[1]; o[2] = quad[3]; o[3] = quad[1]; o[4] = quad[2]; o[5] = quad[3];
            }
        }
        count++;
    }
    return count;
}

__global__ void __launch_bounds__(256) quad_count_kernel(MeshVolume v, uint32_t cells, const int32_t *__restrict__ cell_vertex,
                                                         uint32_t *__restrict__ qcount) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const uint32_t s = pool_voxel(v, t);
    qcount[s] = quads_of(v, s, cell_vertex, nullptr, nullptr);
}

__global__ void __launch_bounds__(256) quad_emit_kernel(MeshVolume v, uint32_t cells, const int32_t *__restrict__ cell_vertex,
                                                        const double *__restrict__ pos, const uint32_t *__restrict__ qoff,
                                                        uint32_t *__restrict__ indices) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const uint32_t s = pool_voxel(v, t);
    (void)quads_of(v, s, cell_vertex, pos, indices + (size_t)qoff[s] * 6);
}

// ---- computeNormals (MS:9-32) -------------------------------------------------------------------

__global__ void degree_kernel(const uint32_t *__restrict__ indices, uint32_t n_indices, uint32_t *__restrict__ deg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_indices) atomicAdd(&deg[indices[i]], 1u);
}

__global__ void adjacency_kernel(const uint32_t *__restrict__ indices, uint32_t n_indices, const uint32_t *__restrict__ off,
                                 uint32_t *__restrict__ cursor, uint32_t *__restrict__ adj) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_indices) return;
    const uint32_t v = indices[i];
    adj[off[v] + atomicAdd(&cursor[v], 1u)] = i / 3u;  // triangle number (order restored below)
}

__global__ void normal_kernel(const float *__restrict__ p, const uint32_t *__restrict__ indices, const uint32_t *__restrict__ off,
                              const uint32_t *__restrict__ deg, const uint32_t *__restrict__ adj, uint32_t n_vertices,
                              float *__restrict__ normals) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vertices) return;
    const uint32_t *list = adj + off[v];
    const uint32_t d = deg[v];
    float n[3] = {0.0f, 0.0f, 0.0f};
    int64_t prev = -1;
    for (uint32_t step = 0; step < d; step++) {
        // next triangle in increasing order (a vertex appears once per triangle)
        uint32_t f = 0xffffffffu;
        for (uint32_t j = 0; j < d; j++)
            if ((int64_t)list[j] > prev && list[j] < f) f = list[j];
        prev = f;
        const size_t a = (size_t)indices[f * 3] * 3, b = (size_t)indices[f * 3 + 1] * 3, c = (size_t)indices[f * 3 + 2] * 3;
        const double e1[3] = {(double)p[b] - p[a], (double)p[b + 1] - p[a + 1], (double)p[b + 2] - p[a + 2]};
        const double e2[3] = {(double)p[c] - p[a], (double)p[c + 1] - p[a + 1], (double)p[c + 2] - p[a + 2]};
        n[0] = (float)((double)n[0] + (e1[1] * e2[2] - e1[2] * e2[1]));
        n[1] = (float)((double)n[1] + (e1[2] * e2[0] - e1[0] * e2[2]));
        n[2] = (float)((double)n[2] + (e1[0] * e2[1] - e1[1] * e2[0]));
    }
    const double x = n[0], y = n[1], z = n[2];
    double length = sqrt(x * x + y * y + z * z);
    if (length == 0.0) length = 1.0;
    normals[(size_t)v * 3] = (float)(x / length);
    normals[(size_t)v * 3 + 1] = (float)(y / length);
    normals[(size_t)v * 3 + 2] = (float)(z / length);
}

// ---- removeSmallComponents (MS:38-90) -----------------------------------------------------------

__global__ void iota_kernel(uint32_t *__restrict__ a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// Hook the larger of two labels under the smaller (labels only decrease).
__global__ void hook_kernel(const uint32_t *__restrict__ indices, uint32_t n_faces, uint32_t *__restrict__ parent,
                            uint32_t *__restrict__ changed) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces) return;
    const uint32_t a = parent[indices[f * 3]];
    for (int k = 1; k < 3; k++) {
        const uint32_t b = parent[indices[f * 3 + k]];
        if (a == b) continue;
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        if (atomicMin(&parent[hi], lo) > lo) *changed = 1u;
    }
}

__global__ void jump_kernel(uint32_t *__restrict__ parent, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = parent[i];
    while (parent[p] != p) p = parent[p];
    parent[i] = p;
}

// Faces per component root.  Most faces of a mesh belong to one component, so one counter would take
// nearly every addition: the lanes of a wavefront whose root is that of its first face add their
// count once (integer additions: the totals do not depend on the grouping).
__global__ void component_faces_kernel(const uint32_t *__restrict__ indices, uint32_t n_faces,
                                       const uint32_t *__restrict__ parent, uint32_t *__restrict__ faces) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = f < n_faces;
    const uint32_t root = active ? parent[indices[f * 3]] : 0u;
    const uint64_t live = __ballot(active);
    if (live == 0) return;  // the whole wavefront
    const int first = __ffsll((unsigned long long)live) - 1;
    const uint32_t common = __shfl(root, first, 64);
    const bool same = active && root == common;
    const uint64_t group = __ballot(same);
    if (same) {
        if ((int)(threadIdx.x & 63u) == first) atomicAdd(&faces[common], (uint32_t)__popcll(group));
    } else if (active) {
        atomicAdd(&faces[root], 1u);
    }
}

__global__ void keep_kernel(const uint32_t *__restrict__ indices, uint32_t n_faces, const uint32_t *__restrict__ parent,
                            const uint32_t *__restrict__ faces, double limit, uint32_t *__restrict__ keep) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n_faces) keep[f] = (double)faces[parent[indices[f * 3]]] < limit ? 0u : 1u;
}

// first position of every vertex in the kept index stream
__global__ void first_use_kernel(const uint32_t *__restrict__ indices, uint32_t n_faces, const uint32_t *__restrict__ keep,
                                 const uint32_t *__restrict__ koff, uint32_t *__restrict__ first) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces || !keep[f]) return;
    for (int k = 0; k < 3; k++) atomicMin(&first[indices[f * 3 + k]], koff[f] * 3u + (uint32_t)k);
}

__global__ void first_flag_kernel(const uint32_t *__restrict__ indices, uint32_t n_faces, const uint32_t *__restrict__ keep,
                                  const uint32_t *__restrict__ koff, const uint32_t *__restrict__ first,
                                  uint32_t *__restrict__ flag) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces || !keep[f]) return;
    for (int k = 0; k < 3; k++) {
        const uint32_t p = koff[f] * 3u + (uint32_t)k;
        flag[p] = first[indices[f * 3 + k]] == p ? 1u : 0u;
    }
}

__global__ void compact_kernel(const uint32_t *__restrict__ indices, uint32_t n_faces, const uint32_t *__restrict__ keep,
                               const uint32_t *__restrict__ koff, const uint32_t *__restrict__ first,
                               const uint32_t *__restrict__ rank, const float *__restrict__ pos, const float *__restrict__ nrm,
                               uint32_t *__restrict__ out_idx, float *__restrict__ out_pos, float *__restrict__ out_nrm) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces || !keep[f]) return;
    for (int k = 0; k < 3; k++) {
        const uint32_t v = indices[f * 3 + k];
        const uint32_t p = koff[f] * 3u + (uint32_t)k;
        const uint32_t r = rank[first[v]];
        out_idx[p] = r;
        if (first[v] == p)
            for (int j = 0; j < 3; j++) {
                out_pos[(size_t)r * 3 + j] = pos[(size_t)v * 3 + j];
                out_nrm[(size_t)r * 3 + j] = nrm[(size_t)v * 3 + j];
            }
    }
}

inline dim3 grid_for(uint64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t mesh_launch_live_bricks(hipStream_t s, const MeshVolume &v, uint32_t n_bricks, uint32_t *live) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_bricks == 0) return hipSuccess;
    hipLaunchKernelGGL(live_brick_kernel, dim3(n_bricks), dim3(256), 0, s, v, n_bricks, live);
    return hipGetLastError();
}

hipError_t mesh_launch_brick_list(hipStream_t s, const uint32_t *live, const uint32_t *pos, uint32_t n_bricks,
                                  uint32_t *list) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_bricks == 0) return hipSuccess;
    hipLaunchKernelGGL(brick_list_kernel, grid_for(n_bricks), dim3(256), 0, s, live, pos, n_bricks, list);
    return hipGetLastError();
}

hipError_t mesh_launch_active(hipStream_t s, const MeshVolume &v, uint32_t cells, uint32_t *flag) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (cells == 0) return hipSuccess;
    hipLaunchKernelGGL(active_kernel, grid_for(cells), dim3(256), 0, s, v, cells, flag);
    return hipGetLastError();
}

hipError_t mesh_launch_vertices(hipStream_t s, const MeshVolume &v, uint32_t cells, const uint32_t *vidx,
                                int32_t *cell_vertex, double *pos_d, float *pos_f) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (cells == 0) return hipSuccess;
    hipLaunchKernelGGL(vertex_kernel, grid_for(cells), dim3(256), 0, s, v, cells, vidx, cell_vertex, pos_d, pos_f);
    return hipGetLastError();
}

hipError_t mesh_launch_quads(hipStream_t s, const MeshVolume &v, uint32_t cells, const int32_t *cell_vertex,
                             const double *pos, const uint32_t *qoff, uint32_t *qcount, uint32_t *indices) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (cells == 0) return hipSuccess;
    if (!indices)
        hipLaunchKernelGGL(quad_count_kernel, grid_for(cells), dim3(256), 0, s, v, cells, cell_vertex, qcount);
    else
        hipLaunchKernelGGL(quad_emit_kernel, grid_for(cells), dim3(256), 0, s, v, cells, cell_vertex, pos, qoff, indices);
    return hipGetLastError();
}

hipError_t mesh_launch_degree(hipStream_t s, const uint32_t *indices, uint32_t n_indices, uint32_t *deg) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_indices) hipLaunchKernelGGL(degree_kernel, grid_for(n_indices), dim3(256), 0, s, indices, n_indices, deg);
    return hipGetLastError();
}

hipError_t mesh_launch_adjacency(hipStream_t s, const uint32_t *indices, uint32_t n_indices, const uint32_t *off,
                                 uint32_t *cursor, uint32_t *adj) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_indices)
        hipLaunchKernelGGL(adjacency_kernel, grid_for(n_indices), dim3(256), 0, s, indices, n_indices, off, cursor, adj);
    return hipGetLastError();
}

hipError_t mesh_launch_normal_sum(hipStream_t s, const float *pos, const uint32_t *indices, const uint32_t *off,
                                  const uint32_t *deg, const uint32_t *adj, uint32_t n_vertices, float *normals) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_vertices)
        hipLaunchKernelGGL(normal_kernel, grid_for(n_vertices), dim3(256), 0, s, pos, indices, off, deg, adj, n_vertices,
                           normals);
    return hipGetLastError();
}

hipError_t mesh_launch_components(hipStream_t s, const uint32_t *indices, uint32_t n_faces, uint32_t n_vertices,
                                  uint32_t *parent, uint32_t *changed_dev, uint32_t *changed_host) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_vertices == 0) return hipSuccess;
    hipLaunchKernelGGL(iota_kernel, grid_for(n_vertices), dim3(256), 0, s, parent, n_vertices);
    hipError_t e = hipGetLastError();
    // labels strictly decrease, so this ends; a path of n vertices needs at most n rounds
    for (uint32_t round = 0; e == hipSuccess && round <= n_vertices; round++) {
        if ((e = hipMemsetAsync(changed_dev, 0, 4, s)) != hipSuccess) break;
        if (n_faces) hipLaunchKernelGGL(hook_kernel, grid_for(n_faces), dim3(256), 0, s, indices, n_faces, parent, changed_dev);
        hipLaunchKernelGGL(jump_kernel, grid_for(n_vertices), dim3(256), 0, s, parent, n_vertices);
        if ((e = hipMemcpyAsync(changed_host, changed_dev, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
        if (*changed_host == 0u) break;
    }
    return e;
}

hipError_t mesh_launch_keep(hipStream_t s, const uint32_t *indices, uint32_t n_faces, const uint32_t *parent,
                            uint32_t *faces, double limit, uint32_t *keep) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_faces == 0) return hipSuccess;
    hipLaunchKernelGGL(component_faces_kernel, grid_for(n_faces), dim3(256), 0, s, indices, n_faces, parent, faces);
    hipLaunchKernelGGL(keep_kernel, grid_for(n_faces), dim3(256), 0, s, indices, n_faces, parent, faces, limit, keep);
    return hipGetLastError();
}

hipError_t mesh_launch_first_use(hipStream_t s, const uint32_t *indices, uint32_t n_faces, const uint32_t *keep,
                                 const uint32_t *koff, uint32_t *first, uint32_t *flag) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_faces == 0) return hipSuccess;
    hipLaunchKernelGGL(first_use_kernel, grid_for(n_faces), dim3(256), 0, s, indices, n_faces, keep, koff, first);
    hipLaunchKernelGGL(first_flag_kernel, grid_for(n_faces), dim3(256), 0, s, indices, n_faces, keep, koff, first, flag);
    return hipGetLastError();
}

hipError_t mesh_launch_compact(hipStream_t s, const uint32_t *indices, uint32_t n_faces, const uint32_t *keep,
                               const uint32_t *koff, const uint32_t *first, const uint32_t *rank, const float *pos,
                               const float *nrm, uint32_t *out_idx, float *out_pos, float *out_nrm) {
    (void)hipGetLastError();  // report this launch's status, not an earlier call's
    if (n_faces == 0) return hipSuccess;
    hipLaunchKernelGGL(compact_kernel, grid_for(n_faces), dim3(256), 0, s, indices, n_faces, keep, koff, first, rank,
                       pos, nrm, out_idx, out_pos, out_nrm);
    return hipGetLastError();
}

}  // namespace lr3d
