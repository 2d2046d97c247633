// Internal declarations shared by the lr3d HIP translation units.
//
// Device data layout (all views of a stereo set share one pool per kind, each layer at its own
// pixel offset `off[layer]` = sum of the previous layers' width*height):
//   quad   u32 per pixel: the 2x2 block I(x,y) | I(x+1,y)<<8 | I(x,y+1)<<16 | I(x+1,y+1)<<24,
//          clamped to the image, so one 32-bit gather feeds one bilinear tap.  A layer (at its
//          own offset `qoff[layer]`, a whole number of 128-byte lines) is stored in column blocks
//          of 16 texels: block b holds columns 16b..16b+15 of every row, row after row (64 bytes
//          per row, `qrows` rows per block), so one 128-byte line holds a 16 x 2 tile and the
//          rows of the ~16 x 8 source patch that one tap of a wavefront gathers sit in
//          consecutive lines (quad_index);
//   hyp    float4 per pixel (normal xyz, depth), cost f32 per pixel;
//   depth  packed u32 per pixel (f16 depth | 6+6-bit octahedral normal | 4-bit cost);
//   filt   packed u32 per pixel (consistency-filtered depth);
//   prior  f32 per pixel (only for solves that pass one).
// TSDF: brick pool, voxel (x, y, z) of pool brick b at b*512 + x + 8y + 64z, tsdf and weight in
// two f32 arrays (the reference layout, consumed unchanged by the surface-nets mesher).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace lr3d {

constexpr int kMaxSources = 8;
constexpr int kSamples = 36;

// Bilateral window weights (PM:90-110) in u8 units: w = exp(-(dc^2) kColorK - |o|^2 kSpaceK) with
// kColorK = 1 / (2 (0.06 * 255)^2) (PM:39) and kSpaceK = 1 / (2 * 5^2) (PM:40).  dc is an integer
// in [-255, 255] and the squared tap offset |o|^2 one of six values, so every weight is an entry of
// a kWindowClasses x 256 table (row = offset class, column = |dc|), formed once on the host: the
// argument in f32 as written, exp in double, rounded to f32.
constexpr float kColorK = (float)(1.0 / (2.0 * (0.06 * 255.0) * (0.06 * 255.0)));
constexpr float kSpaceK = 0.02f;
constexpr int kWindowClasses = 6;
// squared offsets |o|^2 = ox^2 + oy^2 with |ox|, |oy| in {1, 3, 5}, in table-row order
constexpr int kWindowClassDist[kWindowClasses] = {2, 10, 18, 26, 34, 50};
// row of the class of tap offset (ox, oy)
__host__ __device__ constexpr int window_class(int ox, int oy) {
    const int d = ox * ox + oy * oy;
    return d == 2 ? 0 : d == 10 ? 1 : d == 18 ? 2 : d == 26 ? 3 : d == 34 ? 4 : 5;
}

// One source view of a reference view's solve (the WGSL `View`, PM:4-10, after DS:345-353).
struct PmSrc {
    float k[4];     // fx, fy, cx, cy
    float R[9];     // relative rotation ref -> src, row-major
    float t[3];
    int32_t W, H;
    uint32_t off;   // pixel offset of the source layer in the per-pixel pools
    int32_t layer;
    int32_t qrows;  // rows per column block of the layer in the quad pool (quad_rows)
    uint32_t qoff;  // texel offset of the layer in the quad pool
    // The plane homography H = Ks (R + t n^T / rho) Kr^-1 (PM:116-128) split into its parts that do
    // not depend on the plane, formed in double on the host and rounded: G = Ks R Kr^-1, e = Ks t;
    // then H = G + (e / rho) m^T with m = Kr^-T n per hypothesis.
    float G[9];
    float e[3];
};

// One reference view's solve (the WGSL `Params`, PM:12-26, minus the per-launch state).
struct PmDesc {
    int32_t W, H, ref, n_src;
    float k[4];
    float ikx, iky, kx2, ky2;  // 1/fx, 1/fy, -cx/fx, -cy/fy of the reference (f32 divisions)
    float inv_min, inv_max, geo_weight;
    uint32_t flags;            // bit 0: geometric pass
    int32_t topk;              // effective k (1..n_src)
    int32_t iterations;
    uint32_t seed_base;        // (ref + 1) * 0x9e3779b1 + 7 * flags  (DS:333)
    uint32_t off;              // pixel offset of the reference layer
    int32_t has_prior;
    uint32_t prior_off;
    int32_t from_coarse;
    int32_t coarse_W, coarse_H;
    uint32_t coarse_off;       // pixel offset of this layer in the coarse stereo's hyp pool
    float ck[4];               // coarse intrinsics (scaleIntrinsics, CAM:28-36)
    int32_t tiles_x;           // tiles per row for the launch's tile shape (set per launch plan)
    int32_t qrows;             // quad-pool block rows and offset of the reference layer
    uint32_t qoff;
    int32_t pad_;
    PmSrc src[kMaxSources];
};

// Pools a PatchMatch launch reads and writes.
struct PmPools {
    const uint32_t *quad;
    float4 *hyp;
    float *cost;
    uint32_t *depth;
    uint32_t *filt;
    const float *prior;
    const float4 *coarse_hyp;
    const float *wtab;  // bilateral weight table, kWindowClasses x 256
};

// Launch plan: view v owns logical blocks [block_start[v], block_start[v+1]).
struct PmPlan {
    const PmDesc *descs;        // device
    const uint32_t *block_start;  // device, n+1 entries
    int32_t n;
    uint32_t blocks;            // total
    int32_t tx, ty;             // threads per block (x, y)
    int32_t kmax;               // largest top-k of the batch (selects the 3- or 8-slot network)
    int32_t geometric;          // any solve of the batch is a geometric pass
};

// FP64 camera geometry of one solve, read by the adjoint's backward passes: reference intrinsics
// and, per source, intrinsics and the relative pose (CAM:73-77) before rounding to FP32.
struct AdjGeom {
    double kr[4];
    double k[kMaxSources][4];
    double R[kMaxSources][9];
    double t[kMaxSources][3];
};

// Doubles per block in the adjoint's partial sums: loss before, loss after the plane step, then
// dL/dR (9) and dL/dt (3) of each source.
constexpr int32_t kAdjPartial = 2 + kMaxSources * 12;

// One view of a TSDF integration pass (the 128-byte uniform of TV:244-253 plus culling data).
struct TsdfView {
    float k[4];
    float R[9];
    float t[3];
    int32_t W, H;
    float confidence;
    uint32_t map_off;   // word offset of the packed map
    int32_t has_prior;
    uint32_t prior_off;
    // depth tiles of the view's map (max usable depth per kTsdfTile^2 pixels, and per
    // kTsdfTile2^2 of those): word offsets and tile-grid sizes
    uint32_t tile_off, tile2_off;
    int32_t tw, th, tw2, th2;
    double kd[4];       // double copies for the brick culling (TV:220-234 runs in double)
    double Rd[9];
    double td[3];
    double max_depth;
};

constexpr int32_t kTsdfTile = 16;   // pixels per depth-tile side
constexpr int32_t kTsdfTile2 = 8;   // depth tiles per coarse-tile side (128 pixels)

struct TsdfGridDev {
    float origin[3];
    float voxel;
    float mu;
    float inv_mu;  // RN(1 / mu), formed on the host
    int32_t n_bricks;
    double origin_d[3];
    double voxel_d;
    double mu_d;
};

// Kernel launchers (stream-ordered; no synchronisation).
hipError_t pm_build_quad(hipStream_t s, const uint8_t *gray, int32_t W, int32_t H, int32_t qrows, uint32_t *quad);

// Hardware block i of an n-block launch runs on XCD i % 8: the logical block it takes, such that
// each XCD gets a contiguous range of logical blocks (neighbouring blocks then share that XCD's L2).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    uint32_t xcd = i % kXcds, local = i / kXcds;
    uint32_t per = n / kXcds, rem = n % kXcds;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Rows per column block of an H-tall layer in the quad pool: H rounded up so that a block is an
// odd number of 128-byte lines (qrows / 2 odd), so that the blocks a gather straddles start in
// different L1 tag banks.
__host__ __device__ inline int32_t quad_rows(int32_t H) {
    const int32_t lines = (H + 1) / 2;
    return ((lines & 1) ? lines : lines + 1) * 2;
}
// Texels of a W x H layer in the quad pool (whole column blocks; a multiple of 32 texels).
__host__ __device__ inline size_t quad_texels(int32_t W, int32_t H) { return (size_t)((W + 15) / 16) * 16 * (size_t)quad_rows(H); }
hipError_t pm_launch_init(hipStream_t s, const PmPlan &plan, const PmPools &pools);
hipError_t pm_launch_propagate(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                               uint32_t iteration, uint32_t phase, float step);
hipError_t pm_launch_pack(hipStream_t s, const PmPlan &plan, const PmPools &pools);
hipError_t pm_launch_consistency(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                                 uint32_t needed);
// NCC adjoint: costs, dL/d(depth, normal) per pixel and per-block pose partials (kAdjPartial
// doubles per block); apply != 0 also takes one gradient step on the planes, kept where it lowers
// the pixel's cost.
hipError_t ncc_launch_backward(hipStream_t s, const PmPlan &plan, const PmPools &pools,
                               const AdjGeom *geom, const float *weights, float *cost, float *gd,
                               float *gn, double *pose_partials, float step_depth, float step_normal,
                               int32_t apply);
// Depth tiles of every view of a pass (tiles: sum of tw*th + tw2*th2 floats, at the views'
// offsets); max_tw/max_th: the largest tile grid among the views.
hipError_t tsdf_launch_depth_tiles(hipStream_t s, const TsdfView *views, int32_t n_views, int32_t max_tw,
                                   int32_t max_th, const uint32_t *maps, const float *priors, float *tiles);
// Culling bits: the reference's brick test, then a depth test against the view's depth tiles
// (tiles == nullptr: the reference's test alone).
// Culling of every (brick, view) pair into vis, then, when list is not NULL, the list of the bricks
// with a kept view (list: n_bricks entries; ctrs[0]: its length).
hipError_t tsdf_launch_visibility(hipStream_t s, const TsdfGridDev &g, const int4 *coords,
                                  const TsdfView *views, int32_t n_views, uint32_t *vis,
                                  int32_t words, const float *tiles, uint32_t *ctrs, uint32_t *list);
// Words of the pass counters (ctrs): the brick list's length.
constexpr int32_t kTsdfCounters = 1;
// Passes of at most this many views walk the brick list (tsdf_launch_integrate with a list);
// larger ones launch a wavefront per pool brick.  Measured on the 200-view room set integrated in
// passes of n views (integration + culling, all 200 views): the list walk takes 16.1 / 11.9 / 9.6 /
// 8.6 ms for n = 1 / 2 / 4 / 8, a wavefront per pool brick 19.4 / 12.4 / 8.8 / 7.4 ms: past two
// views the list's static striding balances worse than the hardware's dispatch.
constexpr int32_t kListPassViews = 2;
// ntab: the decoded normal of each packed octahedral code (tsdf_launch_normal_table); fast: the
// pass lies in the ranges of the correctly rounded short forms (see integrate_kernel).
hipError_t tsdf_launch_integrate(hipStream_t s, const TsdfGridDev &g, const TsdfView *views,
                                 int32_t n_views, const uint32_t *vis, int32_t words,
                                 const int4 *coords, const uint32_t *maps, const float *priors,
                                 const float4 *ntab, bool fast, float *tsdf, float *weight,
                                 unsigned long long *counter, const uint32_t *ctrs, const uint32_t *list,
                                 int32_t bricks_per_block, int32_t chunk);
// Largest chunk of consecutive blocks dealt to one XCD by tsdf_launch_integrate.
constexpr int32_t kMaxXcdChunk = 4096;
// The 4096-entry normal table of tsdf_launch_integrate.
hipError_t tsdf_launch_normal_table(hipStream_t s, float4 *tab);
// TSDF adjoint over listed voxels (pool indices); n_voxels > 0.  ckpt holds
// ceil(n_views / 16) * n_voxels float2; grad_depth (per pixel of each view, at pixel_off[v]) and
// grad_pose (n_views * 12) are accumulated (zero them first); loss_partials: one per wavefront.
hipError_t tsdf_launch_backward(hipStream_t s, const TsdfGridDev &g, const TsdfView *views, int32_t n_views,
                                const uint32_t *vis, int32_t words, const int4 *coords, const uint32_t *maps,
                                const float *priors, const uint32_t *voxels, const float *targets,
                                uint32_t n_voxels, const uint32_t *pixel_off, float2 *ckpt, float *tsdf_out,
                                double *grad_depth, double *grad_pose, double *loss_partials);
hipError_t launch_to_float(hipStream_t s, const double *in, float *out, size_t n);

// Brick TSDF as the surface extractor reads it (the reference's SdfBricks, surface-nets.ts:4-12).
struct MeshVolume {
    double origin[3];
    double voxel;
    double min_weight;
    int32_t dims[3];            // grid size in bricks
    const int32_t *lookup;      // grid index -> pool brick or -1
    const int4 *coords;         // pool brick -> grid coordinates
    const float *tsdf, *weight;
    const uint32_t *bricks = nullptr;  // the pool bricks the cell kernels run over (nullptr: all)
};

hipError_t mesh_launch_live_bricks(hipStream_t s, const MeshVolume &v, uint32_t n_bricks, uint32_t *live);
hipError_t mesh_launch_brick_list(hipStream_t s, const uint32_t *live, const uint32_t *pos, uint32_t n_bricks,
                                  uint32_t *list);
hipError_t mesh_launch_active(hipStream_t s, const MeshVolume &v, uint32_t cells, uint32_t *flag);
hipError_t mesh_launch_vertices(hipStream_t s, const MeshVolume &v, uint32_t cells, const uint32_t *vidx,
                                int32_t *cell_vertex, double *pos_d, float *pos_f);
// indices == NULL: count the quads of every voxel into qcount; else write their triangles at qoff
hipError_t mesh_launch_quads(hipStream_t s, const MeshVolume &v, uint32_t cells, const int32_t *cell_vertex,
                             const double *pos, const uint32_t *qoff, uint32_t *qcount, uint32_t *indices);
hipError_t mesh_launch_degree(hipStream_t s, const uint32_t *indices, uint32_t n_indices, uint32_t *deg);
hipError_t mesh_launch_adjacency(hipStream_t s, const uint32_t *indices, uint32_t n_indices, const uint32_t *off,
                                 uint32_t *cursor, uint32_t *adj);
hipError_t mesh_launch_normal_sum(hipStream_t s, const float *pos, const uint32_t *indices, const uint32_t *off,
                                  const uint32_t *deg, const uint32_t *adj, uint32_t n_vertices, float *normals);
// union-find labels (synchronises the stream once per round)
hipError_t mesh_launch_components(hipStream_t s, const uint32_t *indices, uint32_t n_faces, uint32_t n_vertices,
                                  uint32_t *parent, uint32_t *changed_dev, uint32_t *changed_host);
hipError_t mesh_launch_keep(hipStream_t s, const uint32_t *indices, uint32_t n_faces, const uint32_t *parent,
                            uint32_t *faces, double limit, uint32_t *keep);
hipError_t mesh_launch_first_use(hipStream_t s, const uint32_t *indices, uint32_t n_faces, const uint32_t *keep,
                                 const uint32_t *koff, uint32_t *first, uint32_t *flag);
hipError_t mesh_launch_compact(hipStream_t s, const uint32_t *indices, uint32_t n_faces, const uint32_t *keep,
                               const uint32_t *koff, const uint32_t *first, const uint32_t *rank, const float *pos,
                               const float *nrm, uint32_t *out_idx, float *out_pos, float *out_nrm);

// One view of a dense-cloud fusion pass (DF:44-74): camera in double (JavaScript numbers).
// Dense-cloud fusion takes the views in chunks: at most this many map pixels per chunk (one flag and
// one sample number each), and at most this many views.
constexpr uint64_t kFusionChunkPixels = (uint64_t)1 << 26;
constexpr int32_t kFusionMaxChunkViews = 1024;

struct FusionView {
    int32_t W, H;
    double k[4];
    double R[9], t[3];
    uint32_t map_off;      // pixel offset of the layer's packed map
    int32_t tw, th;        // thumbnail size
    uint64_t thumb_off;    // byte offset of the thumbnail (RGB)
};

hipError_t fusion_launch_new_cell_flags(hipStream_t s, const uint32_t *idx, const uint32_t *start, const uint32_t *is_new,
                                        uint32_t segments, uint32_t lo, uint32_t hi, uint32_t *flag);
hipError_t fusion_launch_select(hipStream_t s, const uint32_t *flag, const uint32_t *scan, uint32_t n, uint32_t k,
                                uint32_t *out);
hipError_t fusion_launch_chunk_valid(hipStream_t s, const FusionView *views, const uint32_t *vstart, uint32_t nv,
                                     uint32_t pixels, uint32_t from, const uint32_t *maps, uint32_t *flag);
hipError_t fusion_launch_chunk_keys(hipStream_t s, const FusionView *views, const uint32_t *vstart, uint32_t nv,
                                    uint32_t pixels, const uint32_t *maps, const uint8_t *thumbs, const uint32_t *flag,
                                    const uint32_t *pos, double cell, uint64_t *keys, uint32_t *idx,
                                    uint32_t *overflow);
hipError_t fusion_launch_first_views(hipStream_t s, const uint32_t *idx, const uint32_t *start, const uint32_t *is_new,
                                     uint32_t segments, const uint32_t *vstart, uint32_t nv, uint32_t *hist);
hipError_t fusion_launch_limit(hipStream_t s, const uint32_t *idx, const uint32_t *start, uint32_t segments,
                               uint32_t limit, uint32_t *is_new);
hipError_t fusion_launch_heads(hipStream_t s, const uint64_t *keys, uint32_t n, uint32_t *head);
hipError_t fusion_launch_starts(hipStream_t s, const uint32_t *head, const uint32_t *seg, uint32_t n, uint32_t *start);
hipError_t fusion_launch_lookup(hipStream_t s, const uint64_t *keys, const uint32_t *start, uint32_t segments,
                                const uint64_t *state_keys, uint32_t state_n, int32_t *exist, uint32_t *is_new);
hipError_t fusion_launch_reduce(hipStream_t s, const FusionView *views, const uint32_t *vstart, uint32_t nv,
                                uint32_t limit, const uint32_t *maps, const uint8_t *thumbs, const uint64_t *keys,
                                const uint32_t *idx, const uint32_t *start, uint32_t segments, uint32_t n,
                                const int32_t *exist, const uint32_t *is_new, const uint32_t *new_pos,
                                float *state_sums, uint32_t *state_counts, uint64_t *new_keys, float *new_sums,
                                uint32_t *new_counts, uint32_t *big, uint32_t *n_big);
hipError_t fusion_launch_gather(hipStream_t s, const uint32_t *perm, uint32_t n, const uint64_t *keys, const float *sums,
                                const uint32_t *counts, uint64_t *out_keys, float *out_sums, uint32_t *out_counts);
hipError_t fusion_launch_rekey(hipStream_t s, const float *sums, const uint32_t *counts, uint32_t n, double cell,
                               uint64_t *keys, uint32_t *idx, uint32_t *overflow);
hipError_t fusion_launch_merge(hipStream_t s, const uint64_t *keys, const uint32_t *idx, const uint32_t *start,
                               uint32_t segments, uint32_t n, const float *sums, const uint32_t *counts,
                               uint64_t *out_keys, float *out_sums, uint32_t *out_counts);
hipError_t fusion_launch_support(hipStream_t s, const uint32_t *counts, uint32_t n, uint32_t min_support, uint32_t *flag);
hipError_t fusion_launch_cloud(hipStream_t s, const float *sums, const uint32_t *counts, uint32_t n, const uint32_t *flag,
                               const uint32_t *pos, float *positions, float *normals, uint8_t *colors,
                               uint16_t *support);

// Launch shapes (x threads, y threads) picked per (kernel, image size).
struct TileShape {
    int32_t tx, ty;
};
TileShape pm_tile_for(const char *kernel, int32_t W, int32_t H, int32_t n_src);

}  // namespace lr3d
