/*
 * CPU oracle for the dense-reconstruction hot path (PatchMatch MVS + TSDF integration).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's WGSL kernels and
 * their host drivers, used as the parity checker for the HIP library and as the timed CPU
 * baseline.  Nothing in the product library links, loads or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to the reference
 * repository root):
 *   PM = frontend/src/app/vision/mvs/patchmatch.wgsl.ts
 *   PK = frontend/src/app/vision/mvs/depth-packing.wgsl.ts
 *   DS = frontend/src/app/vision/mvs/dense-stereo.ts
 *   TV = frontend/src/app/vision/fusion/tsdf-volume.ts
 *   CAM = frontend/src/app/vision/geometry/camera.ts
 *
 * Arithmetic definition.  WGSL leaves the precision of division, sqrt, exp/sin/cos, texture
 * filtering and FMA contraction to the implementation, so the reference itself has no single
 * bit pattern.  The oracle fixes one (and the HIP kernels follow the same one):
 *   - FP32, no contraction (-ffp-contract=off); fmaf() only where written out (a homography
 *     tap is q = fma(H_0, X, fma(H_1, Y, H_2)) per row);
 *   - division and sqrt are IEEE correctly rounded;
 *   - exp, sin, cos are evaluated in double and rounded to float;
 *   - images are sampled in u8 units (0..255) by an FP32 bilinear filter with clamp-to-edge;
 *     NCC is invariant to that scale, and the two scale-dependent constants are rescaled
 *     (sigma_colour 0.06 -> 15.3, variance floor 1e-5 -> 1e-5*255^2);
 *   - f32 -> f16 rounds to nearest even.
 * Pinning of each function (closed forms, brute force, invariants) lives in tests/.
 */
#ifndef LR3D_ORACLE_H
#define LR3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SOURCES 8

/* One camera: intrinsics + world->camera pose (x_cam = R X + t), CV axes.  CAM:12-26 */
typedef struct {
    int32_t width, height;
    double fx, fy, cx, cy;
    double R[9]; /* row-major */
    double t[3];
} orc_view;

/* One source view in a PatchMatch dispatch: the `View` struct (PM:4-10) after the host packed
 * it (DS:345-353): intrinsics, relative rotation ref->src (rows), translation, layer. */
typedef struct {
    float k[4];
    float R[9];
    float t[3];
    int32_t layer;
    int32_t width, height; /* source image size (equal to the reference size in the reference) */
    int64_t map_off;       /* pixel offset of the layer's packed map in the `depths` pool */
    /* plane-independent parts of the homography (PM:116-128), formed in double and rounded:
     * G = Ks R Kr^-1, e = Ks t, so that H = G + (e / rho) (Kr^-T n)^T */
    float G[9];
    float e[3];
} orc_src;

/* The `Params` uniform (PM:12-26) as DenseStereo.params() fills it (DS:314-355). */
typedef struct {
    int32_t width, height, ref;
    float k[4];
    float inv_min, inv_max, step, geo_weight; /* range */
    uint32_t iteration, phase, seed, flags;   /* state */
    uint32_t needed;                          /* band.z (consistency filter) */
    int32_t n_src;
    int32_t topk; /* 0 = reference rule max(1, min(3, n-1)) (PM:218) */
    int64_t ref_off; /* pixel offset of the reference layer's packed map in `depths` */
    orc_src src[ORC_MAX_SOURCES];
} orc_params;

/* Images: one u8 plane per layer, each with its own size. */
typedef struct {
    int32_t n_layers;
    const uint8_t *const *pixels;
    const int32_t *widths;
    const int32_t *heights;
} orc_images;

/* ---- RNG (PM:56-65, PM:251-253, DS:332-334) ---- */
uint32_t orc_pcg(uint32_t v);
uint32_t orc_host_seed(int32_t ref, uint32_t iteration, uint32_t phase, uint32_t flags);
uint32_t orc_pixel_rng(uint32_t seed, uint32_t iteration, uint32_t phase, uint32_t index);
float orc_random(uint32_t *rng);

/* ---- packing (PK:3-33) ---- */
uint32_t orc_pack_sample(float depth, const float normal[3], float cost);
float orc_unpack_depth(uint32_t packed);
void orc_unpack_normal(uint32_t packed, float out[3]);
void orc_decode_map(const uint32_t *packed, int64_t n, double *depth, double *normal, uint8_t *cost);
uint32_t orc_unpack_cost(uint32_t packed);
uint16_t orc_f32_to_f16(float v);
float orc_f16_to_f32(uint16_t h);

/* ---- host parameter packing (DS:314-355, CAM:73-77) ---- */
void orc_make_params(const orc_view *views, int32_t n_views, int32_t ref, const int32_t *sources,
                     int32_t n_src, double min_depth, double max_depth, uint32_t iteration,
                     uint32_t phase, double step, uint32_t flags, uint32_t needed, int32_t topk,
                     orc_params *out);

/* ---- cost pieces, exposed for the pinning tests ---- */
/* Bilateral window of pixel (x, y) (PM:90-110): values[36], weights[36], stats[3] = sum, mean, var */
void orc_window(const orc_params *p, const orc_images *img, int32_t x, int32_t y, float *values,
                float *weights, float *stats);
/* Plane homography of source s (PM:116-128), row-major 3x3: H = G + (e / rho) m^T with
 * m = Kr^-T n, in the fixed order of orc_homography */
void orc_homography(const orc_params *p, int32_t s, const float n[3], float rho, float H[9]);
/* Bilateral NCC cost of one source view (PM:130-159) */
float orc_view_cost(const orc_params *p, const orc_images *img, int32_t s, const float H[9],
                    int32_t x, int32_t y, const float *values, const float *weights,
                    const float *stats);
/* Packed depth maps of all layers live in one pool, `depths`: layer l's map starts at pixel offset
 * orc_map_offset(views, l) = sum of width*height over the layers before it (the reference keeps
 * one W*H*4-byte slice per layer, DS:97-98, which is this layout when all sizes are equal). */
int64_t orc_map_offset(const orc_view *views, int32_t layer);

/* Forward-backward reprojection of reference point x (camera coordinates, seen at pixel (px, py))
 * through source s's packed map (PM:166-190): out = (pixel distance, depth in the reference,
 * 0, 1), or all -1 when the point is behind the source, projects outside its image, or meets no
 * usable depth there (depth <= 0 or cost code > 7). */
void orc_reproject(const orc_params *p, const uint32_t *depths, int32_t s, float px, float py,
                   const float x[3], float out[4]);

/* Aggregated cost of a plane hypothesis at pixel (x, y) (PM:192-233).  `depths` (the map pool)
 * is read only when flags & 1. */
float orc_hypothesis_cost(const orc_params *p, const orc_images *img, const uint32_t *depths,
                          int32_t x, int32_t y, const float *values, const float *weights,
                          const float *stats, const float n[3], float d);

/* ---- kernels (whole-image dispatches; bands do not change results) ---- */
/* PM:255-289.  hyp: W*H*4 floats (n.xyz, d); cost: W*H.  prior may be NULL (= zeros). */
void orc_init(const orc_params *p, const orc_images *img, const float *prior,
              const uint32_t *depths, float *hyp, float *cost);
/* PM:349-423, one red-black phase (p->phase). */
void orc_propagate(const orc_params *p, const orc_images *img, const uint32_t *depths, float *hyp,
                   float *cost);
/* PM:425-433 */
void orc_pack(const orc_params *p, const float *hyp, const float *cost, uint32_t *out);
/* PM:436-469 */
void orc_consistency(const orc_params *p, const uint32_t *depths, uint32_t *filtered);
/* PM:296-333: coarse-to-fine start from the coarse level's planes (coarse_k, coarse W x H). */
void orc_init_coarse(const orc_params *p, const orc_images *img, const float coarse_k[4],
                     int32_t coarse_w, int32_t coarse_h, const float *coarse_hyp, float *hyp,
                     float *cost);

/* DenseStereo.solve (DS:185-239): init, `iterations` x {phase 0, 1}, pack into depths[ref].
 * geometric: flags bit 0.  Returns 0, or -1 when there are no sources (a no-op, DS:187). */
int32_t orc_solve(const orc_view *views, int32_t n_views, const orc_images *img, int32_t ref,
                  const int32_t *sources, int32_t n_src, double min_depth, double max_depth,
                  const float *prior, int32_t geometric, int32_t iterations, int32_t topk,
                  int32_t pack, uint32_t *depths, float *hyp, float *cost);

/* DenseStereo.filter (DS:250-263) */
void orc_filter(const orc_view *views, int32_t n_views, int32_t ref, const int32_t *checks,
                int32_t n_checks, int32_t needed, const uint32_t *depths, uint32_t *filtered);

/* ---- TSDF (TV:13-65, TV:103-154, TV:171-268) ---- */
typedef struct {
    double origin[3];
    double voxel;
    int32_t dims[3];       /* grid size in bricks */
    int32_t n_bricks;
    const int32_t *coords; /* n_bricks * 4: (bx, by, bz, 0) */
    double truncation;     /* metres (reference: 4 * voxel, TV:179) */
} orc_tsdf_grid;

typedef struct {
    int32_t width, height;
    double fx, fy, cx, cy;
    double R[9];
    double t[3];
    double max_depth;
    double confidence;
} orc_integration_view;

/* BrickGrid.build over all cells (dense grid) (TV:54-64); coords must hold cells*4 ints */
int32_t orc_dense_coords(const int32_t dims[3], int32_t *coords);
/* TsdfVolume.visibleBricks (TV:220-234): flags[b] = 1 if brick b is visible */
int32_t orc_visible_bricks(const orc_tsdf_grid *g, const orc_integration_view *v, uint8_t *flags);
/* TsdfVolume.integrate (TV:240-268) + the WGSL `integrate` (TV:103-154) for one view.
 * depth: packed map W*H; prior may be NULL.  Returns the number of voxel updates. */
int64_t orc_tsdf_integrate(const orc_tsdf_grid *g, const orc_integration_view *v,
                           const uint32_t *depth, const float *prior, float *tsdf, float *weight);

/* ---- surface nets and mesh post-processing (mesh_oracle.c) ---- */
typedef struct {
    double origin[3];
    double voxel;
    int32_t dims[3];        /* grid size in bricks */
    int32_t n_bricks;
    const int32_t *lookup;  /* dims product: grid index -> pool brick or -1 */
    const float *tsdf;      /* n_bricks * 512 */
    const float *weight;
} orc_sdf_bricks;

typedef struct {
    float *positions;  /* n_vertices * 3 */
    float *normals;    /* n_vertices * 3 */
    uint32_t *indices; /* n_indices */
    int64_t n_vertices, n_indices;
} orc_mesh;

/* extractSurface (surface-nets.ts:47-166) with computeNormals (mesh.ts:9-32) */
int32_t orc_extract_surface(const orc_sdf_bricks *vol, double min_weight, orc_mesh *out);
/* removeSmallComponents (mesh.ts:38-90) */
int32_t orc_remove_small_components(const orc_mesh *in, double fraction, int32_t min_faces, orc_mesh *out);
void orc_compute_normals(const float *positions, const uint32_t *indices, int64_t n_indices,
                         int64_t n_vertices, float *normals);
void orc_mesh_free(orc_mesh *m);

/* ---- dense-cloud fusion (fusion_oracle.c) ---- */
typedef struct {
    int32_t width, height;
    double fx, fy, cx, cy;
    double R[9], t[3];       /* world -> camera */
    const uint32_t *packed;  /* width*height packed map */
    int32_t thumb_width, thumb_height;
    const uint8_t *rgb;      /* thumb_width*thumb_height*3 */
} orc_fusion_view;

typedef struct {
    int64_t n;
    float *positions, *normals;
    uint8_t *colors;
    uint16_t *support;
} orc_cloud;

/* depth-maps.ts:65-85: out = depth, nx, ny, nz (camera frame); 0 for an empty sample */
void orc_upsample_depth(const float *meters, int32_t gw, int32_t gh, int32_t width, int32_t height, float *out);
int32_t orc_unpack_depth_sample(uint32_t value, double out[4]);
/* fuseDepthMaps (depth-fusion.ts:23-74) with PointAccumulator (point-accumulator.ts), cloud(min_support) */
int32_t orc_fuse_depth_maps(const orc_fusion_view *views, int32_t n_views, double cell, int64_t cell_budget,
                            int32_t min_support, orc_cloud *out, double *final_cell);
void orc_cloud_free(orc_cloud *c);
void *orc_acc_create(double cell);
void orc_acc_add(void *acc, const double v[9]);
void orc_acc_coarsen(void *acc, double factor);
int64_t orc_acc_size(void *acc);
int64_t orc_acc_cloud(void *acc, int32_t min_support, float *positions, uint8_t *colors, uint16_t *support);
void orc_acc_destroy(void *acc);

/* ---- FP64 TSDF integration adjoint (tsdf_adjoint_oracle.c) ---- */
typedef struct {
    const double *depth;   /* W*H decoded map depth, 0 where the map has none */
    const double *normal;  /* W*H*3 decoded normals */
    const uint8_t *cost;   /* W*H packed cost codes (0..15) */
    const double *prior;   /* W*H or NULL */
} orc64_depth_input;

/* Integrate views[0..n) in order into a fresh volume (tsdf 1, weight 0) at the listed voxels
 * (pool indices) in FP64; L = sum (tsdf - target)^2.  Outputs (any may be NULL): tsdf_out and
 * kink_out per voxel (distance to the nearest decision boundary over all views: the nearest-pixel
 * rounding in pixels, the cam.z > 0.01 and sdf >= -mu tests in units of mu); grad_depth
 * per pixel of every view, view-major (dL/d depth used); grad_pose n*12 (dL/dR row-major, dL/dt). */
int32_t orc64_tsdf_backward(const orc_tsdf_grid *g, const orc_integration_view *views,
                            const orc64_depth_input *inputs, int32_t n_views, const uint32_t *voxels,
                            const double *targets, int64_t n_voxels, double *tsdf_out,
                            double *grad_depth, double *grad_pose, double *kink_out, double *loss);
int32_t orc64_tsdf_loss(const orc_tsdf_grid *g, const orc_integration_view *views, const orc64_depth_input *inputs,
                        int32_t n_views, const uint32_t *voxels, const double *targets, int64_t n_voxels,
                        const uint8_t *vis, double *loss, double *terms);

/* ---- FP64 cost and its gradient (adjoint_oracle.c) ---- */
typedef struct {
    double k[4];
    double R[9];
    double t[3];
    int32_t layer, width, height;
} orc64_src;

typedef struct {
    double k[4];
    double inv_min, inv_max;
    int32_t n_src, topk;
    orc64_src src[ORC_MAX_SOURCES];
} orc64_params;

typedef struct {
    double cost;
    double d_depth;
    double d_normal[3];
    double d_R[ORC_MAX_SOURCES][9];
    double d_t[ORC_MAX_SOURCES][3];
    int32_t selected[ORC_MAX_SOURCES];
    double view_cost[ORC_MAX_SOURCES];
    double min_kink;
} orc64_result;

void orc64_homography(const orc64_params *p, int32_t s, const double n[3], double rho, double H[9]);
/* Aggregated cost of plane (n, d) at reference pixel (x, y) (layer `ref`) and its gradient. */
int32_t orc64_cost_grad(const orc64_params *p, const orc_images *img, int32_t ref, int32_t x, int32_t y,
                        const double n[3], double d, orc64_result *out);

/* Thread count used by the OpenMP loops (0 = runtime default). */
void orc_set_threads(int32_t n);
int32_t orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
