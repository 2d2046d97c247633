/*
 * lr3d — MI355X (gfx950) dense-reconstruction hot path: PatchMatch multi-view stereo and TSDF
 * depth-map integration, behind a plain C ABI.
 *
 * Reference paths are relative to the reference repository root:
 *   DS = frontend/src/app/vision/mvs/dense-stereo.ts
 *   PM = frontend/src/app/vision/mvs/patchmatch.wgsl.ts
 *   TV = frontend/src/app/vision/fusion/tsdf-volume.ts
 *   SM = frontend/src/app/vision/mvs/methods/stereo-method.ts
 *   GD = frontend/src/app/vision/gpu/gpu-device.ts
 *
 * Conventions
 *   - Every call returns an int32_t status: LR3D_OK (0) or a negative LR3D_ERR_* code; the
 *     message of the last failure on the calling thread is lr3d_last_error().
 *   - Host memory passed in is copied before the call returns (like queue.writeBuffer /
 *     writeTexture); read calls copy device data into caller-owned memory.
 *   - Device memory is owned by the handle that allocated it and freed by its destroy call.
 *   - Each context owns one HIP stream; calls on one context are serialised by a mutex.  Calls
 *     return when their device work has finished (the reference awaits every submit).
 *   - Poses are world -> camera (x_cam = R X + t), R row-major, computer-vision axes, exactly the
 *     reference's `Pose` (frontend/src/app/vision/geometry/camera.ts:19-26).
 */
#ifndef LR3D_H
#define LR3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LR3D_OK 0
#define LR3D_ERR_INVALID (-1) /* bad argument */
#define LR3D_ERR_DEVICE (-2)  /* HIP runtime / device failure, or no device */
#define LR3D_ERR_NOMEM (-3)   /* device allocation failed */
#define LR3D_ERR_STATE (-4)   /* call not valid in the handle's current state */

#define LR3D_MAX_SOURCES 8 /* DS:44 MAX_SOURCES */

typedef struct lr3d_ctx lr3d_ctx;
typedef struct lr3d_stereo lr3d_stereo;
typedef struct lr3d_tsdf lr3d_tsdf;

/* One camera of a stereo set: the reference's StereoView {intrinsics, pose} (DS:11-14) plus its
 * own image size (the reference uses one size for all views; per-view sizes allow batches of
 * mixed resolution). */
typedef struct {
    int32_t width, height;
    double fx, fy, cx, cy;
    double R[9];
    double t[3];
} lr3d_view;

/* One PatchMatch solve: the reference's SolveOptions (DS:16-29) for reference view `ref`. */
typedef struct {
    int32_t ref;
    int32_t n_sources;                  /* sliced to LR3D_MAX_SOURCES like DS:186 */
    int32_t sources[LR3D_MAX_SOURCES];
    double min_depth, max_depth;
    const float *prior;                 /* width*height depth guess (0 = none), or NULL */
    int32_t geometric;                  /* second pass: geometric-consistency term */
    int32_t iterations;
    int32_t from_coarse;                /* start from the coarse level's planes */
    int32_t pack;                       /* store the packed depth map (reference default 1) */
    int32_t topk;                       /* views averaged per cost; 0 = reference rule (PM:218) */
} lr3d_solve_desc;

/* One depth map to integrate: the reference's IntegrationView (TV:67-79) plus the image size. */
typedef struct {
    int32_t width, height;
    double fx, fy, cx, cy;
    double R[9];
    double t[3];
    double max_depth;
    const float *prior;  /* width*height phone depth, or NULL (reference: zeros) */
    double confidence;   /* typical scene depth for confidence weighting; 0 = off */
} lr3d_integration_view;

/* Depth-map selectors for the TSDF calls (which map of a stereo layer is integrated). */
#define LR3D_MAP_RAW 0      /* the packed map written by solve (DS `depths[layer]`) */
#define LR3D_MAP_FILTERED 1 /* the consistency-filtered map written by filter (DS `filtered`) */

/* ---- runtime (replaces requestGpu / readBuffer / storageBuffer, GD:29-101) ---------------- */

const char *lr3d_last_error(void);
/* Open device `device` and create its stream. */
int32_t lr3d_ctx_create(int32_t device, lr3d_ctx **out);
/* Stereo sets, volumes, meshes and clouds made on a context keep it alive: if any remain, the
 * context is released by the destroy of the last of them.  The caller must not use `ctx` itself
 * after this call. */
int32_t lr3d_ctx_destroy(lr3d_ctx *ctx);
/* Block until the context's stream is idle. */
int32_t lr3d_ctx_synchronize(lr3d_ctx *ctx);
/* Kernel timing with HIP events recorded around each launch on the context's stream.
 * enable: 1 on, 0 off (statistics are reset when turned on). */
int32_t lr3d_ctx_set_profiling(lr3d_ctx *ctx, int32_t enable);
/* Launch count and total milliseconds of kernel `name` since profiling was enabled.  Profiled
 * calls launch eagerly (HIP cannot time event records inside graphs); the name "graph_launch"
 * returns the number of graph replays of the context instead. */
int32_t lr3d_ctx_kernel_stats(lr3d_ctx *ctx, const char *name, int64_t *launches, double *total_ms);

/* ---- PatchMatch stereo (replaces DenseStereo, DS:61-371) -------------------------------- */

/* Flags of lr3d_stereo_create. */
#define LR3D_STEREO_PRIOR_POOL 1u /* allocate the prior pool now (else on the first solve with a
                                     prior); lr3d_stereo_solve_batched_noalloc needs it for priors */

/* new DenseStereo(device, width, height, views) (DS:80-172).  flags: 0 or LR3D_STEREO_PRIOR_POOL. */
int32_t lr3d_stereo_create(lr3d_ctx *ctx, const lr3d_view *views, int32_t n_views, uint32_t flags,
                           lr3d_stereo **out);
int32_t lr3d_stereo_destroy(lr3d_stereo *st);                    /* DS:357-370 */
/* DenseStereo.setImage (DS:175-182): u8 gray, width x height of that layer. */
int32_t lr3d_stereo_set_image(lr3d_stereo *st, int32_t layer, const uint8_t *gray, int32_t width,
                              int32_t height);
/* DenseStereo.solve (DS:185-239): init, `iterations` red-black rounds, pack.  One launch per
 * kernel per phase, as the reference dispatches it (the "plain" path). */
int32_t lr3d_stereo_solve(lr3d_stereo *st, const lr3d_solve_desc *desc);
/* All solves of `descs` at once, the launch sequence replayed from a captured HIP graph.  Results
 * equal n calls of lr3d_stereo_solve in array order.  Photometric solves share every launch; a
 * geometric pass that reads (as a source) the packed map of an earlier solve of the batch starts a
 * new group of launches after that solve's pack, so it sees the refined map, as the reference's
 * layer-by-layer geometric pass does (…/pipeline/reconstructor.ts:96-108). */
int32_t lr3d_stereo_solve_batched(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n);
/* No counterpart in the reference: lr3d_stereo_solve_batched for callers that cannot allocate
 * device memory (code under their own HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same solves, launches and results (equal to n calls of lr3d_stereo_solve in array
 * order), but the launch plan lives in a fixed per-GPU scratch array of 1 MiB instead of a buffer
 * the set allocates, the launches are issued eagerly (no graph is captured or instantiated), and
 * kernel timing (lr3d_ctx_set_profiling) does not cover them.  At most LR3D_NOALLOC_MAX_SOLVES
 * solves per call (LR3D_ERR_INVALID beyond); a solve with a prior needs a set created with
 * LR3D_STEREO_PRIOR_POOL (LR3D_ERR_STATE otherwise).  Every call on one GPU reuses that GPU's
 * scratch array: calls that overlap on one GPU (sets of different contexts used from several
 * threads) race on it; calls on one context are serialised. */
#define LR3D_NOALLOC_MAX_SOLVES 1024
int32_t lr3d_stereo_solve_batched_noalloc(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n);
/* DenseStereo.filter (DS:250-263): consistency check of `ref` into its filtered map. */
int32_t lr3d_stereo_filter(lr3d_stereo *st, int32_t ref, const int32_t *checks, int32_t n_checks,
                           int32_t needed);
/* DenseStereo.depthMap (DS:309-312): packed u32 map of a layer (width*height words). */
int32_t lr3d_stereo_read_depth(lr3d_stereo *st, int32_t layer, uint32_t *out);
/* DenseStereo.filteredMap readback (DS:301-306) of the last filter of `layer`. */
int32_t lr3d_stereo_read_filtered(lr3d_stereo *st, int32_t layer, uint32_t *out);
/* DenseStereo.hypotheses (DS:242-244): planes (n.xyz, depth) and costs of a layer's last solve.
 * hyp: width*height*4 floats, cost: width*height floats (either may be NULL). */
int32_t lr3d_stereo_read_hyp(lr3d_stereo *st, int32_t layer, float *hyp, float *cost);
/* Overwrite a layer's packed depth map (e.g. maps computed elsewhere, or restored state). */
int32_t lr3d_stereo_write_depth(lr3d_stereo *st, int32_t layer, const uint32_t *packed);
/* Attach a coarse level for from_coarse solves (coarse-to-fine, DS:31-36 CoarseLevel): the
 * coarse stereo must hold one layer per layer of `st` and be another set of the same context;
 * its planes are read, not copied.  NULL detaches.  Destroying the coarse level detaches it from
 * every set using it (their from_coarse solves then fail with LR3D_ERR_STATE). */
int32_t lr3d_stereo_set_coarse(lr3d_stereo *st, lr3d_stereo *coarse);

/* Replace a layer's pose (world -> camera) for later solves, filters and adjoints (pose
 * refinement; the reference keeps poses in its StereoView list, DS:11-14). */
int32_t lr3d_stereo_set_pose(lr3d_stereo *st, int32_t layer, const double R[9], const double t[3]);
/* Overwrite a layer's planes (n.xyz, depth per pixel, width*height*4 floats). */
int32_t lr3d_stereo_write_hyp(lr3d_stereo *st, int32_t layer, const float *hyp);

/* ---- Launch-shape autotuner (no counterpart in the reference) -----------------------------
 * Times the candidate block shapes of init, propagate (second iteration) and pack/consistency on
 * this batch (`repeats` launches each, 1..16) and keeps the fastest per kernel for later calls on
 * the set.  The set is left as found (planes, costs and packed maps are restored); maps do not
 * depend on the shapes.  kernel names: "init", "propagate", "pack"; shapes are threads (x, y)
 * with x*y a multiple of 64 up to 256. */
int32_t lr3d_stereo_autotune(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n, int32_t repeats);
int32_t lr3d_stereo_get_tile(lr3d_stereo *st, const char *kernel, int32_t *tx, int32_t *ty);
int32_t lr3d_stereo_set_tile(lr3d_stereo *st, const char *kernel, int32_t tx, int32_t ty);

/* ---- Adjoint of the aggregated NCC cost (no counterpart in the reference) -------------------
 * C_p = the aggregated cost (PM:192-233) of the current plane of pixel p of desc->ref, evaluated as
 * the forward kernels do; L = sum_p weights[p] * C_p (weights NULL = 1).  Windows and taps are
 * recomputed, nothing is stored by the forward pass.  Outputs (each may be NULL): cost (width*height),
 * grad_depth = dL/dd (width*height), grad_normal = dL/dn (width*height*3), grad_pose = dL/dR (9,
 * row-major) and dL/dt (3) of the relative pose ref -> source of each of the 8 source slots
 * (8*12 doubles, summed over pixels in a fixed order), loss = L over pixels with a valid plane.
 * INVALID, degenerate and clamped view costs and the top-k choice carry no gradient. */
int32_t lr3d_ncc_backward(lr3d_stereo *st, const lr3d_solve_desc *desc, const float *weights,
                          float *cost, float *grad_depth, float *grad_normal, double *grad_pose,
                          double *loss);
/* One gradient step on the planes of every desc's reference view (weights 1): depth moves by
 * -step_depth * dL/dd (at most 0.5 %, kept in the depth range), the normal by -step_normal times
 * the tangential part of dL/dn (kept facing the camera); a pixel keeps the step only where its
 * aggregated cost becomes lower, and the packed depth maps of descs with pack set are rewritten.
 * Returns, per desc, the relative-pose gradients at the planes before the step (n*8*12 doubles, as
 * lr3d_ncc_backward) and the loss before and after the step (n doubles each; any may be NULL).
 * With both steps 0 it evaluates only: planes and packed maps are left unchanged. */
int32_t lr3d_ncc_refine_step(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n,
                             float step_depth, float step_normal, double *grad_pose,
                             double *loss_before, double *loss_after);

/* ---- TSDF integration (replaces TsdfVolume, TV:158-291) ---------------------------------- */

/* new TsdfVolume(device, grid, layout, depth, W, H) (TV:171-217): brick pool of n_bricks 8^3
 * bricks at `coords` (n_bricks*4 ints: bx, by, bz, 0, as BrickGrid.build returns them),
 * tsdf := 1, weight := 0.  truncation_voxels: the reference uses 4 (TV:179). */
int32_t lr3d_tsdf_create(lr3d_ctx *ctx, const double origin[3], double voxel,
                         const int32_t dims[3], const int32_t *coords, int32_t n_bricks,
                         double truncation_voxels, lr3d_tsdf **out);
int32_t lr3d_tsdf_destroy(lr3d_tsdf *ts);                        /* TV:279-290 */
/* TsdfVolume.integrate (TV:240-268) of one map of a stereo layer (LR3D_MAP_*). */
int32_t lr3d_tsdf_integrate(lr3d_tsdf *ts, lr3d_stereo *st, int32_t layer, int32_t map,
                            const lr3d_integration_view *view);
/* Same, for a packed map given in host memory (view->width * view->height words). */
int32_t lr3d_tsdf_integrate_host(lr3d_tsdf *ts, const uint32_t *packed,
                                 const lr3d_integration_view *view);
/* n views in one pass over the volume; per voxel the views apply in array order, so the result
 * equals n calls of lr3d_tsdf_integrate in that order. */
int32_t lr3d_tsdf_integrate_batched(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers,
                                    int32_t map, const lr3d_integration_view *views, int32_t n);
/* Adjoint of TSDF integration (no counterpart in the reference).  tsdf_q is the value voxel q
 * (pool index brick*512 + x + 8y + 64z) holds after integrating views[0..n) of the given layers'
 * maps in order into a fresh volume (tsdf 1, weight 0), with the inputs, culling and arithmetic of
 * lr3d_tsdf_integrate_batched (the volume itself is neither read nor modified);
 * L = sum_q (tsdf_q - targets[q])^2.  Outputs (each may be NULL): tsdf_out (n_voxels), grad_depth
 * = dL/d(depth used at each pixel: the map's depth where the map has one, else the prior), per
 * view in list order (sum of width*height floats), grad_pose = dL/dR (9, row-major) and dL/dt (3)
 * of each view's world -> camera pose (n*12 doubles), loss.  The nearest-pixel choice, the skips
 * and the clamps carry no gradient.  Gradients are accumulated with FP64 atomics: runs agree to
 * FP64 rounding, not bit for bit. */
int32_t lr3d_tsdf_backward(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers, int32_t map,
                           const lr3d_integration_view *views, int32_t n, const uint32_t *voxels,
                           const float *targets, int64_t n_voxels, float *tsdf_out, float *grad_depth,
                           double *grad_pose, double *loss);

/* Times integrate passes of these views with 1, 2 and 4 bricks (wavefronts) per block, each with
 * blocks dealt to the GPU's 8 XCDs one at a time or in chunks of 16, 32, 64, 128 and one grid row of
 * consecutive blocks, and keeps the fastest pair for later passes; the volume is restored.  No
 * counterpart in the reference. */
int32_t lr3d_tsdf_autotune(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers, int32_t map,
                           const lr3d_integration_view *views, int32_t n, int32_t repeats);
int32_t lr3d_tsdf_get_bricks_per_block(lr3d_tsdf *ts, int32_t *bricks);
int32_t lr3d_tsdf_set_bricks_per_block(lr3d_tsdf *ts, int32_t bricks);
/* Consecutive blocks of the integration pass given to one XCD (neighbouring bricks then share the
 * depth-map lines of their footprints in that XCD's L2): 1 (round robin, the default) to 4096.
 * Results do not depend on it.  No counterpart in the reference. */
int32_t lr3d_tsdf_get_xcd_chunk(lr3d_tsdf *ts, int32_t *chunk);
int32_t lr3d_tsdf_set_xcd_chunk(lr3d_tsdf *ts, int32_t chunk);

/* Depth culling (on by default): before a pass, every view's map is reduced to the largest usable
 * depth per 16x16 and 128x128 pixels, and a (brick, view) pair that the reference's culling keeps
 * (TV:220-234) is dropped when no voxel of the brick can be updated by that view (all projections
 * outside the image, or all voxels more than the truncation behind every depth they can see).
 * Volumes, update counts and adjoints are identical with it on or off.  No counterpart in the
 * reference. */
int32_t lr3d_tsdf_set_depth_culling(lr3d_tsdf *ts, int32_t on);

/* Overwrite the volume's tsdf and weight pools (n_bricks*512 floats each). */
int32_t lr3d_tsdf_write(lr3d_tsdf *ts, const float *tsdf, const float *weight);

/* ---- Surface extraction (replaces extractSurface, surface-nets.ts:47-166, with computeNormals,
 * mesh.ts:9-32; and removeSmallComponents, mesh.ts:38-90) --------------------------------------
 * The mesh equals the reference's array for array: vertices numbered in pool order, triangles in
 * voxel / axis order, normals accumulated face by face into floats.  min_weight as the
 * reference's minWeight (0.6 in tsdf-meshing.ts:23).  Meshes live on the device until read. */
typedef struct lr3d_mesh lr3d_mesh;
int32_t lr3d_tsdf_extract_surface(lr3d_tsdf *ts, double min_weight, lr3d_mesh **out);
int32_t lr3d_mesh_remove_small_components(lr3d_mesh *in, double fraction, int32_t min_faces, lr3d_mesh **out);
int32_t lr3d_mesh_size(lr3d_mesh *m, int64_t *vertices, int64_t *indices);
/* positions, normals: vertices*3 floats; indices: indices uint32 (any may be NULL) */
int32_t lr3d_mesh_read(lr3d_mesh *m, float *positions, float *normals, uint32_t *indices);
int32_t lr3d_mesh_destroy(lr3d_mesh *m);

/* ---- Dense-cloud fusion (replaces fuseDepthMaps, depth-fusion.ts:23-74, and PointAccumulator,
 * point-accumulator.ts) ----------------------------------------------------------------------
 * Decodes the layers' packed maps (map selector) with the stereo set's cameras, colours each
 * sample from its view's RGB thumbnail, and averages position, normal and colour per cell of
 * `cell` metres, growing the cell by 1.4 while more than cell_budget cells are occupied (after
 * each view).  Like the reference's accumulator, it re-bins every cell by its mean position when
 * its hash table (65536 slots at first, doubling) would pass 60 % load, at the same sample.
 * Per-cell sums equal the reference's (float sums in sample order); points come out in cell-key
 * order (the reference's follow its hash table), and cells that a re-binning merges are merged in
 * key order (the reference: table order).  final_cell (may be NULL) receives the cell size used.
 * The views are processed in chunks (LR3D_FUSION_CHUNK_VIEWS=k in the environment fixes them at
 * k views; the result does not depend on it). */
typedef struct {
    int32_t width, height;
    const uint8_t *rgb;  /* width*height*3 */
} lr3d_thumb;
typedef struct lr3d_cloud lr3d_cloud;
int32_t lr3d_fuse_depth_maps(lr3d_stereo *st, const int32_t *layers, int32_t map, int32_t n,
                             const lr3d_thumb *thumbs, double cell, int64_t cell_budget, int32_t min_support,
                             lr3d_cloud **out, double *final_cell);
int32_t lr3d_cloud_size(lr3d_cloud *c, int64_t *points);
/* positions, normals: points*3 floats; colors: points*3 bytes; support: points (any may be NULL) */
int32_t lr3d_cloud_read(lr3d_cloud *c, float *positions, float *normals, uint8_t *colors, uint16_t *support);
int32_t lr3d_cloud_destroy(lr3d_cloud *c);

/* TsdfVolume.read (TV:270-277): n_bricks*512 floats each (either may be NULL). */
int32_t lr3d_tsdf_read(lr3d_tsdf *ts, float *tsdf, float *weight);
/* Voxel updates ((voxel, view) pairs written) of the integrate calls since the last reset. */
int32_t lr3d_tsdf_set_counting(lr3d_tsdf *ts, int32_t enable);
int32_t lr3d_tsdf_update_count(lr3d_tsdf *ts, int64_t *out);
/* The same counting passes' totals: out[0] voxel updates, out[1] (voxel, view) pairs evaluated
 * (512 per brick and view that the culling keeps), out[2] bricks read (at least one view kept),
 * out[3] voxels written (changed by the pass).  No counterpart in the reference (measurement). */
int32_t lr3d_tsdf_pass_counts(lr3d_tsdf *ts, int64_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
