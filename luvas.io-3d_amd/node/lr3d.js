'use strict';
/*
 * Host side of the lr3d path for Node / TypeScript (types: lr3d.d.ts).
 *
 * Drop-in counterparts of the reference's stage objects, over the N-API addon lr3d.node:
 *   HipDenseStereo  ~ DenseStereo   (frontend/src/app/vision/mvs/dense-stereo.ts)
 *   HipTsdfVolume   ~ TsdfVolume    (frontend/src/app/vision/fusion/tsdf-volume.ts)
 *   BrickGrid       ~ BrickGrid     (same file, :13-65)
 *   STEREO_METHODS  ~ the StereoMethod registry (frontend/src/app/vision/mvs/methods/index.ts)
 * Same member names, argument meaning and await points; GPU errors reject the promises.
 * Written for Node 12 (CommonJS, no optional chaining).
 */
const path = require('path');

const native = require(path.join(__dirname, '..', 'lib', 'lr3d.node'));

const MAP_RAW = 0;
const MAP_FILTERED = 1;
const MAX_SOURCES = 8;
const BRICK = 8;

class Context {
    /** Opens HIP device `device` (the reference's requestGpu()). */
    constructor(device) {
        this.handle = native.ctxCreate(device === undefined ? 0 : device);
    }
    /** Waits for the context's queued work. */
    synchronize() {
        native.ctxSynchronize(this.handle);
    }
    /** Kernel timing with events around every launch (statistics reset when turned on). */
    setProfiling(on) {
        native.ctxSetProfiling(this.handle, !!on);
    }
    /** {launches, ms} of kernel `name` since profiling was turned on. */
    kernelStats(name) {
        return native.ctxKernelStats(this.handle, name);
    }
    destroy() {
        if (this.handle) native.ctxDestroy(this.handle);
        this.handle = null;
    }
}

function cameraOf(view, width, height) {
    const k = view.intrinsics;
    return {
        width: view.width !== undefined ? view.width : width,
        height: view.height !== undefined ? view.height : height,
        fx: k.fx, fy: k.fy, cx: k.cx, cy: k.cy,
        R: Array.from(view.pose.R), t: Array.from(view.pose.t),
    };
}

class HipDenseStereo {
    /**
     * new DenseStereo(device, width, height, views, options) (DS:80-172).  `views` may carry
     * their own width/height (mixed-resolution sets); options.coarse is the coarse level;
     * options.priorPool allocates the prior pool now (solveBatchedNoalloc needs it for priors).
     */
    constructor(ctx, width, height, views, options) {
        this.ctx = ctx;
        this.width = width;
        this.height = height;
        this.views = views.slice();  // own copy: setPose replaces entries here, not in the caller's array
        this.cameras = views.map((v) => cameraOf(v, width, height));
        this.handle = native.stereoCreate(ctx.handle, this.cameras, options && options.priorPool ? 1 : 0);
        this.lastFiltered = -1;
        this.filteredMapBinding = new FilteredMap(this);
        const coarse = options && options.coarse ? options.coarse : null;
        if (coarse) native.stereoSetCoarse(this.handle, coarse.hypotheses.handle);
    }

    /** DS:175-182 */
    setImage(layer, gray) {
        const c = this.cameras[layer];
        native.stereoSetImage(this.handle, layer, gray, c.width, c.height);
    }

    /** SolveOptions of layer `ref` for the addon; `pixels` lets it check a prior's length. */
    descOf(ref, options) {
        const c = this.cameras[ref];
        return {
            ref,
            pixels: c ? c.width * c.height : -1,
            sources: options.sources.slice(0, MAX_SOURCES),
            minDepth: options.minDepth,
            maxDepth: options.maxDepth,
            prior: options.prior || null,
            geometric: !!options.geometric,
            iterations: options.iterations,
            fromCoarse: !!options.fromCoarse,
            pack: options.pack === undefined ? true : options.pack,
            topk: options.topk || 0,
        };
    }

    /** DS:185-239 */
    async solve(ref, options) {
        if (options.sources.length === 0) return;
        await native.stereoSolve(this.handle, this.descOf(ref, options));
    }

    /**
     * Every (ref, options) pair in shared launches; the same result as solving them in order (a
     * geometric pass that reads an earlier solve's map runs after that solve's pack).
     */
    async solveBatched(solves) {
        const descs = solves.filter((s) => s[1].sources.length > 0)
            .map((s) => this.descOf(s[0], s[1]));
        if (descs.length) await native.stereoSolve(this.handle, descs);
    }

    /**
     * solveBatched for callers that cannot allocate device memory: the launch plan goes to a fixed
     * per-GPU scratch array (at most 1024 solves), launches are eager, priors need a set created
     * with options.priorPool.
     */
    async solveBatchedNoalloc(solves) {
        const descs = solves.filter((s) => s[1].sources.length > 0)
            .map((s) => this.descOf(s[0], s[1]));
        if (descs.length) await native.stereoSolveNoalloc(this.handle, descs);
    }

    /**
     * Times the launch-shape candidates on these solves and keeps the fastest per kernel; the set's
     * planes, costs and maps are restored afterwards.  Returns the chosen shapes.
     */
    async autotune(solves, repeats) {
        const descs = solves.filter((s) => s[1].sources.length > 0)
            .map((s) => this.descOf(s[0], s[1]));
        if (descs.length) await native.stereoAutotune(this.handle, descs, repeats === undefined ? 3 : repeats);
        return this.tiles();
    }

    /** Launch shapes [threads x, threads y] of 'init', 'propagate' and 'pack'. */
    tiles() {
        const out = {};
        for (const k of ['init', 'propagate', 'pack']) out[k] = native.stereoGetTile(this.handle, k);
        return out;
    }

    setTile(kernel, tx, ty) {
        native.stereoSetTile(this.handle, kernel, tx, ty);
    }

    /** Overwrites a layer's packed depth map (width*height words). */
    writeDepth(layer, packed) {
        const c = this.cameras[layer];
        native.stereoWriteDepth(this.handle, layer, packed, c.width * c.height);
    }

    /** Overwrites a layer's planes (n.xyz, depth per pixel: 4*width*height floats). */
    writeHyp(layer, planes) {
        const c = this.cameras[layer];
        native.stereoWriteHyp(this.handle, layer, planes, c.width * c.height);
    }

    /** The planes of the last solve (DS:242-244): this level, usable as a coarse level. */
    get hypotheses() {
        return this;
    }

    /**
     * The filtered map (DS:67 `filtered`): the map of the layer filtered last, as a binding that
     * HipTsdfVolume accepts in place of the reference's GPUBuffer (tsdf-meshing.ts:71).
     */
    get filtered() {
        return this.filteredMapBinding;
    }

    /** DS:250-263 */
    async filter(ref, sources, needed) {
        await native.stereoFilter(this.handle, ref, sources.slice(0, MAX_SOURCES), needed);
        this.lastFiltered = ref;
    }

    /** DS:301-306 */
    async filteredMap(ref, sources, needed) {
        await this.filter(ref, sources, needed);
        const c = this.cameras[ref];
        return native.stereoReadFiltered(this.handle, ref, c.width, c.height);
    }

    /** DS:309-312 */
    async depthMap(ref) {
        const c = this.cameras[ref];
        return native.stereoReadDepth(this.handle, ref, c.width, c.height);
    }

    /** Planes (n.xyz, depth) and costs of a layer's last solve. */
    async planes(ref) {
        const c = this.cameras[ref];
        return native.stereoReadHyp(this.handle, ref, c.width, c.height);
    }

    /**
     * Adjoint of L = sum_p weights[p] * C_p (the aggregated NCC cost of the current planes of `ref`):
     * {cost, gradDepth, gradNormal (3 per pixel), gradPose (8 sources x (dL/dR 9, dL/dt 3)), loss}.
     */
    async nccBackward(ref, options, weights) {
        const c = this.cameras[ref];
        const w = weights === undefined || weights === null ? null : Float32Array.from(weights);
        return native.nccBackward(this.handle, this.descOf(ref, options), w, c.width * c.height);
    }

    /**
     * One gradient step on the planes of every [ref, options] (kept per pixel where the cost drops):
     * {gradPose (per solve, 8 x 12), lossBefore, lossAfter}.  Steps of 0 only evaluate.
     */
    async refineStep(solves, stepDepth, stepNormal) {
        return native.nccRefineStep(this.handle, solves.map((s) => this.descOf(s[0], s[1])),
            stepDepth, stepNormal);
    }

    /** Replaces a layer's world -> camera pose {R (9, row-major), t (3)}. */
    setPose(layer, pose) {
        native.stereoSetPose(this.handle, layer, Array.from(pose.R), Array.from(pose.t));
        this.views[layer] = Object.assign({}, this.views[layer], { pose });
        this.cameras[layer] = cameraOf(this.views[layer], this.width, this.height);
    }

    /**
     * fuseDepthMaps(maps, thumbs, cell, cellBudget) (depth-fusion.ts:23-74) on the GPU, over these
     * layers' maps (MAP_FILTERED by default): {positions, normals, colors, support, cell}.  Points
     * come in cell-key order; per-cell values equal the reference accumulator's.
     */
    async fuseDepthMaps(layers, thumbs, cell, cellBudget, map, minSupport) {
        return native.stereoFuseDepthMaps(this.handle, layers, thumbs, cell, cellBudget,
            map === undefined ? MAP_FILTERED : map, minSupport === undefined ? 1 : minSupport);
    }

    destroy() {
        if (this.handle) native.stereoDestroy(this.handle);
        this.handle = null;
    }
}

/**
 * What `stereo.filtered` hands to a volume: the reference binds the one `filtered` buffer, which
 * holds the map of the layer filtered last (DS:250-263); integrate() reads that layer's map.
 */
class FilteredMap {
    constructor(stereo) {
        this.stereo = stereo;
    }
}

class BrickGrid {
    /** TV:17-29 */
    constructor(origin, voxel, extent) {
        this.origin = origin;
        this.voxel = voxel;
        const side = voxel * BRICK;
        this.dims = extent.map((e) => Math.max(1, Math.ceil(e / side)));
        this.marks = new Uint8Array(this.dims[0] * this.dims[1] * this.dims[2]);
    }

    get cells() {
        return this.marks.length;
    }

    /** TV:36-51 */
    mark(x, y, z, radius) {
        const side = this.voxel * BRICK;
        const o = this.origin;
        const d = this.dims;
        const lo = [x, y, z].map((v, a) => Math.max(0, Math.floor((v - radius - o[a]) / side)));
        const hi = [x, y, z].map((v, a) => Math.min(d[a] - 1, Math.floor((v + radius - o[a]) / side)));
        for (let bz = lo[2]; bz <= hi[2]; bz++) {
            for (let by = lo[1]; by <= hi[1]; by++) {
                for (let bx = lo[0]; bx <= hi[0]; bx++) this.marks[(bz * d[1] + by) * d[0] + bx] = 1;
            }
        }
    }

    /** TV:54-64 */
    build() {
        const lookup = new Int32Array(this.marks.length).fill(-1);
        const coords = [];
        const dx = this.dims[0];
        const dy = this.dims[1];
        for (let i = 0; i < this.marks.length; i++) {
            if (!this.marks[i]) continue;
            lookup[i] = coords.length / 4;
            coords.push(i % dx, Math.floor(i / dx) % dy, Math.floor(i / (dx * dy)), 0);
        }
        return { coords: Int32Array.from(coords), lookup, count: coords.length / 4 };
    }
}

class HipTsdfVolume {
    /**
     * new TsdfVolume(device, grid, layout, depth, width, height) (TV:171-217); `depth` is
     * `stereo.filtered`, as the reference passes it (tsdf-meshing.ts:71), or the stereo level
     * itself.  truncationVoxels: 4 in the reference (TV:179).
     */
    constructor(ctx, grid, layout, depth, width, height, truncationVoxels) {
        if (!(depth instanceof FilteredMap) && !(depth instanceof HipDenseStereo)) {
            throw new TypeError('HipTsdfVolume: depth must be stereo.filtered or a HipDenseStereo');
        }
        this.ctx = ctx;
        this.grid = grid;
        this.layout = layout;
        this.stereo = depth instanceof FilteredMap ? depth.stereo : depth;
        this.width = width;
        this.height = height;
        this.truncation = grid.voxel * (truncationVoxels || 4);
        this.handle = native.tsdfCreate(ctx.handle, grid.origin, grid.voxel, grid.dims, layout.coords,
            truncationVoxels || 4);
    }

    static viewOf(view, width, height) {
        const c = cameraOf(view, width, height);
        c.maxDepth = view.maxDepth;
        c.prior = view.prior || null;
        c.confidence = view.confidence || 0;
        return c;
    }

    /** TV:240-268: integrates the map the stereo level filtered last. */
    async integrate(view) {
        const layer = this.stereo.lastFiltered;
        if (layer < 0) throw new Error('HipTsdfVolume.integrate: no filtered map yet (call stereo.filter first)');
        await native.tsdfIntegrate(this.handle, this.stereo.handle, [layer], MAP_FILTERED,
            [HipTsdfVolume.viewOf(view, this.width, this.height)]);
    }

    /** Several layers' maps (MAP_RAW or MAP_FILTERED) in one pass, applied in array order. */
    async integrateLayers(layers, views, map) {
        await native.tsdfIntegrate(this.handle, this.stereo.handle, layers, map === undefined ? MAP_RAW : map,
            views.map((v) => HipTsdfVolume.viewOf(v, this.width, this.height)));
    }

    /** A packed map from host memory (view.width * view.height words) in one pass. */
    async integrateHost(packed, view) {
        await native.tsdfIntegrateHost(this.handle, packed, HipTsdfVolume.viewOf(view, this.width, this.height));
    }

    /** Times bricks per block (1, 2, 4) and chunks of consecutive blocks per XCD on a pass of these
     * layers and keeps the fastest pair; resolves to the bricks per block (the chunk is xcdChunk). */
    async autotune(layers, views, map, repeats) {
        await native.tsdfAutotune(this.handle, this.stereo.handle, layers, map === undefined ? MAP_RAW : map,
            views.map((v) => HipTsdfVolume.viewOf(v, this.width, this.height)), repeats === undefined ? 3 : repeats);
        return this.bricksPerBlock;
    }

    get bricksPerBlock() {
        return native.tsdfGetBricksPerBlock(this.handle);
    }

    set bricksPerBlock(n) {
        native.tsdfSetBricksPerBlock(this.handle, n);
    }

    /** Consecutive blocks of the integration pass given to one XCD (1: round robin). */
    get xcdChunk() {
        return native.tsdfGetXcdChunk(this.handle);
    }

    set xcdChunk(n) {
        native.tsdfSetXcdChunk(this.handle, n);
    }

    /** Culling of (brick, view) pairs against the views' depth tiles (on by default; no result changes). */
    setDepthCulling(on) {
        native.tsdfSetDepthCulling(this.handle, !!on);
    }

    /** Counting of the passes that follow (off by default): updateCount(), passCounts(). */
    setCounting(on) {
        native.tsdfSetCounting(this.handle, !!on);
    }

    updateCount() {
        return native.tsdfUpdateCount(this.handle);
    }

    /** {updates, pairs, bricksRead, voxelsWritten} of the counted passes. */
    passCounts() {
        return native.tsdfPassCounts(this.handle);
    }

    /** Overwrites the tsdf and weight pools (layout.count * 512 floats each). */
    write(tsdf, weight) {
        native.tsdfWrite(this.handle, tsdf, weight, this.layout.count);
    }

    /** TV:270-277 */
    async read() {
        return native.tsdfRead(this.handle, this.layout.count);
    }

    /**
     * extractSurface(read(), minWeight) followed by removeSmallComponents(mesh, fraction, minFaces)
     * (surface-nets.ts:47-166, mesh.ts:38-90; tsdf-meshing.ts:84-90 uses 0.6, 0.01, 200), on the
     * GPU: {positions, normals, indices}.  fraction < 0 keeps every component.
     */
    async extractSurface(minWeight, fraction, minFaces) {
        return native.tsdfExtractSurface(this.handle, minWeight === undefined ? 0.6 : minWeight,
            fraction === undefined ? 0.01 : fraction, minFaces === undefined ? 200 : minFaces);
    }

    /**
     * Adjoint of integrating `views` (of `layers`' maps, in order, into a fresh volume) at the listed
     * voxels (pool indices) for L = sum (tsdf - target)^2: {tsdf, gradDepth (per view, concatenated),
     * gradPose (per view: dL/dR 9, dL/dt 3), loss}.  The volume is not modified.
     */
    async backward(layers, views, voxels, targets, map) {
        return native.tsdfBackward(this.handle, this.stereo.handle, layers, map === undefined ? MAP_RAW : map,
            views.map((v) => HipTsdfVolume.viewOf(v, this.width, this.height)), Uint32Array.from(voxels),
            Float32Array.from(targets));
    }

    destroy() {
        if (this.handle) native.tsdfDestroy(this.handle);
        this.handle = null;
    }
}

/**
 * The aggregated NCC cost as a forward/backward pair over the planes and relative poses of one
 * reference view.  forward() evaluates L = sum_p w_p C_p; backward() returns its gradients.  Both
 * come from one GPU pass that recomputes the NCC windows (nothing is stored between the calls
 * on the device; the pass's results are kept for backward()).
 */
class NccCostFunction {
    constructor(stereo, ref, options) {
        this.stereo = stereo;
        this.ref = ref;
        this.options = options;
        this.saved = null;
    }

    async forward(weights) {
        this.saved = await this.stereo.nccBackward(this.ref, this.options, weights);
        return { loss: this.saved.loss, cost: this.saved.cost };
    }

    async backward() {
        if (!this.saved) throw new Error('NccCostFunction.backward: call forward first');
        const { gradDepth, gradNormal, gradPose } = this.saved;
        return { gradDepth, gradNormal, gradPose };
    }
}

/**
 * TSDF integration as a forward/backward pair: forward() integrates the layers' maps into the
 * volume (as integrateLayers); backward() replays that integration from a fresh volume at the
 * listed voxels and returns the gradients of sum (tsdf - target)^2 with respect to the depths used
 * and the view poses.
 */
class TsdfIntegrationFunction {
    constructor(volume, layers, views, map) {
        this.volume = volume;
        this.layers = layers;
        this.views = views;
        this.map = map === undefined ? MAP_FILTERED : map;
    }

    async forward() {
        await this.volume.integrateLayers(this.layers, this.views, this.map);
    }

    async backward(voxels, targets) {
        return this.volume.backward(this.layers, this.views, voxels, targets, this.map);
    }
}

async function uploadImages(stereo, input) {
    for (let layer = 0; layer < input.views.length; layer++) {
        stereo.setImage(layer, await input.gray(layer, stereo.width, stereo.height));
    }
}

/** StereoMethod entries (methods/index.ts:7-10): create(ctx, input, settings) -> StereoRun. */
const STEREO_METHODS = {
    'hip-single-scale': {
        async create(ctx, input, settings) {
            const stereo = new HipDenseStereo(ctx, input.width, input.height, input.views);
            try {
                await uploadImages(stereo, input);
            } catch (error) {
                stereo.destroy();
                throw error;
            }
            return {
                stereo,
                solve: (layer) => stereo.solve(layer, {
                    sources: input.sources[layer],
                    minDepth: input.ranges[layer][0],
                    maxDepth: input.ranges[layer][1],
                    prior: input.prior(layer, input.width, input.height),
                    geometric: false,
                    iterations: settings.iterations,
                }),
                destroy: () => stereo.destroy(),
            };
        },
    },
    'hip-coarse-to-fine': {
        async create(ctx, input, settings) {
            const cw = Math.round(input.width / 2);
            const ch = Math.round(input.height / 2);
            const scale = (k) => ({
                fx: k.fx * cw / input.width, fy: k.fy * ch / input.height,
                cx: (k.cx + 0.5) * cw / input.width - 0.5, cy: (k.cy + 0.5) * ch / input.height - 0.5,
            });
            const coarse = new HipDenseStereo(ctx, cw, ch,
                input.views.map((v) => ({ pose: v.pose, intrinsics: scale(v.intrinsics) })));
            const stereo = new HipDenseStereo(ctx, input.width, input.height, input.views, { coarse });
            try {
                await uploadImages(coarse, input);
                await uploadImages(stereo, input);
            } catch (error) {
                stereo.destroy();
                coarse.destroy();
                throw error;
            }
            return {
                stereo,
                async solve(layer) {
                    const range = {
                        sources: input.sources[layer],
                        minDepth: input.ranges[layer][0],
                        maxDepth: input.ranges[layer][1],
                        geometric: false,
                    };
                    await coarse.solve(layer, Object.assign({}, range, {
                        prior: input.prior(layer, cw, ch), iterations: settings.iterations, pack: false,
                    }));
                    await stereo.solve(layer, Object.assign({}, range, {
                        prior: null, iterations: settings.fineIterations, fromCoarse: true,
                    }));
                },
                destroy() {
                    stereo.destroy();
                    coarse.destroy();
                },
            };
        },
    },
};

// ---- gradient refinement of planes and poses (no counterpart in the reference) -----------------
//
// The host side of the NCC adjoint's refinement, as in the Python host (lr3d/refine.py): pose
// perturbation on the left, R' = exp([w]x) R, t' = exp([w]x) t + tau; relative pose of reference r
// and source s R = R_s R_r^T, t = t_s - R t_r (camera.ts:73-77); matrices row-major arrays of 9.

function mat3Mul(A, B) {
    const C = new Array(9);
    for (let i = 0; i < 3; i++)
        for (let j = 0; j < 3; j++) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
    return C;
}
function mat3Vec(A, v) {
    return [0, 1, 2].map((i) => A[i * 3] * v[0] + A[i * 3 + 1] * v[1] + A[i * 3 + 2] * v[2]);
}
function mat3T(A) {
    return [A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]];
}
function cross(a, b) {
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
}
function norm3(v) {
    return Math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/** Rodrigues: exp([w]x). */
function expSo3(w) {
    const theta = norm3(w);
    const K = [0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0];
    const I = [1, 0, 0, 0, 1, 0, 0, 0, 1];
    if (theta < 1e-12) return I.map((x, i) => x + K[i]);
    const KK = mat3Mul(K, K);
    const a = Math.sin(theta) / theta, b = (1 - Math.cos(theta)) / (theta * theta);
    return I.map((x, i) => x + a * K[i] + b * KK[i]);
}

/** R' = exp([w]x) R, t' = exp([w]x) t + tau. */
function perturbPose(pose, w, tau) {
    const E = expSo3(w);
    const t = mat3Vec(E, Array.from(pose.t));
    return { R: mat3Mul(E, Array.from(pose.R)), t: [t[0] + tau[0], t[1] + tau[1], t[2] + tau[2]] };
}

/**
 * World-pose gradients {layer: {w: dL/dw, tau: dL/dtau}} from the relative-pose gradients of one
 * reference view's sources (gradRel: per source slot dL/dR (9, row-major) then dL/dt (3)).
 *   source:    dL/dw = sum_k R[:,k] x G[:,k] + t x g,  dL/dtau = g
 *   reference: dL/dw = -vee(B - B^T), B = R^T G,        dL/dtau = -R^T g
 */
function worldPoseGradients(views, ref, sources, gradRel, out) {
    out = out || new Map();
    const add = (layer, w, tau) => {
        const cur = out.get(layer) || { w: [0, 0, 0], tau: [0, 0, 0] };
        out.set(layer, { w: cur.w.map((x, i) => x + w[i]), tau: cur.tau.map((x, i) => x + tau[i]) });
    };
    const Rr = Array.from(views[ref].pose.R), tr = Array.from(views[ref].pose.t);
    sources.forEach((layer, s) => {
        const G = Array.from(gradRel.slice(s * 12, s * 12 + 9));
        const g = Array.from(gradRel.slice(s * 12 + 9, s * 12 + 12));
        const R = mat3Mul(Array.from(views[layer].pose.R), mat3T(Rr));
        const Rt = mat3Vec(R, tr);
        const t = Array.from(views[layer].pose.t).map((x, i) => x - Rt[i]);
        let wSrc = [0, 0, 0];
        for (let k = 0; k < 3; k++) {
            const c = cross([R[k], R[3 + k], R[6 + k]], [G[k], G[3 + k], G[6 + k]]);
            wSrc = wSrc.map((x, i) => x + c[i]);
        }
        const tg = cross(t, g);
        wSrc = wSrc.map((x, i) => x + tg[i]);
        const B = mat3Mul(mat3T(R), G);
        const wRef = [-(B[7] - B[5]), -(B[2] - B[6]), -(B[3] - B[1])];
        const RTg = mat3Vec(mat3T(R), g);
        add(layer, wSrc, g);
        add(ref, wRef, RTg.map((x) => -x));
    });
    return out;
}

function poseTotals(stereo, solves, gradPose) {
    const out = new Map();
    solves.forEach(([ref, o], i) => {
        const srcs = o.sources.slice(0, MAX_SOURCES);
        worldPoseGradients(stereo.views, ref, srcs, gradPose.subarray(i * MAX_SOURCES * 12, (i + 1) * MAX_SOURCES * 12), out);
    });
    return out;
}

const sum = (a) => a.reduce((x, y) => x + y, 0);

/**
 * `steps` joint steps on the planes of every [ref, options] and the poses of every view not in
 * options.fixed (default [0], the gauge): a plane step on the GPU (refineStep), then each view
 * turns by stepRot radians and moves by stepTrans metres along its negative gradient direction,
 * averaged with the previous steps' (momentum); the planes are solved again at the new poses and
 * given a plane step, and the pose step is kept if the loss is lower (steps x 1.5), else undone
 * (poses restored, planes solved again, which reproduces them) and halved.  The planes must be
 * solved at the current poses.  Returns the loss history (non-increasing).
 */
async function refine(stereo, solves, options) {
    const o = Object.assign({ steps: 50, stepDepth: 2e-4, stepNormal: 2e-3, stepRot: 1e-3, stepTrans: 1e-3,
        fixed: [0], momentum: 0.9 }, options || {});
    let rot = o.stepRot, trans = o.stepTrans;
    let r = await stereo.refineStep(solves, o.stepDepth, o.stepNormal);
    const history = [sum(r.lossBefore)];
    let loss = sum(r.lossAfter);
    let grads = r.gradPose;
    let velocity = new Map();
    for (let it = 0; it < o.steps; it++) {
        const totals = poseTotals(stereo, solves, grads);
        const saved = new Map();
        for (const [layer, gr] of totals) {
            if (o.fixed.includes(layer)) continue;
            saved.set(layer, stereo.views[layer].pose);
            const v = velocity.get(layer) || { w: [0, 0, 0], tau: [0, 0, 0] };
            const nw = norm3(gr.w) + 1e-30, nt = norm3(gr.tau) + 1e-30;
            const vw = v.w.map((x, i) => o.momentum * x + gr.w[i] / nw);
            const vt = v.tau.map((x, i) => o.momentum * x + gr.tau[i] / nt);
            velocity.set(layer, { w: vw, tau: vt });
            const sw = -rot / (norm3(vw) + 1e-30), st = -trans / (norm3(vt) + 1e-30);
            stereo.setPose(layer, perturbPose(stereo.views[layer].pose, vw.map((x) => x * sw), vt.map((x) => x * st)));
        }
        await stereo.solveBatched(solves);
        r = await stereo.refineStep(solves, o.stepDepth, o.stepNormal);
        if (sum(r.lossAfter) < loss) {
            loss = sum(r.lossAfter);
            grads = r.gradPose;
            rot *= 1.5;
            trans *= 1.5;
        } else {
            for (const [layer, pose] of saved) stereo.setPose(layer, pose);
            await stereo.solveBatched(solves);
            r = await stereo.refineStep(solves, o.stepDepth, o.stepNormal);
            grads = r.gradPose;
            loss = Math.min(loss, sum(r.lossAfter));
            rot *= 0.5;
            trans *= 0.5;
            velocity = new Map();
        }
        history.push(loss);
    }
    return history;
}

/** Every view's frame box-filtered 2 x 2 (rounded), intrinsics scaled by 0.5 with pixel centres
 * aligned (camera.ts:28-36), the same poses. */
function halfFrames(views, sizes, images) {
    const outViews = [], frames = [];
    views.forEach((v, i) => {
        const W = sizes[i].width, H = sizes[i].height;
        const w = Math.floor(W / 2), h = Math.floor(H / 2);
        const sx = w / W, sy = h / H, k = v.intrinsics;
        outViews.push({ intrinsics: { fx: k.fx * sx, fy: k.fy * sy, cx: (k.cx + 0.5) * sx - 0.5, cy: (k.cy + 0.5) * sy - 0.5 },
                        pose: v.pose, width: w, height: h });
        const im = images[i], out = new Uint8Array(w * h);
        for (let y = 0; y < h; y++)
            for (let x = 0; x < w; x++) {
                const a = (2 * y) * W + 2 * x;
                out[y * w + x] = (im[a] + im[a + W] + im[a + 1] + im[a + W + 1] + 2) >> 2;
            }
        frames.push(out);
    });
    return { views: outViews, frames };
}

function stereoOf(ctx, views, frames) {
    const st = new HipDenseStereo(ctx, views[0].width, views[0].height, views);
    frames.forEach((f, i) => st.setImage(i, f));
    return st;
}

/**
 * A half-resolution copy of a stereo set for the coarse stage of refinePyramid: frames box-filtered
 * 2 x 2 (rounded), intrinsics scaled by 0.5 with pixel centres aligned (camera.ts:28-36).
 */
function halfResolution(stereo, images) {
    const h = halfFrames(stereo.views, stereo.cameras, images);
    return stereoOf(stereo.ctx, h.views, h.frames);
}

/** `levels` coarse copies of a stereo set, each half the resolution of the one before (finest first). */
function pyramidLevels(stereo, images, levels) {
    let views = stereo.views, sizes = stereo.cameras, frames = images;
    const out = [];
    for (let l = 0; l < (levels === undefined ? 3 : levels); l++) {
        const h = halfFrames(views, sizes, frames);
        views = h.views;
        sizes = h.views;
        frames = h.frames;
        out.push(stereoOf(stereo.ctx, h.views, h.frames));
    }
    return out;
}

/**
 * Coarse to fine, as the Python host's lr3d.refine.refine_pyramid: `coarse` is one coarse copy or an
 * array of them (finest first), options.coarseSteps a count or one per copy; from the coarsest copy
 * up each level starts from the poses the level below reached, then options.steps steps run on
 * `stereo`.  Returns { coarse: histories (coarsest first; one history for a single copy), fine }.
 */
async function refinePyramid(stereo, coarse, solves, options) {
    const o = Object.assign({ coarseSteps: [50, 50, 50], steps: 25, fixed: [0] }, options || {});
    const single = !Array.isArray(coarse);
    const levels = single ? [coarse] : coarse;
    const counts = Array.isArray(o.coarseSteps) ? o.coarseSteps : levels.map(() => o.coarseSteps);
    if (counts.length !== levels.length) throw new RangeError('refinePyramid: one coarseSteps entry per coarse level');
    let poses = stereo.views.map((v) => v.pose);
    const histories = [];
    for (let l = levels.length - 1; l >= 0; l--) {
        const level = levels[l];
        for (let layer = 0; layer < stereo.views.length; layer++) level.setPose(layer, poses[layer]);
        await level.solveBatched(solves);
        histories.push(await refine(level, solves, Object.assign({}, o, { steps: counts[l] })));
        poses = level.views.map((v) => v.pose);
    }
    for (let layer = 0; layer < stereo.views.length; layer++)
        if (!o.fixed.includes(layer)) stereo.setPose(layer, poses[layer]);
    await stereo.solveBatched(solves);
    const hf = await refine(stereo, solves, o);
    return { coarse: single ? histories[0] : histories, fine: hf };
}

module.exports = {
    Context, HipDenseStereo, HipTsdfVolume, BrickGrid, STEREO_METHODS, MAP_RAW, MAP_FILTERED,
    MAX_SOURCES, BRICK, NccCostFunction, TsdfIntegrationFunction, FilteredMap, native,
    refine, refinePyramid, halfResolution, pyramidLevels, worldPoseGradients, perturbPose, expSo3,
};
