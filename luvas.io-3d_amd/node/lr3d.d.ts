// TypeScript declarations of lr3d.js: the reference's stage API (DenseStereo, TsdfVolume,
// BrickGrid, StereoMethod) on the MI355X path.  Types follow
// frontend/src/app/vision/geometry/camera.ts, mvs/dense-stereo.ts, fusion/tsdf-volume.ts and
// mvs/methods/stereo-method.ts.

export interface Intrinsics { fx: number; fy: number; cx: number; cy: number; }
/** World -> camera, x_cam = R X + t, R row-major (camera.ts:19-26). */
export interface Pose { R: ArrayLike<number>; t: ArrayLike<number>; }
export interface StereoView { intrinsics: Intrinsics; pose: Pose; width?: number; height?: number; }

export interface SolveOptions {
    sources: number[];
    minDepth: number;
    maxDepth: number;
    prior: Float32Array | null;
    geometric: boolean;
    iterations: number;
    fromCoarse?: boolean;
    pack?: boolean;
    /** Views averaged per cost; 0 / absent = the reference rule max(1, min(3, n - 1)). */
    topk?: number;
}

export declare class Context {
    constructor(device?: number);
    readonly handle: unknown;
    synchronize(): void;
    /** Kernel timing with events around every launch (statistics reset when turned on). */
    setProfiling(on: boolean): void;
    kernelStats(name: string): { launches: number; ms: number };
    destroy(): void;
}

/** Launch shapes [threads x, threads y] per PatchMatch kernel. */
export interface TileShapes { init: [number, number]; propagate: [number, number]; pack: [number, number]; }

export declare class HipDenseStereo {
    constructor(ctx: Context, width: number, height: number, views: StereoView[],
                options?: { coarse?: HipDenseStereo; priorPool?: boolean });
    readonly width: number;
    readonly height: number;
    readonly hypotheses: HipDenseStereo;
    /** The filtered map (DS:67), as TsdfVolume binds it: the map of the layer filtered last. */
    readonly filtered: FilteredMap;
    readonly lastFiltered: number;
    setImage(layer: number, gray: Uint8Array): void;
    solve(ref: number, options: SolveOptions): Promise<void>;
    solveBatched(solves: Array<[number, SolveOptions]>): Promise<void>;
    /** solveBatched without device allocation (fixed per-GPU plan scratch; at most 1024 solves). */
    solveBatchedNoalloc(solves: Array<[number, SolveOptions]>): Promise<void>;
    /** Times the launch-shape candidates on these solves and keeps the fastest (results unchanged). */
    autotune(solves: Array<[number, SolveOptions]>, repeats?: number): Promise<TileShapes>;
    tiles(): TileShapes;
    setTile(kernel: 'init' | 'propagate' | 'pack', tx: number, ty: number): void;
    writeDepth(layer: number, packed: Uint32Array): void;
    /** n.xyz, depth per pixel (4 * width * height floats). */
    writeHyp(layer: number, planes: Float32Array): void;
    filter(ref: number, sources: number[], needed: number): Promise<void>;
    filteredMap(ref: number, sources: number[], needed: number): Promise<Uint32Array>;
    depthMap(ref: number): Promise<Uint32Array>;
    planes(ref: number): Promise<{ hyp: Float32Array; cost: Float32Array }>;
    nccBackward(ref: number, options: SolveOptions, weights?: ArrayLike<number> | null): Promise<NccGradients>;
    refineStep(solves: Array<[number, SolveOptions]>, stepDepth: number, stepNormal: number):
        Promise<{ gradPose: Float64Array; lossBefore: Float64Array; lossAfter: Float64Array }>;
    setPose(layer: number, pose: Pose): void;
    /** fuseDepthMaps on the GPU over these layers' maps (MAP_FILTERED by default, minSupport 1). */
    fuseDepthMaps(layers: number[], thumbs: Thumbnail[], cell: number, cellBudget: number, map?: number,
                  minSupport?: number): Promise<FusedCloud>;
    destroy(): void;
}

/** What `HipDenseStereo.filtered` returns; HipTsdfVolume reads the stereo level through it. */
export declare class FilteredMap {
    readonly stereo: HipDenseStereo;
}

/** A view's colour thumbnail (depth-fusion.ts: `Thumbnail`). */
export interface Thumbnail { width: number; height: number; rgb: Uint8Array; }
/** The fused cloud: 3 floats / 3 bytes per point, points in cell-key order; `cell` = final cell size. */
export interface FusedCloud {
    positions: Float32Array;
    normals: Float32Array;
    colors: Uint8Array;
    support: Uint16Array;
    cell: number;
}

/** Adjoint of the aggregated NCC cost of one reference view's planes. */
export interface NccGradients {
    cost: Float32Array;
    gradDepth: Float32Array;
    /** 3 per pixel */
    gradNormal: Float32Array;
    /** 8 source slots x (dL/dR row-major 9, dL/dt 3) of the relative pose ref -> source */
    gradPose: Float64Array;
    loss: number;
}

export declare class BrickGrid {
    constructor(origin: [number, number, number], voxel: number, extent: [number, number, number]);
    readonly origin: [number, number, number];
    readonly voxel: number;
    readonly dims: [number, number, number];
    readonly cells: number;
    mark(x: number, y: number, z: number, radius: number): void;
    build(): { coords: Int32Array; lookup: Int32Array; count: number };
}

export interface IntegrationView {
    intrinsics: Intrinsics;
    pose: Pose;
    maxDepth: number;
    prior: Float32Array | null;
    confidence: number;
    width?: number;
    height?: number;
}

export declare class HipTsdfVolume {
    constructor(ctx: Context, grid: BrickGrid, layout: { coords: Int32Array; lookup: Int32Array; count: number },
                depth: FilteredMap | HipDenseStereo, width: number, height: number, truncationVoxels?: number);
    readonly truncation: number;
    integrate(view: IntegrationView): Promise<void>;
    integrateLayers(layers: number[], views: IntegrationView[], map?: number): Promise<void>;
    /** A packed map from host memory (width * height words). */
    integrateHost(packed: Uint32Array, view: IntegrationView): Promise<void>;
    /** Times bricks per block (1, 2, 4) and chunks of consecutive blocks per XCD on a pass of these
     * layers and keeps the fastest pair; resolves to the bricks per block. */
    autotune(layers: number[], views: IntegrationView[], map?: number, repeats?: number): Promise<number>;
    bricksPerBlock: number;
    /** Consecutive blocks of the integration pass given to one XCD (1: round robin). */
    xcdChunk: number;
    setDepthCulling(on: boolean): void;
    setCounting(on: boolean): void;
    updateCount(): number;
    passCounts(): { updates: number; pairs: number; bricksRead: number; voxelsWritten: number };
    write(tsdf: Float32Array, weight: Float32Array): void;
    read(): Promise<{ tsdf: Float32Array; weight: Float32Array }>;
    backward(layers: number[], views: IntegrationView[], voxels: ArrayLike<number>, targets: ArrayLike<number>,
             map?: number): Promise<TsdfGradients>;
    /** extractSurface + removeSmallComponents on the GPU (defaults 0.6, 0.01, 200; fraction < 0 keeps all). */
    extractSurface(minWeight?: number, fraction?: number, minFaces?: number):
        Promise<{ positions: Float32Array; normals: Float32Array; indices: Uint32Array }>;
    destroy(): void;
}

/** Adjoint of integrating views into a fresh volume, at listed voxels, for sum (tsdf - target)^2. */
export interface TsdfGradients {
    tsdf: Float32Array;
    /** per view in list order, width*height each, concatenated */
    gradDepth: Float32Array;
    /** per view: dL/dR (9, row-major), dL/dt (3) of the world -> camera pose */
    gradPose: Float64Array;
    loss: number;
}

export declare class NccCostFunction {
    constructor(stereo: HipDenseStereo, ref: number, options: SolveOptions);
    forward(weights?: ArrayLike<number> | null): Promise<{ loss: number; cost: Float32Array }>;
    backward(): Promise<{ gradDepth: Float32Array; gradNormal: Float32Array; gradPose: Float64Array }>;
}

export declare class TsdfIntegrationFunction {
    constructor(volume: HipTsdfVolume, layers: number[], views: IntegrationView[], map?: number);
    forward(): Promise<void>;
    backward(voxels: ArrayLike<number>, targets: ArrayLike<number>): Promise<TsdfGradients>;
}

export interface StereoInput {
    width: number;
    height: number;
    views: StereoView[];
    sources: number[][];
    ranges: [number, number][];
    gray(layer: number, width: number, height: number): Promise<Uint8Array>;
    prior(layer: number, width: number, height: number): Float32Array | null;
}
export interface StereoSettings { iterations: number; fineIterations: number; }
export interface StereoRun { readonly stereo: HipDenseStereo; solve(layer: number): Promise<void>; destroy(): void; }
export interface StereoMethod { create(ctx: Context, input: StereoInput, settings: StereoSettings): Promise<StereoRun>; }

export declare const STEREO_METHODS: Record<'hip-single-scale' | 'hip-coarse-to-fine', StereoMethod>;
export declare const MAP_RAW: 0;
export declare const MAP_FILTERED: 1;
export declare const MAX_SOURCES: 8;
export declare const BRICK: 8;

/** Options of refine / refinePyramid (defaults as in the Python host's lr3d.refine.refine). */
export interface RefineOptions {
    steps?: number;
    /** refinePyramid: steps on each coarse copy (one count, or one per copy, finest first) */
    coarseSteps?: number | number[];
    stepDepth?: number;
    stepNormal?: number;
    stepRot?: number;
    stepTrans?: number;
    /** layers whose poses stay fixed (the gauge); default [0] */
    fixed?: number[];
    momentum?: number;
}
/** Joint gradient refinement of planes and poses through the NCC adjoint; returns the loss history. */
export declare function refine(stereo: HipDenseStereo, solves: Array<[number, SolveOptions]>,
                               options?: RefineOptions): Promise<number[]>;
/** Half-resolution copy of a stereo set (frames box-filtered 2 x 2) for refinePyramid. */
export declare function halfResolution(stereo: HipDenseStereo, images: Uint8Array[]): HipDenseStereo;
/** `levels` (default 3) coarse copies of a stereo set, each half the resolution of the one before,
 * finest first, for refinePyramid. */
export declare function pyramidLevels(stereo: HipDenseStereo, images: Uint8Array[], levels?: number): HipDenseStereo[];
/** Refinement coarse to fine: coarseSteps on each coarse copy (coarsest first), then steps on
 * `stereo`.  Histories: one per coarse copy (coarsest first) when `coarse` is an array. */
export declare function refinePyramid(stereo: HipDenseStereo, coarse: HipDenseStereo | HipDenseStereo[],
                                      solves: Array<[number, SolveOptions]>, options?: RefineOptions):
    Promise<{ coarse: number[] | number[][]; fine: number[] }>;
