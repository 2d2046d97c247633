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
    destroy(): void;
}

export declare class HipDenseStereo {
    constructor(ctx: Context, width: number, height: number, views: StereoView[],
                options?: { coarse?: HipDenseStereo });
    readonly width: number;
    readonly height: number;
    readonly hypotheses: HipDenseStereo;
    readonly lastFiltered: number;
    setImage(layer: number, gray: Uint8Array): void;
    solve(ref: number, options: SolveOptions): Promise<void>;
    solveBatched(solves: Array<[number, SolveOptions]>): Promise<void>;
    filter(ref: number, sources: number[], needed: number): Promise<void>;
    filteredMap(ref: number, sources: number[], needed: number): Promise<Uint32Array>;
    depthMap(ref: number): Promise<Uint32Array>;
    planes(ref: number): Promise<{ hyp: Float32Array; cost: Float32Array }>;
    destroy(): void;
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
                stereo: HipDenseStereo, width: number, height: number, truncationVoxels?: number);
    readonly truncation: number;
    integrate(view: IntegrationView): Promise<void>;
    integrateLayers(layers: number[], views: IntegrationView[], map?: number): Promise<void>;
    read(): Promise<{ tsdf: Float32Array; weight: Float32Array }>;
    destroy(): void;
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
