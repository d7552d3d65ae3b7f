/**
 * droplet_hip.ts — bun:ffi binding of libdroplet_hip.so (include/droplet_hip.h).
 *
 * A drop-in for the GPU layer behind droplet's public API (src/index.ts:16-38):
 * `createDevice` (src/gpu/device.ts:27-41), `Renderer` with `render` and
 * `destroy` (src/render/pipeline.ts:76-84,181-185,581-591), `voxelize`
 * (src/voxel/voxelize.ts:59-63) returning a `BlockMaskBuffer`
 * (src/voxel/block-mask-buffer.ts:23-58), and `GpuRadixSorter.sort`
 * (src/gpu/radix-sort.ts:142-215, stable here). Beside them: resident
 * scenes and PLY files, the differentiable forward/backward pair, batched
 * views, the autotuner and the collision stages.
 *
 * Every entry point returns a status code; a non-zero code becomes
 * `throw new Error(dl_last_error())`, the reference's error convention
 * (src/gpu/device.ts:30,35). Calls are synchronous on the device; the async
 * methods keep the reference's Promise signatures. Struct layouts are the C
 * layouts of include/droplet_hip.h (x86-64 / aarch64: 8-byte pointers and
 * doubles aligned to 8), written into little-endian byte buffers.
 */
import { dlopen, FFIType, ptr, type Pointer } from 'bun:ffi';

// ---- the reference's types (src/types.ts:23-82), structurally ---------------
export type Projection = 'pinhole' | 'equirect';
export interface SplatData {
    count: number;
    positions: Float32Array;
    rotations: Float32Array;
    logScales: Float32Array;
    opacityLogits: Float32Array;
    colorsDC: Float32Array;
    colorsSH?: Float32Array;
    shBands: 0 | 1 | 2 | 3;
}
export interface RenderCamera {
    width: number;
    height: number;
    viewMatrix: Float32Array;       // column-major world -> camera
    fovY: number;
    near: number;
    far: number;
    projection?: Projection;
}
export interface BackgroundRGBA { r: number; g: number; b: number; a: number; }
export interface Bounds { min: [number, number, number]; max: [number, number, number]; }
export interface RendererOptions {
    chunkCap?: number; maxCoveragePerSplat?: number;
    /** 8, 16 or 32 px (single-view renderer; the image is the 16-px one) */
    tileSize?: number;
    /** 1: one wave per tile, 2: 2x2 tiles per workgroup */
    rasterGroup?: number;
}
export interface VoxelizeOptions { voxelResolution?: number; opacityCutoff?: number; bounds?: Bounds; }
export interface VoxelBlock { blockIdx: number; maskLo: number; maskHi: number; }

const P = FFIType.ptr, I32 = FFIType.i32, U32 = FFIType.u32, U64 = FFIType.u64, F32 = FFIType.f32;
const VOID = FFIType.void;

// ---- every entry point of include/droplet_hip.h ------------------------------
export const symbols = {
    dl_context_create: { args: [P, P], returns: I32 },
    dl_context_destroy: { args: [P], returns: VOID },
    dl_last_error: { args: [], returns: FFIType.cstring },
    dl_context_device_name: { args: [P, P, U64], returns: I32 },
    dl_context_synchronize: { args: [P], returns: I32 },
    dl_renderer_create: { args: [P, P, P], returns: I32 },
    dl_renderer_destroy: { args: [P], returns: VOID },
    dl_render: { args: [P, P, P, P, P], returns: I32 },
    dl_scene_create: { args: [P, P, P], returns: I32 },
    dl_scene_destroy: { args: [P], returns: VOID },
    dl_scene_create_from_ply: { args: [P, P, U64, I32, P], returns: I32 },
    dl_scene_info: { args: [P, P, P, P], returns: I32 },
    dl_scene_read: { args: [P, P, P, P, P, P, P], returns: I32 },
    dl_scene_bounds: { args: [P, P, P], returns: I32 },
    dl_render_scene: { args: [P, P, P, P, P], returns: I32 },
    dl_renderer_read_image: { args: [P, P], returns: I32 },
    dl_renderer_set_timing: { args: [P, I32], returns: I32 },
    dl_renderer_set_counting: { args: [P, I32], returns: I32 },
    dl_renderer_stats: { args: [P, P], returns: I32 },
    dl_renderer_set_pair_budget: { args: [P, U64], returns: I32 },
    dl_renderer_row_iters: { args: [P, P, U32, P], returns: I32 },
    dl_renderer_set_taps: { args: [P, I32], returns: I32 },
    dl_renderer_tap: { args: [P, I32, P, U64, P], returns: I32 },
    dl_render_forward: { args: [P, P, P, P, P], returns: I32 },
    dl_render_backward: { args: [P, P, P], returns: I32 },
    dl_renderer_autotune: { args: [P, P, P, P], returns: I32 },
    dl_renderer_tune_candidates: { args: [P, U64, P, U32, P], returns: I32 },
    dl_batch_create: { args: [P, P, P], returns: I32 },
    dl_batch_destroy: { args: [P], returns: VOID },
    dl_batch_set_graph: { args: [P, I32], returns: I32 },
    dl_batch_render: { args: [P, P, U32, P, U32], returns: I32 },
    dl_batch_set_pair_budget: { args: [P, U64], returns: I32 },
    dl_batch_read_image: { args: [P, U32, P], returns: I32 },
    dl_batch_get_stats: { args: [P, P], returns: I32 },
    dl_batch_autotune: { args: [P, P, U32, P, U32, P], returns: I32 },
    dl_sort_pairs_u32: { args: [P, P, P, U32, U32], returns: I32 },
    dl_sort_pairs_u32_noalloc: { args: [P, P, P, U32, U32], returns: I32 },
    dl_sort_pairs_u32_noalloc_max: { args: [], returns: U64 },
    dl_scan_u32: { args: [P, P, P, U64, U32, P], returns: I32 },
    dl_voxelize: { args: [P, P, P, P], returns: I32 },
    dl_voxel_result_count: { args: [P], returns: U64 },
    dl_voxel_result_blocks: { args: [P, P, P, P], returns: I32 },
    dl_voxel_result_bounds: { args: [P, P, P, P], returns: I32 },
    dl_voxel_result_device_ms: { args: [P], returns: F32 },
    dl_voxel_result_stats: { args: [P, P], returns: I32 },
    dl_voxel_result_free: { args: [P], returns: VOID },
    dl_voxel_result_device: { args: [P, P], returns: I32 },
    dl_voxel_result_create: { args: [P, P, P, P, FFIType.f64, P, P, P, U64, P], returns: I32 },
    dl_voxel_fill: { args: [P, P, P], returns: I32 },
    dl_voxel_carve: { args: [P, P, P, U64, P], returns: I32 },
    dl_voxel_query: { args: [P, P, P, U64, P], returns: I32 },
    dl_voxel_mesh_build: { args: [P, P, P], returns: I32 },
    dl_voxel_mesh_quads: { args: [P], returns: U64 },
    dl_voxel_mesh_read: { args: [P, P, P], returns: I32 },
    dl_voxel_mesh_device_ms: { args: [P], returns: F32 },
    dl_voxel_mesh_free: { args: [P], returns: VOID },
    dl_debug_math: { args: [P, I32, P, P, P, U32], returns: I32 },
    dl_debug_valu_probe: { args: [P, U32, U32, P], returns: I32 },
} as const;

const LIB = process.env.DROPLET_HIP_LIB ?? 'libdroplet_hip.so';
const { symbols: L } = dlopen(LIB, symbols);

const check = (rc: number): void => {
    if (rc !== 0) throw new Error(String(L.dl_last_error()));
};
type Handle = Pointer;
const handleOut = () => new BigUint64Array(1);          // receives an opaque handle
const asHandle = (h: BigUint64Array): Handle => Number(h[0]) as unknown as Handle;
const addr = (a?: ArrayBufferView | null): bigint => (a && a.byteLength ? BigInt(ptr(a)) : 0n);

// ---- C struct layouts (bytes) ----------------------------------------------
export const LAYOUT = {
    dl_splats: 64,          // u32 count | 6 pointers at 8..48 | u32 sh_bands at 56
    dl_camera: 104,         // u32 w, h | f32 view[16] at 8 | f64 fov_y 72, near 80, far 88 | u32 projection 96
    dl_background: 16,      // f32 r, g, b, a
    dl_renderer_opts: 36,   // 9 x u32: ... sort_digit_bits 24, tile_size 28, raster_group 32
    dl_device_opts: 4,      // i32 device_index
    dl_voxel_opts: 72,      // f64 res 0, cutoff 8 | i32 has_bounds 16 | f64 min[3] 24, max[3] 48
    dl_grads: 48,           // 6 pointers
    dl_view: 128,           // u32 scene | dl_camera at 8 | dl_background at 112
    dl_tune_result: 48,     // dl_renderer_opts chosen | f32 chosen_ms 36, default_ms 40 | u32 candidates 44
    dl_render_stats: 136,
    dl_batch_stats: 32,
    dl_voxel_stats: 24,
} as const;

const splatsStruct = (d: SplatData): Uint8Array => {
    const b = new Uint8Array(LAYOUT.dl_splats), v = new DataView(b.buffer);
    v.setUint32(0, d.count, true);
    const arrs = [d.positions, d.rotations, d.logScales, d.opacityLogits, d.colorsDC, d.shBands > 0 ? d.colorsSH : undefined];
    arrs.forEach((a, i) => v.setBigUint64(8 + 8 * i, addr(a), true));
    v.setUint32(56, d.shBands, true);
    return b;
};

const writeCamera = (v: DataView, at: number, c: RenderCamera): void => {
    v.setUint32(at + 0, c.width, true);
    v.setUint32(at + 4, c.height, true);
    for (let k = 0; k < 16; k++) v.setFloat32(at + 8 + 4 * k, c.viewMatrix[k], true);
    v.setFloat64(at + 72, c.fovY, true);
    v.setFloat64(at + 80, c.near, true);
    v.setFloat64(at + 88, c.far, true);
    v.setUint32(at + 96, c.projection === 'equirect' ? 1 : 0, true);
};
const cameraStruct = (c: RenderCamera): Uint8Array => {
    const b = new Uint8Array(LAYOUT.dl_camera);
    writeCamera(new DataView(b.buffer), 0, c);
    return b;
};
const DEFAULT_BG: BackgroundRGBA = { r: 0, g: 0, b: 0, a: 1 };   // pipeline.ts:184
const bgStruct = (bg: BackgroundRGBA = DEFAULT_BG) => new Float32Array([bg.r, bg.g, bg.b, bg.a]);

// ---- device (createDevice, src/gpu/device.ts:27-41) ---------------------------
export class Device {
    constructor(readonly handle: Handle) {}
    get name(): string {
        const buf = new Uint8Array(64);
        check(L.dl_context_device_name(this.handle, buf, 64n));
        return new TextDecoder().decode(buf.subarray(0, buf.indexOf(0)));
    }
    synchronize(): void { check(L.dl_context_synchronize(this.handle)); }
    destroy(): void { L.dl_context_destroy(this.handle); }
}

export const createDevice = async (opts: { deviceIndex?: number } = {}): Promise<Device> => {
    const h = handleOut();
    check(L.dl_context_create(new Int32Array([opts.deviceIndex ?? 0]), h));
    return new Device(asHandle(h));
};

// ---- resident scenes and PLY (loadPlySplats, src/io/ply.ts:120-239) -----------
const SH_COEFFS = [0, 3, 8, 15] as const;

export class Scene {
    private constructor(readonly handle: Handle, readonly count: number, readonly shBands: 0 | 1 | 2 | 3,
                        readonly shCoeffs: number) {}
    private static info(h: Handle): [number, 0 | 1 | 2 | 3, number] {
        const c = new Uint32Array(3);
        check(L.dl_scene_info(h, c.subarray(0, 1), c.subarray(1, 2), c.subarray(2, 3)));
        return [c[0], c[1] as 0 | 1 | 2 | 3, c[2]];
    }
    static create(device: Device, data: SplatData): Scene {
        const h = handleOut();
        check(L.dl_scene_create(device.handle, splatsStruct(data), h));
        return new Scene(asHandle(h), ...Scene.info(asHandle(h)));
    }
    static fromPly(device: Device, bytes: Uint8Array, flipY = true): Scene {
        const h = handleOut();
        check(L.dl_scene_create_from_ply(device.handle, bytes, BigInt(bytes.byteLength), flipY ? 1 : 0, h));
        return new Scene(asHandle(h), ...Scene.info(asHandle(h)));
    }
    /** loadPlySplats's SplatData, read back from the device */
    read(): SplatData {
        const n = this.count;
        const d: SplatData = {
            count: n, positions: new Float32Array(3 * n), rotations: new Float32Array(4 * n),
            logScales: new Float32Array(3 * n), opacityLogits: new Float32Array(n), colorsDC: new Float32Array(3 * n),
            colorsSH: this.shCoeffs ? new Float32Array(3 * this.shCoeffs * n) : undefined, shBands: this.shBands,
        };
        check(L.dl_scene_read(this.handle, d.positions, d.rotations, d.logScales, d.opacityLogits, d.colorsDC,
                              d.colorsSH ?? null));
        return d;
    }
    /** splatBounds (src/io/ply.ts:242-255) */
    bounds(): Bounds {
        const lo = new Float64Array(3), hi = new Float64Array(3);
        check(L.dl_scene_bounds(this.handle, lo, hi));
        return { min: [lo[0], lo[1], lo[2]], max: [hi[0], hi[1], hi[2]] };
    }
    destroy(): void { L.dl_scene_destroy(this.handle); }
}

// ---- Renderer (src/render/pipeline.ts:50-592) --------------------------------
export interface RenderStats {
    numVisible: number; numChunks: number; numTiles: number; numPairs: bigint;
    stageMs: number[]; totalMs: number; pixelIters: bigint; listEntries: bigint;
    tileSize: number; rasterGroup: number;
}

export class Renderer {
    private h: Handle;
    private trainScene: Scene | null = null;
    private trainSize: [number, number] = [0, 0];

    constructor(device: Device, opts: RendererOptions = {}) {
        const o = new Uint32Array([opts.chunkCap ?? 0, opts.maxCoveragePerSplat ?? 0, 0, 0, 0, 0, 0,
                                   opts.tileSize ?? 0, opts.rasterGroup ?? 0]);
        const h = handleOut();
        check(L.dl_renderer_create(device.handle, o, h));
        this.h = asHandle(h);
    }

    /** Renderer.render (pipeline.ts:181-426): RGBA8, W*H*4 bytes, a fresh array */
    async render(data: SplatData, camera: RenderCamera, background: BackgroundRGBA = DEFAULT_BG): Promise<Uint8Array> {
        const rgba = new Uint8Array(camera.width * camera.height * 4);
        check(L.dl_render(this.h, splatsStruct(data), cameraStruct(camera), bgStruct(background), rgba));
        return rgba;
    }

    /** the same for a scene resident in device memory */
    async renderScene(scene: Scene, camera: RenderCamera, background: BackgroundRGBA = DEFAULT_BG): Promise<Uint8Array> {
        const rgba = new Uint8Array(camera.width * camera.height * 4);
        check(L.dl_render_scene(this.h, scene.handle, cameraStruct(camera), bgStruct(background), rgba));
        return rgba;
    }

    /** differentiable forward: clamp(C + T*bg, 0, 1) before quantisation, W*H*3 floats */
    async forward(scene: Scene, camera: RenderCamera, background: BackgroundRGBA = DEFAULT_BG): Promise<Float32Array> {
        const img = new Float32Array(camera.width * camera.height * 3);
        check(L.dl_render_forward(this.h, scene.handle, cameraStruct(camera), bgStruct(background), img));
        this.trainScene = scene;
        this.trainSize = [camera.width, camera.height];
        return img;
    }

    /** dL/d(image) of the last forward in, dL/d(every SplatData float) out */
    async backward(gradRgb: Float32Array): Promise<Omit<SplatData, 'count' | 'shBands'>> {
        const s = this.trainScene;
        if (!s) throw new Error('backward needs a preceding forward');
        if (gradRgb.length !== this.trainSize[0] * this.trainSize[1] * 3)
            throw new Error('gradRgb must hold width*height*3 floats');
        const n = s.count, m = SH_COEFFS[s.shBands];
        const g = {
            positions: new Float32Array(3 * n), rotations: new Float32Array(4 * n), logScales: new Float32Array(3 * n),
            opacityLogits: new Float32Array(n), colorsDC: new Float32Array(3 * n),
            colorsSH: m ? new Float32Array(3 * m * n) : undefined,
        };
        const b = new Uint8Array(LAYOUT.dl_grads), v = new DataView(b.buffer);
        [g.positions, g.rotations, g.logScales, g.opacityLogits, g.colorsDC, g.colorsSH]
            .forEach((a, i) => v.setBigUint64(8 * i, addr(a), true));
        check(L.dl_render_backward(this.h, gradRgb, b));
        return g;
    }

    /** times candidate options on this workload and keeps the fastest */
    autotune(scene: Scene, camera: RenderCamera): { chosenMs: number; defaultMs: number; candidates: number } {
        const r = new Uint8Array(LAYOUT.dl_tune_result), v = new DataView(r.buffer);
        check(L.dl_renderer_autotune(this.h, scene.handle, cameraStruct(camera), r));
        return { chosenMs: v.getFloat32(36, true), defaultMs: v.getFloat32(40, true), candidates: v.getUint32(44, true) };
    }

    setTiming(on: boolean): void { check(L.dl_renderer_set_timing(this.h, on ? 1 : 0)); }

    stats(): RenderStats {
        const b = new Uint8Array(LAYOUT.dl_render_stats), v = new DataView(b.buffer);
        check(L.dl_renderer_stats(this.h, b));
        const stageMs: number[] = [];
        for (let k = 0; k < 8; k++) stageMs.push(v.getFloat32(32 + 4 * k, true));
        return {
            numVisible: v.getUint32(0, true), numChunks: v.getUint32(4, true), numTiles: v.getUint32(8, true),
            numPairs: v.getBigUint64(24, true), stageMs, totalMs: v.getFloat32(64, true),
            pixelIters: v.getBigUint64(72, true), listEntries: v.getBigUint64(88, true),
            tileSize: v.getUint32(124, true), rasterGroup: v.getUint32(128, true),
        };
    }

    destroy(): void { L.dl_renderer_destroy(this.h); }
}

// ---- batched views ----------------------------------------------------------
export interface View { scene: number; camera: RenderCamera; background?: BackgroundRGBA; }

export class BatchRenderer {
    private h: Handle;
    private sizes: [number, number][] = [];
    constructor(device: Device, opts: RendererOptions = {}) {
        const o = new Uint32Array([opts.chunkCap ?? 0, opts.maxCoveragePerSplat ?? 0, 0, 0, 0, 0, 0,
                                   opts.tileSize ?? 0, opts.rasterGroup ?? 0]);
        const h = handleOut();
        check(L.dl_batch_create(device.handle, o, h));
        this.h = asHandle(h);
    }
    async render(scenes: Scene[], views: View[]): Promise<void> {
        const sc = new BigUint64Array(scenes.map((s) => BigInt(s.handle as unknown as number)));
        const b = new Uint8Array(LAYOUT.dl_view * views.length), v = new DataView(b.buffer);
        views.forEach((w, i) => {
            const at = LAYOUT.dl_view * i;
            v.setUint32(at, w.scene, true);
            writeCamera(v, at + 8, w.camera);
            const bg = w.background ?? DEFAULT_BG;
            [bg.r, bg.g, bg.b, bg.a].forEach((x, k) => v.setFloat32(at + 112 + 4 * k, x, true));
        });
        check(L.dl_batch_render(this.h, sc, scenes.length, b, views.length));
        this.sizes = views.map((w) => [w.camera.width, w.camera.height]);
    }
    image(view: number): Uint8Array {
        const [w, hh] = this.sizes[view];
        const rgba = new Uint8Array(w * hh * 4);
        check(L.dl_batch_read_image(this.h, view, rgba));
        return rgba;
    }
    destroy(): void { L.dl_batch_destroy(this.h); }
}

// ---- voxelize (src/voxel/voxelize.ts:59-248) and BlockMaskBuffer ---------------
export class BlockMaskBuffer {
    readonly blocks: VoxelBlock[] = [];
    private index = new Map<number, number>();
    /** addBlock (block-mask-buffer.ts:30-41): OR-merges a repeated block */
    addBlock(blockIdx: number, maskLo: number, maskHi: number): void {
        if (maskLo === 0 && maskHi === 0) return;
        const at = this.index.get(blockIdx);
        if (at === undefined) {
            this.index.set(blockIdx, this.blocks.length);
            this.blocks.push({ blockIdx, maskLo: maskLo >>> 0, maskHi: maskHi >>> 0 });
        } else {
            this.blocks[at].maskLo = (this.blocks[at].maskLo | maskLo) >>> 0;
            this.blocks[at].maskHi = (this.blocks[at].maskHi | maskHi) >>> 0;
        }
    }
    get(blockIdx: number): VoxelBlock | undefined {
        const at = this.index.get(blockIdx);
        return at === undefined ? undefined : this.blocks[at];
    }
    get count(): number { return this.blocks.length; }
}

export interface VoxelizeResult {
    buffer: BlockMaskBuffer; bounds: Bounds; voxelResolution: number;
    numBlocksX: number; numBlocksY: number; numBlocksZ: number;
}

export const voxelize = async (device: Device, data: SplatData, opts: VoxelizeOptions = {}): Promise<VoxelizeResult> => {
    const o = new Uint8Array(LAYOUT.dl_voxel_opts), v = new DataView(o.buffer);
    const res = opts.voxelResolution ?? 0.05;
    v.setFloat64(0, res, true);
    v.setFloat64(8, opts.opacityCutoff ?? 0.1, true);
    if (opts.bounds) {
        v.setInt32(16, 1, true);
        for (let k = 0; k < 3; k++) {
            v.setFloat64(24 + 8 * k, opts.bounds.min[k], true);
            v.setFloat64(48 + 8 * k, opts.bounds.max[k], true);
        }
    }
    const h = handleOut();
    check(L.dl_voxelize(device.handle, splatsStruct(data), o, h));
    const r = asHandle(h);
    try {
        const n = Number(L.dl_voxel_result_count(r));
        const idx = new Uint32Array(n), lo = new Uint32Array(n), hi = new Uint32Array(n);
        if (n) check(L.dl_voxel_result_blocks(r, idx, lo, hi));
        const bmin = new Float64Array(3), bmax = new Float64Array(3), nb = new Float64Array(3);
        check(L.dl_voxel_result_bounds(r, bmin, bmax, nb));
        const buffer = new BlockMaskBuffer();
        for (let i = 0; i < n; i++) buffer.addBlock(idx[i], lo[i], hi[i]);   // the reference's insertion order
        return {
            buffer, bounds: { min: [bmin[0], bmin[1], bmin[2]], max: [bmax[0], bmax[1], bmax[2]] },
            voxelResolution: res, numBlocksX: nb[0], numBlocksY: nb[1], numBlocksZ: nb[2],
        };
    } finally {
        L.dl_voxel_result_free(r);
    }
};

// ---- GpuRadixSorter (src/gpu/radix-sort.ts:91-216), stable ------------------------
export class GpuRadixSorter {
    constructor(private device: Device) {}
    /** sorts host keys/vals by the low keyBits bits, stably; returns them sorted */
    async sort(keys: Uint32Array, vals: Uint32Array, n: number, keyBits = 32): Promise<{ keys: Uint32Array; vals: Uint32Array }> {
        const k = keys.slice(0, n), v = vals.slice(0, n);
        check(L.dl_sort_pairs_u32(this.device.handle, k, v, n, keyBits));
        return { keys: k, vals: v };
    }
}
