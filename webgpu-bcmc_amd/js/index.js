'use strict';
/*
 * JavaScript facade with the reference's classes (Twinklebear/webgpu-bcmc):
 *   CompressedMarchingCubes   js/compressed_marching_cubes.js
 *   LRUCache                  js/lru_cache.js
 *   ExclusiveScanPipeline     js/exclusive_scan.js:14-122
 *   ExclusiveScanner          js/exclusive_scan.js:124-295
 *   StreamCompact             js/stream_compact.js (also exported as StreamCompactor)
 *   RadixSorter               js/radix_sort_by_key.js
 * over the N-API addon (bcmc_addon.node), which calls the gfx950 HIP library
 * through its C ABI.  Methods keep the reference's names, arguments and
 * Promise-returning shape; `device` is a Device from requestDevice() instead
 * of a GPUDevice and buffers are DeviceBuffers instead of GPUBuffers.  The
 * GPU work of every async method runs on a worker thread (the addon's *Async
 * calls), so the event loop keeps running while it is pending; each device
 * runs its operations in the order they were called, as a GPU queue does.
 */
const path = require('path');

const native = require(path.join(__dirname, '..', 'lib', 'bcmc_addon.node'));

const VertexFormat = Object.freeze({LOCAL_F32: 0, GLOBAL_F32: 1, PACKED_U32X2: 2});
const McArray = Object.freeze(
    {RANGES: 0, ACTIVE_FLAGS: 1, ACTIVE_IDS: 2, NEW_ITEMS: 3, WITH_VERTS: 4, VERT_OFFSETS: 5});

class DeviceBuffer {
    constructor(device, handle, owned = true) {
        this.device = device;
        this.handle = handle;
        this.owned = owned;
    }
    get size() {
        return native.bufSize(this.handle);
    }
    // writeBuffer equivalent: copy a TypedArray in at a byte offset
    write(typed, byteOffset = 0) {
        native.bufUpload(this.handle, byteOffset, typed);
    }
    // mapAsync(READ) + getMappedRange equivalent: read `count` elements
    read(TypedArrayType, count, byteOffset = 0) {
        const out = new TypedArrayType(count);
        if (count > 0) native.bufDownload(this.handle, byteOffset, out);
        return out;
    }
    destroy() {
        if (this.owned && this.handle) native.bufFree(this.handle);
        this.handle = null;
    }
}

class Device {
    constructor(index = 0) {
        this.handle = native.ctxCreate(index);
        this.info = native.ctxInfo(this.handle);
        this._tail = Promise.resolve();
    }
    // Runs start() (which returns the addon's Promise) after every operation
    // submitted before it on this device; resolves to its result.
    _submit(start) {
        const p = this._tail.then(start);
        this._tail = p.catch(() => {});
        return p;
    }
    createBuffer({size}) {
        return new DeviceBuffer(this, native.bufAlloc(this.handle, Math.max(size, 4)));
    }
    bufferFrom(typed) {
        const b = this.createBuffer({size: typed.byteLength});
        b.write(typed);
        return b;
    }
    async onSubmittedWorkDone() {
        await this._tail;
        native.ctxSync(this.handle);
    }
    // Times launch-shape candidates of the scan/sort/decode kernels and keeps
    // the fastest (results never depend on the shape).
    autotune() {
        return native.ctxAutotune(this.handle);
    }
    // Query chains launched directly (false, the default) or replayed from
    // captured HIP graphs (true); results are identical either way.
    setGraphReplay(enable) {
        native.ctxSetGraphReplay(this.handle, !!enable);
    }
    destroy() {
        if (this.handle) native.ctxDestroy(this.handle);
        this.handle = null;
    }
}

async function requestDevice(index = 0) {
    return new Device(index);
}

const alignTo = (val, align) => Math.floor((val + align - 1) / align) * align;

class ExclusiveScanPipeline {
    constructor(device) {
        this.device = device;
    }
    getAlignedSize(size) {
        return native.scanAlignedSize(size);
    }
    prepareInput(cpuArray) {
        const alignedSize = this.getAlignedSize(cpuArray.length);
        const padded = new Uint32Array(alignedSize);
        padded.set(cpuArray);
        const buf = this.device.bufferFrom(padded);
        return new ExclusiveScanner(this, buf, alignedSize, cpuArray.length);
    }
    prepareGPUInput(gpuBuffer, alignedSize, dataSize) {
        if (this.getAlignedSize(alignedSize) != alignedSize) {
            throw new Error('Error: GPU input must be aligned to getAlignedSize');
        }
        return new ExclusiveScanner(this, gpuBuffer, alignedSize, dataSize);
    }
}

class ExclusiveScanner {
    constructor(scanPipeline, gpuBuffer, alignedSize) {
        this.scanPipeline = scanPipeline;
        this.inputSize = alignedSize;
        this.inputBuf = gpuBuffer;
    }
    async scan(dataSize) {
        const dev = this.scanPipeline.device;
        return dev._submit(
            () => native.scanExclusiveAsync(dev.handle, this.inputBuf.handle, this.inputSize, dataSize));
    }
}

class StreamCompact {
    constructor(device) {
        this.device = device;
    }
    async compactActiveIDs(numElements, isActiveBuffer, offsetsBuffer, outputBuffer) {
        await this.device._submit(() => native.compactIdsAsync(
            this.device.handle, numElements, isActiveBuffer.handle, offsetsBuffer.handle,
            outputBuffer.handle));
    }
    // Fused scan + compaction of 0/1 flags in one pass; resolves to the count.
    async compactFlags(numElements, flagsBuffer, outputBuffer) {
        return this.device._submit(() => native.compactFlagsAsync(
            this.device.handle, numElements, flagsBuffer.handle, outputBuffer.handle));
    }
}

class RadixSorter {
    constructor(device) {
        this.device = device;
    }
    getAlignedSize(size) {
        return native.sortAlignedSize(size);
    }
    async sort(keys, values, size, reverse) {
        await this.device._submit(() => native.sortByKeyAsync(
            this.device.handle, keys.handle, values.handle, size, !!reverse));
    }
}

const LRU_PERF_KEYS = [
    'lruMarkNewItems', 'lruScanNeededSlots', 'lruNumNewItems', 'lruCompactNewItems',
    'lruScanSlotsAvailable', 'lruNumSlotsAvailable', 'lruGrowCache', 'lruCompactAvailableSlots',
    'lruPrepKeyValue', 'lruTotalSortTime', 'lruWriteNewItems'
];

class LRUCache {
    constructor(device, scanPipeline, streamCompact, initialSize, elementSize, totalElements,
                _handle) {
        this.device = device;
        this.scanPipeline = scanPipeline;
        this.streamCompact = streamCompact;
        this._owned = !_handle;
        this.handle = _handle || native.lruCreate(device.handle, initialSize, elementSize,
                                                  totalElements);
    }
    get cacheSize() {
        return native.lruInfo(this.handle).cacheSize;
    }
    get totalElements() {
        return native.lruInfo(this.handle).totalElements;
    }
    get elementSize() {
        return native.lruInfo(this.handle).elementSize;
    }
    get displayNumSlotsAvailable() {
        return native.lruInfo(this.handle).displayNumSlotsAvailable;
    }
    get cache() {
        return new DeviceBuffer(this.device, native.lruCacheBuffer(this.handle), false);
    }
    // Returns [numNewItems, newItemIDs (DeviceBuffer, caller destroys) | undefined].
    // With a perfTracker, each of the reference's 11 lru* keys gets one entry
    // per update: device milliseconds of the stage, or the count.
    async update(itemNeeded, perfTracker) {
        if (!(itemNeeded instanceof DeviceBuffer)) {
            throw new TypeError('itemNeeded info must be a DeviceBuffer');
        }
        const [n, ids, perf] = await this.device._submit(
            () => native.lruUpdateAsync(this.handle, itemNeeded.handle, !!perfTracker));
        if (perfTracker) {
            if (perfTracker.lruMarkNewItems === undefined) {
                for (const k of LRU_PERF_KEYS) perfTracker[k] = [];
            }
            for (const k of LRU_PERF_KEYS) perfTracker[k].push(perf[k]);
        }
        return [n, ids ? new DeviceBuffer(this.device, ids) : undefined];
    }
    async reset() {
        await this.device._submit(async () => native.lruReset(this.handle));
    }
    // Slot state for inspection: {age, available, item, itemSlot}
    state() {
        const info = native.lruInfo(this.handle);
        const c = info.cacheSize, t = info.totalElements;
        const s = {
            age: new Uint32Array(c), available: new Uint32Array(c), item: new Int32Array(c),
            itemSlot: new Int32Array(t)
        };
        native.lruState(this.handle, s.age, s.available, s.item, s.itemSlot);
        return s;
    }
    destroy() {
        if (this._owned && this.handle) native.lruDestroy(this.handle);
        this.handle = null;
    }
}

const MC_PERF_KEYS = [
    'computeActiveBlocks', 'numActiveBlocks', 'cacheUpdate', 'numBlocksDecompressed',
    'decompression', 'compactActiveIDs', 'computeBlockHasVertices', 'numBlocksWithVertices',
    'compactBlocksWithVertices', 'numVertices', 'computeVertices'
];

class CompressedMarchingCubes {
    constructor(device) {
        this.device = device;
        this.scanPipeline = new ExclusiveScanPipeline(device);
        this.streamCompact = new StreamCompact(device);
        this.handle = native.mcCreate(device.handle);
        this.numActiveBlocks = 0;
        this.numVertices = 0;
        this.numBlocksWithVertices = 0;
        this.newDecompressed = 0;
        this.totalBlocks = 0;
    }
    // blockRanges (optional): ranges persisted from blockRanges() -- a
    // Float32Array of 2 (min, max) or 4 (and the tile range) values per block
    // -- in place of the full-volume decode at load.
    async setCompressedVolume(volume, compressionRate, volumeDims, volumeScale,
                              initialCacheSlots = 0, blockRanges = undefined) {
        this.volumeDims = volumeDims;
        this.paddedDims = [alignTo(volumeDims[0], 4), alignTo(volumeDims[1], 4),
                           alignTo(volumeDims[2], 4)];
        this.totalBlocks = (this.paddedDims[0] * this.paddedDims[1] * this.paddedDims[2]) / 64;
        native.mcSetVolume(this.handle, volume, compressionRate, volumeDims,
                           volumeScale || [1, 1, 1], initialCacheSlots, blockRanges);
        this.compressedDataSize = volume.byteLength;
        this.lruCache = new LRUCache(this.device, this.scanPipeline, this.streamCompact, 0, 0, 0,
                                     native.mcLru(this.handle));
    }
    // the loaded volume's block ranges for persisting beside the stream:
    // (min, max) per block, followed by the tile ranges when withTile
    blockRanges(withTile = true) {
        const out = new Float32Array(this.totalBlocks * (withTile ? 4 : 2));
        native.mcGetRanges(this.handle, withTile, out);
        return out;
    }
    setVertexFormat(fmt) {
        native.mcSetFormat(this.handle, fmt);
    }
    async computeSurface(isovalue, perfTracker) {
        // without a perfTracker the chain runs in the device's launch mode (no stage times)
        const r = await this.device._submit(
            () => native.mcComputeSurfaceAsync(this.handle, isovalue, !!perfTracker));
        if (perfTracker) {
            if (perfTracker.computeActiveBlocks === undefined) {
                perfTracker.totalBlocks = this.totalBlocks;
                for (const k of MC_PERF_KEYS) perfTracker[k] = [];
            }
            for (const k of MC_PERF_KEYS) perfTracker[k].push(r[k]);
        }
        this.numActiveBlocks = r.numActiveBlocks;
        this.numBlocksWithVertices = r.numBlocksWithVertices;
        this.numVertices = r.numVertices;
        this.newDecompressed = r.numBlocksDecompressed;
        return r.numVertices;
    }
    get vertexBuffer() {
        return new DeviceBuffer(this.device, native.mcVertexBuffer(this.handle), false);
    }
    array(which) {
        return new DeviceBuffer(this.device, native.mcArray(this.handle, which), false);
    }
    reportMemoryUse() {
        const fmt = (count) => {
            if (count > 1e9) return (count / 1e9).toFixed(2) + ' GB';
            if (count > 1e6) return (count / 1e6).toFixed(2) + ' MB';
            if (count > 1e3) return (count / 1e3).toFixed(2) + ' KB';
            return count + ' B';
        };
        const m = native.mcMemoryUse(this.handle);
        const memUse = {mc: {total: m.total - m.cache}, cache: {cache: m.cache}};
        return [`Marching Cubes Data: ${fmt(m.total - m.cache)}`, `LRU Cache Data: ${fmt(m.cache)}`,
                fmt(m.total), memUse];
    }
    destroy() {
        if (this.handle) native.mcDestroy(this.handle);
        this.handle = null;
    }
}

// Batched and served forms (no reference counterpart: the reference answers
// one isovalue per computeSurface).  Several compressed volumes and one
// decoded-block pool stay resident; run() answers many (volume, isovalue)
// queries in one step, serve() replays timestamped requests with continuous
// batching.  Each query's vertices equal a single computeSurface's.
class Engine {
    constructor(device, poolBytes) {
        this.device = device;
        this.handle = native.engCreate(device.handle, poolBytes);
    }
    addVolume(volume, compressionRate, volumeDims, blockRanges = undefined) {
        return native.engAddVolume(this.handle, volume, compressionRate, volumeDims, blockRanges);
    }
    setVertexFormat(fmt) {
        native.engSetFormat(this.handle, fmt);
    }
    // resolves to [{numVertices, vertexOffset, numActiveBlocks, stepDecodedBlocks}]
    async run(volumeIds, isovalues) {
        const ids = Uint32Array.from(volumeIds), isos = Float32Array.from(isovalues);
        return this.device._submit(() => native.engRunAsync(this.handle, ids, isos));
    }
    // the last step's vertices; query i's run from results[i].vertexOffset
    get vertexBuffer() {
        return new DeviceBuffer(this.device, native.engVertexBuffer(this.handle), false);
    }
    stats() {
        return native.engStats(this.handle);
    }
    // arrivals in seconds from the call; resolves to per-request
    // {arrival, start, finish, numVertices, batchSize, step}.  Requests of one
    // step with the same volume and isovalue share one query unless
    // coalesce is false.
    async serve(arrivals, volumeIds, isovalues, {maxBatch = 0, vertexBudgetBytes = 0, coalesce = true} = {}) {
        const a = Float64Array.from(arrivals), ids = Uint32Array.from(volumeIds);
        const isos = Float32Array.from(isovalues);
        return this.device._submit(
            () => native.engServeAsync(this.handle, a, ids, isos, maxBatch, vertexBudgetBytes, coalesce ? 0 : 1));
    }
    destroy() {
        if (this.handle) native.engDestroy(this.handle);
        this.handle = null;
    }
}

// Fixed-rate zfp encoding on the GPU (no reference counterpart: the reference
// loads streams compressed offline).  field: DeviceBuffer of float32, x
// fastest; resolves to a DeviceBuffer with the stream plus two zero words.
async function zfpEncode(device, field, dims, compressionRate) {
    const maxBits = Math.round(64 * compressionRate);
    const nb = Math.ceil(dims[0] / 4) * Math.ceil(dims[1] / 4) * Math.ceil(dims[2] / 4);
    const bytes = Math.ceil((nb * maxBits) / 64) * 8;
    const out = device.createBuffer({size: bytes + 16});
    native.zfpEncode(device.handle, field.handle, dims, maxBits, out.handle);
    return out;
}

// World-space float xyz of vertices [first, first + n) of a block-local
// vertex buffer, placed as the reference's vertex shader places them
// (mc_isosurface.vert:51-55); resolves to a DeviceBuffer of 12 n bytes.
async function worldPositions(device, vertexBuffer, n, dims, scale = [1, 1, 1], first = 0) {
    const out = device.createBuffer({size: Math.max(12 * n, 4)});
    native.verticesWorld(device.handle, vertexBuffer.handle, first, n, dims, scale, out.handle);
    return out;
}

module.exports = {
    requestDevice,
    Device,
    DeviceBuffer,
    ExclusiveScanPipeline,
    ExclusiveScanner,
    StreamCompact,
    StreamCompactor: StreamCompact,
    RadixSorter,
    LRUCache,
    CompressedMarchingCubes,
    Engine,
    zfpEncode,
    worldPositions,
    VertexFormat,
    McArray,
    version: native.version,
    ...require('./benchmark.js'),
};
