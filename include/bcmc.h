/* bcmc.h - C ABI of the MI355X (gfx950) block-compressed marching-cubes library.
 *
 * Drop-in boundary for the hot path of Twinklebear/webgpu-bcmc.  The reference
 * exposes JS classes over a WebGPU device; each entry point below replaces one
 * of their methods (cited as file:line) and is what a Node N-API addon or any
 * other FFI binds.  Conventions:
 *   - every call returns BCMC_OK (0) or a negative bcmc_status; the message of
 *     the last failure on the calling thread is in bcmc_last_error();
 *   - a context owns one HIP device and one stream; all work on a context runs
 *     in issue order, like the reference's single GPUQueue;
 *   - buffers are device allocations (the reference's GPUBuffer); objects own
 *     their internal buffers; buffers returned to the caller are the caller's
 *     to free (as compressed_marching_cubes.js:693 destroys newItemIDs);
 *   - no exceptions cross the ABI; objects are not thread-safe (one object,
 *     one caller at a time, as with the reference's async methods).
 */
#ifndef BCMC_H
#define BCMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define BCMC_API __attribute__((visibility("default")))
#else
#define BCMC_API
#endif

typedef enum {
    BCMC_OK = 0,
    BCMC_ERR_INVALID = -1,   /* bad argument (alert() in the reference) */
    BCMC_ERR_HIP = -2,       /* HIP runtime error */
    BCMC_ERR_NOMEM = -3,     /* device or host allocation failed */
    BCMC_ERR_OVERFLOW = -4,  /* a count exceeds what the output can hold */
    BCMC_ERR_STATE = -5,     /* call order violated (e.g. no volume set) */
    BCMC_ERR_NO_DEVICE = -6  /* no gfx950 device */
} bcmc_status;

typedef struct bcmc_ctx bcmc_ctx; /* GPUDevice + GPUQueue */
typedef struct bcmc_buf bcmc_buf; /* GPUBuffer */
typedef struct bcmc_lru bcmc_lru; /* LRUCache */
typedef struct bcmc_mc bcmc_mc;   /* CompressedMarchingCubes */

BCMC_API const char *bcmc_last_error(void);
BCMC_API const char *bcmc_version(void);

/* ---- context (navigator.gpu.requestAdapter/requestDevice, js/bcmc.js:1-20) */
BCMC_API int bcmc_ctx_create(int device, bcmc_ctx **out);
BCMC_API int bcmc_ctx_destroy(bcmc_ctx *ctx);
/* queue.onSubmittedWorkDone(): wait for all work issued on the context. */
BCMC_API int bcmc_ctx_sync(bcmc_ctx *ctx);
/* Number of compute units and the device name (for reports). */
BCMC_API int bcmc_ctx_info(bcmc_ctx *ctx, int *num_cus, char *name, size_t name_len);

/* ---- buffers (device.createBuffer / mapAsync / copyBufferToBuffer) -------- */
BCMC_API int bcmc_buf_alloc(bcmc_ctx *ctx, uint64_t bytes, bcmc_buf **out);
/* Wrap caller-owned device memory (not freed by bcmc_buf_free). */
BCMC_API int bcmc_buf_wrap(bcmc_ctx *ctx, void *device_ptr, uint64_t bytes, bcmc_buf **out);
BCMC_API int bcmc_buf_free(bcmc_buf *buf);
BCMC_API int bcmc_buf_upload(bcmc_buf *buf, uint64_t offset, const void *src, uint64_t bytes);
BCMC_API int bcmc_buf_download(bcmc_buf *buf, uint64_t offset, void *dst, uint64_t bytes);
BCMC_API int bcmc_buf_fill_u32(bcmc_buf *buf, uint64_t offset, uint32_t value, uint64_t count);
/* commandEncoder.copyBufferToBuffer (used throughout compressed_marching_cubes.js
 * and lru_cache.js): device copy by a kernel on the context stream; offsets
 * 16-byte aligned.  Also the calibration kernel for HBM byte counters. */
BCMC_API int bcmc_buf_copy(bcmc_buf *dst, uint64_t dst_offset, bcmc_buf *src, uint64_t src_offset,
                           uint64_t bytes);
BCMC_API uint64_t bcmc_buf_size(const bcmc_buf *buf);
BCMC_API void *bcmc_buf_device_ptr(const bcmc_buf *buf);

/* ---- ExclusiveScanPipeline / ExclusiveScanner (js/exclusive_scan.js) ------ */
/* getAlignedSize (exclusive_scan.js:95-97): n rounded up to 512. */
BCMC_API uint32_t bcmc_scan_aligned_size(uint32_t n);
/* ExclusiveScanner.scan (exclusive_scan.js:178-295): in-place exclusive scan of
 * the first n u32 of `data` (wrapping), total = sum of those n.  If n <
 * aligned_n, data[n] receives the total, as in the reference (:271-276).
 * aligned_n must equal bcmc_scan_aligned_size(aligned_n) (:116-119). */
BCMC_API int bcmc_scan_exclusive_u32(bcmc_ctx *ctx, bcmc_buf *data, uint32_t aligned_n, uint32_t n,
                                     uint32_t *total);

/* The same scan as bcmc_scan_exclusive_u32, allocating nothing: no reference
 * counterpart.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its look-back state lives in a fixed 1 MiB + 256 B device array that the
 * library holds on each device from bcmc_ctx_create until the process ends,
 * so n is limited to 2^30 (BCMC_ERR_INVALID beyond).  All calls on one device
 * share that array, whichever context issues them: calls that overlap on one
 * device (two contexts or threads at once) race and must be serialised by the
 * caller. */
BCMC_API int bcmc_scan_exclusive_u32_noalloc(bcmc_ctx *ctx, bcmc_buf *data, uint32_t aligned_n,
                                             uint32_t n, uint32_t *total);

/* ---- StreamCompact / StreamCompactor (js/stream_compact.js:56-165) -------- */
/* out[offsets[i]] = i for every i < n with flags[i] != 0 (stream_compact.comp). */
BCMC_API int bcmc_compact_ids(bcmc_ctx *ctx, uint32_t n, bcmc_buf *flags, bcmc_buf *offsets,
                              bcmc_buf *out);
/* Fused single-pass scan + compact of 0/1 flags: out = ascending ids of the
 * non-zero flags; *count = their number.  (No reference counterpart: it is the
 * scan + compactActiveIDs pair of compressed_marching_cubes.js:815-825,704-707
 * as one pass.) */
BCMC_API int bcmc_compact_flags(bcmc_ctx *ctx, uint32_t n, bcmc_buf *flags, bcmc_buf *out,
                                uint32_t *count);

/* ---- RadixSorter (js/radix_sort_by_key.js) --------------------------------- */
/* getAlignedSize (radix_sort_by_key.js:155-158): nextPow2(ceil(n/64))*64. */
BCMC_API uint32_t bcmc_sort_aligned_size(uint32_t n);
/* RadixSorter.sort (radix_sort_by_key.js:161-400): sorts the first n (key,
 * value) u32 pairs; keys/vals hold bcmc_sort_aligned_size(n) entries.  On
 * return vals[0..n) holds exactly what the reference leaves there (ascending
 * keys, or descending with reverse != 0, with the reference's order among equal
 * keys); the contents of keys are unspecified, as in the reference. */
BCMC_API int bcmc_sort_u32_by_key(bcmc_ctx *ctx, bcmc_buf *keys, bcmc_buf *vals, uint32_t n,
                                  int reverse);

/* ---- LRUCache (js/lru_cache.js) -------------------------------------------- */
/* new LRUCache(device, scan, compact, initialSize, elementSize, totalElements)
 * (lru_cache.js:4-305).  The slot contents are stored for the logical size
 * plus room for the next growth and moved to a 1.5x larger allocation when
 * an update grows the cache past it (the reference re-creates its buffers the
 * same way, lru_cache.js:509-672); the per-slot bookkeeping (age,
 * availability, item: 24 bytes a slot) is sized for totalElements once.  The
 * logical size follows the reference exactly. */
BCMC_API int bcmc_lru_create(bcmc_ctx *ctx, uint32_t initial_size, uint32_t elem_bytes,
                             uint32_t total, bcmc_lru **out);
BCMC_API int bcmc_lru_destroy(bcmc_lru *lru);
/* LRUCache.update (lru_cache.js:312-843).  needed: `total` u32 flags (0/1).
 * Returns the number of new items and, when non-zero, a new buffer with their
 * ids in ascending order (caller frees); *new_items = NULL when zero. */
BCMC_API int bcmc_lru_update(bcmc_lru *lru, bcmc_buf *needed, uint32_t *num_new,
                             bcmc_buf **new_items);
/* LRUCache.update with the reference's perfTracker entries
 * (lru_cache.js:320-333): device milliseconds of each stage and the two
 * counts.  Stages the reference runs apart run fused here: the new-item pass
 * marks, counts and compacts (lruMarkNewItems; lruScanNeededSlots and
 * lruCompactNewItems are 0), the aging pass scans and compacts the available
 * slots (lruScanSlotsAvailable; lruCompactAvailableSlots 0), and the sort
 * forms its keys in its first pass (lruPrepKeyValue 0).  With no new items
 * every stage after lruNumNewItems is 0, as in the reference. */
typedef struct {
    double lruMarkNewItems, lruScanNeededSlots, lruNumNewItems, lruCompactNewItems;
    double lruScanSlotsAvailable, lruNumSlotsAvailable, lruGrowCache, lruCompactAvailableSlots;
    double lruPrepKeyValue, lruTotalSortTime, lruWriteNewItems;
} bcmc_lru_perf;
BCMC_API int bcmc_lru_update_timed(bcmc_lru *lru, bcmc_buf *needed, uint32_t *num_new,
                                   bcmc_buf **new_items, bcmc_lru_perf *perf);
/* LRUCache.reset (lru_cache.js:846-903). */
BCMC_API int bcmc_lru_reset(bcmc_lru *lru);
typedef struct {
    uint32_t cache_size;      /* this.cacheSize (slots) */
    uint32_t total_elements;  /* this.totalElements */
    uint32_t element_size;    /* this.elementSize (bytes) */
    uint32_t slots_available; /* this.displayNumSlotsAvailable after the last update */
    uint64_t reserved_bytes;  /* device memory of the slot contents: follows cacheSize (plus
                                 room for growth), not totalElements */
    uint64_t logical_bytes;   /* cacheSize * elementSize */
} bcmc_lru_info;
BCMC_API int bcmc_lru_get_info(bcmc_lru *lru, bcmc_lru_info *info);
/* Slot state for inspection: age[C], available[C], item[C] (-1 = empty) and
 * item_slot[total] (-1 = not cached); any pointer may be NULL. */
BCMC_API int bcmc_lru_read_state(bcmc_lru *lru, uint32_t *age, uint32_t *available, int32_t *item,
                                 int32_t *item_slot);
/* The cache contents (this.cache): cache_size * element_size bytes. */
BCMC_API bcmc_buf *bcmc_lru_cache_buffer(bcmc_lru *lru);

/* ---- CompressedMarchingCubes (js/compressed_marching_cubes.js) ------------- */
typedef enum {
    /* 16 B per vertex: float x, y, z in block-local voxel units (the
     * reference's lerp + voxel + 0.5, compute_block_vertices.comp:141-164)
     * and the u32 block id; global position = 4*block_pos + local. */
    BCMC_VERTS_LOCAL_F32 = 0,
    /* 12 B per vertex: global float xyz = 4*block_pos + local. */
    BCMC_VERTS_GLOBAL_F32 = 1,
    /* 8 B per vertex: the reference's uvec2(compress_position(p), block_id)
     * (block_vertex_position.comp:9-17), for its renderer. */
    BCMC_VERTS_PACKED_U32X2 = 2
} bcmc_vertex_format;

/* World-space float xyz of vertices first..first+n of a BCMC_VERTS_LOCAL_F32
 * buffer, as the reference's renderer places them (shaders/
 * mc_isosurface.vert:51-55): scale * ((4*block_pos + local) / dims - 0.5),
 * dims the volume's dims.  out: at least 12*n bytes.  Enqueued on the
 * context stream. */
BCMC_API int bcmc_vertices_world(bcmc_ctx *ctx, bcmc_buf *verts, uint64_t first, uint64_t n,
                                 const uint32_t dims[3], const float scale[3], bcmc_buf *out);

typedef struct {
    /* perfTracker keys of compressed_marching_cubes.js:593-610, milliseconds of
     * device time per stage: computeBlockHasVertices is the count launch,
     * compactBlocksWithVertices the scan of the chunk totals, computeVertices
     * the emit launch; compactActiveIDs is 0 (it runs inside
     * computeActiveBlocks). */
    double computeActiveBlocks, cacheUpdate, decompression, compactActiveIDs,
        computeBlockHasVertices, compactBlocksWithVertices, computeVertices;
    double total_ms; /* wall time of the call */
    uint32_t numActiveBlocks, numBlocksDecompressed, numBlocksWithVertices, cacheSize;
    uint64_t numVertices;
} bcmc_perf;

/* new CompressedMarchingCubes(device) (compressed_marching_cubes.js:1-364). */
BCMC_API int bcmc_mc_create(bcmc_ctx *ctx, bcmc_mc **out);
BCMC_API int bcmc_mc_destroy(bcmc_mc *mc);
/* setCompressedVolume(volume, compressionRate, volumeDims, volumeScale)
 * (compressed_marching_cubes.js:432-520, ranges :522-578).  `stream` is a
 * headerless zfp fixed-rate float32 stream; rate in bits per value (64*rate
 * must be an integer >= 9).  initial_cache_slots = 0 selects the reference's
 * ceil(0.05 * totalBlocks) (:443-448). */
BCMC_API int bcmc_mc_set_compressed_volume(bcmc_mc *mc, const void *stream, uint64_t bytes,
                                           double rate, const uint32_t dims[3],
                                           const float scale[3], uint32_t initial_cache_slots);
/* Block value ranges persisted beside a stream, so that a load need not decode
 * the whole volume (the reference always does, computeBlockRanges,
 * compressed_marching_cubes.js:522-578, and notes at :523-524 that the server
 * could compute them).
 * BCMC_RANGES_BLOCK: float2 (min, max) per block over its voxels inside the
 *   volume, x-fastest block order: the reference's blockRangesBuffer layout,
 *   as bcmc_zfp_block_ranges writes it.
 * BCMC_RANGES_BLOCK_TILE: that array followed by a second float2 per block,
 *   the range of every decoded voxel the block's cells read (its own 64,
 *   padding included, and the ghost faces, edges and corner of its +x/+y/+z
 *   neighbours). */
typedef enum { BCMC_RANGES_BLOCK = 0, BCMC_RANGES_BLOCK_TILE = 1 } bcmc_ranges_format;
/* setCompressedVolume with persisted ranges instead of the full-volume decode.
 * `stream` and `ranges` may be host or device memory; ranges_bytes must be 8
 * (BLOCK) or 16 (BLOCK_TILE) times the block count.  With BLOCK the tile
 * ranges are bounded by the union of the block's and its +x/+y/+z neighbours'
 * ranges ((-inf, inf) for blocks that read padding voxels): results are the
 * same, the count pass examines more blocks.  The ranges must be this
 * stream's: others give another active set, as they would in the reference. */
BCMC_API int bcmc_mc_set_compressed_volume_ranges(bcmc_mc *mc, const void *stream, uint64_t bytes,
                                                  double rate, const uint32_t dims[3],
                                                  const float scale[3], uint32_t initial_cache_slots,
                                                  const void *ranges, bcmc_ranges_format fmt,
                                                  uint64_t ranges_bytes);
/* The loaded volume's ranges in `fmt` (host or device memory of at least
 * 8 or 16 bytes per block), for persisting beside the stream. */
BCMC_API int bcmc_mc_get_block_ranges(bcmc_mc *mc, bcmc_ranges_format fmt, void *out, uint64_t bytes);
BCMC_API int bcmc_mc_set_vertex_format(bcmc_mc *mc, bcmc_vertex_format fmt);
/* computeSurface(isovalue, perfTracker) (compressed_marching_cubes.js:580-809).
 * The isovalue is used as FP32 (:613-619).  perf may be NULL: the query chain
 * is then launched in the context's launch mode (bcmc_ctx_set_graph_replay:
 * direct launches, or a HIP graph captured on first use and captured again
 * when its buffers change); with perf the stages are launched directly and
 * timed with events.  Results are identical either way. */
BCMC_API int bcmc_mc_compute_surface(bcmc_mc *mc, float isovalue, bcmc_perf *perf,
                                     uint64_t *num_vertices);
/* Public fields of the reference object. */
typedef struct {
    uint32_t total_blocks, num_active_blocks, num_blocks_with_vertices, new_decompressed;
    uint64_t num_vertices, vertex_capacity;
    uint32_t dims[3], padded_dims[3], max_bits;
    float scale[3];
    bcmc_vertex_format format;
} bcmc_mc_info;
BCMC_API int bcmc_mc_get_info(bcmc_mc *mc, bcmc_mc_info *info);
/* this.vertexBuffer: num_vertices entries in the selected format. */
BCMC_API bcmc_buf *bcmc_mc_vertex_buffer(bcmc_mc *mc);
BCMC_API bcmc_lru *bcmc_mc_lru(bcmc_mc *mc);
/* Device arrays of the last query, for inspection: block ranges (float2[N]),
 * active flags (u32[N]), active ids (u32[A]), decoded ids (u32[D]), indices
 * into the active list of blocks with vertices (u32[H]) and their vertex
 * offsets (u64[H]). */
typedef enum {
    BCMC_ARRAY_RANGES = 0,
    BCMC_ARRAY_ACTIVE_FLAGS = 1,
    BCMC_ARRAY_ACTIVE_IDS = 2,
    BCMC_ARRAY_NEW_ITEMS = 3,
    BCMC_ARRAY_WITH_VERTS = 4,
    BCMC_ARRAY_VERT_OFFSETS = 5
} bcmc_mc_array;
BCMC_API bcmc_buf *bcmc_mc_array_buffer(bcmc_mc *mc, bcmc_mc_array which);
/* reportMemoryUse() (compressed_marching_cubes.js:366-430): total bytes of
 * device memory held by the object (out), with the cache share (cache_out). */
BCMC_API int bcmc_mc_memory_use(bcmc_mc *mc, uint64_t *total_out, uint64_t *cache_out);

/* ---- Batched and served forms (no reference counterpart) ------------------- */
/* An engine keeps compressed volumes and a decoded-block pool shared by all of
 * them resident on the device; one bcmc_engine_run call answers many
 * (volume, isovalue) queries with one launch per stage for all of them.  Each
 * query's vertices equal those of a single computeSurface on that volume
 * (same blocks, order and FP32 positions); pool slots are the engine's own
 * (the least recently used blocks are evicted when the pool is full). */
typedef struct bcmc_engine bcmc_engine;
typedef struct {
    uint64_t num_vertices;        /* this query's vertices */
    uint64_t vertex_offset;       /* its first vertex in the step's vertex buffer */
    uint32_t num_active_blocks;   /* this query's active blocks */
    uint32_t step_decoded_blocks; /* blocks the whole step decoded */
} bcmc_query_result;
typedef struct {
    uint32_t num_volumes, pool_slots, pool_used, epoch;
    uint64_t pool_bytes, device_bytes;
    uint32_t steps;       /* bcmc_engine_run calls that launched a step */
    uint32_t graph_steps; /* of which replayed a captured HIP graph */
} bcmc_engine_info;
/* pool_bytes: HBM for decoded blocks (256 B per block). */
BCMC_API int bcmc_engine_create(bcmc_ctx *ctx, uint64_t pool_bytes, bcmc_engine **out);
BCMC_API int bcmc_engine_destroy(bcmc_engine *e);
BCMC_API int bcmc_engine_add_volume(bcmc_engine *e, const void *stream, uint64_t bytes, double rate,
                                    const uint32_t dims[3], uint32_t *volume_id);
/* add_volume with persisted block ranges (bcmc_ranges_format, as
 * bcmc_mc_set_compressed_volume_ranges takes them). */
BCMC_API int bcmc_engine_add_volume_ranges(bcmc_engine *e, const void *stream, uint64_t bytes, double rate,
                                           const uint32_t dims[3], const void *ranges,
                                           bcmc_ranges_format fmt, uint64_t ranges_bytes,
                                           uint32_t *volume_id);
BCMC_API int bcmc_engine_set_vertex_format(bcmc_engine *e, bcmc_vertex_format fmt);
/* One step over nq queries; results[i] for query i (may be NULL).  Fails with
 * BCMC_ERR_NOMEM when the step's blocks do not fit the pool. */
BCMC_API int bcmc_engine_run(bcmc_engine *e, const uint32_t *volume_ids, const float *isovalues,
                             uint32_t nq, bcmc_query_result *results);
/* The last step's vertices (results[i].vertex_offset .. + num_vertices). */
BCMC_API bcmc_buf *bcmc_engine_vertex_buffer(bcmc_engine *e);
BCMC_API int bcmc_engine_stats(bcmc_engine *e, bcmc_engine_info *info);
/* Pool slots that one step of queries at isovalues[0..n) on volume_id may
 * need (the served loop's admission bound): the volume's blocks except those
 * whose dilated range (the active test's) lies strictly between two of the
 * isovalues or strictly below or above all of them, counted from a histogram
 * of the ranges built at add_volume.  Never below the blocks a step of these
 * queries makes active; 0 for n == 0. */
BCMC_API int bcmc_engine_block_bound(bcmc_engine *e, uint32_t volume_id, const float *isovalues, uint32_t n,
                                     uint64_t *blocks);

/* Served form: replay timestamped requests in real time through one engine
 * with continuous batching (each step takes the requests that have arrived,
 * FIFO, up to max_batch).  Admission waits for room rather than failing:
 * a step takes requests while their bound on touched blocks fits the pool
 * slots the previous step left evictable (an empty step frees the rest when
 * the head request needs them), and emits only the leading requests whose
 * exact vertex counts fit vertex_budget_bytes; the others stay at the head of
 * the queue.  Fails only when one request alone exceeds the pool or the
 * budget.  out[i] receives request i's arrival, step start and finish times
 * (seconds from the call) and its vertex count; the vertices of the last step
 * stay in the engine buffer.  Equal requests in one step are answered by one
 * query unless opts->no_coalesce is set; they never add to the pool bound. */
typedef struct {
    uint32_t max_batch;           /* 0: 256 */
    uint64_t vertex_budget_bytes; /* HBM for one step's vertices; 0: 64 GiB */
    /* 0: requests of one step with the same volume and isovalue (bitwise) are
     * coalesced: the step computes that query once and every such request is
     * answered by its result (same vertex count and vertex range).
     * 1: every request is a query of its own. */
    uint32_t no_coalesce;
} bcmc_serve_opts;
typedef struct {
    double arrival_s, start_s, finish_s;
    uint64_t num_vertices;
    uint32_t batch_size, step;
} bcmc_serve_result;
BCMC_API int bcmc_serve_trace(bcmc_engine *e, const double *arrival_s, const uint32_t *volume_ids,
                              const float *isovalues, uint32_t n, const bcmc_serve_opts *opts,
                              bcmc_serve_result *out);

/* ---- ZFP decode (shaders/zfp_decompress.comp, zfp_decompress_block.comp) -- */
/* Decode blocks ids[0..n) of a fixed-rate stream into out[i*64 .. i*64+64)
 * (float32, x fastest).  ids may be NULL for blocks 0..n-1. */
BCMC_API int bcmc_zfp_decode_blocks(bcmc_ctx *ctx, bcmc_buf *stream, uint32_t max_bits,
                                    bcmc_buf *ids, uint32_t n, bcmc_buf *out);
/* Block value ranges (zfp_compute_block_range.comp): float2 per block. */
BCMC_API int bcmc_zfp_block_ranges(bcmc_ctx *ctx, bcmc_buf *stream, uint32_t max_bits,
                                   const uint32_t dims[3], bcmc_buf *ranges);
/* Fixed-rate zfp encoder on the device: the ingest side of the stream format
 * the decoders read.  No reference counterpart (the reference loads streams
 * compressed offline with zfp 0.5.5, whose zfp_compress in fixed-rate mode at
 * rate max_bits/64 this matches byte for byte).  field: finite float32
 * values, x fastest, dims[0]*dims[1]*dims[2] of them; stream: at least
 * ceil(nblocks*max_bits/64)*8 bytes, all of it overwritten (zeroed, then the
 * blocks' bits; leave two extra words for the decoders' read-ahead). */
BCMC_API int bcmc_zfp_encode(bcmc_ctx *ctx, bcmc_buf *field, const uint32_t dims[3],
                             uint32_t max_bits, bcmc_buf *stream);

/* With timing enabled, bcmc_zfp_decode_blocks and bcmc_zfp_block_ranges wait
 * for their kernel and record its device time (HIP events on the context
 * stream); bcmc_ctx_last_kernel_ms returns it in ms (0 if none recorded). */
BCMC_API int bcmc_ctx_enable_timing(bcmc_ctx *ctx, int enable);
BCMC_API double bcmc_ctx_last_kernel_ms(bcmc_ctx *ctx);

/* How the query chains of the pipelines and engines on this context are
 * launched.  0 (the default): every kernel is launched directly.  1: a chain
 * is captured into a HIP graph the second time its shape is seen (a
 * pipeline's untimed computeSurface; an engine step's front and emit parts)
 * and replayed after that.  Results are identical either way; on MI355X with
 * ROCm 7.2 the direct launches measured faster (plain queries 2%, a served
 * request alone in its step 5%).  No reference counterpart (WebGPU command
 * buffers are recorded per call in compressed_marching_cubes.js). */
BCMC_API int bcmc_ctx_set_graph_replay(bcmc_ctx *ctx, int enable);
BCMC_API int bcmc_ctx_get_graph_replay(bcmc_ctx *ctx, int *enabled);

/* Launch shapes of the context's persistent kernels: workgroups per compute
 * unit for the single-pass scans, the radix-sort passes, the zfp decoders and
 * the marching-cubes emit and count launches.  No reference counterpart (WebGPU dispatch
 * sizes are fixed in compressed_marching_cubes.js / exclusive_scan.js /
 * radix_sort.js); results are identical for every shape, only speed changes.
 * Defaults: scan 2, sort 8, decode 8, mc (emit) 5, count 8; tiles 16, 16, 256.  The *_ms fields are filled by
 * bcmc_ctx_autotune (device time of the chosen shape on its probe workload)
 * and ignored by bcmc_ctx_set_tuning. */
typedef struct bcmc_tuning {
    uint32_t scan_wg_per_cu;
    uint32_t sort_wg_per_cu;
    uint32_t decode_wg_per_cu;
    uint32_t mc_wg_per_cu;
    uint32_t count_wg_per_cu;
    float scan_ms;   /* 16M-element exclusive scan */
    float sort_ms;   /* 2M pairs, 24-bit keys (3 digit passes) */
    float decode_ms; /* 256K blocks of a synthetic rate-4 stream into a block cache */
    float mc_ms;     /* not measured: the MC launches keep their occupancy-derived shapes */
    /* Tile shapes (replace the reference's build-time BLOCK_SIZE and
     * SORT_CHUNK_SIZE, shaders/embed_shaders.py:30-31); 0 in
     * bcmc_ctx_set_tuning keeps the current one. */
    uint32_t scan_items;     /* scan tile = 512 x scan_items elements: 8, 16 or 32 */
    uint32_t sort_items;     /* radix-sort tile = 256 x sort_items: 8, 16 or 32 */
    uint32_t decode_threads; /* zfp decoder blocks per workgroup: 64, 128 or 256 */
} bcmc_tuning;
BCMC_API int bcmc_ctx_get_tuning(bcmc_ctx *ctx, bcmc_tuning *out);
/* Each *_wg_per_cu must be in 1..64; pipelines re-capture their graphs. */
BCMC_API int bcmc_ctx_set_tuning(bcmc_ctx *ctx, const bcmc_tuning *tuning);
/* Times the scan, sort and decode candidates (tile shape x workgroups per
 * CU) on synthetic device data at several sizes (scans of 256K, 2M and 16M
 * elements, sorts of 64K and 2M pairs, decodes of 32K and 256K blocks; about
 * 200 MiB of temporary device memory, about a second) and keeps per family
 * the shape with the least time per element summed over the sizes; out
 * (nullable) receives the result, its *_ms at the largest size.  Call it
 * before the first query: it synchronises the stream. */
BCMC_API int bcmc_ctx_autotune(bcmc_ctx *ctx, bcmc_tuning *out);

#ifdef __cplusplus
}
#endif
#endif
