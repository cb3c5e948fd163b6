/* Node.js N-API addon over the C ABI in include/bcmc.h.
 *
 * Thin by design: each function converts JS values to plain C arguments, calls
 * one bcmc_* entry point and throws a JS Error carrying bcmc_last_error() on a
 * negative status.  Handles cross as N-API externals.  The JS facade
 * (index.js) builds the reference's classes (CompressedMarchingCubes,
 * LRUCache, ExclusiveScanPipeline/ExclusiveScanner, StreamCompact(or),
 * RadixSorter) on top of these functions.
 *
 * The GPU work of the reference's async methods (computeSurface,
 * LRUCache.update, scan, compaction, sort, Engine.run / serve) also has a
 * *Async form: a napi_async_work runs the bcmc_* call on a libuv worker
 * thread and resolves a Promise on the main thread, so the event loop keeps
 * running meanwhile (the reference awaits onSubmittedWorkDone /
 * mapAsync the same way).  Every library call, synchronous or not, holds one
 * mutex: the library's objects are not meant for concurrent calls, and the
 * facade issues each device's operations in order.  The typed arrays an
 * async call reads are kept alive by references until it completes.
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bcmc.h"

static pthread_mutex_t g_lib_mu = PTHREAD_MUTEX_INITIALIZER;

#define NAPI_CALL(env, call)                                                                    \
    do {                                                                                        \
        if ((call) != napi_ok) {                                                                \
            napi_throw_error((env), NULL, "N-API call failed: " #call);                          \
            return NULL;                                                                        \
        }                                                                                       \
    } while (0)

/* One library call under the mutex; a negative status throws. */
#define BCMC_CALL(env, call)                                                                    \
    do {                                                                                        \
        char msg_[640];                                                                         \
        pthread_mutex_lock(&g_lib_mu);                                                          \
        int rc_ = (call);                                                                       \
        if (rc_ != BCMC_OK)                                                                     \
            snprintf(msg_, sizeof(msg_), "bcmc error %d: %s", rc_, bcmc_last_error());          \
        pthread_mutex_unlock(&g_lib_mu);                                                        \
        if (rc_ != BCMC_OK) {                                                                   \
            napi_throw_error((env), NULL, msg_);                                                \
            return NULL;                                                                        \
        }                                                                                       \
    } while (0)

#define GET_ARGS(n)                                                                             \
    size_t argc = (n);                                                                          \
    napi_value argv[(n) > 0 ? (n) : 1];                                                         \
    NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL))

static void *ext(napi_env env, napi_value v)
{
    void *p = NULL;
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok || t != napi_external)
        return NULL;
    napi_get_value_external(env, v, &p);
    return p;
}

static napi_value mk_ext(napi_env env, void *p)
{
    napi_value r;
    if (!p) {
        napi_get_null(env, &r);
        return r;
    }
    napi_create_external(env, p, NULL, NULL, &r);
    return r;
}

static double num(napi_env env, napi_value v)
{
    double d = 0;
    napi_get_value_double(env, v, &d);
    return d;
}

static napi_value mk_num(napi_env env, double d)
{
    napi_value r;
    napi_create_double(env, d, &r);
    return r;
}

static void set_num(napi_env env, napi_value obj, const char *k, double d)
{
    napi_set_named_property(env, obj, k, mk_num(env, d));
}

/* typed array / buffer data and byte length */
static int bytes_of(napi_env env, napi_value v, void **data, size_t *len)
{
    bool is_ta = false, is_buf = false;
    napi_is_typedarray(env, v, &is_ta);
    if (is_ta) {
        napi_typedarray_type t;
        size_t n, off;
        napi_value ab;
        void *d;
        if (napi_get_typedarray_info(env, v, &t, &n, &d, &ab, &off) != napi_ok)
            return -1;
        size_t es = 1;
        switch (t) {
        case napi_int16_array: case napi_uint16_array: es = 2; break;
        case napi_int32_array: case napi_uint32_array: case napi_float32_array: es = 4; break;
        case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: es = 8; break;
        default: es = 1;
        }
        *data = d;
        *len = n * es;
        return 0;
    }
    napi_is_buffer(env, v, &is_buf);
    if (is_buf)
        return napi_get_buffer_info(env, v, data, len) == napi_ok ? 0 : -1;
    return -1;
}

static void *ext_arg(napi_env env, napi_value v, const char *what)
{
    void *p = ext(env, v);
    if (!p) {
        char m[128];
        snprintf(m, sizeof(m), "expected a %s handle", what);
        napi_throw_type_error(env, NULL, m);
    }
    return p;
}

#define EXT(i, T, what)                                                                         \
    T *what = (T *)ext_arg(env, argv[i], #what);                                                \
    if (!what)                                                                                  \
        return NULL

/* ---- context & buffers ---------------------------------------------------- */
static napi_value ctx_create(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    bcmc_ctx *c = NULL;
    BCMC_CALL(env, bcmc_ctx_create(argc > 0 ? (int)num(env, argv[0]) : 0, &c));
    return mk_ext(env, c);
}

static napi_value ctx_destroy(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_ctx, ctx);
    BCMC_CALL(env, bcmc_ctx_destroy(ctx));
    return NULL;
}

static napi_value ctx_sync(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_ctx, ctx);
    BCMC_CALL(env, bcmc_ctx_sync(ctx));
    return NULL;
}

static napi_value ctx_info(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_ctx, ctx);
    int cus = 0;
    char name[128];
    BCMC_CALL(env, bcmc_ctx_info(ctx, &cus, name, sizeof(name)));
    napi_value o, s;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "numCUs", cus);
    NAPI_CALL(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s));
    napi_set_named_property(env, o, "name", s);
    return o;
}

/* ctxAutotune(ctx) -> {scan, sort, decode, mc, count, scanItems, sortItems,
 * decodeThreads, scanMs, sortMs, decodeMs} */
static napi_value ctx_autotune(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_ctx, ctx);
    bcmc_tuning t;
    BCMC_CALL(env, bcmc_ctx_autotune(ctx, &t));
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "scan", t.scan_wg_per_cu);
    set_num(env, o, "sort", t.sort_wg_per_cu);
    set_num(env, o, "decode", t.decode_wg_per_cu);
    set_num(env, o, "mc", t.mc_wg_per_cu);
    set_num(env, o, "count", t.count_wg_per_cu);
    set_num(env, o, "scanItems", t.scan_items);
    set_num(env, o, "sortItems", t.sort_items);
    set_num(env, o, "decodeThreads", t.decode_threads);
    set_num(env, o, "scanMs", t.scan_ms);
    set_num(env, o, "sortMs", t.sort_ms);
    set_num(env, o, "decodeMs", t.decode_ms);
    return o;
}

/* ctxSetGraphReplay(ctx, enable): query chains launched directly (false) or
 * replayed from captured HIP graphs (true) */
static napi_value ctx_set_graph_replay(napi_env env, napi_callback_info info)
{
    GET_ARGS(2);
    EXT(0, bcmc_ctx, ctx);
    bool on = false;
    NAPI_CALL(env, napi_get_value_bool(env, argv[1], &on));
    BCMC_CALL(env, bcmc_ctx_set_graph_replay(ctx, on ? 1 : 0));
    return NULL;
}

static napi_value buf_alloc(napi_env env, napi_callback_info info)
{
    GET_ARGS(2);
    EXT(0, bcmc_ctx, ctx);
    bcmc_buf *b = NULL;
    BCMC_CALL(env, bcmc_buf_alloc(ctx, (uint64_t)num(env, argv[1]), &b));
    return mk_ext(env, b);
}

static napi_value buf_free(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_buf, buf);
    BCMC_CALL(env, bcmc_buf_free(buf));
    return NULL;
}

static napi_value buf_size(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_buf, buf);
    return mk_num(env, (double)bcmc_buf_size(buf));
}

/* bufUpload(buf, offset, typedArray) / bufDownload(buf, offset, typedArray) */
static napi_value buf_upload(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_buf, buf);
    void *d;
    size_t n;
    if (bytes_of(env, argv[2], &d, &n)) {
        napi_throw_type_error(env, NULL, "expected a TypedArray or Buffer");
        return NULL;
    }
    BCMC_CALL(env, bcmc_buf_upload(buf, (uint64_t)num(env, argv[1]), d, n));
    return NULL;
}

static napi_value buf_download(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_buf, buf);
    void *d;
    size_t n;
    if (bytes_of(env, argv[2], &d, &n)) {
        napi_throw_type_error(env, NULL, "expected a TypedArray or Buffer");
        return NULL;
    }
    BCMC_CALL(env, bcmc_buf_download(buf, (uint64_t)num(env, argv[1]), d, n));
    return NULL;
}

/* ---- primitives ------------------------------------------------------------ */
static napi_value scan_aligned_size(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    return mk_num(env, bcmc_scan_aligned_size((uint32_t)num(env, argv[0])));
}

static napi_value scan_exclusive(napi_env env, napi_callback_info info)
{
    GET_ARGS(4);
    EXT(0, bcmc_ctx, ctx);
    EXT(1, bcmc_buf, buf);
    uint32_t total = 0;
    BCMC_CALL(env, bcmc_scan_exclusive_u32(ctx, buf, (uint32_t)num(env, argv[2]),
                                           (uint32_t)num(env, argv[3]), &total));
    return mk_num(env, total);
}

static napi_value compact_ids(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_ctx, ctx);
    EXT(2, bcmc_buf, flags);
    EXT(3, bcmc_buf, offsets);
    EXT(4, bcmc_buf, out);
    BCMC_CALL(env, bcmc_compact_ids(ctx, (uint32_t)num(env, argv[1]), flags, offsets, out));
    return NULL;
}

static napi_value compact_flags(napi_env env, napi_callback_info info)
{
    GET_ARGS(4);
    EXT(0, bcmc_ctx, ctx);
    EXT(2, bcmc_buf, flags);
    EXT(3, bcmc_buf, out);
    uint32_t count = 0;
    BCMC_CALL(env, bcmc_compact_flags(ctx, (uint32_t)num(env, argv[1]), flags, out, &count));
    return mk_num(env, count);
}

static napi_value sort_aligned_size(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    return mk_num(env, bcmc_sort_aligned_size((uint32_t)num(env, argv[0])));
}

static napi_value sort_by_key(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_ctx, ctx);
    EXT(1, bcmc_buf, keys);
    EXT(2, bcmc_buf, vals);
    bool rev = false;
    napi_get_value_bool(env, argv[4], &rev);
    BCMC_CALL(env, bcmc_sort_u32_by_key(ctx, keys, vals, (uint32_t)num(env, argv[3]), rev ? 1 : 0));
    return NULL;
}

/* ---- LRU ------------------------------------------------------------------- */
static napi_value lru_create(napi_env env, napi_callback_info info)
{
    GET_ARGS(4);
    EXT(0, bcmc_ctx, ctx);
    bcmc_lru *l = NULL;
    BCMC_CALL(env, bcmc_lru_create(ctx, (uint32_t)num(env, argv[1]), (uint32_t)num(env, argv[2]),
                                   (uint32_t)num(env, argv[3]), &l));
    return mk_ext(env, l);
}

static napi_value lru_destroy(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_lru, lru);
    BCMC_CALL(env, bcmc_lru_destroy(lru));
    return NULL;
}

/* [numNew, idsBuffer | null, perf | undefined] */
static napi_value lru_update_result(napi_env env, uint32_t n, bcmc_buf *ids, const bcmc_lru_perf *p)
{
    napi_value arr;
    NAPI_CALL(env, napi_create_array_with_length(env, 3, &arr));
    napi_set_element(env, arr, 0, mk_num(env, n));
    napi_set_element(env, arr, 1, mk_ext(env, ids));
    if (p) {
        napi_value o;
        NAPI_CALL(env, napi_create_object(env, &o));
        set_num(env, o, "lruMarkNewItems", p->lruMarkNewItems);
        set_num(env, o, "lruScanNeededSlots", p->lruScanNeededSlots);
        set_num(env, o, "lruNumNewItems", p->lruNumNewItems);
        set_num(env, o, "lruCompactNewItems", p->lruCompactNewItems);
        set_num(env, o, "lruScanSlotsAvailable", p->lruScanSlotsAvailable);
        set_num(env, o, "lruNumSlotsAvailable", p->lruNumSlotsAvailable);
        set_num(env, o, "lruGrowCache", p->lruGrowCache);
        set_num(env, o, "lruCompactAvailableSlots", p->lruCompactAvailableSlots);
        set_num(env, o, "lruPrepKeyValue", p->lruPrepKeyValue);
        set_num(env, o, "lruTotalSortTime", p->lruTotalSortTime);
        set_num(env, o, "lruWriteNewItems", p->lruWriteNewItems);
        napi_set_element(env, arr, 2, o);
    }
    return arr;
}

/* lruUpdate(lru, needed, timed) */
static napi_value lru_update(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_lru, lru);
    EXT(1, bcmc_buf, needed);
    bool timed = false;
    if (argc > 2)
        napi_get_value_bool(env, argv[2], &timed);
    uint32_t n = 0;
    bcmc_buf *ids = NULL;
    bcmc_lru_perf p;
    BCMC_CALL(env, bcmc_lru_update_timed(lru, needed, &n, &ids, timed ? &p : NULL));
    return lru_update_result(env, n, ids, timed ? &p : NULL);
}

static napi_value lru_reset(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_lru, lru);
    BCMC_CALL(env, bcmc_lru_reset(lru));
    return NULL;
}

static napi_value lru_info(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_lru, lru);
    bcmc_lru_info li;
    BCMC_CALL(env, bcmc_lru_get_info(lru, &li));
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "cacheSize", li.cache_size);
    set_num(env, o, "totalElements", li.total_elements);
    set_num(env, o, "elementSize", li.element_size);
    set_num(env, o, "displayNumSlotsAvailable", li.slots_available);
    set_num(env, o, "reservedBytes", (double)li.reserved_bytes);
    set_num(env, o, "logicalBytes", (double)li.logical_bytes);
    return o;
}

/* lruState(lru, age:Uint32Array, avail:Uint32Array, item:Int32Array, slot:Int32Array) */
static napi_value lru_state(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_lru, lru);
    void *p[4];
    size_t n;
    for (int i = 0; i < 4; ++i)
        if (bytes_of(env, argv[1 + i], &p[i], &n)) {
            napi_throw_type_error(env, NULL, "expected typed arrays");
            return NULL;
        }
    BCMC_CALL(env, bcmc_lru_read_state(lru, p[0], p[1], p[2], p[3]));
    return NULL;
}

static napi_value lru_cache_buffer(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_lru, lru);
    return mk_ext(env, bcmc_lru_cache_buffer(lru));
}

/* ---- CompressedMarchingCubes ---------------------------------------------- */
static napi_value mc_create(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_ctx, ctx);
    bcmc_mc *m = NULL;
    BCMC_CALL(env, bcmc_mc_create(ctx, &m));
    return mk_ext(env, m);
}

static napi_value mc_destroy(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_mc, mc);
    BCMC_CALL(env, bcmc_mc_destroy(mc));
    return NULL;
}

static int three(napi_env env, napi_value arr, double out[3])
{
    for (uint32_t i = 0; i < 3; ++i) {
        napi_value e;
        if (napi_get_element(env, arr, i, &e) != napi_ok)
            return -1;
        out[i] = num(env, e);
    }
    return 0;
}

/* mcSetVolume(mc, bytes, rate, dims[3], scale[3], initialCacheSlots) */
/* Optional persisted block ranges (a typed array of 8 or 16 bytes per block):
 * *data = NULL when v is undefined or null.  Returns -1 for anything else
 * that is not a typed array. */
static int ranges_of(napi_env env, napi_value v, const uint32_t dims[3], void **data, size_t *len,
                     bcmc_ranges_format *fmt)
{
    napi_valuetype t;
    *data = NULL;
    *len = 0;
    *fmt = BCMC_RANGES_BLOCK;
    if (napi_typeof(env, v, &t) != napi_ok)
        return -1;
    if (t == napi_undefined || t == napi_null)
        return 0;
    if (bytes_of(env, v, data, len))
        return -1;
    const uint64_t nb = (uint64_t)((dims[0] + 3) / 4) * ((dims[1] + 3) / 4) * ((dims[2] + 3) / 4);
    *fmt = *len == 16 * nb ? BCMC_RANGES_BLOCK_TILE : BCMC_RANGES_BLOCK;
    return 0;
}

/* mcSetVolume(mc, bytes, rate, dims[3], scale[3], initialSlots, ranges?) */
static napi_value mc_set_volume(napi_env env, napi_callback_info info)
{
    GET_ARGS(7);
    EXT(0, bcmc_mc, mc);
    void *d, *r;
    size_t n, rn;
    bcmc_ranges_format fmt;
    double dd[3], sc[3];
    if (bytes_of(env, argv[1], &d, &n) || three(env, argv[3], dd) || three(env, argv[4], sc)) {
        napi_throw_type_error(env, NULL, "expected (mc, Uint8Array, rate, dims[3], scale[3], n, ranges?)");
        return NULL;
    }
    const uint32_t dims[3] = {(uint32_t)dd[0], (uint32_t)dd[1], (uint32_t)dd[2]};
    const float scale[3] = {(float)sc[0], (float)sc[1], (float)sc[2]};
    if (ranges_of(env, argv[6], dims, &r, &rn, &fmt)) {
        napi_throw_type_error(env, NULL, "block ranges must be a typed array");
        return NULL;
    }
    if (r)
        BCMC_CALL(env, bcmc_mc_set_compressed_volume_ranges(mc, d, n, num(env, argv[2]), dims, scale,
                                                            (uint32_t)num(env, argv[5]), r, fmt, rn));
    else
        BCMC_CALL(env, bcmc_mc_set_compressed_volume(mc, d, n, num(env, argv[2]), dims, scale,
                                                     (uint32_t)num(env, argv[5])));
    return NULL;
}

/* mcGetRanges(mc, withTile, out Float32Array) */
static napi_value mc_get_ranges(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_mc, mc);
    bool tile = false;
    void *d;
    size_t n;
    NAPI_CALL(env, napi_get_value_bool(env, argv[1], &tile));
    if (bytes_of(env, argv[2], &d, &n)) {
        napi_throw_type_error(env, NULL, "expected (mc, withTile, Float32Array)");
        return NULL;
    }
    BCMC_CALL(env, bcmc_mc_get_block_ranges(mc, tile ? BCMC_RANGES_BLOCK_TILE : BCMC_RANGES_BLOCK, d, n));
    return NULL;
}

static napi_value mc_set_format(napi_env env, napi_callback_info info)
{
    GET_ARGS(2);
    EXT(0, bcmc_mc, mc);
    BCMC_CALL(env, bcmc_mc_set_vertex_format(mc, (bcmc_vertex_format)(int)num(env, argv[1])));
    return NULL;
}

/* computeSurface: with `timed` the per-stage device times are recorded
 * (direct launches); otherwise the query chain runs in the context's launch
 * mode (ctxSetGraphReplay) and only the counts are filled in (from
 * bcmc_mc_get_info). */
static int mc_surface_call(bcmc_mc *mc, float iso, int timed, bcmc_perf *p, uint64_t *nv)
{
    memset(p, 0, sizeof(*p));
    int rc = bcmc_mc_compute_surface(mc, iso, timed ? p : NULL, nv);
    if (rc == BCMC_OK && !timed) {
        bcmc_mc_info mi;
        rc = bcmc_mc_get_info(mc, &mi);
        p->numActiveBlocks = mi.num_active_blocks;
        p->numBlocksDecompressed = mi.new_decompressed;
        p->numBlocksWithVertices = mi.num_blocks_with_vertices;
        p->numVertices = mi.num_vertices;
    }
    return rc;
}

static napi_value mc_surface_result(napi_env env, const bcmc_perf *p, uint64_t nv)
{
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "numVertices", (double)nv);
    set_num(env, o, "computeActiveBlocks", p->computeActiveBlocks);
    set_num(env, o, "numActiveBlocks", p->numActiveBlocks);
    set_num(env, o, "cacheUpdate", p->cacheUpdate);
    set_num(env, o, "numBlocksDecompressed", p->numBlocksDecompressed);
    set_num(env, o, "decompression", p->decompression);
    set_num(env, o, "compactActiveIDs", p->compactActiveIDs);
    set_num(env, o, "computeBlockHasVertices", p->computeBlockHasVertices);
    set_num(env, o, "numBlocksWithVertices", p->numBlocksWithVertices);
    set_num(env, o, "compactBlocksWithVertices", p->compactBlocksWithVertices);
    set_num(env, o, "computeVertices", p->computeVertices);
    set_num(env, o, "cacheSize", p->cacheSize);
    set_num(env, o, "totalMs", p->total_ms);
    return o;
}

/* mcComputeSurface(mc, iso, timed) */
static napi_value mc_compute_surface(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_mc, mc);
    bool timed = true;
    if (argc > 2)
        napi_get_value_bool(env, argv[2], &timed);
    bcmc_perf p;
    uint64_t nv = 0;
    BCMC_CALL(env, mc_surface_call(mc, (float)num(env, argv[1]), timed, &p, &nv));
    return mc_surface_result(env, &p, nv);
}

static napi_value mc_info(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_mc, mc);
    bcmc_mc_info mi;
    BCMC_CALL(env, bcmc_mc_get_info(mc, &mi));
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "totalBlocks", mi.total_blocks);
    set_num(env, o, "numActiveBlocks", mi.num_active_blocks);
    set_num(env, o, "numBlocksWithVertices", mi.num_blocks_with_vertices);
    set_num(env, o, "newDecompressed", mi.new_decompressed);
    set_num(env, o, "numVertices", (double)mi.num_vertices);
    set_num(env, o, "vertexCapacity", (double)mi.vertex_capacity);
    set_num(env, o, "maxBits", mi.max_bits);
    set_num(env, o, "format", mi.format);
    return o;
}

static napi_value mc_vertex_buffer(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_mc, mc);
    return mk_ext(env, bcmc_mc_vertex_buffer(mc));
}

static napi_value mc_lru(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_mc, mc);
    return mk_ext(env, bcmc_mc_lru(mc));
}

static napi_value mc_array(napi_env env, napi_callback_info info)
{
    GET_ARGS(2);
    EXT(0, bcmc_mc, mc);
    return mk_ext(env, bcmc_mc_array_buffer(mc, (bcmc_mc_array)(int)num(env, argv[1])));
}

static napi_value mc_memory_use(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_mc, mc);
    uint64_t t = 0, c = 0;
    BCMC_CALL(env, bcmc_mc_memory_use(mc, &t, &c));
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "total", (double)t);
    set_num(env, o, "cache", (double)c);
    return o;
}

/* ---- batched / served engine ---------------------------------------------- */
static napi_value eng_create(napi_env env, napi_callback_info info)
{
    GET_ARGS(2);
    EXT(0, bcmc_ctx, ctx);
    bcmc_engine *e = NULL;
    BCMC_CALL(env, bcmc_engine_create(ctx, (uint64_t)num(env, argv[1]), &e));
    return mk_ext(env, e);
}

static napi_value eng_destroy(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_engine, eng);
    BCMC_CALL(env, bcmc_engine_destroy(eng));
    return NULL;
}

/* engAddVolume(eng, bytes, rate, dims[3], ranges?) -> volume id */
static napi_value eng_add_volume(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_engine, eng);
    void *d, *r;
    size_t n, rn;
    bcmc_ranges_format fmt;
    double dd[3];
    if (bytes_of(env, argv[1], &d, &n) || three(env, argv[3], dd)) {
        napi_throw_type_error(env, NULL, "expected (engine, Uint8Array, rate, dims[3], ranges?)");
        return NULL;
    }
    const uint32_t dims[3] = {(uint32_t)dd[0], (uint32_t)dd[1], (uint32_t)dd[2]};
    if (ranges_of(env, argv[4], dims, &r, &rn, &fmt)) {
        napi_throw_type_error(env, NULL, "block ranges must be a typed array");
        return NULL;
    }
    uint32_t id = 0;
    if (r)
        BCMC_CALL(env, bcmc_engine_add_volume_ranges(eng, d, n, num(env, argv[2]), dims, r, fmt, rn, &id));
    else
        BCMC_CALL(env, bcmc_engine_add_volume(eng, d, n, num(env, argv[2]), dims, &id));
    return mk_num(env, id);
}

static napi_value eng_run_result(napi_env env, const bcmc_query_result *r, uint32_t nq)
{
    napi_value arr;
    NAPI_CALL(env, napi_create_array_with_length(env, nq, &arr));
    for (uint32_t i = 0; i < nq; ++i) {
        napi_value o;
        napi_create_object(env, &o);
        set_num(env, o, "numVertices", (double)r[i].num_vertices);
        set_num(env, o, "vertexOffset", (double)r[i].vertex_offset);
        set_num(env, o, "numActiveBlocks", r[i].num_active_blocks);
        set_num(env, o, "stepDecodedBlocks", r[i].step_decoded_blocks);
        napi_set_element(env, arr, i, o);
    }
    return arr;
}

/* (engine, Uint32Array, Float32Array) of one length -> ids, isos, nq */
static int eng_run_args(napi_env env, napi_value *argv, void **ids, void **isos, uint32_t *nq)
{
    size_t ni, nf;
    if (bytes_of(env, argv[1], ids, &ni) || bytes_of(env, argv[2], isos, &nf) || ni / 4 != nf / 4) {
        napi_throw_type_error(env, NULL, "expected (engine, Uint32Array, Float32Array) of one length");
        return -1;
    }
    *nq = (uint32_t)(ni / 4);
    return 0;
}

/* engRun(eng, Uint32Array volumeIds, Float32Array isovalues) -> [{...}] */
static napi_value eng_run(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_engine, eng);
    void *ids, *isos;
    uint32_t nq;
    if (eng_run_args(env, argv, &ids, &isos, &nq))
        return NULL;
    bcmc_query_result *r = (bcmc_query_result *)calloc(nq ? nq : 1, sizeof(*r));
    if (!r) {
        napi_throw_error(env, NULL, "out of host memory");
        return NULL;
    }
    char msg[640];
    pthread_mutex_lock(&g_lib_mu);
    const int rc = bcmc_engine_run(eng, (const uint32_t *)ids, (const float *)isos, nq, r);
    if (rc != BCMC_OK)
        snprintf(msg, sizeof(msg), "bcmc error %d: %s", rc, bcmc_last_error());
    pthread_mutex_unlock(&g_lib_mu);
    if (rc != BCMC_OK) {
        free(r);
        napi_throw_error(env, NULL, msg);
        return NULL;
    }
    napi_value out = eng_run_result(env, r, nq);
    free(r);
    return out;
}

static napi_value eng_vertex_buffer(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_engine, eng);
    return mk_ext(env, bcmc_engine_vertex_buffer(eng));
}

static napi_value eng_set_format(napi_env env, napi_callback_info info)
{
    GET_ARGS(2);
    EXT(0, bcmc_engine, eng);
    BCMC_CALL(env, bcmc_engine_set_vertex_format(eng, (bcmc_vertex_format)(int)num(env, argv[1])));
    return NULL;
}

static napi_value eng_stats(napi_env env, napi_callback_info info)
{
    GET_ARGS(1);
    EXT(0, bcmc_engine, eng);
    bcmc_engine_info s;
    BCMC_CALL(env, bcmc_engine_stats(eng, &s));
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "numVolumes", s.num_volumes);
    set_num(env, o, "poolSlots", s.pool_slots);
    set_num(env, o, "poolUsed", s.pool_used);
    set_num(env, o, "epoch", s.epoch);
    set_num(env, o, "poolBytes", (double)s.pool_bytes);
    set_num(env, o, "deviceBytes", (double)s.device_bytes);
    set_num(env, o, "steps", s.steps);
    set_num(env, o, "graphSteps", s.graph_steps);
    return o;
}

static napi_value eng_serve_result(napi_env env, const bcmc_serve_result *r, uint32_t n)
{
    napi_value out;
    NAPI_CALL(env, napi_create_array_with_length(env, n, &out));
    for (uint32_t i = 0; i < n; ++i) {
        napi_value o;
        napi_create_object(env, &o);
        set_num(env, o, "arrival", r[i].arrival_s);
        set_num(env, o, "start", r[i].start_s);
        set_num(env, o, "finish", r[i].finish_s);
        set_num(env, o, "numVertices", (double)r[i].num_vertices);
        set_num(env, o, "batchSize", r[i].batch_size);
        set_num(env, o, "step", r[i].step);
        napi_set_element(env, out, i, o);
    }
    return out;
}

/* (engine, Float64Array, Uint32Array, Float32Array, maxBatch, budget[, noCoalesce]) */
static int eng_serve_args(napi_env env, napi_value *argv, void **arr, void **ids, void **isos, uint32_t *n,
                          bcmc_serve_opts *opts)
{
    size_t na, ni, nf;
    if (bytes_of(env, argv[1], arr, &na) || bytes_of(env, argv[2], ids, &ni) ||
        bytes_of(env, argv[3], isos, &nf) || na / 8 != ni / 4 || ni / 4 != nf / 4) {
        napi_throw_type_error(env, NULL,
                              "expected (engine, Float64Array, Uint32Array, Float32Array) of one length");
        return -1;
    }
    *n = (uint32_t)(ni / 4);
    opts->max_batch = (uint32_t)num(env, argv[4]);
    opts->vertex_budget_bytes = (uint64_t)num(env, argv[5]);
    opts->no_coalesce = num(env, argv[6]) != 0 ? 1u : 0u; /* undefined: 0 */
    return 0;
}

/* engServe(eng, Float64Array arrivals, Uint32Array ids, Float32Array isos,
 * maxBatch, vertexBudgetBytes[, noCoalesce]) -> [{arrival, start, finish,
 * numVertices, batchSize, step}] (seconds from the call) */
static napi_value eng_serve(napi_env env, napi_callback_info info)
{
    GET_ARGS(7);
    EXT(0, bcmc_engine, eng);
    void *arr, *ids, *isos;
    uint32_t n;
    bcmc_serve_opts opts;
    if (eng_serve_args(env, argv, &arr, &ids, &isos, &n, &opts))
        return NULL;
    bcmc_serve_result *r = (bcmc_serve_result *)calloc(n ? n : 1, sizeof(*r));
    if (!r) {
        napi_throw_error(env, NULL, "out of host memory");
        return NULL;
    }
    char msg[640];
    pthread_mutex_lock(&g_lib_mu);
    const int rc = bcmc_serve_trace(eng, (const double *)arr, (const uint32_t *)ids, (const float *)isos, n,
                                    &opts, r);
    if (rc != BCMC_OK)
        snprintf(msg, sizeof(msg), "bcmc error %d: %s", rc, bcmc_last_error());
    pthread_mutex_unlock(&g_lib_mu);
    if (rc != BCMC_OK) {
        free(r);
        napi_throw_error(env, NULL, msg);
        return NULL;
    }
    napi_value out = eng_serve_result(env, r, n);
    free(r);
    return out;
}

/* ---- asynchronous forms ----------------------------------------------------- */
enum { J_SCAN, J_COMPACT_IDS, J_COMPACT_FLAGS, J_SORT, J_LRU_UPDATE, J_MC, J_ENG_RUN, J_ENG_SERVE };

typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    napi_ref keep[4];
    int nkeep;
    int kind;
    int rc;
    char err[640];
    /* arguments */
    void *h;
    void *p[4];
    uint32_t u[3];
    int flag;
    float f;
    bcmc_serve_opts opts;
    /* results */
    uint32_t count;
    uint64_t nv;
    bcmc_buf *buf;
    void *res;
    bcmc_perf perf;
    bcmc_lru_perf lperf;
} job;

static void job_exec(napi_env env, void *data)
{
    (void)env;
    job *j = (job *)data;
    pthread_mutex_lock(&g_lib_mu);
    switch (j->kind) {
    case J_SCAN:
        j->rc = bcmc_scan_exclusive_u32((bcmc_ctx *)j->h, (bcmc_buf *)j->p[0], j->u[0], j->u[1], &j->count);
        break;
    case J_COMPACT_IDS:
        j->rc = bcmc_compact_ids((bcmc_ctx *)j->h, j->u[0], (bcmc_buf *)j->p[0], (bcmc_buf *)j->p[1],
                                 (bcmc_buf *)j->p[2]);
        break;
    case J_COMPACT_FLAGS:
        j->rc = bcmc_compact_flags((bcmc_ctx *)j->h, j->u[0], (bcmc_buf *)j->p[0], (bcmc_buf *)j->p[1],
                                   &j->count);
        break;
    case J_SORT:
        j->rc = bcmc_sort_u32_by_key((bcmc_ctx *)j->h, (bcmc_buf *)j->p[0], (bcmc_buf *)j->p[1], j->u[0],
                                     j->flag);
        break;
    case J_LRU_UPDATE:
        j->rc = bcmc_lru_update_timed((bcmc_lru *)j->h, (bcmc_buf *)j->p[0], &j->count, &j->buf,
                                      j->flag ? &j->lperf : NULL);
        break;
    case J_MC:
        j->rc = mc_surface_call((bcmc_mc *)j->h, j->f, j->flag, &j->perf, &j->nv);
        break;
    case J_ENG_RUN:
        j->rc = bcmc_engine_run((bcmc_engine *)j->h, (const uint32_t *)j->p[0], (const float *)j->p[1], j->u[0],
                                (bcmc_query_result *)j->res);
        break;
    case J_ENG_SERVE:
        j->rc = bcmc_serve_trace((bcmc_engine *)j->h, (const double *)j->p[0], (const uint32_t *)j->p[1],
                                 (const float *)j->p[2], j->u[0], &j->opts, (bcmc_serve_result *)j->res);
        break;
    default:
        j->rc = BCMC_ERR_INVALID;
    }
    if (j->rc != BCMC_OK)
        snprintf(j->err, sizeof(j->err), "bcmc error %d: %s", j->rc, bcmc_last_error());
    pthread_mutex_unlock(&g_lib_mu);
}

static void job_done(napi_env env, napi_status status, void *data)
{
    job *j = (job *)data;
    if (status != napi_ok && j->rc == BCMC_OK) {
        j->rc = BCMC_ERR_STATE;
        snprintf(j->err, sizeof(j->err), "async work did not run (status %d)", (int)status);
    }
    if (j->rc != BCMC_OK) {
        napi_value msg, e;
        napi_create_string_utf8(env, j->err, NAPI_AUTO_LENGTH, &msg);
        napi_create_error(env, NULL, msg, &e);
        napi_reject_deferred(env, j->deferred, e);
    } else {
        napi_value v = NULL;
        switch (j->kind) {
        case J_SCAN: case J_COMPACT_FLAGS: v = mk_num(env, j->count); break;
        case J_LRU_UPDATE: v = lru_update_result(env, j->count, j->buf, j->flag ? &j->lperf : NULL); break;
        case J_MC: v = mc_surface_result(env, &j->perf, j->nv); break;
        case J_ENG_RUN: v = eng_run_result(env, (const bcmc_query_result *)j->res, j->u[0]); break;
        case J_ENG_SERVE: v = eng_serve_result(env, (const bcmc_serve_result *)j->res, j->u[0]); break;
        default: break;
        }
        if (!v)
            napi_get_undefined(env, &v);
        napi_resolve_deferred(env, j->deferred, v);
    }
    for (int i = 0; i < j->nkeep; ++i)
        napi_delete_reference(env, j->keep[i]);
    napi_delete_async_work(env, j->work);
    free(j->res);
    free(j);
}

/* Queues j; `keep` are the JS values whose memory the work reads. */
static napi_value job_start(napi_env env, job *j, napi_value *keep, int nkeep)
{
    napi_value promise, name;
    if (napi_create_promise(env, &j->deferred, &promise) != napi_ok) {
        free(j->res);
        free(j);
        napi_throw_error(env, NULL, "napi_create_promise failed");
        return NULL;
    }
    for (int i = 0; i < nkeep; ++i)
        napi_create_reference(env, keep[i], 1, &j->keep[j->nkeep++]);
    napi_create_string_utf8(env, "bcmc", NAPI_AUTO_LENGTH, &name);
    napi_create_async_work(env, NULL, name, job_exec, job_done, j, &j->work);
    napi_queue_async_work(env, j->work);
    return promise;
}

static job *job_new(napi_env env, int kind)
{
    job *j = (job *)calloc(1, sizeof(job));
    if (!j)
        napi_throw_error(env, NULL, "out of host memory");
    else
        j->kind = kind;
    return j;
}

/* scanExclusiveAsync(ctx, buf, alignedSize, dataSize) -> Promise<total> */
static napi_value scan_exclusive_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(4);
    EXT(0, bcmc_ctx, ctx);
    EXT(1, bcmc_buf, buf);
    job *j = job_new(env, J_SCAN);
    if (!j)
        return NULL;
    j->h = ctx;
    j->p[0] = buf;
    j->u[0] = (uint32_t)num(env, argv[2]);
    j->u[1] = (uint32_t)num(env, argv[3]);
    return job_start(env, j, NULL, 0);
}

/* compactIdsAsync(ctx, n, flags, offsets, out) -> Promise */
static napi_value compact_ids_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_ctx, ctx);
    EXT(2, bcmc_buf, flags);
    EXT(3, bcmc_buf, offsets);
    EXT(4, bcmc_buf, out);
    job *j = job_new(env, J_COMPACT_IDS);
    if (!j)
        return NULL;
    j->h = ctx;
    j->u[0] = (uint32_t)num(env, argv[1]);
    j->p[0] = flags;
    j->p[1] = offsets;
    j->p[2] = out;
    return job_start(env, j, NULL, 0);
}

/* compactFlagsAsync(ctx, n, flags, out) -> Promise<count> */
static napi_value compact_flags_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(4);
    EXT(0, bcmc_ctx, ctx);
    EXT(2, bcmc_buf, flags);
    EXT(3, bcmc_buf, out);
    job *j = job_new(env, J_COMPACT_FLAGS);
    if (!j)
        return NULL;
    j->h = ctx;
    j->u[0] = (uint32_t)num(env, argv[1]);
    j->p[0] = flags;
    j->p[1] = out;
    return job_start(env, j, NULL, 0);
}

/* sortByKeyAsync(ctx, keys, vals, n, reverse) -> Promise */
static napi_value sort_by_key_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_ctx, ctx);
    EXT(1, bcmc_buf, keys);
    EXT(2, bcmc_buf, vals);
    bool rev = false;
    napi_get_value_bool(env, argv[4], &rev);
    job *j = job_new(env, J_SORT);
    if (!j)
        return NULL;
    j->h = ctx;
    j->p[0] = keys;
    j->p[1] = vals;
    j->u[0] = (uint32_t)num(env, argv[3]);
    j->flag = rev ? 1 : 0;
    return job_start(env, j, NULL, 0);
}

/* lruUpdateAsync(lru, needed, timed) -> Promise<[n, ids, perf]> */
static napi_value lru_update_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_lru, lru);
    EXT(1, bcmc_buf, needed);
    bool timed = false;
    if (argc > 2)
        napi_get_value_bool(env, argv[2], &timed);
    job *j = job_new(env, J_LRU_UPDATE);
    if (!j)
        return NULL;
    j->h = lru;
    j->p[0] = needed;
    j->flag = timed ? 1 : 0;
    return job_start(env, j, NULL, 0);
}

/* mcComputeSurfaceAsync(mc, iso, timed) -> Promise<{...}> */
static napi_value mc_compute_surface_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_mc, mc);
    bool timed = true;
    if (argc > 2)
        napi_get_value_bool(env, argv[2], &timed);
    job *j = job_new(env, J_MC);
    if (!j)
        return NULL;
    j->h = mc;
    j->f = (float)num(env, argv[1]);
    j->flag = timed ? 1 : 0;
    return job_start(env, j, NULL, 0);
}

/* engRunAsync(eng, Uint32Array, Float32Array) -> Promise<[{...}]> */
static napi_value eng_run_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(3);
    EXT(0, bcmc_engine, eng);
    void *ids, *isos;
    uint32_t nq;
    if (eng_run_args(env, argv, &ids, &isos, &nq))
        return NULL;
    job *j = job_new(env, J_ENG_RUN);
    if (!j)
        return NULL;
    j->res = calloc(nq ? nq : 1, sizeof(bcmc_query_result));
    if (!j->res) {
        free(j);
        napi_throw_error(env, NULL, "out of host memory");
        return NULL;
    }
    j->h = eng;
    j->p[0] = ids;
    j->p[1] = isos;
    j->u[0] = nq;
    return job_start(env, j, argv + 1, 2);
}

/* engServeAsync(eng, Float64Array, Uint32Array, Float32Array, maxBatch,
 * vertexBudgetBytes[, noCoalesce]) -> Promise<[{...}]> */
static napi_value eng_serve_async(napi_env env, napi_callback_info info)
{
    GET_ARGS(7);
    EXT(0, bcmc_engine, eng);
    void *arr, *ids, *isos;
    uint32_t n;
    bcmc_serve_opts opts;
    if (eng_serve_args(env, argv, &arr, &ids, &isos, &n, &opts))
        return NULL;
    job *j = job_new(env, J_ENG_SERVE);
    if (!j)
        return NULL;
    j->res = calloc(n ? n : 1, sizeof(bcmc_serve_result));
    if (!j->res) {
        free(j);
        napi_throw_error(env, NULL, "out of host memory");
        return NULL;
    }
    j->h = eng;
    j->p[0] = arr;
    j->p[1] = ids;
    j->p[2] = isos;
    j->u[0] = n;
    j->opts = opts;
    return job_start(env, j, argv + 1, 3);
}

/* zfpEncode(ctx, fieldBuf, dims[3], maxBits, streamBuf) */
static napi_value zfp_encode(napi_env env, napi_callback_info info)
{
    GET_ARGS(5);
    EXT(0, bcmc_ctx, ctx);
    EXT(1, bcmc_buf, field);
    EXT(4, bcmc_buf, stream);
    double dd[3];
    if (three(env, argv[2], dd)) {
        napi_throw_type_error(env, NULL, "expected dims[3]");
        return NULL;
    }
    const uint32_t dims[3] = {(uint32_t)dd[0], (uint32_t)dd[1], (uint32_t)dd[2]};
    BCMC_CALL(env, bcmc_zfp_encode(ctx, field, dims, (uint32_t)num(env, argv[3]), stream));
    return NULL;
}

/* verticesWorld(ctx, vertsBuf, first, n, dims[3], scale[3], outBuf) */
static napi_value vertices_world(napi_env env, napi_callback_info info)
{
    GET_ARGS(7);
    EXT(0, bcmc_ctx, ctx);
    EXT(1, bcmc_buf, verts);
    EXT(6, bcmc_buf, out);
    double dd[3], sc[3];
    if (three(env, argv[4], dd) || three(env, argv[5], sc)) {
        napi_throw_type_error(env, NULL, "expected dims[3] and scale[3]");
        return NULL;
    }
    const uint32_t dims[3] = {(uint32_t)dd[0], (uint32_t)dd[1], (uint32_t)dd[2]};
    const float scale[3] = {(float)sc[0], (float)sc[1], (float)sc[2]};
    BCMC_CALL(env, bcmc_vertices_world(ctx, verts, (uint64_t)num(env, argv[2]), (uint64_t)num(env, argv[3]),
                                       dims, scale, out));
    return NULL;
}

static napi_value version(napi_env env, napi_callback_info info)
{
    napi_value s;
    NAPI_CALL(env, napi_create_string_utf8(env, bcmc_version(), NAPI_AUTO_LENGTH, &s));
    return s;
}

#define FN(name, fn) {name, NULL, fn, NULL, NULL, NULL, napi_default, NULL}

static napi_value init(napi_env env, napi_value exports)
{
    napi_property_descriptor d[] = {
        FN("version", version),
        FN("ctxCreate", ctx_create), FN("ctxDestroy", ctx_destroy), FN("ctxSync", ctx_sync),
        FN("ctxInfo", ctx_info), FN("ctxAutotune", ctx_autotune),
        FN("ctxSetGraphReplay", ctx_set_graph_replay),
        FN("bufAlloc", buf_alloc), FN("bufFree", buf_free), FN("bufSize", buf_size),
        FN("bufUpload", buf_upload), FN("bufDownload", buf_download),
        FN("scanAlignedSize", scan_aligned_size), FN("scanExclusive", scan_exclusive),
        FN("compactIds", compact_ids), FN("compactFlags", compact_flags),
        FN("sortAlignedSize", sort_aligned_size), FN("sortByKey", sort_by_key),
        FN("lruCreate", lru_create), FN("lruDestroy", lru_destroy), FN("lruUpdate", lru_update),
        FN("lruReset", lru_reset), FN("lruInfo", lru_info), FN("lruState", lru_state),
        FN("lruCacheBuffer", lru_cache_buffer),
        FN("mcCreate", mc_create), FN("mcDestroy", mc_destroy), FN("mcGetRanges", mc_get_ranges),
        FN("mcSetVolume", mc_set_volume),
        FN("mcSetFormat", mc_set_format), FN("mcComputeSurface", mc_compute_surface),
        FN("mcInfo", mc_info), FN("mcVertexBuffer", mc_vertex_buffer), FN("mcLru", mc_lru),
        FN("mcArray", mc_array), FN("mcMemoryUse", mc_memory_use),
        FN("engCreate", eng_create), FN("engDestroy", eng_destroy), FN("engAddVolume", eng_add_volume),
        FN("engRun", eng_run), FN("engVertexBuffer", eng_vertex_buffer), FN("engSetFormat", eng_set_format),
        FN("engStats", eng_stats), FN("engServe", eng_serve),
        FN("zfpEncode", zfp_encode), FN("verticesWorld", vertices_world),
        FN("scanExclusiveAsync", scan_exclusive_async), FN("compactIdsAsync", compact_ids_async),
        FN("compactFlagsAsync", compact_flags_async), FN("sortByKeyAsync", sort_by_key_async),
        FN("lruUpdateAsync", lru_update_async), FN("mcComputeSurfaceAsync", mc_compute_surface_async),
        FN("engRunAsync", eng_run_async), FN("engServeAsync", eng_serve_async),
    };
    napi_define_properties(env, exports, sizeof(d) / sizeof(d[0]), d);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
