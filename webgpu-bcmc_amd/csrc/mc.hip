// CompressedMarchingCubes on gfx950 (js/compressed_marching_cubes.js).
//
// One query (computeSurface, :580-809) is a fixed chain of kernels on the
// context's stream with a single host read-back at the end:
//   active     block active test (compute_block_active.comp) against a
//              per-volume dilated range, in a single-pass scan + compaction
//              of the flags -> active ids and LRU misses;
//   lru        LRUCache.update over the active list (lru.hip);
//   decode     zfp decode of the misses into their cache slots (zfp.hip);
//   mc         count: per block a 5^3 tile with the ghost layer
//              (compute_vertex_values.comp:96-252) and the per-cell triangle
//              counts (compute_block_has_vertices.comp,
//              compute_block_voxel_num_verts.comp); prefix: a two-channel scan
//              of the chunk totals giving the list of blocks with vertices and
//              their 64-bit vertex offsets; emit: the vertex writes
//              (compute_block_vertices.comp:76-174).
// The vertex buffer grows like the reference (x1.5, :774-785); when a query
// needs more room than it has, the emit launch runs again once it has grown.
#include "common.h"
#include "kernels.h"
#include "lru.h"
#include "mc_device.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace bcmc {

// ---- active blocks ---------------------------------------------------------
// One single-pass scan over the N blocks with two channels: the active flags
// (one comparison against the dilated range, see block_active) compacted to
// the ascending active ids, and the LRU misses among them (active and not
// cached, lru_cache_mark_new_items.comp:28-43) compacted to the ascending new
// item ids (lru_cache.js:463-492).
struct ActiveOp {
    const float2 *dilated;
    const int32_t *item_slot;
    uint32_t *flags, *active_ids, *new_items;
    QState *q;
    const float *iso_src;  // the query's isovalue, in host-coherent pinned memory
    uint32_t n;
    BCMC_DEV uint32_t size() const { return n; }
    BCMC_DEV void begin() const {
        q->iso = *iso_src;  // for the later launches of the query
        q->num_with = 0;
        q->num_verts = 0;
        q->overflow = 0;
        q->cache_short = 0;
    }
    struct Tile {
        float iso;
    };
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {*iso_src}; }
    BCMC_DEV void value(const Tile &tc, uint32_t b, uint32_t (&v)[2], Item &) const {
        const float2 d = dilated[b];
        const int32_t s = item_slot[b];
        const bool act = block_active(d, tc.iso);
        v[0] = act ? 1u : 0u;
        v[1] = (uint32_t)act & (uint32_t)(s < 0);
    }
    BCMC_DEV void emit(const Tile &, uint32_t b, const uint32_t (&v)[2], const Item &,
                       const unsigned long long (&e)[2]) const {
        flags[b] = v[0];
        if (v[0])
            active_ids[e[0]] = b;
        if (v[1])
            new_items[e[1]] = b;
    }
    BCMC_DEV void total(const unsigned long long (&t)[2]) const {
        q->num_active = (uint32_t)t[0];
        q->num_new = (uint32_t)t[1];
    }
};

// ---- counts, offsets and vertices -------------------------------------------
// Three launches over chunks of 256 consecutive active blocks.  Count: each
// wave counts the vertices of 4 blocks at a time (one per 16-lane row); every
// block's count and every chunk's two totals (blocks with vertices, vertices)
// go to memory.  Prefix: a two-channel scan of the chunk totals.  Emit: each
// chunk's list of blocks with vertices (its slice of the reference's
// compacted list) and their 64-bit vertex offsets, then each wave takes 4
// consecutive blocks of that list, reloads them (mostly from L2) and writes
// their vertices as one contiguous run.  Vertices past the buffer capacity are
// not written; the host grows the buffer and runs the emit launch again when
// the total exceeds the capacity.  No chunk waits for another.
struct McArgs {
    const float2 *tile;  // per block: tile range (launch_tile_ranges)
    const float *cache;
    const int32_t *item_slot;
    const uint32_t *flags;
    uint32_t nb[3];
    FastDiv div_x, div_y;
    uint32_t cache_cap;
    const uint32_t *active_ids;
    QState *q;
    uint32_t *counts;         // per active block: vertices
    uint32_t *chunk_with;     // per chunk: blocks with vertices
    uint32_t *chunk_verts;    // per chunk: vertices
    uint32_t *chunk_with_pre;           // exclusive prefixes of the chunk totals
    unsigned long long *chunk_verts_pre;
    uint32_t *with_verts;
    unsigned long long *vert_offsets;
    void *verts;
    unsigned long long capacity;
    int fmt;
    QState *q_out;  // host-coherent pinned copy of *q, written by the emit launch
};

BCMC_DEV RowBlock mc_row_base(const McArgs &a) {
    RowBlock b;
    b.cache = a.cache;
    b.item_slot = a.item_slot;
    b.flags = a.flags;
    b.nb[0] = a.nb[0];
    b.nb[1] = a.nb[1];
    b.nb[2] = a.nb[2];
    b.div_x = a.div_x;
    b.div_y = a.div_y;
    b.cap = a.cache_cap;
    b.tag = 1;
    b.iso = a.q->iso;
    b.bid = 0;
    b.live = false;
    return b;
}

__global__ __launch_bounds__(kMcThreads) void mc_count_kernel(McArgs a) {
    constexpr int NW = kMcThreads / 64;
    __shared__ CountShared cs;
    __shared__ uint32_t lds_w[NW + 1];
    __shared__ CountTables tab;
    load_count_tables(&tab);
    const uint32_t n = a.q->num_active;
    const uint32_t nchunks = div_up(n, kChunk);
    const uint32_t wave = threadIdx.x / 64;
    const RowBlock base = mc_row_base(a);
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint32_t k0 = chunk * kChunk;
        const bool in = k0 + threadIdx.x < n;
        const uint32_t bid = in ? a.active_ids[k0 + threadIdx.x] : 0u;
        const uint32_t ncand = chunk_prologue(cs, in, bid, 0u, in ? a.tile[bid] : make_float2(1.0f, 0.0f),
                                              base.iso, lds_w);
        count_sweep(cs, ncand, wave, tab, [&](uint32_t e, RowBlock &b) {
            b = base;
            b.live = k0 + e < n;
            b.bid = cs.bid[e];
        });
        __syncthreads();
        const uint32_t c = cs.cnt[threadIdx.x];
        if (in)
            a.counts[k0 + threadIdx.x] = c;
        uint32_t agg_h, agg_v;
        block_excl_scan<kMcThreads>(c ? 1u : 0u, lds_w, &agg_h);
        block_excl_scan<kMcThreads>(c, lds_w, &agg_v);
        if (threadIdx.x == 0) {
            a.chunk_with[chunk] = agg_h;
            a.chunk_verts[chunk] = agg_v;
        }
    }
}

// Exclusive prefixes of the chunk totals; the grand totals are the query's
// numBlocksWithVertices and numVertices.
struct ChunkPrefixOp {
    const uint32_t *with, *verts;
    uint32_t *with_pre;
    unsigned long long *verts_pre;
    QState *q;
    BCMC_DEV uint32_t size() const { return div_up(q->num_active, kChunk); }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[2], Item &) const {
        v[0] = with[i];
        v[1] = verts[i];
    }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&)[2], const Item &,
                       const unsigned long long (&e)[2]) const {
        with_pre[i] = (uint32_t)e[0];
        verts_pre[i] = e[1];
    }
    BCMC_DEV void total(const unsigned long long (&t)[2]) const {
        q->num_with = (uint32_t)t[0];
        q->num_verts = t[1];
    }
};

template <int FMT>
__global__ __launch_bounds__(kMcThreads) __attribute__((amdgpu_waves_per_eu(5))) void mc_emit_kernel(McArgs a) {
    constexpr int NW = kMcThreads / 64;
    __shared__ EmitShared cs;
    __shared__ uint32_t lds_w[NW + 1];
    __shared__ McTables tab;
    __shared__ EmitScratch es[NW];
    // the query's counters into host-coherent pinned memory for the host's
    // one read-back: every field is final before this launch (the emit
    // writes none), so the first workgroup copies them at once
    static_assert(sizeof(QState) % 4 == 0 && sizeof(QState) <= kMcThreads * 4, "one word per thread");
    if (blockIdx.x == 0 && threadIdx.x < sizeof(QState) / 4)
        reinterpret_cast<uint32_t *>(a.q_out)[threadIdx.x] = reinterpret_cast<const uint32_t *>(a.q)[threadIdx.x];
    load_mc_tables(&tab);
    const uint32_t n = a.q->num_active;
    const uint32_t nchunks = div_up(n, kChunk);
    const uint32_t wave = threadIdx.x / 64;
    const RowBlock base = mc_row_base(a);
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint32_t k0 = chunk * kChunk;
        const bool in = k0 + threadIdx.x < n;
        const uint32_t c = in ? a.counts[k0 + threadIdx.x] : 0u;
        const uint32_t bid = in ? a.active_ids[k0 + threadIdx.x] : 0u;
        uint32_t ex_h, ex_v, agg_v;
        const uint32_t num_with = emit_prologue(cs, c, bid, 0u, lds_w, &ex_h, &ex_v, &agg_v);
        if (num_with == 0)
            continue;
        const uint32_t ph = a.chunk_with_pre[chunk];
        const unsigned long long pv = a.chunk_verts_pre[chunk];
        if (c) {
            a.with_verts[ph + ex_h] = k0 + threadIdx.x;
            a.vert_offsets[ph + ex_h] = pv + ex_v;
        }
        auto setup = [&](uint32_t e, RowBlock &b) {
            b = base;
            b.live = k0 + e < n;
            b.bid = cs.bid[e];
        };
        emit_sweep<FMT>(cs, wave, num_with, pv, tab, es[wave], a.capacity, a.verts, setup);
        __syncthreads();
    }
}

static uint32_t vertex_bytes(int fmt) {
    return fmt == BCMC_VERTS_LOCAL_F32 ? 16u : fmt == BCMC_VERTS_GLOBAL_F32 ? 12u : 8u;
}

}  // namespace bcmc

using namespace bcmc;

// ---------------------------------------------------------------------------
// C ABI: CompressedMarchingCubes
// ---------------------------------------------------------------------------
struct bcmc_mc {
    bcmc_ctx *ctx = nullptr;
    bool has_volume = false;
    uint32_t dims[3] = {0, 0, 0}, pdims[3] = {0, 0, 0}, nb[3] = {0, 0, 0};
    uint32_t total = 0, maxbits = 0;
    float scale[3] = {1, 1, 1};
    uint64_t *stream = nullptr;
    uint64_t stream_bytes = 0, stream_words = 0;
    float2 *ranges = nullptr;
    uint32_t *flags = nullptr, *active_ids = nullptr, *with_verts = nullptr;
    float2 *dilated = nullptr;  // per block: range over its 3x3x3 neighbourhood
    float2 *tile = nullptr;     // per block: range of every voxel its cells can read
    unsigned long long *vert_offsets = nullptr;
    // marching-cubes pass: per active block vertex counts, per chunk totals
    // and their exclusive prefixes
    uint32_t *counts = nullptr, *chunk_with = nullptr, *chunk_verts = nullptr, *chunk_with_pre = nullptr;
    unsigned long long *chunk_verts_pre = nullptr;
    bcmc_lru lru;  // non-owning handle over the cache state held here
    QState *q = nullptr, *qh = nullptr;
    ScanSite site_active, site_c0, site_c1;
    void *verts = nullptr;
    uint64_t vert_cap = 0;
    int fmt = BCMC_VERTS_LOCAL_F32;
    bcmc_buf vert_view, views[6];
    hipEvent_t ev[7] = {};
    uint32_t last_active = 0, last_with = 0, last_new = 0;
    uint64_t last_verts = 0;
    // with the context's graph replay on, the query chain is captured once
    // into a HIP graph and replayed; it is captured again when anything it
    // bakes in (buffers, format, sort passes) changes.  use_graph turns off
    // when a capture fails.  The isovalue is read by the active scan straight from
    // host-coherent pinned memory, and the counters are written back into
    // such memory by the emit launch's first workgroup (no copy nodes).
    float *iso_h = nullptr;
    const float *iso_d = nullptr;  // device address of iso_h
    QState *qh_d = nullptr;        // device address of qh
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t graph_key[6] = {0, 0, 0, 0, 0, 0};
};

static void mc_drop_graph(bcmc_mc *m) {
    if (m->gexec)
        (void)hipGraphExecDestroy(m->gexec);
    if (m->graph)
        (void)hipGraphDestroy(m->graph);
    m->gexec = nullptr;
    m->graph = nullptr;
}

static void mc_free_volume(bcmc_mc *m) {
    if (m->ctx)
        (void)hipStreamSynchronize(m->ctx->stream);
    mc_drop_graph(m);
    dev_free(m->stream);
    dev_free(m->ranges);
    dev_free(m->flags);
    dev_free(m->active_ids);
    dev_free(m->dilated);
    dev_free(m->tile);
    dev_free(m->with_verts);
    dev_free(m->vert_offsets);
    dev_free(m->counts);
    dev_free(m->chunk_with);
    dev_free(m->chunk_verts);
    dev_free(m->chunk_with_pre);
    dev_free(m->chunk_verts_pre);
    dev_free(m->verts);
    scan_site_free(&m->site_active);
    scan_site_free(&m->site_c0);
    scan_site_free(&m->site_c1);
    lru_state_free(&m->lru.st);
    m->stream = nullptr;
    m->ranges = nullptr;
    m->flags = m->active_ids = m->with_verts = nullptr;
    m->dilated = nullptr;
    m->tile = nullptr;
    m->vert_offsets = nullptr;
    m->counts = m->chunk_with = m->chunk_verts = m->chunk_with_pre = nullptr;
    m->chunk_verts_pre = nullptr;
    m->verts = nullptr;
    m->vert_cap = 0;
    m->has_volume = false;
}

static void mc_refresh_views(bcmc_mc *m) {
    auto set = [&](bcmc_buf &b, void *p, uint64_t bytes) {
        b.ctx = m->ctx;
        b.ptr = p;
        b.bytes = bytes;
        b.owned = false;
    };
    set(m->vert_view, m->verts, m->last_verts * vertex_bytes(m->fmt));
    set(m->views[BCMC_ARRAY_RANGES], m->ranges, (uint64_t)m->total * 8);
    set(m->views[BCMC_ARRAY_ACTIVE_FLAGS], m->flags, (uint64_t)m->total * 4);
    set(m->views[BCMC_ARRAY_ACTIVE_IDS], m->active_ids, (uint64_t)m->last_active * 4);
    set(m->views[BCMC_ARRAY_NEW_ITEMS], m->lru.st.new_items, (uint64_t)m->last_new * 4);
    set(m->views[BCMC_ARRAY_WITH_VERTS], m->with_verts, (uint64_t)m->last_with * 4);
    set(m->views[BCMC_ARRAY_VERT_OFFSETS], m->vert_offsets, (uint64_t)m->last_with * 8);
    m->lru.cache_view.ctx = m->ctx;
    m->lru.cache_view.ptr = m->lru.st.data;
    m->lru.cache_view.bytes = (uint64_t)m->lru.st.cache_size * m->lru.st.elem_bytes;
    m->lru.cache_view.owned = false;
}

static McArgs mc_args(bcmc_mc *m) {
    McArgs a;
    a.tile = m->tile;
    a.cache = (const float *)m->lru.st.data;
    a.item_slot = m->lru.st.v.item_slot;
    a.flags = m->flags;
    a.nb[0] = m->nb[0];
    a.nb[1] = m->nb[1];
    a.nb[2] = m->nb[2];
    a.div_x = make_fast_div(m->nb[0]);
    a.div_y = make_fast_div(m->nb[1]);
    a.cache_cap = m->lru.st.slots;  // slot `slots` is the zero slot
    a.active_ids = m->active_ids;
    a.q = m->q;
    a.counts = m->counts;
    a.chunk_with = m->chunk_with;
    a.chunk_verts = m->chunk_verts;
    a.chunk_with_pre = m->chunk_with_pre;
    a.chunk_verts_pre = m->chunk_verts_pre;
    a.with_verts = m->with_verts;
    a.vert_offsets = m->vert_offsets;
    a.verts = m->verts;
    a.capacity = m->vert_cap;
    a.fmt = m->fmt;
    a.q_out = m->qh_d;
    return a;
}

static int mc_launch_emit(bcmc_mc *m) {
    static const uint32_t resident =
        resident_per_cu((const void *)mc_emit_kernel<BCMC_VERTS_LOCAL_F32>, kMcThreads);
    const uint32_t grid = std::max<uint32_t>(
        1, std::min<uint32_t>(div_up(m->total, kChunk),
                              persistent_grid(m->ctx, std::min(m->ctx->tune.mc, resident))));
    hipStream_t st = m->ctx->stream;
    if (m->fmt == BCMC_VERTS_LOCAL_F32)
        mc_emit_kernel<BCMC_VERTS_LOCAL_F32><<<grid, kMcThreads, 0, st>>>(mc_args(m));
    else if (m->fmt == BCMC_VERTS_GLOBAL_F32)
        mc_emit_kernel<BCMC_VERTS_GLOBAL_F32><<<grid, kMcThreads, 0, st>>>(mc_args(m));
    else
        mc_emit_kernel<BCMC_VERTS_PACKED_U32X2><<<grid, kMcThreads, 0, st>>>(mc_args(m));
    BCMC_HIP_CHECK(hipGetLastError());
    return BCMC_OK;
}

// count -> chunk prefixes -> emit; `ev` (nullable) receives an event after
// the count and after the prefix scan
static int mc_launch_pass(bcmc_mc *m, hipEvent_t *ev) {
    const McArgs a = mc_args(m);
    const uint32_t max_chunks = div_up(m->total, kChunk);
    static const uint32_t resident = resident_per_cu((const void *)mc_count_kernel, kMcThreads);
    const uint32_t grid = std::max<uint32_t>(
        1, std::min<uint32_t>(max_chunks, persistent_grid(m->ctx, std::min(m->ctx->tune.count, resident))));
    mc_count_kernel<<<grid, kMcThreads, 0, m->ctx->stream>>>(a);
    if (ev)
        BCMC_HIP_CHECK(hipEventRecord(ev[0], m->ctx->stream));
    launch_scan<2>(m->ctx, m->site_c0, m->site_c1,
                   ChunkPrefixOp{a.chunk_with, a.chunk_verts, a.chunk_with_pre, a.chunk_verts_pre, a.q},
                   max_chunks);
    if (ev)
        BCMC_HIP_CHECK(hipEventRecord(ev[1], m->ctx->stream));
    return mc_launch_emit(m);
}

// The query from the slot assignment on: assignment + decode of the misses
// into their slots, the MC pass (counts, offsets, vertices) and the read-back
// of the counters; `timed` records events 3..6.
static int mc_enqueue_tail(bcmc_mc *m, bool timed) {
    hipStream_t st = m->ctx->stream;
    LruState &L = m->lru.st;
    launch_decode_assign(m->ctx, m->stream, m->stream_words, m->maxbits, L.new_items, m->total, L.v.order,
                         L.v.age, L.v.avail, L.v.item, L.v.item_slot, L.total, (float *)L.data, L.slots,
                         m->q);
    if (timed)
        BCMC_HIP_CHECK(hipEventRecord(m->ev[3], st));
    BCMC_TRY(mc_launch_pass(m, timed ? &m->ev[4] : nullptr));
    if (timed)
        BCMC_HIP_CHECK(hipEventRecord(m->ev[6], st));
    return BCMC_OK;
}

// One query as a kernel chain on the context's stream (capturable when not
// timed): the isovalue copy, active scan, LRU update, assignment + decode, MC
// pass, and the read-back of the counters; `timed` records the perfTracker
// events between the stages (direct launches only: the elapsed time of events
// recorded by a graph is not available).
static int mc_enqueue_query(bcmc_mc *m, uint32_t npasses, bool timed) {
    bcmc_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    LruState &L = m->lru.st;
    if (timed)
        BCMC_HIP_CHECK(hipEventRecord(m->ev[0], st));
    // active blocks and the LRU misses among them, one scan
    ActiveOp aop{m->dilated, L.v.item_slot, m->flags, m->active_ids, L.new_items, m->q, m->iso_d, m->total};
    launch_scan<2>(ctx, m->site_active, L.site_miss, aop, m->total);
    if (timed)
        BCMC_HIP_CHECK(hipEventRecord(m->ev[1], st));
    // cache update: age/mark + available slots, growth, eviction order
    lru_enqueue_update_mc(ctx, &L, m->flags, m->q, npasses);
    if (timed)
        BCMC_HIP_CHECK(hipEventRecord(m->ev[2], st));
    return mc_enqueue_tail(m, timed);
}

extern "C" {

BCMC_API int bcmc_mc_create(bcmc_ctx *ctx, bcmc_mc **out) {
    BCMC_CHECK_ARG(ctx && out, "null argument");
    bcmc_mc *m = new bcmc_mc;
    m->ctx = ctx;
    if (dev_alloc((void **)&m->q, sizeof(QState)) != BCMC_OK ||
        hipHostMalloc(&m->qh, sizeof(QState), hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc(&m->iso_h, sizeof(float), hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&m->qh_d, m->qh, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&m->iso_d, m->iso_h, 0) != hipSuccess) {
        dev_free(m->q);
        if (m->qh)
            (void)hipHostFree(m->qh);
        if (m->iso_h)
            (void)hipHostFree(m->iso_h);
        delete m;
        set_error("allocation of the query counters failed");
        return BCMC_ERR_NOMEM;
    }
    for (auto &e : m->ev)
        BCMC_HIP_CHECK(hipEventCreate(&e));
    m->lru.ctx = ctx;
    m->lru.q = m->q;
    m->lru.qh = m->qh;
    m->lru.owns = false;
    *out = m;
    return BCMC_OK;
}

BCMC_API int bcmc_mc_destroy(bcmc_mc *m) {
    if (!m)
        return BCMC_OK;
    mc_free_volume(m);
    dev_free(m->q);
    (void)hipHostFree(m->qh);
    (void)hipHostFree(m->iso_h);
    for (auto &e : m->ev)
        (void)hipEventDestroy(e);
    delete m;
    return BCMC_OK;
}

// setCompressedVolume with the block ranges computed by the full-volume range
// pass (ranges == nullptr) or taken from the caller (persisted ranges).
static int mc_set_volume(bcmc_mc *m, const void *stream, uint64_t bytes, double rate, const uint32_t dims[3],
                         const float scale[3], uint32_t initial_cache_slots, const void *ranges,
                         bcmc_ranges_format fmt, uint64_t ranges_bytes) {
    BCMC_CHECK_ARG(m && stream && dims, "null argument");
    BCMC_CHECK_ARG(dims[0] && dims[1] && dims[2], "volume dims must be positive");
    const double mb = 64.0 * rate;
    BCMC_CHECK_ARG(mb >= 9.0 && mb <= 4096.0 && mb == std::floor(mb),
                   "64*compressionRate must be an integer in [9, 4096]");
    uint32_t nb[3];
    for (int i = 0; i < 3; ++i)
        nb[i] = (dims[i] + 3) / 4;
    const uint64_t total = (uint64_t)nb[0] * nb[1] * nb[2];
    BCMC_CHECK_ARG(total < (1ull << 31), "volume has too many blocks");
    const uint32_t maxbits = (uint32_t)mb;
    const uint64_t need = (total * maxbits + 7) / 8;
    BCMC_CHECK_ARG(bytes >= need, "compressed stream shorter than the volume at this rate");
    if (ranges) {
        BCMC_CHECK_ARG(fmt == BCMC_RANGES_BLOCK || fmt == BCMC_RANGES_BLOCK_TILE, "bad ranges format");
        BCMC_CHECK_ARG(ranges_bytes == total * (fmt == BCMC_RANGES_BLOCK ? 8u : 16u),
                       "ranges must hold 8 (block) or 16 (block and tile) bytes per block");
    }
    mc_free_volume(m);
    bcmc_ctx *ctx = m->ctx;
    for (int i = 0; i < 3; ++i) {
        m->dims[i] = dims[i];
        m->pdims[i] = nb[i] * 4;
        m->nb[i] = nb[i];
        m->scale[i] = scale ? scale[i] : 1.0f;
    }
    m->total = (uint32_t)total;
    m->maxbits = maxbits;
    // stream + zero padding (the reader may look one word past a block)
    m->stream_bytes = bytes;
    m->stream_words = (bytes + 7) / 8 + 2;
    BCMC_TRY(dev_alloc((void **)&m->stream, m->stream_words * 8));
    BCMC_HIP_CHECK(hipMemsetAsync(m->stream, 0, m->stream_words * 8, ctx->stream));
    BCMC_HIP_CHECK(hipMemcpyAsync(m->stream, stream, bytes, hipMemcpyDefault, ctx->stream));
    const size_t n = m->total;
    BCMC_TRY(dev_alloc((void **)&m->ranges, n * 8));
    BCMC_TRY(dev_alloc((void **)&m->flags, n * 4));
    BCMC_TRY(dev_alloc((void **)&m->active_ids, n * 4));
    BCMC_TRY(dev_alloc((void **)&m->dilated, n * 8));
    BCMC_TRY(dev_alloc((void **)&m->tile, n * 8));
    BCMC_TRY(dev_alloc((void **)&m->with_verts, n * 4));
    BCMC_TRY(dev_alloc((void **)&m->vert_offsets, n * 8));
    BCMC_TRY(scan_site_alloc(&m->site_active, n));
    const size_t nch = div_up((uint64_t)n, kChunk);
    BCMC_TRY(dev_alloc((void **)&m->counts, n * 4));
    BCMC_TRY(dev_alloc((void **)&m->chunk_with, nch * 4));
    BCMC_TRY(dev_alloc((void **)&m->chunk_verts, nch * 4));
    BCMC_TRY(dev_alloc((void **)&m->chunk_with_pre, nch * 4));
    BCMC_TRY(dev_alloc((void **)&m->chunk_verts_pre, nch * 8));
    BCMC_TRY(scan_site_alloc(&m->site_c0, nch));
    BCMC_TRY(scan_site_alloc(&m->site_c1, nch));
    const uint32_t init = initial_cache_slots ? initial_cache_slots
                                              : (uint32_t)std::ceil((double)total * 0.05);
    BCMC_TRY(lru_state_alloc(ctx, &m->lru.st, init, 256, m->total));
    QState qi{};
    qi.cache_size = m->lru.st.cache_size;
    qi.cache_size_next = qi.cache_size;
    *m->qh = qi;
    BCMC_HIP_CHECK(hipMemcpyAsync(m->q, m->qh, sizeof(QState), hipMemcpyHostToDevice, ctx->stream));
    if (ranges) {
        // persisted ranges (the server-side computation :523-524 suggests)
        BCMC_HIP_CHECK(hipMemcpyAsync(m->ranges, ranges, n * 8, hipMemcpyDefault, ctx->stream));
        if (fmt == BCMC_RANGES_BLOCK_TILE)
            BCMC_HIP_CHECK(hipMemcpyAsync(m->tile, (const uint8_t *)ranges + n * 8, n * 8, hipMemcpyDefault,
                                          ctx->stream));
        else
            launch_tile_bound_ranges(ctx, m->ranges, m->dims, m->tile);
        launch_dilate_ranges(ctx, m->ranges, m->nb, m->dilated);
        BCMC_HIP_CHECK(hipGetLastError());
        BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } else {
        // computeBlockRanges (:522-578): full-volume decode, min/max per block
        float2 *parts = nullptr;  // 8 part ranges per block, transient
        BCMC_TRY(dev_alloc((void **)&parts, n * 64));
        launch_block_ranges(ctx, m->stream, m->stream_words, maxbits, m->dims, m->ranges, parts);
        launch_dilate_ranges(ctx, m->ranges, m->nb, m->dilated);
        launch_tile_ranges(ctx, parts, m->nb, m->tile);
        const hipError_t err = hipGetLastError();
        const hipError_t serr = hipStreamSynchronize(ctx->stream);
        dev_free(parts);
        BCMC_HIP_CHECK(err);
        BCMC_HIP_CHECK(serr);
    }
    m->has_volume = true;
    m->last_active = m->last_with = m->last_new = 0;
    m->last_verts = 0;
    mc_refresh_views(m);
    return BCMC_OK;
}

BCMC_API int bcmc_mc_set_compressed_volume(bcmc_mc *m, const void *stream, uint64_t bytes, double rate,
                                           const uint32_t dims[3], const float scale[3],
                                           uint32_t initial_cache_slots) {
    return mc_set_volume(m, stream, bytes, rate, dims, scale, initial_cache_slots, nullptr, BCMC_RANGES_BLOCK, 0);
}

BCMC_API int bcmc_mc_set_compressed_volume_ranges(bcmc_mc *m, const void *stream, uint64_t bytes, double rate,
                                                  const uint32_t dims[3], const float scale[3],
                                                  uint32_t initial_cache_slots, const void *ranges,
                                                  bcmc_ranges_format fmt, uint64_t ranges_bytes) {
    BCMC_CHECK_ARG(ranges, "null ranges");
    return mc_set_volume(m, stream, bytes, rate, dims, scale, initial_cache_slots, ranges, fmt, ranges_bytes);
}

BCMC_API int bcmc_mc_get_block_ranges(bcmc_mc *m, bcmc_ranges_format fmt, void *out, uint64_t bytes) {
    BCMC_CHECK_ARG(m && out, "null argument");
    if (!m->has_volume) {
        set_error("getBlockRanges called before setCompressedVolume");
        return BCMC_ERR_STATE;
    }
    BCMC_CHECK_ARG(fmt == BCMC_RANGES_BLOCK || fmt == BCMC_RANGES_BLOCK_TILE, "bad ranges format");
    const uint64_t n = m->total;
    BCMC_CHECK_ARG(bytes >= n * (fmt == BCMC_RANGES_BLOCK ? 8u : 16u), "output shorter than the ranges");
    hipStream_t st = m->ctx->stream;
    BCMC_HIP_CHECK(hipMemcpyAsync(out, m->ranges, n * 8, hipMemcpyDefault, st));
    if (fmt == BCMC_RANGES_BLOCK_TILE)
        BCMC_HIP_CHECK(hipMemcpyAsync((uint8_t *)out + n * 8, m->tile, n * 8, hipMemcpyDefault, st));
    BCMC_HIP_CHECK(hipStreamSynchronize(st));
    return BCMC_OK;
}

BCMC_API int bcmc_mc_set_vertex_format(bcmc_mc *m, bcmc_vertex_format fmt) {
    BCMC_CHECK_ARG(m, "null argument");
    BCMC_CHECK_ARG(fmt >= BCMC_VERTS_LOCAL_F32 && fmt <= BCMC_VERTS_PACKED_U32X2, "bad format");
    if (fmt != m->fmt) {
        if (m->ctx)
            (void)hipStreamSynchronize(m->ctx->stream);
        dev_free(m->verts);
        m->verts = nullptr;
        m->vert_cap = 0;
        m->fmt = fmt;
        m->last_verts = 0;
    }
    return BCMC_OK;
}

BCMC_API int bcmc_mc_compute_surface(bcmc_mc *m, float isovalue, bcmc_perf *perf,
                                     uint64_t *num_vertices) {
    BCMC_CHECK_ARG(m, "null argument");
    if (!m->has_volume) {
        set_error("computeSurface called before setCompressedVolume");
        return BCMC_ERR_STATE;
    }
    const auto t0 = std::chrono::steady_clock::now();
    bcmc_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    const bool timed = perf != nullptr;
    *m->iso_h = isovalue;
    // counted before the launch, so the host's update count never lags the device's
    const uint32_t npasses = lru_begin_update(&m->lru.st);
    const uint64_t key[6] = {(uint64_t)(uintptr_t)m->verts, m->vert_cap,
                             (uint64_t)m->fmt | ctx->tune_gen << 8, npasses,
                             (uint64_t)(uintptr_t)m->lru.st.data, m->lru.st.slots};
    bool launched = false;
    if (ctx->graph_replay && m->use_graph && !timed) {
        if (!m->gexec || std::memcmp(key, m->graph_key, sizeof(key)) != 0) {
            mc_drop_graph(m);
            BCMC_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = mc_enqueue_query(m, npasses, false);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc == BCMC_OK && ec == hipSuccess && g &&
                hipGraphInstantiate(&m->gexec, g, nullptr, nullptr, 0) == hipSuccess) {
                m->graph = g;
                std::memcpy(m->graph_key, key, sizeof(key));
            } else {
                if (g)
                    (void)hipGraphDestroy(g);
                m->gexec = nullptr;
                m->use_graph = false;  // replay is an optimisation: fall back to direct launches
                (void)hipGetLastError();
            }
        }
        if (m->gexec) {
            BCMC_HIP_CHECK(hipGraphLaunch(m->gexec, st));
            launched = true;
        }
    }
    if (!launched)
        BCMC_TRY(mc_enqueue_query(m, npasses, timed));
    BCMC_HIP_CHECK(hipStreamSynchronize(st));
    QState qs = *m->qh;
    float ms[6] = {};
    if (timed)
        for (int i = 0; i < 6; ++i)
            BCMC_HIP_CHECK(hipEventElapsedTime(&ms[i], m->ev[i], m->ev[i + 1]));
    if (qs.cache_short) {
        // the cache grew past its storage, so nothing was assigned: enlarge
        // the storage, then assign, decode, count and emit again (the LRU
        // state up to the assignment is complete)
        BCMC_TRY(lru_reserve(ctx, &m->lru.st, (uint64_t)qs.cache_size_next + qs.num_new / 2 + 32));
        mc_refresh_views(m);
        BCMC_HIP_CHECK(hipMemsetAsync(&m->q->cache_short, 0, 4, st));
        if (timed)
            BCMC_HIP_CHECK(hipEventRecord(m->ev[2], st));
        BCMC_TRY(mc_enqueue_tail(m, timed));
        BCMC_HIP_CHECK(hipStreamSynchronize(st));
        qs = *m->qh;
        if (timed)
            for (int i = 2; i < 6; ++i)
                BCMC_HIP_CHECK(hipEventElapsedTime(&ms[i], m->ev[i], m->ev[i + 1]));
    }
    if (qs.num_verts > m->vert_cap) {
        // grow like compressed_marching_cubes.js:774-785, then run the pass again
        const uint64_t want = std::max<uint64_t>(qs.num_verts, (uint64_t)std::floor(m->vert_cap * 1.5));
        dev_free(m->verts);
        m->verts = nullptr;
        m->vert_cap = 0;
        BCMC_TRY(dev_alloc(&m->verts, want * vertex_bytes(m->fmt)));
        m->vert_cap = want;
        if (timed)
            BCMC_HIP_CHECK(hipEventRecord(m->ev[5], st));
        BCMC_TRY(mc_launch_emit(m));
        if (timed)
            BCMC_HIP_CHECK(hipEventRecord(m->ev[6], st));
        BCMC_HIP_CHECK(hipMemcpyAsync(m->qh, m->q, sizeof(QState), hipMemcpyDeviceToHost, st));
        BCMC_HIP_CHECK(hipStreamSynchronize(st));
        qs = *m->qh;
        if (timed)
            BCMC_HIP_CHECK(hipEventElapsedTime(&ms[5], m->ev[5], m->ev[6]));
    }
    m->lru.st.cache_size = qs.cache_size;
    // storage for some growth of the next query (misses a third of this
    // one's, which a sweep's later queries stay under); a larger growth takes
    // the path above
    BCMC_TRY(lru_reserve(ctx, &m->lru.st, (uint64_t)qs.cache_size + qs.num_new / 2 + 32));
    m->last_active = qs.num_active;
    m->last_with = qs.num_with;
    m->last_new = qs.num_new;
    m->last_verts = qs.num_verts;
    mc_refresh_views(m);
    if (num_vertices)
        *num_vertices = qs.num_verts;
    if (perf) {
        perf->computeActiveBlocks = ms[0];
        perf->cacheUpdate = ms[1];
        perf->decompression = ms[2];
        perf->compactActiveIDs = 0.0;            // fused into computeActiveBlocks
        perf->computeBlockHasVertices = ms[3];   // count launch
        perf->compactBlocksWithVertices = ms[4]; // prefix scan of the chunk totals
        perf->computeVertices = ms[5];           // emit launch (after a growth: its rerun)
        perf->numActiveBlocks = qs.num_active;
        perf->numBlocksDecompressed = qs.num_new;
        perf->numBlocksWithVertices = qs.num_with;
        perf->cacheSize = qs.cache_size;
        perf->numVertices = qs.num_verts;
        perf->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return BCMC_OK;
}

BCMC_API int bcmc_mc_get_info(bcmc_mc *m, bcmc_mc_info *info) {
    BCMC_CHECK_ARG(m && info, "null argument");
    memset(info, 0, sizeof(*info));
    info->total_blocks = m->total;
    info->num_active_blocks = m->last_active;
    info->num_blocks_with_vertices = m->last_with;
    info->new_decompressed = m->last_new;
    info->num_vertices = m->last_verts;
    info->vertex_capacity = m->vert_cap;
    for (int i = 0; i < 3; ++i) {
        info->dims[i] = m->dims[i];
        info->padded_dims[i] = m->pdims[i];
        info->scale[i] = m->scale[i];
    }
    info->max_bits = m->maxbits;
    info->format = (bcmc_vertex_format)m->fmt;
    return BCMC_OK;
}

BCMC_API bcmc_buf *bcmc_mc_vertex_buffer(bcmc_mc *m) { return m ? &m->vert_view : nullptr; }
BCMC_API bcmc_lru *bcmc_mc_lru(bcmc_mc *m) { return m && m->has_volume ? &m->lru : nullptr; }

BCMC_API bcmc_buf *bcmc_mc_array_buffer(bcmc_mc *m, bcmc_mc_array which) {
    if (!m || which < BCMC_ARRAY_RANGES || which > BCMC_ARRAY_VERT_OFFSETS)
        return nullptr;
    return &m->views[which];
}

BCMC_API int bcmc_mc_memory_use(bcmc_mc *m, uint64_t *total_out, uint64_t *cache_out) {
    BCMC_CHECK_ARG(m, "null argument");
    const uint64_t n = m->total;
    const LruState &L = m->lru.st;
    const uint64_t cache = (uint64_t)L.slots * L.elem_bytes + 256;  // the storage, not all blocks
    const uint64_t lru_meta = (uint64_t)L.capacity * 4 * 5 + (uint64_t)L.aligned_total * 8;
    const uint64_t mc = m->stream_words * 8 + n * (8 + 4 + 4 + 4 + 4 + 8) +
                        m->vert_cap * vertex_bytes(m->fmt);
    if (total_out)
        *total_out = mc + cache + lru_meta;
    if (cache_out)
        *cache_out = cache;
    return BCMC_OK;
}

}  // extern "C"
