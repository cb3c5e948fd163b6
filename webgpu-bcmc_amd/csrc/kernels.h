// Internal launchers and the single-pass scan kernel template shared by the
// library's translation units.
#pragma once

#include "common.h"

namespace bcmc {

// ---------------------------------------------------------------------------
// Device-resident per-query counters (one struct per pipeline object).  All
// sizes that the reference reads back to the host between stages
// (compressed_marching_cubes.js:580-809, lru_cache.js:312-843) live here and
// are consumed by later kernels directly; the host reads the struct once at
// the end of a query.
// ---------------------------------------------------------------------------
struct QState {
    float iso;
    uint32_t num_active;       // A
    uint32_t num_new;          // D: LRU misses
    uint32_t num_avail;        // S: available slots before growth
    uint32_t cache_size;       // C: logical LRU size before this update
    uint32_t cache_size_next;  // C after this update (growth)
    uint32_t num_with;         // H: active blocks with vertices
    uint32_t overflow;         // emit skipped: V > vertex capacity
    unsigned long long num_verts;     // V
    unsigned long long vert_capacity; // vertices the output buffer holds
    uint32_t avail_display;    // S' - D (displayNumSlotsAvailable)
    uint32_t num_avail_total;  // S': available slots after growth
    uint32_t cache_short;      // the assignment did not run: the cache grew past its storage
    uint32_t pad[3];
};

// ---------------------------------------------------------------------------
// Single-pass (decoupled look-back) scan over a device-sized range with one or
// two u32 channels.  Op provides:
//   struct Tile, struct Item              per-tile context, per-element payload
//   uint32_t size() const                 number of elements (may read memory)
//   void begin() const                    once, by thread 0 of workgroup 0
//   Tile tile(uint32_t t) const           context of tile t (every thread)
//   void value(const Tile &, uint32_t i, uint32_t v[NCH], Item &) const
//   void emit(const Tile &, uint32_t i, const uint32_t v[NCH], const Item &,
//             const unsigned long long excl[NCH]) const
//   void total(const unsigned long long t[NCH]) const  (one thread, once)
// value() only reads memory (no stores) and takes no branch on what it reads,
// so the loads of all ITEMS elements of a thread are in flight together; it is
// also evaluated for the positions past the end (at element n - 1, with the
// values then zeroed).  Stores belong in emit(), which the payload feeds.
// Each wave scans a contiguous run of 64 x ITEMS elements of its tile in
// ITEMS coalesced rounds with DPP; wave totals combine through LDS; tiles
// chain through the look-back status words of site[c].  Results do not depend
// on ITEMS (the tile shape the autotuner picks).
// ---------------------------------------------------------------------------
struct NoTile {};
struct NoItem {};

template <int NCH, class Op, int ITEMS>
__global__ __launch_bounds__(kScanThreads) void scan_kernel(ScanSite s0, ScanSite s1, Op op) {
    constexpr uint32_t kTile = kScanThreads * ITEMS;
    constexpr int NW = kScanThreads / 64;
    __shared__ uint32_t lds_w[NCH][NW + 1];
    __shared__ LookbackLds<kScanThreads> lds_lb;
    __shared__ uint32_t lds_tile;
    const ScanSite *sites[2] = {&s0, &s1};
    const uint32_t n = op.size();
    const uint32_t ntiles = div_up(n, kTile);
    uint32_t epoch[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        epoch[c] = scan_site_epoch(*sites[c]);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        op.begin();
    if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        unsigned long long z[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            z[c] = 0;
        op.total(z);
    }
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
    for (;;) {
        const uint32_t tile = next_tile(s0, &lds_tile);
        if (tile >= ntiles)
            break;
        const typename Op::Tile tc = op.tile(tile);
        const uint32_t wbase = tile * kTile + wave * (64 * ITEMS);
        uint32_t vals[ITEMS][NCH], excl[ITEMS][NCH], run[NCH];
        typename Op::Item items[ITEMS];
        // all loads of the tile first (independent), then the scans
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t i = wbase + j * 64 + lane;
            op.value(tc, min(i, n - 1), vals[j], items[j]);
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool in = wbase + j * 64 + lane < n;
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                vals[j][c] = in ? vals[j][c] : 0u;
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            run[c] = 0;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const uint32_t v = vals[j][c];
                const uint32_t inc = wave_incl_scan(v);
                excl[j][c] = run[c] + inc - v;
                run[c] += wave_read_lane(inc, 63);
            }
        }
        unsigned long long prefix[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (lane == 0)
                lds_w[c][wave] = run[c];
        }
        __syncthreads();
        uint32_t before[NCH], agg[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            before[c] = 0;
            agg[c] = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint32_t t = lds_w[c][w];
                before[c] += w < (int)wave ? t : 0u;
                agg[c] += t;
            }
        }
        __syncthreads();
        {
            unsigned long long q[2];
            block_lookback<kScanThreads, NCH>(s0, NCH == 2 ? s1 : s0, tile, epoch[0], epoch[NCH - 1], agg[0],
                                              agg[NCH - 1], lds_lb, q[0], q[1]);
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                prefix[c] = q[c];
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t i = wbase + j * 64 + lane;
            if (i < n) {
                unsigned long long e[NCH];
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    e[c] = prefix[c] + before[c] + excl[j][c];
                op.emit(tc, i, vals[j], items[j], e);
            }
        }
        if (tile == ntiles - 1 && threadIdx.x == 0) {
            unsigned long long t[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                t[c] = prefix[c] + agg[c];
            op.total(t);
        }
        // the grid's first claims are tiles 0 .. gridDim.x - 1: with no more
        // tiles than that, no workgroup claims a second one
        if (ntiles <= gridDim.x)
            break;
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        scan_site_finish(*sites[c], epoch[c]);
}

template <int NCH, int ITEMS, class Op>
inline void launch_scan_items(bcmc_ctx *ctx, const ScanSite &s0, const ScanSite &s1, const Op &op,
                              uint64_t n_max) {
    const uint32_t grid = std::max<uint32_t>(
        1, std::min<uint32_t>(div_up(n_max, kScanThreads * ITEMS), persistent_grid(ctx, ctx->tune.scan)));
    scan_kernel<NCH, Op, ITEMS><<<grid, kScanThreads, 0, ctx->stream>>>(s0, s1, op);
}

// The scan at the context's tile shape, at most MAX_ITEMS elements per
// thread (an Op with a large per-element payload would spill registers at 32).
template <int NCH, class Op, int MAX_ITEMS = 32>
inline void launch_scan(bcmc_ctx *ctx, const ScanSite &s0, const ScanSite &s1, const Op &op,
                        uint64_t n_max) {
    const uint32_t items = std::min<uint32_t>(ctx->tune.scan_items, MAX_ITEMS);
    if (items == 8)
        launch_scan_items<NCH, 8>(ctx, s0, s1, op, n_max);
    else if (items == 32)
        launch_scan_items<NCH, (MAX_ITEMS >= 32 ? 32 : 16)>(ctx, s0, s1, op, n_max);
    else
        launch_scan_items<NCH, 16>(ctx, s0, s1, op, n_max);
}

// ---------------------------------------------------------------------------
// Launchers implemented in the .hip files
// ---------------------------------------------------------------------------
void launch_block_ranges(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords, uint32_t maxbits,
                         const uint32_t dims[3], float2 *ranges, float2 *part_ranges = nullptr);
void launch_decode_to_cache(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords,
                            uint32_t maxbits, const uint32_t *ids, const uint32_t *n_dev,
                            uint32_t n_max, const int32_t *item_slot, float *cache,
                            uint32_t capacity);
// LRU assignment fused with the decode: new item w (w < D, 0 when no slot is
// available) evicts the previous item of slot order[w], takes the slot
// {age 0, unavailable} and is decoded into it; commits the cache size.
// `capacity` is the number of slots `cache` holds: when the grown cache size
// exceeds it nothing is assigned and q->cache_short is set (the host enlarges
// the storage and launches it again).
void launch_decode_assign(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords, uint32_t maxbits,
                          const uint32_t *new_items, uint32_t n_max, const uint32_t *order,
                          uint32_t *slot_age, uint32_t *slot_avail, int32_t *slot_item,
                          int32_t *item_slot, uint32_t total, float *cache, uint32_t capacity,
                          QState *q);
// Per block its tile range: [min, max] of every decoded voxel its cells can
// read (its own 64, padding included, and the ghost faces, edges and corner of
// its in-grid +x/+y/+z neighbours), from the range pass's 8 part ranges per
// block.  A block has vertices only if lo <= iso < hi: otherwise all its cells
// have case 0 or 255.
void launch_tile_ranges(bcmc_ctx *ctx, const float2 *part_ranges, const uint32_t nb[3], float2 *out);
// A wider tile range from block ranges alone (no part ranges): the union of the
// ranges of the block and its in-grid +x/+y/+z, edge and corner neighbours, or
// (-inf, inf) for a block whose cells read padding voxels (last on an axis
// whose size is not a multiple of 4).  Every other voxel a cell reads lies
// inside the volume in one of those blocks, so the bound skips only blocks
// without vertices, as the tile range does.
void launch_tile_bound_ranges(bcmc_ctx *ctx, const float2 *ranges, const uint32_t dims[3], float2 *out);
// Per block the range [min lo, max hi] over its in-bounds 3x3x3 neighbourhood
// (the active test's interval, see block_active).
void launch_dilate_ranges(bcmc_ctx *ctx, const float2 *ranges, const uint32_t nb[3], float2 *out);
void launch_decode_dense(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords, uint32_t maxbits,
                         const uint32_t *ids, uint32_t n, float *out);

// Radix sort of (key, value) pairs, see sort.hip.
struct SortScratch {
    uint32_t *keys[2] = {nullptr, nullptr};
    uint32_t *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;   // 256 * max_tiles
    uint32_t max_elems = 0, max_tiles = 0;
    ScanSite site;
};
int sort_scratch_alloc(SortScratch *s, uint32_t max_elems);
void sort_scratch_free(SortScratch *s);

// Update chains since an LRU reset below which the eviction order sorts
// compressed ages.  After a reset every slot is fresh (age 100000); in the
// u-th update chain since then (u counts the current one), a slot used since
// the reset has an age below u and a fresh slot an age in [100000,
// 100000 + u].  So age -> age (used) or u + age - 100000 (fresh) keeps the
// order and every tie, and maps into [0, 2u]: one 8-bit radix pass for the
// first 127 updates after a reset instead of three.
constexpr uint32_t kCompressedUpdates = 0x8000;

// Sort key of a slot's age: ~(compressed age) while u < kCompressedUpdates,
// ~age otherwise (u = 0xFFFFFFFF before the first reset: constructor ages).
BCMC_DEV uint32_t lru_order_key(uint32_t age, uint32_t u) {
    const uint32_t c = age >= 100000u ? u + (age - 100000u) : age;
    return ~(u < kCompressedUpdates ? c : age);
}

// Orders the `count` available cache slots for LRU assignment exactly as the
// reference's RadixSorter.sort(ages, slot ids, S, reverse=true) does, writing
// slot ids to out[0..count).  avail_keys[i] is the order key of slot
// avail_ids[i] (lru_order_key).  count is read from device memory.  npasses:
// the 8-bit digits of the keys that vary (those above every key are all 0xFF).
void launch_lru_order(bcmc_ctx *ctx, SortScratch *s, const uint32_t *avail_ids,
                      const uint32_t *avail_keys, const uint32_t *count_dev, uint32_t count_max,
                      uint32_t *out, uint32_t npasses);
// Stable ascending sort of `count` (key, value) pairs (count read from device
// memory) on the low 8 * npasses key bits; sorted values to out[0..count).
void launch_sort_pairs(bcmc_ctx *ctx, SortScratch *s, const uint32_t *keys, const uint32_t *vals,
                       const uint32_t *count_dev, uint32_t count_max, uint32_t *out,
                       uint32_t npasses);
// RadixSorter.sort semantics for host-known n over an aligned buffer.
void launch_sort_by_key(bcmc_ctx *ctx, SortScratch *s, uint32_t *keys, uint32_t *vals, uint32_t n,
                        int reverse);

}  // namespace bcmc
