// Stable LSD radix sort of (u32 key, u32 value) pairs on gfx950, used for the
// LRU slot ordering and the RadixSorter drop-in.
//
// The reference (js/radix_sort_by_key.js:161-400) sorts 64-element chunks
// with 32 one-bit LDS split passes (radix_sort_chunk.comp), merges runs with
// merge_sorted_chunks.comp, whose upper/lower bound pair places equal keys of
// the later run first, reverses the value buffer (reverse_buffer.comp) and
// reads back a slice (:382-391).  The resulting order is a total order:
// ascending: key asc, ties by chunk (position/64) descending, then position
// ascending; reversed: key desc, ties by chunk ascending, then position
// descending, over the UINT_MAX-padded aligned array, of which the reference
// keeps the first n (ascending) or the last n (reversed).  This file produces
// exactly that order with a stable 8-bit LSD radix sort whose first pass
// reads the input in that tie order (keys complemented for the descending
// case).  Each pass: per-tile LDS histogram, a single-pass scan of the
// digit-major histogram table, and a stable scatter that ranks equal digits
// inside a tile with wave64 ballots.
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace bcmc {

// Tiles of 256 threads x ITEMS elements, ITEMS = 8, 16 or 32 (2048, 4096,
// 8192 elements; the autotuner picks one).  The sort is stable at every tile
// shape, so the output does not depend on it.
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr uint32_t kSortTileMin = kSortThreads * 8;

// Position of item j of a lane in tile t: each wave owns a contiguous run of
// 64 * ITEMS positions, read in ITEMS coalesced rounds.
template <int ITEMS>
BCMC_DEV uint32_t sort_pos(uint32_t tile, uint32_t wave, int j, uint32_t lane) {
    return tile * (kSortThreads * ITEMS) + wave * (64u * ITEMS) + (uint32_t)j * 64u + lane;
}

// Lanes of the wave holding the same digit as this lane (among `valid` ones).
BCMC_DEV uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    return peers;
}

// ---- element sources -------------------------------------------------------
struct BufSource {  // passes >= 1
    const uint32_t *keys, *vals;
    const uint32_t *n_dev;
    uint32_t n_host;
    BCMC_DEV uint32_t size() const { return n_dev ? *n_dev : n_host; }
    BCMC_DEV void get(uint32_t p, uint32_t &k, uint32_t &v) const {
        k = keys[p];
        v = vals[p];
    }
};

// LRU: position p of the available-slot list in the reference's tie order
// (64-chunks ascending, descending inside a chunk), key = ~age (descending).
struct LruSource {
    const uint32_t *avail_ids, *avail_keys;
    const uint32_t *n_dev;
    uint32_t capacity;
    BCMC_DEV uint32_t size() const { return *n_dev; }
    BCMC_DEV void get(uint32_t p, uint32_t &k, uint32_t &v) const {
        const uint32_t n = *n_dev;
        const uint32_t c0 = p & ~63u;
        const uint32_t m = min(64u, n - c0);
        const uint32_t j = c0 + (m - 1 - (p - c0));
        v = avail_ids[j];
        const uint32_t kk = avail_keys[j];
        k = v < capacity ? kk : 0u;
    }
};

// RadixSorter.sort: all `aligned` positions including UINT_MAX padding.
struct GenericSource {
    const uint32_t *keys, *vals;
    uint32_t n, aligned, reverse;
    BCMC_DEV uint32_t size() const { return aligned; }
    BCMC_DEV void get(uint32_t p, uint32_t &k, uint32_t &v) const {
        const uint32_t nch = aligned / 64, c = p / 64, i = p % 64;
        const uint32_t j = reverse ? c * 64 + (63 - i) : (nch - 1 - c) * 64 + i;
        const uint32_t kk = j < n ? keys[j] : 0xFFFFFFFFu;
        v = j < n ? vals[j] : 0xFFFFFFFFu;
        k = reverse ? ~kk : kk;
    }
};

// ---- kernels ---------------------------------------------------------------
// Elements are read for every position of a tile (positions past n read
// element n - 1 and are ignored), so all loads of a thread are in flight
// together.
template <class Src, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void sort_hist_kernel(Src src, uint32_t shift,
                                                                 uint32_t *hist) {
    __shared__ uint32_t h[256];
    const uint32_t n = src.size(), ntiles = div_up(n, kSortThreads * ITEMS);
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t key[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            uint32_t v;
            src.get(min(sort_pos<ITEMS>(tile, wave, j, lane), n - 1), key[j], v);
        }
        h[threadIdx.x] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool valid = sort_pos<ITEMS>(tile, wave, j, lane) < n;
            const uint32_t d = (key[j] >> shift) & 255u;
            const uint64_t peers = digit_peers(d, valid);
            if (valid && (peers & lt) == 0)
                atomicAdd(&h[d], (uint32_t)__popcll(peers));
        }
        __syncthreads();
        hist[threadIdx.x * ntiles + tile] = h[threadIdx.x];
        __syncthreads();
    }
}

struct HistScanOp {
    uint32_t *hist;
    const uint32_t *n_dev;
    uint32_t n_host;
    uint32_t tile_elems;  // elements per sort tile
    BCMC_DEV uint32_t size() const {
        const uint32_t n = n_dev ? *n_dev : n_host;
        return 256u * div_up(n, tile_elems);
    }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[1], Item &) const { v[0] = hist[i]; }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        hist[i] = (uint32_t)e[0];
    }
    BCMC_DEV void total(const unsigned long long (&)[1]) const {}
};

// Stable scatter.  Each wave ranks its run in order: an element's rank is the
// count of its digit in the wave's earlier rounds (a per-wave LDS counter
// that the digit's lowest lane advances) plus its lower lanes of the same
// digit in this round; then every digit's start in the tile (from the scanned
// histogram table) is split among the four waves by their counts.
template <class Src, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(Src src, uint32_t shift,
                                                                    const uint32_t *hist,
                                                                    uint32_t *okeys, uint32_t *ovals,
                                                                    uint32_t out_skip, uint32_t out_cap) {
    __shared__ uint32_t cnt[kSortWaves][256];
    const uint32_t n = src.size(), ntiles = div_up(n, kSortThreads * ITEMS);
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            src.get(min(sort_pos<ITEMS>(tile, wave, j, lane), n - 1), key[j], val[j]);
        const uint32_t start = hist[threadIdx.x * ntiles + tile];  // digit threadIdx.x
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w)
            cnt[w][threadIdx.x] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool valid = sort_pos<ITEMS>(tile, wave, j, lane) < n;
            const uint32_t d = (key[j] >> shift) & 255u;
            const uint64_t peers = digit_peers(d, valid);
            const uint32_t before = cnt[wave][d];
            rank[j] = before + (uint32_t)__popcll(peers & lt);
            if (valid && (peers & lt) == 0)
                cnt[wave][d] = before + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        {
            uint32_t run = start;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const uint32_t c = cnt[w][threadIdx.x];
                cnt[w][threadIdx.x] = run;
                run += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if (sort_pos<ITEMS>(tile, wave, j, lane) >= n)
                continue;
            const uint32_t dst = cnt[wave][(key[j] >> shift) & 255u] + rank[j];
            if (dst >= out_skip && dst - out_skip < out_cap) {
                if (okeys)
                    okeys[dst - out_skip] = key[j];
                ovals[dst - out_skip] = val[j];
            }
        }
        __syncthreads();  // the counters are cleared for the next tile
    }
}

// `npasses` 8-bit digit passes, LSB first; the final pass writes only the
// values (to final_vals, from position out_skip on).  Pass p reads the source
// (p = 0) or buffer (p - 1) & 1, and writes buffer p & 1 or the final output.
template <int ITEMS, class FirstSrc>
static void sort_passes_items(bcmc_ctx *ctx, SortScratch *s, uint32_t n_max, const uint32_t *n_dev,
                              uint32_t n_host, const FirstSrc &first_src, uint32_t npasses,
                              uint32_t *final_vals, uint32_t out_skip) {
    constexpr uint32_t kTile = kSortThreads * ITEMS;
    const uint32_t tiles = std::max<uint32_t>(1, div_up(n_max, kTile));
    const uint32_t grid = std::min<uint32_t>(tiles, persistent_grid(ctx, ctx->tune.sort));
    hipStream_t st = ctx->stream;
    HistScanOp hop{s->hist, n_dev, n_host, kTile};
    for (uint32_t pass = 0; pass < npasses; ++pass) {
        const uint32_t shift = 8 * pass;
        const bool last = pass + 1 == npasses;
        uint32_t *ok = last ? nullptr : s->keys[pass & 1];
        uint32_t *ov = last ? final_vals : s->vals[pass & 1];
        const uint32_t skip = last ? out_skip : 0;
        if (pass == 0) {
            sort_hist_kernel<FirstSrc, ITEMS><<<grid, kSortThreads, 0, st>>>(first_src, shift, s->hist);
            launch_scan<1>(ctx, s->site, s->site, hop, 256ull * tiles);
            sort_scatter_kernel<FirstSrc, ITEMS><<<grid, kSortThreads, 0, st>>>(first_src, shift, s->hist, ok,
                                                                                ov, skip, s->max_elems);
        } else {
            BufSource b{s->keys[(pass - 1) & 1], s->vals[(pass - 1) & 1], n_dev, n_host};
            sort_hist_kernel<BufSource, ITEMS><<<grid, kSortThreads, 0, st>>>(b, shift, s->hist);
            launch_scan<1>(ctx, s->site, s->site, hop, 256ull * tiles);
            sort_scatter_kernel<BufSource, ITEMS><<<grid, kSortThreads, 0, st>>>(b, shift, s->hist, ok, ov,
                                                                                 skip, s->max_elems);
        }
    }
}

// The passes at the context's tile shape.
template <class FirstSrc>
static void sort_passes(bcmc_ctx *ctx, SortScratch *s, uint32_t n_max, const uint32_t *n_dev,
                        uint32_t n_host, const FirstSrc &first_src, uint32_t npasses,
                        uint32_t *final_vals, uint32_t out_skip) {
    if (ctx->tune.sort_items == 8)
        sort_passes_items<8>(ctx, s, n_max, n_dev, n_host, first_src, npasses, final_vals, out_skip);
    else if (ctx->tune.sort_items == 32)
        sort_passes_items<32>(ctx, s, n_max, n_dev, n_host, first_src, npasses, final_vals, out_skip);
    else
        sort_passes_items<16>(ctx, s, n_max, n_dev, n_host, first_src, npasses, final_vals, out_skip);
}

int sort_scratch_alloc(SortScratch *s, uint32_t max_elems) {
    s->max_elems = std::max<uint32_t>(max_elems, 64);
    s->max_tiles = div_up(s->max_elems, kSortTileMin);  // tiles of the smallest shape
    for (int i = 0; i < 2; ++i) {
        BCMC_TRY(dev_alloc((void **)&s->keys[i], (size_t)s->max_elems * 4));
        BCMC_TRY(dev_alloc((void **)&s->vals[i], (size_t)s->max_elems * 4));
    }
    BCMC_TRY(dev_alloc((void **)&s->hist, (size_t)256 * s->max_tiles * 4));
    BCMC_TRY(scan_site_alloc(&s->site, (uint64_t)256 * s->max_tiles));
    return BCMC_OK;
}

void sort_scratch_free(SortScratch *s) {
    for (int i = 0; i < 2; ++i) {
        dev_free(s->keys[i]);
        dev_free(s->vals[i]);
        s->keys[i] = s->vals[i] = nullptr;
    }
    dev_free(s->hist);
    s->hist = nullptr;
    scan_site_free(&s->site);
}

void launch_lru_order(bcmc_ctx *ctx, SortScratch *s, const uint32_t *avail_ids,
                      const uint32_t *avail_keys, const uint32_t *count_dev, uint32_t count_max,
                      uint32_t *out, uint32_t npasses) {
    LruSource src{avail_ids, avail_keys, count_dev, count_max};
    sort_passes(ctx, s, count_max, count_dev, 0, src, npasses, out, 0);
}

void launch_sort_pairs(bcmc_ctx *ctx, SortScratch *s, const uint32_t *keys, const uint32_t *vals,
                       const uint32_t *count_dev, uint32_t count_max, uint32_t *out,
                       uint32_t npasses) {
    BufSource src{keys, vals, count_dev, 0};
    sort_passes(ctx, s, count_max, count_dev, 0, src, npasses, out, 0);
}

void launch_sort_by_key(bcmc_ctx *ctx, SortScratch *s, uint32_t *keys, uint32_t *vals, uint32_t n,
                        int reverse) {
    uint32_t aligned = 64;  // nextPow2(ceil(n/64)) * 64
    while (aligned < (uint64_t)((n + 63) / 64) * 64)
        aligned *= 2;
    GenericSource src{keys, vals, n, aligned, (uint32_t)(reverse != 0)};
    // reversed: the reference keeps the last n of the descending order
    sort_passes(ctx, s, aligned, nullptr, aligned, src, 4, s->vals[1], reverse ? aligned - n : 0);
}

}  // namespace bcmc
