// Launch-shape autotuner: times the tile shapes and persistent-grid
// candidates of the scan, radix-sort and zfp-decode kernel families on
// synthetic device data and keeps the fastest in the context.  Tiles: the
// scan's elements per thread (tiles of 4096, 8192 or 16384 elements), the
// sort's (2048, 4096 or 8192) and the decoders' blocks per workgroup (64, 128
// or 256).  Results of every kernel are independent of both (tiles are
// claimed in order, per-tile histograms and look-back status words are
// indexed by tile, the sort is stable, a decoded block is a pure function of
// its bits), so tuning changes speed only.
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <initializer_list>
#include <utility>
#include <vector>

namespace bcmc {
namespace {

// 32-bit mix of (index, seed): deterministic synthetic payloads.
__device__ inline uint32_t mix32(uint64_t i, uint32_t seed) {
    uint64_t z = i * 0x9E3779B97F4A7C15ull + seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

__global__ void tune_fill_kernel(uint32_t *p, uint64_t n, uint32_t seed, uint32_t mask,
                                 int identity) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = identity ? (uint32_t)i : (mix32(i, seed) & mask);
}

struct TuneScanOp {
    uint32_t *data;
    uint32_t n;
    unsigned long long *total_out;
    BCMC_DEV uint32_t size() const { return n; }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[1], Item &) const { v[0] = data[i] & 0xFF; }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        data[i] = (uint32_t)e[0];
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { *total_out = t[0]; }
};

// Device buffers of one tuning run, freed on every exit path.
struct TuneBufs {
    std::vector<void *> ptrs;
    ~TuneBufs() {
        for (void *p : ptrs)
            dev_free(p);
    }
    template <class T>
    int alloc(T **p, size_t bytes) {
        BCMC_TRY(dev_alloc((void **)p, bytes));
        ptrs.push_back(*p);
        return BCMC_OK;
    }
};

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
        if (a)
            (void)hipEventDestroy(a);
        if (b)
            (void)hipEventDestroy(b);
    }
};

void fill(bcmc_ctx *ctx, uint32_t *p, uint64_t n, uint32_t seed, uint32_t mask, int identity) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(n, 256), persistent_grid(ctx, 8));
    tune_fill_kernel<<<std::max(grid, 1u), 256, 0, ctx->stream>>>(p, n, seed, mask, identity);
}

// Median device time (ms) of `reps` runs of f after one warm-up run.
template <class F>
int time_median(bcmc_ctx *ctx, Events &ev, int reps, F &&f, float *ms) {
    f();
    BCMC_HIP_CHECK(hipGetLastError());
    std::vector<float> t;
    for (int r = 0; r < reps; ++r) {
        BCMC_HIP_CHECK(hipEventRecord(ev.a, ctx->stream));
        f();
        BCMC_HIP_CHECK(hipEventRecord(ev.b, ctx->stream));
        BCMC_HIP_CHECK(hipEventSynchronize(ev.b));
        float x = 0.f;
        BCMC_HIP_CHECK(hipEventElapsedTime(&x, ev.a, ev.b));
        t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return BCMC_OK;
}

// Tries each candidate (tile, workgroups per CU) for the pair of slots on
// every problem size (f(n) runs one), leaves the one with the least summed
// time per element in them (each size weighs the same, so a shape that only
// suits the largest size loses); *best_ms receives its time at the last size.
template <class F>
int pick(bcmc_ctx *ctx, Events &ev, uint32_t &tile_slot, uint32_t &grid_slot,
         const std::vector<std::pair<uint32_t, uint32_t>> &cands, const std::vector<uint32_t> &sizes, F &&f,
         float *best_ms) {
    std::pair<uint32_t, uint32_t> best{tile_slot, grid_slot};
    double best_score = 0.0;
    float best_t = 0.f;
    for (size_t i = 0; i < cands.size(); ++i) {
        tile_slot = cands[i].first;
        grid_slot = cands[i].second;
        double score = 0.0;
        float t = 0.f;
        for (uint32_t n : sizes) {
            BCMC_TRY(time_median(ctx, ev, 5, [&] { f(n); }, &t));
            score += (double)t / n;
        }
        if (i == 0 || score < best_score) {
            best_score = score;
            best_t = t;
            best = cands[i];
        }
    }
    tile_slot = best.first;
    grid_slot = best.second;
    *best_ms = best_t;
    return BCMC_OK;
}

// Every tile shape with every grid size.
std::vector<std::pair<uint32_t, uint32_t>> grid_of(std::initializer_list<uint32_t> tiles,
                                                   std::initializer_list<uint32_t> grids) {
    std::vector<std::pair<uint32_t, uint32_t>> v;
    for (uint32_t t : tiles)
        for (uint32_t g : grids)
            v.push_back({t, g});
    return v;
}

}  // namespace
}  // namespace bcmc

using namespace bcmc;

static void to_public(const bcmc_ctx *ctx, bcmc_tuning *t) {
    t->scan_wg_per_cu = ctx->tune.scan;
    t->sort_wg_per_cu = ctx->tune.sort;
    t->decode_wg_per_cu = ctx->tune.decode;
    t->mc_wg_per_cu = ctx->tune.mc;
    t->count_wg_per_cu = ctx->tune.count;
    t->scan_items = ctx->tune.scan_items;
    t->sort_items = ctx->tune.sort_items;
    t->decode_threads = ctx->tune.decode_threads;
}

static bool one_of(uint32_t x, uint32_t a, uint32_t b, uint32_t c) { return x == a || x == b || x == c; }

extern "C" {

BCMC_API int bcmc_ctx_get_tuning(bcmc_ctx *ctx, bcmc_tuning *out) {
    BCMC_CHECK_ARG(ctx && out, "null argument");
    *out = bcmc_tuning{};
    to_public(ctx, out);
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_set_tuning(bcmc_ctx *ctx, const bcmc_tuning *t) {
    BCMC_CHECK_ARG(ctx && t, "null argument");
    const uint32_t v[5] = {t->scan_wg_per_cu, t->sort_wg_per_cu, t->decode_wg_per_cu,
                           t->mc_wg_per_cu, t->count_wg_per_cu};
    for (uint32_t x : v)
        BCMC_CHECK_ARG(x >= 1 && x <= 64, "workgroups per CU must be in 1..64");
    BCMC_CHECK_ARG(t->scan_items == 0 || one_of(t->scan_items, 8, 16, 32), "scan_items must be 8, 16 or 32");
    BCMC_CHECK_ARG(t->sort_items == 0 || one_of(t->sort_items, 8, 16, 32), "sort_items must be 8, 16 or 32");
    BCMC_CHECK_ARG(t->decode_threads == 0 || one_of(t->decode_threads, 64, 128, 256),
                   "decode_threads must be 64, 128 or 256");
    if (t->scan_items)
        ctx->tune.scan_items = t->scan_items;
    if (t->sort_items)
        ctx->tune.sort_items = t->sort_items;
    if (t->decode_threads)
        ctx->tune.decode_threads = t->decode_threads;
    ctx->tune.scan = v[0];
    ctx->tune.sort = v[1];
    ctx->tune.decode = v[2];
    ctx->tune.mc = v[3];
    ctx->tune.count = v[4];
    ++ctx->tune_gen;
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_autotune(bcmc_ctx *ctx, bcmc_tuning *out) {
    BCMC_CHECK_ARG(ctx, "ctx is null");
    BCMC_HIP_CHECK(hipSetDevice(ctx->device));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    Events ev;
    BCMC_HIP_CHECK(hipEventCreate(&ev.a));
    BCMC_HIP_CHECK(hipEventCreate(&ev.b));
    TuneBufs B;
    bcmc_tuning res{};

    // scan: 16M elements through the single-pass scan
    constexpr uint32_t kScanN = 1u << 24;
    uint32_t *scan_data = nullptr;
    unsigned long long *total = nullptr;
    BCMC_TRY(B.alloc(&scan_data, (size_t)kScanN * 4));
    BCMC_TRY(B.alloc(&total, 8));
    ScanSite site;
    BCMC_TRY(scan_site_alloc(&site, kScanN));
    struct SiteGuard {
        ScanSite *s;
        ~SiteGuard() { scan_site_free(s); }
    } sg{&site};
    fill(ctx, scan_data, kScanN, 1, 0xFF, 0);
    // the sizes of a per-query active scan (a 512^3 volume has 2M blocks),
    // of small ones and of a batched step
    BCMC_TRY(pick(ctx, ev, ctx->tune.scan_items, ctx->tune.scan, grid_of({8, 16, 32}, {1, 2, 4, 8}),
                  {1u << 18, 1u << 21, kScanN},
                  [&](uint32_t n) { launch_scan<1>(ctx, site, site, TuneScanOp{scan_data, n, total}, n); },
                  &res.scan_ms));

    // radix sort: 2M pairs with 24-bit keys (3 digit passes, as an LRU order)
    constexpr uint32_t kSortN = 1u << 21;
    uint32_t *keys = nullptr, *vals = nullptr, *sorted = nullptr, *count = nullptr;
    BCMC_TRY(B.alloc(&keys, (size_t)kSortN * 4));
    BCMC_TRY(B.alloc(&vals, (size_t)kSortN * 4));
    BCMC_TRY(B.alloc(&sorted, (size_t)kSortN * 4));
    BCMC_TRY(B.alloc(&count, 8));
    fill(ctx, keys, kSortN, 2, 0xFFFFFF, 0);
    fill(ctx, vals, kSortN, 0, 0, 1);
    constexpr uint32_t kSortSmall = 1u << 16;
    const uint32_t counts[2] = {kSortSmall, kSortN};
    BCMC_HIP_CHECK(hipMemcpyAsync(count, counts, 8, hipMemcpyHostToDevice, ctx->stream));
    SortScratch ss;
    struct SortGuard {
        SortScratch *s;
        ~SortGuard() { sort_scratch_free(s); }
    } sog{&ss};
    BCMC_TRY(sort_scratch_alloc(&ss, kSortN));
    BCMC_TRY(pick(ctx, ev, ctx->tune.sort_items, ctx->tune.sort, grid_of({8, 16, 32}, {2, 4, 8, 16}),
                  {kSortSmall, kSortN},
                  [&](uint32_t n) {
                      launch_sort_pairs(ctx, &ss, keys, vals, count + (n == kSortN ? 1 : 0), n, sorted, 3);
                  },
                  &res.sort_ms));

    // zfp decode: 256K blocks of a synthetic rate-4 stream into a block cache
    constexpr uint32_t kDecN = 1u << 18, kMaxBits = 256;
    const uint64_t nwords = (uint64_t)kDecN * kMaxBits / 64;
    uint32_t *stream = nullptr, *slots = nullptr;
    float *cache = nullptr;
    BCMC_TRY(B.alloc(&stream, nwords * 8));
    BCMC_TRY(B.alloc(&slots, (size_t)kDecN * 4));
    BCMC_TRY(B.alloc(&cache, (size_t)kDecN * 64 * 4));
    fill(ctx, stream, nwords * 2, 3, 0xFFFFFFFFu, 0);
    fill(ctx, slots, kDecN, 0, 0, 1);
    // workgroups per CU scaled with the workgroup size: the same waves per CU
    std::vector<std::pair<uint32_t, uint32_t>> dec;
    for (uint32_t threads : {64u, 128u, 256u})
        for (uint32_t waves : {8u, 16u, 32u, 64u})
            dec.push_back({threads, waves / (threads / 64)});
    // the misses of a query with a cold cache, and of a mostly warm one
    BCMC_TRY(pick(ctx, ev, ctx->tune.decode_threads, ctx->tune.decode, dec, {1u << 15, kDecN},
                  [&](uint32_t n) {
                      launch_decode_to_cache(ctx, (const uint64_t *)stream, nwords, kMaxBits, nullptr,
                                             nullptr, n, (const int32_t *)slots, cache, kDecN);
                  },
                  &res.decode_ms));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ++ctx->tune_gen;
    if (out) {
        to_public(ctx, &res);
        *out = res;
    }
    return BCMC_OK;
}

}  // extern "C"
