// Shared host/device helpers for the gfx950 kernels of the BCMC library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "bcmc.h"

// ---------------------------------------------------------------------------
// Host error handling
// ---------------------------------------------------------------------------
namespace bcmc {

void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *last_error();

}  // namespace bcmc

#define BCMC_HIP_CHECK(expr)                                                               \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            bcmc::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                            __FILE__, __LINE__);                                           \
            return BCMC_ERR_HIP;                                                           \
        }                                                                                  \
    } while (0)

#define BCMC_CHECK_ARG(cond, msg)                                                          \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            bcmc::set_error("invalid argument: %s", msg);                                  \
            return BCMC_ERR_INVALID;                                                       \
        }                                                                                  \
    } while (0)

#define BCMC_TRY(expr)                                                                     \
    do {                                                                                   \
        int rc_ = (expr);                                                                  \
        if (rc_ != BCMC_OK)                                                                \
            return rc_;                                                                    \
    } while (0)

// ---------------------------------------------------------------------------
// Host objects
// ---------------------------------------------------------------------------
struct bcmc_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    char name[128] = {0};
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_ms = 0.0;
    // small pinned staging area for counters read back after a query
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    // launch shapes: persistent workgroups per CU of each kernel family
    // (bcmc_ctx_autotune measures and sets them; results never depend on them)
    struct {
        uint32_t scan = 2, sort = 8, decode = 8, mc = 5, count = 8;
        // tile shapes: elements per thread of the scan (tile = 512 x items)
        // and of the radix sort (tile = 256 x items), blocks per workgroup
        // of the zfp decoders (one block per lane)
        uint32_t scan_items = 16, sort_items = 16, decode_threads = 256;
    } tune;
    // bumped on every change of `tune`: captured graphs keyed on it re-capture
    uint64_t tune_gen = 0;
    // query chains replayed from captured HIP graphs (bcmc_ctx_set_graph_replay)
    bool graph_replay = false;
};

struct bcmc_buf {
    bcmc_ctx *ctx = nullptr;
    void *ptr = nullptr;
    uint64_t bytes = 0;
    bool owned = true;
};

namespace bcmc {

// Device allocation helpers (hipMalloc, with an injected-failure hook for the
// failure-handling tests: BCMC_FAIL_ALLOC_AFTER=k fails the k-th allocation).
int dev_alloc(void **p, size_t bytes);
void dev_free(void *p);
int make_buf(bcmc_ctx *ctx, uint64_t bytes, bcmc_buf **out);


// ---------------------------------------------------------------------------
// Single-pass scan state (decoupled look-back).  One per call site; the
// kernels reset it themselves, so launches need no memset (graph friendly).
//   ctrl[0] tile counter, ctrl[1] finished-workgroup counter, ctrl[2] epoch
//   status[t]: bits 63..42 epoch, 41..40 flag (1 aggregate, 2 inclusive),
//              39..0 value
// ---------------------------------------------------------------------------
struct ScanSite {
    uint32_t *ctrl = nullptr;
    unsigned long long *status = nullptr;
    uint32_t max_tiles = 0;
};

// Look-back state for scans of up to max_elems elements at any tile shape.
int scan_site_alloc(ScanSite *s, uint64_t max_elems);
void scan_site_free(ScanSite *s);

// Scan tiles: 512 threads x 8, 16 or 32 elements (4096, 8192, 16384).
constexpr int kScanThreads = 512;
constexpr uint32_t kScanTileMin = kScanThreads * 8, kScanTileMax = kScanThreads * 32;
inline uint32_t scan_tile(const bcmc_ctx *ctx) { return kScanThreads * ctx->tune.scan_items; }

__host__ __device__ inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Persistent grid for device-sized loops: enough workgroups to fill the chip.
inline uint32_t persistent_grid(const bcmc_ctx *ctx, uint32_t per_cu) {
    return (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * per_cu;
}

// Workgroups of `kernel` that one CU holds at once (its registers and LDS);
// a persistent grid larger than that would leave workgroups waiting for
// others to finish.  Callers cache it in a function-local static.
inline uint32_t resident_per_cu(const void *kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return 1;
    }
    return (uint32_t)n;
}

}  // namespace bcmc

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
#define BCMC_DEV __device__ __forceinline__

namespace bcmc {

BCMC_DEV uint32_t lane_id() { return __lane_id(); }

// Inclusive wave64 prefix sum using DPP row shifts and the GFX9 row
// broadcasts (lane 15 -> row 1, lane 31 -> rows 2-3).
BCMC_DEV uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

BCMC_DEV uint32_t wave_read_lane(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

BCMC_DEV uint32_t wave_sum(uint32_t v) { return wave_read_lane(wave_incl_scan(v), 63); }

// Agent-scope relaxed atomics: coherent across the XCDs' L2 caches without
// the cache write-back / invalidate that acquire/release at agent scope
// implies.  Every word accessed this way is self-contained (a counter, or a
// status word carrying its own value and epoch), so no ordering with other
// memory is needed; data written before a kernel ends is visible to the next
// kernel on the stream.
BCMC_DEV unsigned long long ld_agent_u64(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
BCMC_DEV void st_agent_u64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
BCMC_DEV uint32_t ld_agent_u32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
BCMC_DEV void st_agent_u32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
BCMC_DEV uint32_t atomic_add_u32(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr unsigned long long kValueMask = (1ull << 40) - 1;
constexpr uint32_t kFlagAggregate = 1, kFlagInclusive = 2;

BCMC_DEV unsigned long long pack_status(uint32_t epoch, uint32_t flag, unsigned long long value) {
    return ((unsigned long long)(epoch & 0x3FFFFF) << 42) | ((unsigned long long)flag << 40) |
           (value & kValueMask);
}

// Workgroup-wide exclusive scan of one u32 per thread.  Returns the thread's
// exclusive prefix; *total receives the workgroup sum (all threads).
template <int NT>
BCMC_DEV uint32_t block_excl_scan(uint32_t v, uint32_t *lds_waves /* NT/64 + 1 */, uint32_t *total) {
    constexpr int NW = NT / 64;
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63)
        lds_waves[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const uint32_t t = lds_waves[w];
        before += (w < (int)wave) ? t : 0u;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

BCMC_DEV unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Decoupled look-back across tiles by the whole workgroup, for one or two
// channels.  Each step reads the statuses of the NT nearest unvisited
// predecessors at once (one agent-scope load per thread and channel): if
// every predecessor up to and including the closest inclusive one is ready,
// their values are summed and the walk ends at that inclusive one; with no
// inclusive status in the window (all aggregates) the window moves back NT
// tiles; when a needed predecessor has not published yet the step is read
// again.  Positions before tile 0 count as an inclusive zero.  Returns the
// exclusive prefixes of the tile (status values are exact below 2^40) and
// publishes the tile's inclusive values.  Called by every thread of the
// workgroup; all decisions are taken from LDS, so control flow is uniform.
template <int NT>
struct LookbackLds {
    unsigned long long incl[2][NT / 64], wait[2][NT / 64], part[2][NT / 64];
};

template <int NT, int NCH>
BCMC_DEV void block_lookback(const ScanSite &s0, const ScanSite &s1, uint32_t tile, uint32_t e0,
                             uint32_t e1, uint32_t agg0, uint32_t agg1, LookbackLds<NT> &sh,
                             unsigned long long &p0, unsigned long long &p1) {
    static_assert(NCH == 1 || NCH == 2, "one or two channels");
    constexpr int NW = NT / 64;
    constexpr uint32_t kWindow = NT;
    const ScanSite *site[2] = {&s0, &s1};
    const uint32_t ep[2] = {e0 & 0x3FFFFF, e1 & 0x3FFFFF};
    const uint32_t agg[2] = {agg0, agg1};
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64, pos = threadIdx.x;
    unsigned long long prefix[2] = {0, 0};
    if (tile == 0) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                st_agent_u64(&site[c]->status[0], pack_status(ep[c], kFlagInclusive, agg[c]));
        }
        p0 = p1 = 0;
        return;
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            st_agent_u64(&site[c]->status[tile], pack_status(ep[c], kFlagAggregate, agg[c]));
    }
    bool done[2] = {false, NCH < 2};
    int64_t base = (int64_t)tile - 1;  // the tile at window position 0
    while (!(done[0] && done[1])) {
        const int64_t t = base - (int64_t)pos;
        unsigned long long s[2] = {0, 0};
        uint32_t flag[2] = {kFlagInclusive, kFlagInclusive};
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (!done[c] && t >= 0) {
                s[c] = ld_agent_u64(&site[c]->status[t]);
                flag[c] = ((uint32_t)(s[c] >> 42) == ep[c]) ? (uint32_t)(s[c] >> 40) & 3u : 0u;
            }
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const uint64_t incl = __ballot(flag[c] == kFlagInclusive);
            const uint64_t wait = __ballot(flag[c] == 0);
            if (lane == 0) {
                sh.incl[c][wave] = incl;
                sh.wait[c][wave] = wait;
            }
        }
        __syncthreads();
        uint32_t first[2] = {kWindow, kWindow};
        bool retry = false;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (done[c])
                continue;
#pragma unroll
            for (int w = NW - 1; w >= 0; --w) {
                const uint64_t m = sh.incl[c][w];
                if (m)
                    first[c] = 64u * w + (uint32_t)__builtin_ctzll(m);
            }
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint32_t lo = 64u * w;
                const uint64_t need = first[c] >= lo + 64 ? ~0ull
                                      : first[c] < lo     ? 0ull
                                                          : ((2ull << (first[c] - lo)) - 1ull);
                retry |= (sh.wait[c][w] & need) != 0;
            }
        }
        if (retry) {
            __syncthreads();
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (done[c])
                continue;
            const unsigned long long v = (pos <= first[c] && t >= 0) ? (s[c] & kValueMask) : 0ull;
            const unsigned long long ws = wave_sum_u64(v);
            if (lane == 0)
                sh.part[c][wave] = ws;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (done[c])
                continue;
#pragma unroll
            for (int w = 0; w < NW; ++w)
                prefix[c] += sh.part[c][w];
            done[c] = first[c] < kWindow;
        }
        base -= kWindow;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            st_agent_u64(&site[c]->status[tile], pack_status(ep[c], kFlagInclusive, prefix[c] + agg[c]));
    }
    p0 = prefix[0];
    p1 = prefix[1];
}

// Dynamic tile index for persistent workgroups (broadcast through LDS).
BCMC_DEV uint32_t next_tile(const ScanSite &site, uint32_t *lds_tile) {
    if (threadIdx.x == 0)
        *lds_tile = atomic_add_u32(&site.ctrl[0], 1u);
    __syncthreads();
    const uint32_t t = __builtin_amdgcn_readfirstlane(*lds_tile);
    __syncthreads();
    return t;
}

// Called once per workgroup when it has no more tiles: the last workgroup of
// the grid resets the tile counter and advances the epoch (and clears the
// status words when the 22-bit epoch wraps), so the next launch starts clean.
BCMC_DEV void scan_site_finish(const ScanSite &site, uint32_t epoch) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t done = atomic_add_u32(&site.ctrl[1], 1u);
        if (done == gridDim.x - 1) {
            st_agent_u32(&site.ctrl[0], 0u);
            st_agent_u32(&site.ctrl[1], 0u);
            uint32_t next = (epoch + 1) & 0x3FFFFF;
            if (next == 0) {
                for (uint32_t t = 0; t < site.max_tiles; ++t)
                    st_agent_u64(&site.status[t], 0ull);
                next = 1;
            }
            st_agent_u32(&site.ctrl[2], next);
        }
    }
}

BCMC_DEV uint32_t scan_site_epoch(const ScanSite &site) { return ld_agent_u32(&site.ctrl[2]); }

}  // namespace bcmc
