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
        uint32_t scan = 4, sort = 4, decode = 8, mc = 4;
    } tune;
    // bumped on every change of `tune`: captured graphs keyed on it re-capture
    uint64_t tune_gen = 0;
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

int scan_site_alloc(ScanSite *s, uint64_t max_elems, uint32_t tile);
void scan_site_free(ScanSite *s);

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 4096 elements

__host__ __device__ inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Persistent grid for device-sized loops: enough workgroups to fill the chip.
inline uint32_t persistent_grid(const bcmc_ctx *ctx, uint32_t per_cu) {
    return (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * per_cu;
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

// Decoupled look-back across tiles, done by wave 0 of the workgroup.  Each
// round reads the statuses of the 64 nearest unvisited predecessors at once
// (one agent-scope load per lane): if every predecessor up to and including
// the closest inclusive one is ready, their values are summed and the walk
// ends at that inclusive one; with no inclusive status in the window (all
// aggregates) the window moves back 64 tiles.  Returns the exclusive prefix
// of the tile (status values are exact below 2^40) and publishes the tile's
// inclusive value.
// The look-back walk of one channel, run by one whole wave: returns the
// exclusive prefix of `tile` and publishes its inclusive value.
BCMC_DEV unsigned long long lookback_wave(const ScanSite &site, uint32_t tile, uint32_t epoch,
                                          uint32_t aggregate) {
    const uint32_t lane = lane_id();
    const uint32_t ep = epoch & 0x3FFFFF;
    unsigned long long prefix = 0;
    if (tile == 0) {
        if (lane == 0)
            st_agent_u64(&site.status[0], pack_status(ep, kFlagInclusive, aggregate));
        return 0;
    }
    if (lane == 0)
        st_agent_u64(&site.status[tile], pack_status(ep, kFlagAggregate, aggregate));
    int64_t base = (int64_t)tile - 1;
    for (;;) {
        const int64_t t = base - (int64_t)lane;
        unsigned long long s = 0;
        uint32_t flag = kFlagInclusive;  // before tile 0: an inclusive zero
        if (t >= 0) {
            s = ld_agent_u64(&site.status[t]);
            flag = ((uint32_t)(s >> 42) == ep) ? (uint32_t)(s >> 40) & 3u : 0u;
        }
        const uint64_t incl = __ballot(flag == kFlagInclusive);
        const uint64_t wait = __ballot(flag == 0);
        const uint32_t first = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
        const uint64_t need = first == 64 ? ~0ull : ((2ull << first) - 1ull);
        if (wait & need) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const unsigned long long v = (lane <= first && t >= 0) ? (s & kValueMask) : 0ull;
        prefix += wave_sum_u64(v);
        if (first < 64)
            break;
        base -= 64;
    }
    if (lane == 0)
        st_agent_u64(&site.status[tile], pack_status(ep, kFlagInclusive, prefix + aggregate));
    return prefix;
}

// Decoupled look-back across tiles, done by wave 0 of the workgroup.  Each
// round reads the statuses of the 64 nearest unvisited predecessors at once
// (one agent-scope load per lane): if every predecessor up to and including
// the closest inclusive one is ready, their values are summed and the walk
// ends at that inclusive one; with no inclusive status in the window (all
// aggregates) the window moves back 64 tiles.  Returns the exclusive prefix
// of the tile (status values are exact below 2^40) and publishes the tile's
// inclusive value.
template <int NT>
BCMC_DEV unsigned long long tile_lookback(const ScanSite &site, uint32_t tile, uint32_t epoch,
                                          uint32_t aggregate, unsigned long long *lds_prefix) {
    if (threadIdx.x < 64) {
        const unsigned long long p = lookback_wave(site, tile, epoch, aggregate);
        if (threadIdx.x == 0)
            *lds_prefix = p;
    }
    __syncthreads();
    const unsigned long long p = *lds_prefix;
    __syncthreads();
    return p;
}

// Two channels at once: wave 0 walks channel 0, wave 1 channel 1.
template <int NT>
BCMC_DEV void tile_lookback2(const ScanSite &s0, const ScanSite &s1, uint32_t tile, uint32_t e0, uint32_t e1,
                             uint32_t agg0, uint32_t agg1, unsigned long long *lds_prefix /* [2] */,
                             unsigned long long &p0, unsigned long long &p1) {
    static_assert(NT >= 128, "two look-back waves");
    const uint32_t wave = threadIdx.x / 64;
    if (wave < 2) {
        const unsigned long long p = wave == 0 ? lookback_wave(s0, tile, e0, agg0)
                                               : lookback_wave(s1, tile, e1, agg1);
        if ((threadIdx.x & 63u) == 0)
            lds_prefix[wave] = p;
    }
    __syncthreads();
    p0 = lds_prefix[0];
    p1 = lds_prefix[1];
    __syncthreads();
}

// Dynamic tile index for persistent workgroups (broadcast through LDS).
BCMC_DEV uint32_t next_tile(const ScanSite &site, uint32_t *lds_tile) {
    if (threadIdx.x == 0)
        *lds_tile = atomic_add_u32(&site.ctrl[0], 1u);
    __syncthreads();
    const uint32_t t = *lds_tile;
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
