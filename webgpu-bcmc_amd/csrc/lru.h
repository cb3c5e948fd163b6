// Device state of the decoded-block LRU cache (see lru.hip).
#pragma once

#include "kernels.h"

namespace bcmc {

// Kernel-side view (passed by value).
struct LruView {
    uint32_t *age = nullptr;      // [capacity]
    uint32_t *avail = nullptr;    // [capacity]
    int32_t *item = nullptr;      // [capacity], -1 = empty
    int32_t *item_slot = nullptr; // [aligned_total], -1 = not cached (cachedItemSlots)
    uint32_t *avail_ids = nullptr;// [capacity] compacted available slots
    uint32_t *avail_keys = nullptr;// [capacity] their eviction-order keys (lru_order_key)
    uint32_t *order = nullptr;    // [capacity] available slots in eviction order
    uint32_t total = 0;           // totalElements
    uint32_t capacity = 0;        // reserved slots
};

struct LruState {
    LruView v;
    // cache contents: `slots` slots of elem_bytes and a zero slot after them
    // (absent blocks read it).  The storage follows the logical size: it
    // holds the cache size plus room for growth, and is enlarged (contents
    // copied) when a query grows the cache past it, up to `capacity`.
    uint8_t *data = nullptr;
    uint32_t slots = 0;
    uint32_t *new_items = nullptr;// [aligned_total] ids to decode, ascending
    uint32_t total = 0, aligned_total = 0, elem_bytes = 0, capacity = 0;
    uint32_t cache_size = 0;      // host mirror of the logical size
    uint64_t age_bound = 0;       // every slot age is below this (host-side bound)
    // ctl[0]: update chains since the last reset (kUpdatesUnknown before the
    // first one), advanced once per chain by the growth kernel; ctl[1]: the
    // order key of a fresh slot in the current chain.  `updates` is the host
    // mirror of ctl[0], advanced per enqueued update.  See lru_order_key.
    uint32_t *ctl = nullptr;
    uint32_t updates = 0xFFFFFFFFu;
    ScanSite site_miss, site_avail;
    SortScratch sort;
};

int lru_state_alloc(bcmc_ctx *ctx, LruState *L, uint32_t initial, uint32_t elem_bytes,
                    uint32_t total);
// Makes the storage hold at least `slots` slots (at most the capacity):
// a larger one, 1.5x at least, with the contents copied.  Synchronises the
// stream when it moves; callers refresh the pointers they keep.
int lru_reserve(bcmc_ctx *ctx, LruState *L, uint64_t slots);
void lru_state_free(LruState *L);
// LRUCache.update for a marching-cubes query whose misses the active pass has
// already listed (new_items, q->num_new): `flags` are the N active flags; does
// nothing when q->num_active == 0 (compressed_marching_cubes.js:630-641).  Ends
// with the eviction order in v.order; the slot assignment is done by the
// decoder (launch_decode_assign), which decodes each new item into its slot.
// npasses = lru_begin_update(L), called once per update before it is enqueued.
void lru_enqueue_update_mc(bcmc_ctx *ctx, LruState *L, const uint32_t *flags, QState *q,
                           uint32_t npasses);
// Radix passes for the eviction order given the current age bound.
// Counts one more update (age bound, updates since the reset) and returns the
// radix passes its eviction order needs.
uint32_t lru_begin_update(LruState *L);
// LRUCache.update for arbitrary 0/1 `needed` flags over all items.  With `ev`
// (6 events), ev[k] is recorded before stage k and ev[5] at the end: new-item
// pass, aging + available-slot pass, growth, eviction order, assignment.
void lru_enqueue_update_all(bcmc_ctx *ctx, LruState *L, const uint32_t *needed, QState *q,
                            hipEvent_t *ev = nullptr);
void lru_enqueue_reset(bcmc_ctx *ctx, LruState *L);
int lru_copy_state(bcmc_ctx *ctx, LruState *L, uint32_t *age, uint32_t *available, int32_t *item,
                   int32_t *item_slot);

}  // namespace bcmc

// C-ABI handle (include/bcmc.h).  A marching-cubes object hands out a
// non-owning handle over its own cache state (owns = false).
struct bcmc_lru {
    bcmc_ctx *ctx = nullptr;
    bcmc::LruState st;
    bcmc::QState *q = nullptr;   // device counters
    bcmc::QState *qh = nullptr;  // pinned host mirror
    bcmc_buf cache_view;         // non-owning view of the cache contents
    bool owns = true;
    hipEvent_t ev[6] = {};       // stage events of timed updates (created on first use)
};
