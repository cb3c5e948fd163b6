// Decoded-block LRU cache on gfx950 (js/lru_cache.js + shaders/lru_cache_*.comp,
// lru_copy_available_slot_age.comp).
//
// One update runs as a chain of device-sized kernels with no host round trip:
//   age+mark   one pass over the C slots; a slot ages by one, and a slot whose
//              item is needed becomes {age 0, unavailable}, a slot whose item
//              is cached but not needed becomes available
//              (lru_cache_age_slots.comp + lru_cache_mark_new_items.comp; the
//              item->slot map is a bijection on cached items, so visiting
//              slots is the same as visiting cached items);
//   misses     single-pass scan+compact of {needed and not cached} -> new item
//              ids in ascending order (lru_cache.js:463-492);
//   available  single-pass scan+compact of the availability flags (:501,
//              :684-710);
//   plan/grow  the reference's growth rule (:510-672) on the device: slots
//              [C, C') get {100000, 1, -1} and are appended to the available
//              list; storage for alignTo(total, 32) slots is reserved up front;
//   order      the radix sort reproducing RadixSorter.sort(ages, ids, S,
//              reverse) order (:717-769, sort.hip);
//   assign     lru_cache_update.comp: evict the previous item of each chosen
//              slot and install the new item.
// Slot state is kept as three arrays (age, available, item) instead of the
// reference's 12-byte struct.
#include "common.h"
#include "kernels.h"
#include "lru.h"

#include <algorithm>

namespace bcmc {

constexpr uint32_t kUpdatesUnknown = 0xFFFFFFFFu;

__global__ void lru_init_kernel(LruView v, uint32_t first, uint32_t end, uint32_t age0,
                                int init_items, uint32_t *ctl, uint32_t ctl_value) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        ctl[0] = ctl_value;
    for (uint32_t s = first + blockIdx.x * blockDim.x + threadIdx.x; s < end;
         s += gridDim.x * blockDim.x) {
        v.age[s] = age0 == 0 ? 2 + s : age0;  // constructor ages 2+i (lru_cache.js:55-59)
        v.avail[s] = 1;
        v.item[s] = -1;
    }
    if (init_items)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < v.total;
             i += gridDim.x * blockDim.x)
            v.item_slot[i] = -1;
}

// misses over all items: needed (0/1) and not cached (lru_cache_mark_new_items.comp:28-43,
// lru_cache.js:463-492)
struct MissAllOp {
    LruView v;
    const uint32_t *needed;
    uint32_t *new_items;
    QState *q;
    BCMC_DEV uint32_t size() const { return v.total; }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&x)[1], Item &) const {
        const int32_t s = v.item_slot[i];
        const uint32_t nd = needed[i];
        x[0] = (uint32_t)(s < 0) & (uint32_t)(nd != 0);
    }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&x)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        if (x[0])
            new_items[e[0]] = i;
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { q->num_new = (uint32_t)t[0]; }
};

// Age + mark over the C slots (lru_cache_age_slots.comp:14-21,
// lru_cache_mark_new_items.comp:26-44) fused with the extraction of the
// available slots (lru_cache.js:501, :684-710): a slot ages by one; a slot
// whose item is needed becomes {age 0, unavailable}, a slot whose item is
// cached but not needed becomes available (the item->slot map is a bijection
// on cached items, so visiting slots visits the cached items).  Available
// slots are listed only when there is something to insert (q->num_new, known
// before this pass), as the reference returns before extracting them
// otherwise.  `gate`: skip the whole update when *gate == 0.
struct AgeAvailOp {
    LruView v;
    const uint32_t *needed;
    QState *q;
    const uint32_t *gate;
    uint32_t *ctl;  // [0] update chains since the reset (read only here), [1] set by total()
    BCMC_DEV uint32_t size() const { return (gate && *gate == 0) ? 0u : q->cache_size; }
    BCMC_DEV void begin() const {}
    struct Tile {
        bool inserting;  // q->num_new != 0
        uint32_t u;      // this chain's count for the order keys
    };
    struct Item {
        uint32_t age;
        uint32_t avail;  // bit 0: available; bit 1: the slot holds an item (avail is rewritten)
    };
    // this chain's count: the chains counted so far plus this one
    BCMC_DEV uint32_t chain_count() const {
        const uint32_t c = ctl[0];
        return c == kUpdatesUnknown ? c : c + 1;
    }
    BCMC_DEV Tile tile(uint32_t) const { return {q->num_new != 0, chain_count()}; }
    BCMC_DEV void value(const Tile &tc, uint32_t s, uint32_t (&x)[1], Item &it) const {
        const uint32_t age = v.age[s], avail = v.avail[s];
        const int32_t p = v.item[s];
        const bool cached = p >= 0 && (uint32_t)p < v.total;
        const uint32_t nd = needed[cached ? (uint32_t)p : 0u];
        const bool keep = cached & (nd == 1);  // needed: {age 0, unavailable}
        const uint32_t av = cached ? (keep ? 0u : 1u) : avail;
        it.avail = av | (cached ? 2u : 0u);
        it.age = keep ? 0u : age + 1;
        x[0] = (uint32_t)(av != 0) & (uint32_t)tc.inserting;
    }
    BCMC_DEV void emit(const Tile &tc, uint32_t s, const uint32_t (&x)[1], const Item &it,
                       const unsigned long long (&e)[1]) const {
        v.age[s] = it.age;
        if (it.avail & 2u)
            v.avail[s] = it.avail & 1u;
        if (x[0]) {
            v.avail_ids[e[0]] = s;
            v.avail_keys[e[0]] = lru_order_key(it.age, tc.u);
        }
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const {
        q->num_avail = q->num_new ? (uint32_t)t[0] : 0u;
        ctl[1] = lru_order_key(100000, chain_count());  // for the slots the growth appends
    }
};

// Growth (lru_cache.js:510-672) for the available count S of this update:
// when S < D, C' = alignTo(min(C + ceil(1.5 (D - S)), N), 32) (storage for
// alignTo(N, 32) slots is reserved up front); slots [C, C') become
// {100000, 1, -1} and are appended to the available list.  Every thread
// evaluates the rule; thread 0 of workgroup 0 records C' and S'.
__global__ void lru_grow_kernel(LruView v, QState *q, const uint32_t *gate, uint32_t *ctl) {
    const bool run = !(gate && *gate == 0) && q->num_new != 0;
    const uint32_t fresh_key = ctl[1];
    // one more update chain since the reset (counted whether or not the
    // update had work: the count only has to bound the updates from above);
    // nothing else in this launch reads ctl[0]
    if (blockIdx.x == 0 && threadIdx.x == 0 && ctl[0] != kUpdatesUnknown)
        ctl[0] += 1;
    const uint32_t c = q->cache_size, avail = run ? q->num_avail : 0u;
    uint32_t next = c;
    if (run && avail < q->num_new) {
        const uint64_t want = (uint64_t)c + ((uint64_t)(q->num_new - avail) * 3 + 1) / 2;
        uint64_t sz = std::min<uint64_t>(want, v.total);
        sz = (sz + 31) / 32 * 32;
        next = (uint32_t)std::min<uint64_t>(sz, v.capacity);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        q->cache_size_next = next;
        q->num_avail_total = avail + (next - c);
    }
    for (uint32_t s = c + blockIdx.x * blockDim.x + threadIdx.x; s < next;
         s += gridDim.x * blockDim.x) {
        v.age[s] = 100000;
        v.avail[s] = 1;
        v.item[s] = -1;
        v.avail_ids[avail + (s - c)] = s;
        v.avail_keys[avail + (s - c)] = fresh_key;
    }
}

// lru_cache_update.comp:19-37; thread 0 also commits the new size.
__global__ void lru_assign_kernel(LruView v, const uint32_t *new_items, QState *q) {
    const uint32_t n = q->num_avail_total ? q->num_new : 0u;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) {
        const uint32_t it = new_items[w], s = v.order[w];
        if (s >= v.capacity || it >= v.total)
            continue;  // cannot happen for a consistent state; never write out of bounds
        const int32_t prev = v.item[s];
        if (prev >= 0 && (uint32_t)prev < v.total)
            v.item_slot[prev] = -1;
        v.age[s] = 0;
        v.item[s] = (int32_t)it;
        v.avail[s] = 0;
        v.item_slot[it] = (int32_t)s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        q->avail_display = q->num_avail_total - n;
        q->cache_size = q->cache_size_next;
    }
}

static uint32_t grid_for(const bcmc_ctx *ctx, uint64_t n, uint32_t per_cu = 8) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(n, 256),
                                                              persistent_grid(ctx, per_cu)));
}

int lru_state_alloc(bcmc_ctx *ctx, LruState *L, uint32_t initial, uint32_t elem_bytes,
                    uint32_t total) {
    L->total = total;
    L->aligned_total = (total + 31) / 32 * 32;
    L->elem_bytes = elem_bytes;
    L->capacity = std::max<uint32_t>(L->aligned_total, (initial + 31) / 32 * 32);
    L->cache_size = (initial + 31) / 32 * 32;
    const size_t cap = L->capacity;
    BCMC_TRY(dev_alloc((void **)&L->v.age, cap * 4));
    BCMC_TRY(dev_alloc((void **)&L->v.avail, cap * 4));
    BCMC_TRY(dev_alloc((void **)&L->v.item, cap * 4));
    BCMC_TRY(dev_alloc((void **)&L->v.avail_ids, cap * 4));
    BCMC_TRY(dev_alloc((void **)&L->v.avail_keys, cap * 4));
    BCMC_TRY(dev_alloc((void **)&L->v.order, cap * 4));
    BCMC_TRY(dev_alloc((void **)&L->v.item_slot, (size_t)L->aligned_total * 4 + 4));

    BCMC_TRY(dev_alloc((void **)&L->new_items, (size_t)L->aligned_total * 4 + 4));
    BCMC_TRY(dev_alloc((void **)&L->ctl, 8));
    BCMC_TRY(scan_site_alloc(&L->site_miss, L->aligned_total));
    BCMC_TRY(scan_site_alloc(&L->site_avail, cap));
    BCMC_TRY(sort_scratch_alloc(&L->sort, L->capacity));
    L->v.total = total;
    L->v.capacity = L->capacity;
    L->age_bound = std::max<uint64_t>((uint64_t)L->cache_size + 1, 100000);
    L->updates = kUpdatesUnknown;  // constructor ages 2+i: no compressed order keys
    lru_init_kernel<<<grid_for(ctx, cap), 256, 0, ctx->stream>>>(L->v, 0, L->cache_size, 0, 1, L->ctl,
                                                                  kUpdatesUnknown);
    BCMC_HIP_CHECK(hipGetLastError());
    return lru_reserve(ctx, L, L->cache_size);
}

int lru_reserve(bcmc_ctx *ctx, LruState *L, uint64_t slots) {
    slots = std::min<uint64_t>(slots, L->capacity);
    if (L->data && slots <= L->slots)
        return BCMC_OK;
    uint64_t want = std::max<uint64_t>(slots, (uint64_t)L->slots + L->slots / 2);
    want = std::min<uint64_t>(L->capacity, (want + 31) / 32 * 32);
    uint8_t *d = nullptr;
    BCMC_TRY(dev_alloc((void **)&d, want * L->elem_bytes + 256));
    if (L->data && L->slots)
        BCMC_HIP_CHECK(hipMemcpyAsync(d, L->data, (size_t)L->slots * L->elem_bytes, hipMemcpyDeviceToDevice,
                                      ctx->stream));
    BCMC_HIP_CHECK(hipMemsetAsync(d + want * L->elem_bytes, 0, 256, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    dev_free(L->data);
    L->data = d;
    L->slots = (uint32_t)want;
    return BCMC_OK;
}

void lru_state_free(LruState *L) {
    dev_free(L->v.age);
    dev_free(L->v.avail);
    dev_free(L->v.item);
    dev_free(L->v.avail_ids);
    dev_free(L->v.avail_keys);
    dev_free(L->v.order);
    dev_free(L->v.item_slot);
    dev_free(L->data);
    dev_free(L->new_items);
    dev_free(L->ctl);
    scan_site_free(&L->site_miss);
    scan_site_free(&L->site_avail);
    sort_scratch_free(&L->sort);
    *L = LruState{};
}

// Radix-sort passes for the eviction order.  Keys are ~age (or ~compressed
// age after a reset, lru_order_key), and every key is below the bound (raw
// ages: initial 2+s, grown and reset slots 100000, +1 per update), so the
// 8-bit digits above the bound's are all 0xFF and their passes would keep the
// order.
uint32_t lru_begin_update(LruState *L) {
    L->age_bound += 1;
    if (L->updates != kUpdatesUnknown)
        L->updates += 1;
    // The device counts the same chains (lru_grow_kernel) and sorts
    // compressed keys (lru_order_key, at most 2u for u chains) while its count
    // is below kCompressedUpdates; the host counts first and switches to the
    // raw-age bound 0x1000 updates earlier, so its bound always covers the
    // device's keys.
    const uint64_t bound = L->updates < kCompressedUpdates - 0x1000 ? 2 * (uint64_t)L->updates + 1
                                                                     : L->age_bound;
    uint32_t bits = 0;
    while (bits < 32 && (bound >> bits) != 0)
        ++bits;
    return std::max<uint32_t>(1, (bits + 7) / 8);
}

// Age/mark + available list, growth and eviction order, once q->num_new and
// the new items are on the device; `ev` (nullable) receives an event after
// each of the three stages.
static void enqueue_tail(bcmc_ctx *ctx, LruState *L, const uint32_t *needed, QState *q,
                         const uint32_t *gate, uint32_t npasses, hipEvent_t *ev = nullptr) {
    hipStream_t st = ctx->stream;
    launch_scan<1>(ctx, L->site_avail, L->site_avail, AgeAvailOp{L->v, needed, q, gate, L->ctl}, L->capacity);
    if (ev)
        (void)hipEventRecord(ev[0], st);
    lru_grow_kernel<<<grid_for(ctx, L->capacity, 2), 256, 0, st>>>(L->v, q, gate, L->ctl);
    if (ev)
        (void)hipEventRecord(ev[1], st);
    launch_lru_order(ctx, &L->sort, L->v.avail_ids, L->v.avail_keys, &q->num_avail_total, L->capacity,
                     L->v.order, npasses);
    if (ev)
        (void)hipEventRecord(ev[2], st);
}

void lru_enqueue_update_mc(bcmc_ctx *ctx, LruState *L, const uint32_t *flags, QState *q,
                           uint32_t npasses) {
    enqueue_tail(ctx, L, flags, q, &q->num_active, npasses);
}

void lru_enqueue_update_all(bcmc_ctx *ctx, LruState *L, const uint32_t *needed, QState *q,
                            hipEvent_t *ev) {
    if (ev)
        (void)hipEventRecord(ev[0], ctx->stream);
    launch_scan<1>(ctx, L->site_miss, L->site_miss, MissAllOp{L->v, needed, L->new_items, q},
                   L->aligned_total);
    if (ev)
        (void)hipEventRecord(ev[1], ctx->stream);
    enqueue_tail(ctx, L, needed, q, nullptr, lru_begin_update(L), ev ? ev + 2 : nullptr);
    lru_assign_kernel<<<grid_for(ctx, L->capacity, 2), 256, 0, ctx->stream>>>(L->v, L->new_items, q);
    if (ev)
        (void)hipEventRecord(ev[5], ctx->stream);
}

void lru_enqueue_reset(bcmc_ctx *ctx, LruState *L) {
    L->updates = 0;
    lru_init_kernel<<<grid_for(ctx, std::max(L->capacity, L->aligned_total)), 256, 0, ctx->stream>>>(
        L->v, 0, L->cache_size, 100000, 1, L->ctl, 0u);
}

}  // namespace bcmc

using namespace bcmc;

// ---------------------------------------------------------------------------
// C ABI: LRUCache
// ---------------------------------------------------------------------------

static int lru_sync_state(bcmc_lru *l) {
    BCMC_HIP_CHECK(hipMemcpyAsync(l->qh, l->q, sizeof(QState), hipMemcpyDeviceToHost,
                                  l->ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(l->ctx->stream));
    l->st.cache_size = l->qh->cache_size;
    l->cache_view.bytes = (uint64_t)l->st.cache_size * l->st.elem_bytes;
    return BCMC_OK;
}

extern "C" {

BCMC_API int bcmc_lru_create(bcmc_ctx *ctx, uint32_t initial_size, uint32_t elem_bytes,
                             uint32_t total, bcmc_lru **out) {
    BCMC_CHECK_ARG(ctx && out, "null argument");
    BCMC_CHECK_ARG(total > 0 && total < 0x7FFFFFE0u, "totalElements out of range");
    BCMC_CHECK_ARG(elem_bytes > 0 && elem_bytes % 4 == 0, "elementSize must be a multiple of 4");
    *out = nullptr;
    bcmc_lru *l = new bcmc_lru;
    l->ctx = ctx;
    int rc = lru_state_alloc(ctx, &l->st, initial_size, elem_bytes, total);
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&l->q, sizeof(QState));
    if (rc == BCMC_OK && hipHostMalloc(&l->qh, sizeof(QState), hipHostMallocDefault) != hipSuccess) {
        set_error("pinned allocation failed");
        rc = BCMC_ERR_NOMEM;
    }
    if (rc != BCMC_OK) {
        lru_state_free(&l->st);
        dev_free(l->q);
        delete l;
        return rc;
    }
    QState init{};
    init.cache_size = l->st.cache_size;
    init.cache_size_next = l->st.cache_size;
    *l->qh = init;
    BCMC_HIP_CHECK(hipMemcpyAsync(l->q, l->qh, sizeof(QState), hipMemcpyHostToDevice, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    l->cache_view.ctx = ctx;
    l->cache_view.ptr = l->st.data;
    l->cache_view.bytes = (uint64_t)l->st.cache_size * elem_bytes;
    l->cache_view.owned = false;
    *out = l;
    return BCMC_OK;
}

BCMC_API int bcmc_lru_destroy(bcmc_lru *lru) {
    if (!lru)
        return BCMC_OK;
    if (!lru->owns) {
        bcmc::set_error("this LRUCache belongs to a CompressedMarchingCubes object");
        return BCMC_ERR_INVALID;
    }
    (void)hipStreamSynchronize(lru->ctx->stream);
    lru_state_free(&lru->st);
    dev_free(lru->q);
    (void)hipHostFree(lru->qh);
    for (hipEvent_t e : lru->ev)
        if (e)
            (void)hipEventDestroy(e);
    delete lru;
    return BCMC_OK;
}

BCMC_API int bcmc_lru_update(bcmc_lru *lru, bcmc_buf *needed, uint32_t *num_new,
                             bcmc_buf **new_items) {
    return bcmc_lru_update_timed(lru, needed, num_new, new_items, nullptr);
}

BCMC_API int bcmc_lru_update_timed(bcmc_lru *lru, bcmc_buf *needed, uint32_t *num_new,
                                   bcmc_buf **new_items, bcmc_lru_perf *perf) {
    BCMC_CHECK_ARG(lru && needed && num_new, "null argument");
    BCMC_CHECK_ARG(needed->bytes >= (uint64_t)lru->st.total * 4,
                   "itemNeeded must hold totalElements u32 flags");
    if (new_items)
        *new_items = nullptr;
    if (perf && !lru->ev[0])
        for (hipEvent_t &e : lru->ev)
            BCMC_HIP_CHECK(hipEventCreate(&e));
    const uint32_t size_before = lru->st.cache_size;
    lru_enqueue_update_all(lru->ctx, &lru->st, (const uint32_t *)needed->ptr, lru->q,
                           perf ? lru->ev : nullptr);
    BCMC_HIP_CHECK(hipGetLastError());
    BCMC_TRY(lru_sync_state(lru));
    *num_new = lru->qh->num_new;
    // storage for every slot of the (possibly grown) cache: its contents
    // belong to the caller, who may write any of them
    BCMC_TRY(lru_reserve(lru->ctx, &lru->st, lru->st.cache_size));
    lru->cache_view.ptr = lru->st.data;
    if (perf) {
        float t[5];
        for (int k = 0; k < 5; ++k)
            BCMC_HIP_CHECK(hipEventElapsedTime(&t[k], lru->ev[k], lru->ev[k + 1]));
        *perf = bcmc_lru_perf{};
        perf->lruNumNewItems = *num_new;
        if (*num_new == 0) {
            // as the reference: the stages after the count did not run (here
            // they ran as no-ops and aged the slots): their time goes to the
            // mark stage, which also ages in the reference
            perf->lruMarkNewItems = t[0] + t[1] + t[2] + t[3] + t[4];
        } else {
            perf->lruMarkNewItems = t[0];
            perf->lruScanSlotsAvailable = t[1];
            perf->lruNumSlotsAvailable = lru->qh->num_avail;
            perf->lruGrowCache = lru->st.cache_size != size_before ? t[2] : 0.0;
            perf->lruTotalSortTime = t[3];
            perf->lruWriteNewItems = t[4];
        }
    }
    if (*num_new && new_items) {
        bcmc_buf *b = nullptr;
        BCMC_TRY(make_buf(lru->ctx, (uint64_t)*num_new * 4, &b));
        BCMC_HIP_CHECK(hipMemcpyAsync(b->ptr, lru->st.new_items, (size_t)*num_new * 4,
                                      hipMemcpyDeviceToDevice, lru->ctx->stream));
        BCMC_HIP_CHECK(hipStreamSynchronize(lru->ctx->stream));
        *new_items = b;
    }
    return BCMC_OK;
}

BCMC_API int bcmc_lru_reset(bcmc_lru *lru) {
    BCMC_CHECK_ARG(lru, "null argument");
    lru_enqueue_reset(lru->ctx, &lru->st);
    BCMC_HIP_CHECK(hipGetLastError());
    BCMC_HIP_CHECK(hipStreamSynchronize(lru->ctx->stream));
    return BCMC_OK;
}

BCMC_API int bcmc_lru_get_info(bcmc_lru *lru, bcmc_lru_info *info) {
    BCMC_CHECK_ARG(lru && info, "null argument");
    BCMC_TRY(lru_sync_state(lru));
    info->cache_size = lru->st.cache_size;
    info->total_elements = lru->st.total;
    info->element_size = lru->st.elem_bytes;
    info->slots_available = lru->qh->avail_display;
    info->reserved_bytes = (uint64_t)lru->st.slots * lru->st.elem_bytes + 256;
    info->logical_bytes = (uint64_t)lru->st.cache_size * lru->st.elem_bytes;
    return BCMC_OK;
}

BCMC_API int bcmc_lru_read_state(bcmc_lru *lru, uint32_t *age, uint32_t *available, int32_t *item,
                                 int32_t *item_slot) {
    BCMC_CHECK_ARG(lru, "null argument");
    BCMC_TRY(lru_sync_state(lru));
    return lru_copy_state(lru->ctx, &lru->st, age, available, item, item_slot);
}

BCMC_API bcmc_buf *bcmc_lru_cache_buffer(bcmc_lru *lru) { return lru ? &lru->cache_view : nullptr; }

}  // extern "C"

namespace bcmc {
int lru_copy_state(bcmc_ctx *ctx, LruState *L, uint32_t *age, uint32_t *available, int32_t *item,
                   int32_t *item_slot) {
    const size_t c = L->cache_size;
    if (age)
        BCMC_HIP_CHECK(hipMemcpyAsync(age, L->v.age, c * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (available)
        BCMC_HIP_CHECK(hipMemcpyAsync(available, L->v.avail, c * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (item)
        BCMC_HIP_CHECK(hipMemcpyAsync(item, L->v.item, c * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (item_slot)
        BCMC_HIP_CHECK(hipMemcpyAsync(item_slot, L->v.item_slot, (size_t)L->total * 4,
                                      hipMemcpyDeviceToHost, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return BCMC_OK;
}
}  // namespace bcmc
