// Batched and served forms of the marching-cubes path (no reference
// counterpart: the reference answers one query per computeSurface call,
// js/compressed_marching_cubes.js:580-809).
//
// An engine holds device-resident compressed volumes and one decoded-block
// pool shared by all of them.  A step answers many (volume, isovalue) queries
// at once; every stage is one launch whose persistent workgroups pull work
// from lists that concatenate all queries of the step:
//   rows     per query and row segment of kSegBlocks blocks: does the
//            isovalue lie in the segment's union of dilated block ranges?
//            single-pass scan + compaction -> candidate segments
//   active   over the blocks of the candidate segments: the active test
//            (one comparison against the block's dilated range), scan +
//            compaction -> (query, block) pairs in query order, ascending
//            blocks
//   mark     pairs whose block is pooled stamp its slot with the step epoch;
//            a missing block is listed once, by the first query of the step
//            (in query order) that has it active
//   alloc    misses take free pool slots, then the least recently stamped
//            slots not used by this or the previous step (radix-sorted by
//            stamp), evicting their blocks
//   decode   one launch per volume of the step (its zfp rate selects the
//            LDS-staged decoder variant)
//   mc       count and chunk-prefix launches over all pairs in chunks that
//            never span two queries; the host reads the exact vertex counts,
//            admits the leading queries whose vertices fit the output budget
//            and sizes the vertex buffer; then the emit launch over their
//            chunks.  Each query's vertices come out contiguous, in the same
//            order as a single-query computeSurface
// Query results do not depend on which other queries share the step or on
// the pool state: the decoded values of a block are a pure function of the
// stream.
#include "common.h"
#include "kernels.h"
#include "mc_device.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <thread>
#include <unordered_map>
#include <vector>

namespace bcmc {

struct EngVol {                 // device descriptor of a volume
    uint32_t nb[3];
    uint32_t n;
    const float2 *dilated;      // per block: range over its 3x3x3 neighbourhood (block_active)
    const float2 *seg_range;    // per row segment: union of its blocks' dilated ranges
    const float2 *tile;         // per block: tile range (launch_tile_ranges)
    FastDiv div_x, div_y;       // by nb[0], nb[1]
    int32_t *slot_map;          // block -> pool slot; -1 none
    uint32_t *miss_list;        // this step's missing blocks of the volume (decoder input)
};

struct EngQuery {
    uint32_t vol;
    float iso;
    uint32_t flag_base;         // first index of this query in the concatenated block arrays
    uint32_t n;
    uint32_t row_base;          // first index of this query in the concatenated row segments
                                // (a multiple of the scan tile: every tile has one query)
    uint32_t nb0, nsegx;        // the volume's blocks in x and row segments in x
    uint32_t nrows;             // the volume's row segments
    // the volume's arrays, copied here so that per-element work reads one descriptor
    const float2 *dilated;
    const float2 *seg_range;
    int32_t *slot_map;
    // flag bases of the earlier queries of this step on the same volume:
    // same_list[same_off .. same_off + same_cnt)
    uint32_t same_off, same_cnt;
};

struct EngCounters {
    uint32_t epoch;
    uint32_t num_active;
    uint32_t num_with;
    uint32_t num_miss;
    uint32_t used;              // pool slots ever handed out (high-water mark)
    uint32_t num_cand;          // eviction candidates
    uint32_t short_slots;       // misses that found no slot (admission failed)
    uint32_t num_rows;          // candidate row segments of all queries
    uint32_t num_chunks;        // marching-cubes chunks of all queries
    uint32_t pad;
    unsigned long long num_verts;
    uint32_t vol_miss[64];      // misses per volume
};

constexpr uint32_t kMaxVols = 64;

// Word N of the lane's 16-lane row (DPP row_newbcast:N).
template <int N>
BCMC_DEV uint32_t row_share(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + N, 0xF, 0xF, false);
}

constexpr uint32_t kNoPairs = 0xFFFFFFFFu;

__global__ void eng_begin_kernel(EngCounters *c, uint32_t *q_act_start, uint32_t nq) {
    if (threadIdx.x == 0) {
        c->epoch += 1;
        c->num_active = 0;
        c->num_with = 0;
        c->num_miss = 0;
        c->num_cand = 0;
        c->short_slots = 0;
        c->num_rows = 0;
        c->num_chunks = 0;
        c->num_verts = 0;
    }
    for (uint32_t v = threadIdx.x; v < kMaxVols; v += blockDim.x)
        c->vol_miss[v] = 0;
    for (uint32_t q = threadIdx.x; q < nq; q += blockDim.x)
        q_act_start[q] = kNoPairs;
}

// Block-row segments of kSegBlocks blocks along x.  Per volume, once at load:
// the union of the dilated block ranges (block_active) of each segment.  A
// block is active only if the isovalue lies in its dilated range, hence in its
// segment's.
constexpr uint32_t kSegBlocks = 16;

__global__ void eng_seg_range_kernel(const float2 *dilated, uint32_t nbx, uint32_t nsegx, uint32_t nsegs,
                                     float2 *out) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nsegs; s += gridDim.x * blockDim.x) {
        const uint32_t sx = s % nsegx, row = s / nsegx;
        const uint32_t x0 = sx * kSegBlocks, x1 = min(nbx, x0 + kSegBlocks);
        float lo = 1e30f, hi = -1e30f;
        const float2 *r = dilated + (uint64_t)row * nbx;
        for (uint32_t x = x0; x < x1; ++x) {
            const float2 v = r[x];
            lo = fminf(lo, v.x);
            hi = fmaxf(hi, v.y);
        }
        out[s] = make_float2(lo, hi);
    }
}

// Per volume, once at load: a 2-D histogram of the blocks' dilated ranges,
// bin(lo) * kRangeBins + bin(hi), where bin(x) is the number of the kRangeBins
// - 1 ascending edges that are <= x.  Only comparisons, so the host's bin of
// an isovalue agrees with it exactly (eng_union_bound).
constexpr uint32_t kRangeBins = 1024;

__global__ void eng_range_hist_kernel(const float2 *dilated, uint64_t n, const float *edges, uint32_t *hist) {
    __shared__ float e[kRangeBins - 1];
    for (uint32_t i = threadIdx.x; i < kRangeBins - 1; i += blockDim.x)
        e[i] = edges[i];
    __syncthreads();
    auto bin = [&](float x) {
        uint32_t lo = 0, hi = kRangeBins - 1;  // the answer lies in [lo, hi]
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;  // edges[mid] <= x: at least mid + 1 edges
            if (e[mid] <= x)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    };
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const float2 r = dilated[b];
        atomicAdd(&hist[bin(r.x) * kRangeBins + bin(r.y)], 1u);
    }
}

// A candidate row segment with everything its active test and output need,
// so that the per-block work reads one 32-byte descriptor and one range.
// SegActiveOp reads it as 8 words: dilated (0, 1), b0, qi, iso, nx, flag_base.
struct CandRow {
    const float2 *dilated;  // range of the segment's first block
    uint32_t b0;            // the segment's first block
    uint32_t qi;            // query
    float iso;
    uint32_t nx;            // blocks of the segment inside the volume (<= kSegBlocks)
    uint32_t flag_base;     // the query's first index in the flag words
    uint32_t pad;
};
static_assert(sizeof(CandRow) == 32 && offsetof(CandRow, b0) == 8 && offsetof(CandRow, qi) == 12 &&
                  offsetof(CandRow, iso) == 16 && offsetof(CandRow, nx) == 20 && offsetof(CandRow, flag_base) == 24,
              "SegActiveOp reads CandRow by word");

// Candidate segments of every query: iso within the segment's dilated range.
// Each query's segments start on a scan tile (the rows past its nrows are
// padding), so a tile's query is tile_q[tile].
struct CandSegOp {
    const EngQuery *qs;
    const uint32_t *tile_q;
    CandRow *cand;
    EngCounters *c;
    uint32_t nsegs;
    BCMC_DEV uint32_t size() const { return nsegs; }
    BCMC_DEV void begin() const {}
    struct Tile {
        uint32_t qi, row_base, nrows, nb0, nsegx, flag_base;
        float iso;
        const float2 *dilated, *seg_range;
    };
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t t) const {
        const uint32_t qi = tile_q[t];
        const EngQuery &q = qs[qi];
        return {qi, q.row_base, q.nrows, q.nb0, q.nsegx, q.flag_base, q.iso, q.dilated, q.seg_range};
    }
    BCMC_DEV void value(const Tile &tc, uint32_t r, uint32_t (&v)[1], Item &) const {
        const uint32_t seg = r - tc.row_base;
        const bool in = seg < tc.nrows;
        const float2 rr = tc.seg_range[in ? seg : 0u];
        v[0] = (uint32_t)in & (uint32_t)(tc.iso >= rr.x) & (uint32_t)(tc.iso <= rr.y);
    }
    BCMC_DEV void emit(const Tile &tc, uint32_t r, const uint32_t (&v)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        if (v[0]) {
            const uint32_t seg = r - tc.row_base;
            const uint32_t x0 = (seg % tc.nsegx) * kSegBlocks, row = seg / tc.nsegx;
            CandRow cr;
            cr.b0 = x0 + tc.nb0 * row;
            cr.dilated = tc.dilated + cr.b0;
            cr.qi = tc.qi;
            cr.iso = tc.iso;
            cr.nx = min(kSegBlocks, tc.nb0 - x0);
            cr.flag_base = tc.flag_base;
            cr.pad = 0;
            cand[e[0]] = cr;
        }
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { c->num_rows = (uint32_t)t[0]; }
};

// Active test (compute_block_active.comp:30-60, as block_active) over the
// blocks of the candidate segments (kSegBlocks elements each).  Output: the active
// (query, block) pairs in query order, ascending blocks; flag words stamped
// with the step epoch; the first pair of each query.
struct SegActiveOp {
    const CandRow *cand;
    uint32_t *flags, *act_q, *act_b, *q_act_start;
    EngCounters *c;
    BCMC_DEV uint32_t size() const { return c->num_rows * kSegBlocks; }
    BCMC_DEV void begin() const {}
    struct Tile {
        uint32_t epoch;
    };
    struct Item {
        uint32_t qf;  // query | 1 << 31 at the first block of the query's first segment
        uint32_t b;   // block
        uint32_t f;   // its flag word
    };
    BCMC_DEV Tile tile(uint32_t) const { return {c->epoch}; }
    // The 16 lanes of a row hold the 16 blocks of one candidate row k (n and
    // the scan's 64-element rounds are multiples of 16): lanes 0..7 of the row
    // load the 8 words of cand[k], lanes 8..15 those of cand[k - 1], and the
    // fields reach every lane of the row by DPP.  Positions past the end all
    // read the last row (words chosen by lane, so the addresses stay inside).
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[1], Item &it) const {
        const uint32_t k = i / kSegBlocks, j = i % kSegBlocks, l = lane_id() & 15u;
        const uint32_t kk = l < 8 ? k : (k ? k - 1 : 0u);
        const uint32_t w = reinterpret_cast<const uint32_t *>(cand + kk)[l & 7u];
        const uint32_t d_lo = row_share<0>(w), d_hi = row_share<1>(w), b0 = row_share<2>(w),
                       qi = row_share<3>(w), nx = row_share<5>(w), fb = row_share<6>(w),
                       prev_q = row_share<11>(w);
        const float iso = __uint_as_float(row_share<4>(w));
        const float2 *dil = reinterpret_cast<const float2 *>(((uint64_t)d_hi << 32) | d_lo);
        const bool in = j < nx;
        const float2 d = dil[in ? j : 0u];
        // bitwise: a short-circuit would let the load sink under a branch
        v[0] = (uint32_t)in & (uint32_t)block_active(d, iso);
        const bool first = (j == 0) & ((k == 0) | (prev_q != qi));
        it.qf = qi | (first ? 0x80000000u : 0u);
        it.b = b0 + j;
        it.f = fb + b0 + j;
    }
    BCMC_DEV void emit(const Tile &tc, uint32_t, const uint32_t (&v)[1], const Item &it,
                       const unsigned long long (&e)[1]) const {
        const uint32_t qi = it.qf & 0x7FFFFFFFu;
        if (it.qf >> 31)
            q_act_start[qi] = (uint32_t)e[0];
        if (v[0]) {
            act_q[e[0]] = qi;
            act_b[e[0]] = it.b;
            flags[it.f] = tc.epoch;
        }
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { c->num_active = (uint32_t)t[0]; }
};

// eviction candidates: pooled slots not stamped by this or the previous step
struct CandOp {
    const uint32_t *stamp;
    uint32_t *cand;
    EngCounters *c;
    uint32_t capacity;
    // only when the misses exceed the free room; only handed-out slots
    BCMC_DEV uint32_t size() const { return c->num_miss > capacity - c->used ? c->used : 0u; }
    BCMC_DEV void begin() const {}
    struct Tile {
        uint32_t epoch;
    };
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {c->epoch}; }
    BCMC_DEV void value(const Tile &tc, uint32_t s, uint32_t (&v)[1], Item &) const {
        v[0] = stamp[s] + 1 < tc.epoch ? 1u : 0u;
    }
    BCMC_DEV void emit(const Tile &, uint32_t s, const uint32_t (&v)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        if (v[0])
            cand[e[0]] = s;
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { c->num_cand = (uint32_t)t[0]; }
};

// keys for the eviction order: stamp (ascending = least recently used first)
__global__ void eng_cand_keys_kernel(const uint32_t *stamp, const uint32_t *cand,
                                     const EngCounters *c, uint32_t *keys) {
    const uint32_t n = c->num_cand;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        keys[i] = stamp[cand[i]];
}

// miss w takes free slot used+w while the pool has room, then the w'-th
// eviction candidate in LRU order; its old block leaves the pool
__global__ void eng_assign_kernel(const EngVol *vols, EngCounters *c, const unsigned long long *miss,
                                  const uint32_t *order, uint32_t capacity, uint32_t *stamp,
                                  unsigned long long *owner) {
    const uint32_t m = c->num_miss, used = c->used, ep = c->epoch;
    const uint32_t room = capacity - used;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < m; w += gridDim.x * blockDim.x) {
        const unsigned long long pr = miss[w];
        const uint32_t vi = (uint32_t)(pr >> 32), b = (uint32_t)pr;
        uint32_t s;
        if (w < room) {
            s = used + w;
        } else {
            const uint32_t k = w - room;
            if (k >= c->num_cand) {
                atomic_add_u32(&c->short_slots, 1u);
                vols[vi].slot_map[b] = -1;
                continue;
            }
            s = order[k];
            const unsigned long long old = owner[s];
            if (old != ~0ull)
                vols[(uint32_t)(old >> 32)].slot_map[(uint32_t)old] = -1;
        }
        owner[s] = pr;
        stamp[s] = ep;
        vols[vi].slot_map[b] = (int32_t)s;
    }
}

__global__ void eng_commit_kernel(EngCounters *c, uint32_t capacity) {
    const uint32_t room = capacity - c->used;
    c->used += min(room, c->num_miss);
}

struct EngMcArgs {
    const EngQuery *qs;
    const EngVol *vols;
    const uint32_t *act_q, *act_b;
    const uint32_t *flags;
    const float *pool;
    uint32_t capacity;
    EngCounters *c;
    uint32_t nq;
    const uint32_t *same_list;           // see EngQuery::same_off
    const uint32_t *q_act_start;         // per query: first pair (kNoPairs: none)
    const uint32_t *q_pair_end;          // per query: one past its last pair
    const uint32_t *q_chunk_base;        // per query: its first chunk
    uint32_t *counts;                    // per pair: vertices
    uint32_t *chunk_verts;               // per chunk: vertices
    unsigned long long *chunk_verts_pre; // exclusive prefixes of chunk_verts
    // chunks the emit launch writes: set by the host per step, in
    // host-coherent pinned memory (device alias)
    const uint32_t *emit_chunks;
    void *verts;
    unsigned long long vert_cap;
    int fmt;
};

// The RowBlock of a (query, block) pair: its query's volume, flags, isovalue.
BCMC_DEV void eng_row_block(const EngMcArgs &a, uint32_t qi, uint32_t bid, bool live, uint32_t epoch,
                            RowBlock &b) {
    const EngQuery &q = a.qs[qi];
    const EngVol &v = a.vols[q.vol];
    b.cache = a.pool;
    b.item_slot = v.slot_map;
    b.flags = a.flags + q.flag_base;
    b.nb[0] = v.nb[0];
    b.nb[1] = v.nb[1];
    b.nb[2] = v.nb[2];
    b.div_x = v.div_x;
    b.div_y = v.div_y;
    b.cap = a.capacity;
    b.tag = epoch;  // flag words hold the epoch of the step that found the block active
    b.iso = q.iso;
    b.live = live;
    b.bid = bid;
}

constexpr int kEngThreads = kMcThreads;
// Blocks of all queries of one step (u32 indices into the concatenated arrays).
constexpr uint64_t kMaxStepBlocks = 1ull << 31;

// Chunks of the marching-cubes pass never cross a query: query q's pairs
// [start, end) form ceil((end - start) / kChunk) chunks from q_chunk_base[q]
// on.  One workgroup: per query the end of its pairs (the next query's first
// pair, or the step's total) and the exclusive scan of the chunk counts.
__global__ void eng_chunk_table_kernel(EngCounters *c, const uint32_t *q_act_start, uint32_t nq,
                                       uint32_t *q_pair_end, uint32_t *q_chunk_base) {
    constexpr int NT = 256;
    __shared__ uint32_t lds_w[NT / 64 + 1];
    __shared__ uint32_t lds_next;
    const uint32_t t = threadIdx.x, lane = lane_id(), wave = t / 64;
    const uint32_t n = c->num_active;
    if (t == 0)
        lds_next = n;
    __syncthreads();
    // backwards in blocks of NT queries: end = min start of the later queries
    for (uint32_t hi = nq; hi > 0;) {
        const uint32_t lo = hi > NT ? hi - NT : 0u;
        const uint32_t q = lo + t;
        const uint32_t st = q < hi ? q_act_start[q] : kNoPairs;
        // suffix minimum over the block (exclusive), carried from later blocks
        uint32_t v = st;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = (uint32_t)__shfl_down((int)v, o, 64);
            if (lane + o < 64)
                v = min(v, u);
        }
        if (lane == 0)
            lds_w[wave] = v;  // minimum of the wave
        __syncthreads();
        uint32_t later = lds_next;
        for (uint32_t w = wave + 1; w < NT / 64; ++w)
            later = min(later, lds_w[w]);
        const uint32_t after = (uint32_t)__shfl_down((int)v, 1, 64);
        const uint32_t excl = min(later, lane + 1 < 64 ? after : 0xFFFFFFFFu);
        if (q < hi)
            q_pair_end[q] = st == kNoPairs ? st : excl;
        __syncthreads();
        if (t == 0)
            lds_next = min(lds_next, min(min(lds_w[0], lds_w[1]), min(lds_w[2], lds_w[3])));
        __syncthreads();
        hi = lo;
    }
    __syncthreads();
    // forward: chunk bases
    uint32_t carry = 0;
    for (uint32_t lo = 0; lo < nq; lo += NT) {
        const uint32_t q = lo + t;
        uint32_t k = 0;
        if (q < nq) {
            const uint32_t st = q_act_start[q];
            k = st == kNoPairs ? 0u : div_up(q_pair_end[q] - st, kChunk);
        }
        uint32_t total;
        const uint32_t ex = block_excl_scan<NT>(k, lds_w, &total);
        if (q < nq)
            q_chunk_base[q] = carry + ex;
        carry += total;
    }
    if (t == 0)
        c->num_chunks = carry;
}

// The query of chunk ch: the last query whose first chunk is <= ch
// (workgroup-uniform binary search).
BCMC_DEV uint32_t eng_query_of_chunk(const EngMcArgs &a, uint32_t ch) {
    uint32_t lo = 0, hi = a.nq;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.q_chunk_base[mid] <= ch)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Pairs [k0, k1) of chunk ch of query q.
BCMC_DEV void eng_chunk_range(const EngMcArgs &a, uint32_t ch, uint32_t q, uint32_t &k0, uint32_t &k1) {
    k0 = a.q_act_start[q] + (ch - a.q_chunk_base[q]) * kChunk;
    k1 = min(a.q_pair_end[q], k0 + kChunk);
}

// Stamp pooled blocks and list the missing ones, each once: a missing block
// is claimed by the first query of the step (in query order) that has it
// active, i.e. by a pair whose earlier same-volume queries all have the
// block's flag word below this step's epoch.  Each thread takes kMarkPairs
// pairs whose loads are all issued before any is used.  List entries are
// reserved per workgroup and iteration: ranks from LDS counters per volume,
// then one global atomic per volume present (and one for the step list), so
// that cold steps with millions of misses do not serialise on two counters.
constexpr uint32_t kMarkPairs = 4;

__global__ __launch_bounds__(256) void eng_mark_kernel(EngMcArgs a, uint32_t *stamp, unsigned long long *miss,
                                                       uint32_t miss_cap) {
    __shared__ uint32_t hist[kMaxVols], vbase[kMaxVols], lbase[kMaxVols];
    const uint32_t n = a.c->num_active, ep = a.c->epoch;
    const uint32_t span = blockDim.x * kMarkPairs;
    for (uint32_t base = blockIdx.x * span; base < n; base += gridDim.x * span) {
        if (threadIdx.x < kMaxVols)
            hist[threadIdx.x] = 0;
        uint32_t qi[kMarkPairs], b[kMarkPairs], vi[kMarkPairs], rank[kMarkPairs];
        int32_t *sm[kMarkPairs];
        int32_t sl[kMarkPairs];
        uint32_t so[kMarkPairs], sc[kMarkPairs];
        bool claimed[kMarkPairs];
#pragma unroll
        for (uint32_t j = 0; j < kMarkPairs; ++j) {
            const uint32_t k = base + j * blockDim.x + threadIdx.x;
            qi[j] = k < n ? a.act_q[k] : 0u;
            b[j] = k < n ? a.act_b[k] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kMarkPairs; ++j) {
            const EngQuery &q = a.qs[qi[j]];
            sm[j] = q.slot_map;
            so[j] = q.same_off;
            sc[j] = q.same_cnt;
            vi[j] = q.vol;
        }
#pragma unroll
        for (uint32_t j = 0; j < kMarkPairs; ++j) {
            const uint32_t k = base + j * blockDim.x + threadIdx.x;
            // plain load: slots of pooled blocks do not change during the kernel
            sl[j] = k < n ? sm[j][b[j]] : -3;
        }
        __syncthreads();  // hist cleared
#pragma unroll
        for (uint32_t j = 0; j < kMarkPairs; ++j) {
            claimed[j] = false;
            if (sl[j] >= 0) {
                stamp[sl[j]] = ep;
            } else if (sl[j] == -1) {
                claimed[j] = true;
                for (uint32_t t = 0; t < sc[j] && claimed[j]; ++t)
                    claimed[j] = a.flags[a.same_list[so[j] + t] + b[j]] != ep;
            }
            if (claimed[j])
                rank[j] = atomicAdd(&hist[vi[j]], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t v = 0; v < kMaxVols; ++v) {
                lbase[v] = tot;
                tot += hist[v];
            }
            const uint32_t g = tot ? atomic_add_u32(&a.c->num_miss, tot) : 0u;
            for (uint32_t v = 0; v < kMaxVols; ++v)
                lbase[v] += g;
        }
        if (threadIdx.x < kMaxVols && hist[threadIdx.x])
            vbase[threadIdx.x] = atomic_add_u32(&a.c->vol_miss[threadIdx.x], hist[threadIdx.x]);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kMarkPairs; ++j) {
            if (!claimed[j])
                continue;
            const uint32_t w = lbase[vi[j]] + rank[j];
            if (w < miss_cap)
                miss[w] = ((unsigned long long)vi[j] << 32) | b[j];
            const uint32_t wv = vbase[vi[j]] + rank[j];
            if (wv < a.vols[vi[j]].n)
                a.vols[vi[j]].miss_list[wv] = b[j];
        }
        __syncthreads();  // LDS reused by the next iteration
    }
}

// The count / prefix / emit launches of mc.hip over (query, block) pairs of
// many queries, with workgroup-uniform volume views (every chunk lies inside
// one query); vertex offsets are global, and each query's vertices are
// contiguous because pairs and chunks are in query order.
__global__ __launch_bounds__(kEngThreads) void eng_count_kernel(EngMcArgs a) {
    constexpr int NW = kEngThreads / 64;
    __shared__ CountShared cs;
    __shared__ uint32_t lds_w[NW + 1];
    __shared__ CountTables tab;
    load_count_tables(&tab);
    const uint32_t nchunks = a.c->num_chunks, epoch = a.c->epoch;
    const uint32_t wave = threadIdx.x / 64;
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint32_t q = eng_query_of_chunk(a, chunk);
        uint32_t k0, k1;
        eng_chunk_range(a, chunk, q, k0, k1);
        RowBlock ub;
        eng_row_block(a, q, 0, true, epoch, ub);
        const float2 *tile = a.vols[a.qs[q].vol].tile;
        const bool in = k0 + threadIdx.x < k1;
        const uint32_t bid = in ? a.act_b[k0 + threadIdx.x] : 0u;
        const uint32_t ncand = chunk_prologue(cs, in, bid, q, in ? tile[bid] : make_float2(1.0f, 0.0f), ub.iso,
                                              lds_w);
        count_sweep(cs, ncand, wave, tab, [&](uint32_t e, RowBlock &b) {
            b = ub;
            b.live = k0 + e < k1;
            b.bid = cs.bid[e];
        });
        __syncthreads();
        const uint32_t c = cs.cnt[threadIdx.x];
        if (in)
            a.counts[k0 + threadIdx.x] = c;
        uint32_t agg_v;
        block_excl_scan<kEngThreads>(c, lds_w, &agg_v);
        if (threadIdx.x == 0)
            a.chunk_verts[chunk] = agg_v;
    }
}

struct EngChunkPrefixOp {
    const uint32_t *verts;
    unsigned long long *verts_pre;
    EngCounters *c;
    BCMC_DEV uint32_t size() const { return c->num_chunks; }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[1], Item &) const { v[0] = verts[i]; }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        verts_pre[i] = e[0];
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { c->num_verts = t[0]; }
};

// What the host reads after the step's front chain, written by its last
// launch straight into host-coherent pinned memory (device aliases of the
// host buffers), so one stream synchronisation and no copy launches bring it
// to the host.
struct EngReadback {
    EngCounters *c;                   // the step's counters
    uint32_t *act_start;              // per query: first active pair (kNoPairs: none)
    unsigned long long *vfirst;       // per query: its first vertex (~0: no active blocks)
    uint32_t *chunk_base;             // per query: its first chunk
};

// Where each query's vertices start (its first chunk's prefix), with the
// query's first pair and chunk and the step's counters, to the host.
__global__ void eng_vfirst_kernel(EngMcArgs a, EngReadback rb) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < a.nq; q += gridDim.x * blockDim.x) {
        const uint32_t s = a.q_act_start[q], cb = a.q_chunk_base[q];
        rb.vfirst[q] = s == kNoPairs ? ~0ull : a.chunk_verts_pre[cb];
        rb.act_start[q] = s;
        rb.chunk_base[q] = cb;
    }
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < sizeof(EngCounters) / 4; i += blockDim.x)
            reinterpret_cast<uint32_t *>(rb.c)[i] = reinterpret_cast<const uint32_t *>(a.c)[i];
}

template <int FMT>
__global__ __launch_bounds__(kEngThreads) __attribute__((amdgpu_waves_per_eu(5))) void eng_emit_kernel(EngMcArgs a) {
    constexpr int NW = kEngThreads / 64;
    __shared__ EmitShared cs;
    __shared__ uint32_t lds_w[NW + 1];
    __shared__ McTables tab;
    __shared__ EmitScratch es[NW];
    load_mc_tables(&tab);
    // the host admits the step's leading queries whose vertices fit the
    // output budget: only their chunks are written
    const uint32_t nchunks = min(a.c->num_chunks, *a.emit_chunks), epoch = a.c->epoch;
    const uint32_t wave = threadIdx.x / 64;
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint32_t q = eng_query_of_chunk(a, chunk);
        uint32_t k0, k1;
        eng_chunk_range(a, chunk, q, k0, k1);
        const uint32_t k = k0 + threadIdx.x;
        const bool in = k < k1;
        const uint32_t c = in ? a.counts[k] : 0u;
        const uint32_t bid = in ? a.act_b[k] : 0u;
        uint32_t ex_h, ex_v, agg_v;
        const uint32_t num_with = emit_prologue(cs, c, bid, q, lds_w, &ex_h, &ex_v, &agg_v);
        const unsigned long long pv = a.chunk_verts_pre[chunk];
        if (num_with == 0)
            continue;
        RowBlock ub;
        eng_row_block(a, q, 0, true, epoch, ub);
        emit_sweep<FMT>(cs, wave, num_with, pv, tab, es[wave], a.vert_cap, a.verts,
                        [&](uint32_t e, RowBlock &b) {
                            b = ub;
                            b.live = k0 + e < k1;
                            b.bid = cs.bid[e];
                        });
        __syncthreads();
    }
}

static uint32_t vbytes(int fmt) {
    return fmt == BCMC_VERTS_LOCAL_F32 ? 16u : fmt == BCMC_VERTS_GLOBAL_F32 ? 12u : 8u;
}

}  // namespace bcmc

using namespace bcmc;

// ---------------------------------------------------------------------------
// Host objects
// ---------------------------------------------------------------------------
struct EngVolume {
    uint32_t dims[3], nb[3], n, maxbits;
    uint64_t *stream = nullptr;
    uint64_t nwords = 0;
    float2 *ranges = nullptr, *dilated = nullptr, *seg_range = nullptr, *tile = nullptr;
    int32_t *slot_map = nullptr;
    uint32_t *miss_list = nullptr;  // per-step block ids to decode (EngVol::miss_list)
    // bins of the dilated ranges (eng_range_hist_kernel): the kRangeBins - 1
    // edges and below[a * (kRangeBins + 1) + b] = blocks with bin(lo) >= a and
    // bin(hi) < b, a in [0, kRangeBins], b in [0, kRangeBins]
    std::vector<float> edges;
    std::vector<uint32_t> below;
    uint32_t bin(float x) const {
        return (uint32_t)(std::upper_bound(edges.begin(), edges.end(), x) - edges.begin());
    }
    // blocks strictly between the isovalues of bins a and c (a = -1: below
    // every isovalue; c = kRangeBins: above): bin(lo) > a and bin(hi) < c, so
    // lo and hi lie strictly between the two isovalues and no isovalue in
    // bins <= a or >= c makes them active
    uint64_t gap(int a, int c) const {
        if (c - a < 2)
            return 0;
        return below[(uint64_t)(a + 1) * (kRangeBins + 1) + (uint32_t)c];
    }
};

// The blocks that queries at a set of isovalues on one volume can touch
// together: every block but those whose dilated range lies strictly between
// two neighbouring isovalues of the set, or strictly below or above all of
// them (counted from the range bins, so only blocks that are inactive for
// every isovalue are left out).  Kept per bin: isovalues in one bin leave the
// same blocks out.
struct EngUnion {
    std::vector<int> bins;  // ascending, distinct
    uint64_t inactive = 0;  // blocks left out
    bool empty() const { return bins.empty(); }
    // blocks touched after adding isovalue bin b (the set itself unchanged)
    uint64_t with(const EngVolume &v, int b, uint64_t *inactive_after) const {
        const auto it = std::lower_bound(bins.begin(), bins.end(), b);
        uint64_t in = bins.empty() ? v.n : inactive;
        if (it == bins.end() || *it != b) {
            const int a = it == bins.begin() ? -1 : *(it - 1);
            const int c = it == bins.end() ? (int)kRangeBins : *it;
            in = in - (bins.empty() ? v.n : v.gap(a, c)) + v.gap(a, b) + v.gap(b, c);
        }
        *inactive_after = in;
        return v.n - std::min<uint64_t>(v.n, in);
    }
    void add(int b, uint64_t inactive_after) {
        const auto it = std::lower_bound(bins.begin(), bins.end(), b);
        if (it == bins.end() || *it != b)
            bins.insert(it, b);
        inactive = inactive_after;
    }
};

using EngKey = std::array<uint64_t, 6>;
struct EngGraph {
    EngKey key;
    hipGraph_t graph;
    hipGraphExec_t exec;
    uint64_t last_use;
};
constexpr size_t kEngGraphs = 16, kEngSeen = 64;

struct bcmc_engine {
    bcmc_ctx *ctx = nullptr;
    std::vector<EngVolume> vols;
    EngVol *d_vols = nullptr;
    uint32_t pool_cap = 0;
    float *pool = nullptr;
    uint32_t *stamp = nullptr, *cand = nullptr, *cand_keys = nullptr, *order = nullptr;
    unsigned long long *owner = nullptr, *miss = nullptr;
    uint32_t miss_cap = 0;
    SortScratch sort;
    EngCounters *c = nullptr, *ch = nullptr, *ch_d = nullptr;
    // per-step arrays (grow)
    uint64_t blocks_cap = 0;
    uint32_t *flags = nullptr, *act_q = nullptr, *act_b = nullptr;
    uint32_t *counts = nullptr, *chunk_verts = nullptr;
    unsigned long long *chunk_verts_pre = nullptr;
    uint64_t chunks_cap = 0;
    ScanSite site_chunks;
    uint64_t rows_cap = 0, row_elems_cap = 0;
    CandRow *cand_rows = nullptr;
    ScanSite site_rows;
    uint32_t q_cap = 0;
    EngQuery *d_q = nullptr;
    uint32_t *q_act_start = nullptr, *q_pair_end = nullptr, *q_chunk_base = nullptr;
    uint32_t *same_list = nullptr;
    uint64_t same_cap = 0;
    uint32_t *tile_q = nullptr, *h_tile_q = nullptr;  // query of each scan tile's first row segment
    uint64_t tile_q_cap = 0;
    // pinned host staging of the per-step query tables; the per-query
    // results and the counters (ch) are host-coherent, written by the device
    // through their aliases (_d), as is the emit's chunk count (h_emit)
    EngQuery *h_q = nullptr;
    uint32_t *h_same = nullptr, *h_act = nullptr, *h_act_d = nullptr;
    unsigned long long *h_vfirst = nullptr, *h_vfirst_d = nullptr;
    uint32_t *h_chunk_base = nullptr, *h_chunk_base_d = nullptr;
    uint32_t *h_emit = nullptr, *h_emit_d = nullptr;
    ScanSite site_active, site_cand;
    void *verts = nullptr;
    uint64_t vert_cap = 0;
    int fmt = BCMC_VERTS_LOCAL_F32;
    bcmc_buf vert_view;
    uint64_t budget = 0;
    // replay of the step's kernel chain from captured HIP graphs, with the
    // context's graph replay on: a step shape (key in bcmc_engine_run) is
    // captured the second time it is seen among the last kEngSeen steps, and
    // replayed whenever it comes again; the kEngGraphs most recently used
    // graphs are kept.  gen is bumped whenever a buffer the chain reads is
    // replaced, which also drops every graph.  use_graph turns off when a
    // capture fails.
    bool use_graph = true;
    uint64_t gen = 0;
    std::vector<EngGraph> graphs;
    std::vector<EngKey> seen;
    size_t seen_next = 0;
    uint64_t clock = 0;
    uint32_t steps = 0, graph_steps = 0;
    // last step
    std::vector<bcmc_query_result> last;
};

static void eng_drop_graph(EngGraph &g) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
}

static void eng_drop_graphs(bcmc_engine *e) {
    for (EngGraph &g : e->graphs)
        eng_drop_graph(g);
    e->graphs.clear();
    e->seen.clear();
    e->seen_next = 0;
}

// a buffer the step chain reads was replaced (the caller has synchronised)
static void eng_changed(bcmc_engine *e) {
    ++e->gen;
    eng_drop_graphs(e);
}

template <class T> static void eng_release(T *&p) {
    dev_free((void *)p);
    p = nullptr;
}

template <class T> static void eng_release_host(T *&p) {
    if (p)
        (void)hipHostFree((void *)p);
    p = nullptr;
}

// Growing a buffer first synchronises, drops the captured chain and zeroes
// the capacity, so an allocation failure part-way leaves no dangling pointer
// and the next step allocates again.
static int eng_ensure_blocks(bcmc_engine *e, uint64_t total) {
    if (total <= e->blocks_cap)
        return BCMC_OK;
    BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
    eng_changed(e);
    const uint64_t cap = std::max<uint64_t>(total, e->blocks_cap * 3 / 2);
    e->blocks_cap = 0;
    eng_release(e->flags);
    eng_release(e->act_q);
    eng_release(e->act_b);
    eng_release(e->counts);
    BCMC_TRY(dev_alloc((void **)&e->flags, cap * 4));
    // flag words hold the epoch of the step that found the block active;
    // zero is no step's epoch
    BCMC_HIP_CHECK(hipMemsetAsync(e->flags, 0, cap * 4, e->ctx->stream));
    BCMC_TRY(dev_alloc((void **)&e->act_q, cap * 4));
    BCMC_TRY(dev_alloc((void **)&e->act_b, cap * 4));
    BCMC_TRY(dev_alloc((void **)&e->counts, cap * 4));
    e->blocks_cap = cap;
    return BCMC_OK;
}

// candidate-row lists and the padded row-block element space of a step
static int eng_ensure_rows(bcmc_engine *e, uint64_t rows, uint64_t row_elems) {
    if (rows > e->rows_cap) {
        BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
        eng_changed(e);
        const uint64_t cap = std::max<uint64_t>(rows, e->rows_cap * 3 / 2);
        e->rows_cap = 0;
        eng_release(e->cand_rows);
        scan_site_free(&e->site_rows);
        BCMC_TRY(dev_alloc((void **)&e->cand_rows, cap * sizeof(CandRow)));
        BCMC_TRY(scan_site_alloc(&e->site_rows, cap));
        e->rows_cap = cap;
    }
    if (row_elems > e->row_elems_cap) {
        BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
        eng_changed(e);
        const uint64_t cap = std::max<uint64_t>(row_elems, e->row_elems_cap * 3 / 2);
        e->row_elems_cap = 0;
        scan_site_free(&e->site_active);
        BCMC_TRY(scan_site_alloc(&e->site_active, cap));
        e->row_elems_cap = cap;
    }
    return BCMC_OK;
}

static int eng_ensure_queries(bcmc_engine *e, uint32_t nq) {
    if (nq <= e->q_cap)
        return BCMC_OK;
    BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
    eng_changed(e);
    e->q_cap = 0;
    eng_release(e->d_q);
    eng_release(e->q_act_start);
    eng_release(e->q_pair_end);
    eng_release(e->q_chunk_base);
    eng_release_host(e->h_q);
    eng_release_host(e->h_act);
    eng_release_host(e->h_vfirst);
    eng_release_host(e->h_chunk_base);
    const uint32_t cap = std::max<uint32_t>(nq, 64);
    BCMC_HIP_CHECK(hipHostMalloc((void **)&e->h_q, (size_t)cap * sizeof(EngQuery), hipHostMallocDefault));
    BCMC_HIP_CHECK(hipHostMalloc((void **)&e->h_act, (size_t)cap * 4, hipHostMallocCoherent));
    BCMC_HIP_CHECK(hipHostMalloc((void **)&e->h_vfirst, (size_t)cap * 8, hipHostMallocCoherent));
    BCMC_HIP_CHECK(hipHostMalloc((void **)&e->h_chunk_base, (size_t)cap * 4, hipHostMallocCoherent));
    BCMC_HIP_CHECK(hipHostGetDevicePointer((void **)&e->h_act_d, e->h_act, 0));
    BCMC_HIP_CHECK(hipHostGetDevicePointer((void **)&e->h_vfirst_d, e->h_vfirst, 0));
    BCMC_HIP_CHECK(hipHostGetDevicePointer((void **)&e->h_chunk_base_d, e->h_chunk_base, 0));
    BCMC_TRY(dev_alloc((void **)&e->d_q, (size_t)cap * sizeof(EngQuery)));
    BCMC_TRY(dev_alloc((void **)&e->q_act_start, (size_t)cap * 4));
    BCMC_TRY(dev_alloc((void **)&e->q_pair_end, (size_t)cap * 4));
    BCMC_TRY(dev_alloc((void **)&e->q_chunk_base, (size_t)cap * 4));
    e->q_cap = cap;
    return BCMC_OK;
}

static uint32_t eng_grid(const bcmc_ctx *ctx, uint64_t n, uint32_t per_cu = 8) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(n, 256), persistent_grid(ctx, per_cu)));
}

static EngMcArgs eng_mc_args(bcmc_engine *e, uint32_t nq) {
    EngMcArgs a;
    a.nq = nq;
    a.same_list = e->same_list;
    a.q_act_start = e->q_act_start;
    a.q_pair_end = e->q_pair_end;
    a.q_chunk_base = e->q_chunk_base;
    a.qs = e->d_q;
    a.vols = e->d_vols;
    a.act_q = e->act_q;
    a.act_b = e->act_b;
    a.flags = e->flags;
    a.pool = e->pool;
    a.capacity = e->pool_cap;
    a.c = e->c;
    a.counts = e->counts;
    a.chunk_verts = e->chunk_verts;
    a.chunk_verts_pre = e->chunk_verts_pre;
    a.emit_chunks = e->h_emit_d;
    a.verts = e->verts;
    a.vert_cap = e->vert_cap;
    a.fmt = e->fmt;
    return a;
}

// chunks of a step: at most one partial chunk per query
static uint64_t eng_max_chunks(uint64_t total, uint32_t nq) { return div_up(total, kChunk) + nq; }

static int eng_launch_emit(bcmc_engine *e, uint32_t nq, uint64_t total) {
    static const uint32_t resident =
        resident_per_cu((const void *)eng_emit_kernel<BCMC_VERTS_LOCAL_F32>, kEngThreads);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(eng_max_chunks(total, nq),
                              persistent_grid(e->ctx, std::min(e->ctx->tune.mc, resident))));
    hipStream_t st = e->ctx->stream;
    if (e->fmt == BCMC_VERTS_LOCAL_F32)
        eng_emit_kernel<BCMC_VERTS_LOCAL_F32><<<grid, kEngThreads, 0, st>>>(eng_mc_args(e, nq));
    else if (e->fmt == BCMC_VERTS_GLOBAL_F32)
        eng_emit_kernel<BCMC_VERTS_GLOBAL_F32><<<grid, kEngThreads, 0, st>>>(eng_mc_args(e, nq));
    else
        eng_emit_kernel<BCMC_VERTS_PACKED_U32X2><<<grid, kEngThreads, 0, st>>>(eng_mc_args(e, nq));
    BCMC_HIP_CHECK(hipGetLastError());
    return BCMC_OK;
}

static int eng_ensure_chunks(bcmc_engine *e, uint64_t max_chunks) {
    if (max_chunks > e->chunks_cap) {
        BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
        eng_changed(e);
        const uint64_t cap = std::max<uint64_t>(max_chunks, e->chunks_cap * 3 / 2);
        e->chunks_cap = 0;
        eng_release(e->chunk_verts);
        eng_release(e->chunk_verts_pre);
        scan_site_free(&e->site_chunks);
        BCMC_TRY(dev_alloc((void **)&e->chunk_verts, cap * 4));
        BCMC_TRY(dev_alloc((void **)&e->chunk_verts_pre, cap * 8));
        BCMC_TRY(scan_site_alloc(&e->site_chunks, cap));
        e->chunks_cap = cap;
    }
    return BCMC_OK;
}

// count -> chunk prefixes -> each query's first vertex (chunk table built
// before the mark); the emit is launched after the host has seen the counts
static int eng_launch_count(bcmc_engine *e, uint32_t nq, uint64_t total) {
    const uint64_t max_chunks = eng_max_chunks(total, nq);
    const EngMcArgs a = eng_mc_args(e, nq);
    static const uint32_t resident = resident_per_cu((const void *)eng_count_kernel, kEngThreads);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(max_chunks, persistent_grid(e->ctx, std::min(e->ctx->tune.count, resident))));
    eng_count_kernel<<<grid, kEngThreads, 0, e->ctx->stream>>>(a);
    launch_scan<1>(e->ctx, e->site_chunks, e->site_chunks,
                   EngChunkPrefixOp{a.chunk_verts, a.chunk_verts_pre, a.c}, max_chunks);
    eng_vfirst_kernel<<<div_up(nq, 256), 256, 0, e->ctx->stream>>>(
        a, EngReadback{e->ch_d, e->h_act_d, e->h_vfirst_d, e->h_chunk_base_d});
    BCMC_HIP_CHECK(hipGetLastError());
    return BCMC_OK;
}

// What a step's kernel chain bakes in besides the engine's buffers.
struct EngPlan {
    uint32_t nq;
    uint64_t rows, row_elems, total;
    uint64_t used_mask;    // volumes the step touches (bit per volume id)
    uint32_t sort_passes;  // eviction-order sort passes; 0: no eviction chain
};

// The step's kernel chain, from the epoch bump to the count pass and the
// vertex offsets (the emit follows in a launch of its own, once the host has
// sized the output from the counts).  It reads the step's queries, tile table
// and same-volume lists from the device tables uploaded before it, and makes
// no allocation, copy or memset, so it can be captured into a graph and
// replayed.
static int eng_enqueue_step(bcmc_engine *e, const EngPlan &p) {
    bcmc_ctx *ctx = e->ctx;
    hipStream_t st = ctx->stream;
    const uint32_t nq = p.nq;
    eng_begin_kernel<<<1, 256, 0, st>>>(e->c, e->q_act_start, nq);
    // active pairs: candidate row segments first, then their blocks
    launch_scan<1>(ctx, e->site_rows, e->site_rows,
                   CandSegOp{e->d_q, e->tile_q, e->cand_rows, e->c, (uint32_t)p.rows}, p.rows);
    SegActiveOp aop{e->cand_rows, e->flags, e->act_q, e->act_b, e->q_act_start, e->c};
    launch_scan<1, SegActiveOp, 16>(ctx, e->site_active, e->site_active, aop, p.row_elems);
    // marching-cubes chunks of every query (also the work units of the mark)
    eng_chunk_table_kernel<<<1, 256, 0, st>>>(e->c, e->q_act_start, nq, e->q_pair_end, e->q_chunk_base);
    eng_mark_kernel<<<(uint32_t)std::max<uint64_t>(
                          1, std::min<uint64_t>(div_up(p.total, 256 * kMarkPairs), persistent_grid(ctx, 8))),
                      256, 0, st>>>(eng_mc_args(e, nq), e->stamp, e->miss, e->miss_cap);
    if (p.sort_passes) {
        launch_scan<1>(ctx, e->site_cand, e->site_cand, CandOp{e->stamp, e->cand, e->c, e->pool_cap},
                       e->pool_cap);
        eng_cand_keys_kernel<<<eng_grid(ctx, e->pool_cap, 2), 256, 0, st>>>(e->stamp, e->cand, e->c,
                                                                             e->cand_keys);
        launch_sort_pairs(ctx, &e->sort, e->cand_keys, e->cand, &e->c->num_cand, e->pool_cap,
                          e->order, p.sort_passes);
    }
    eng_assign_kernel<<<eng_grid(ctx, p.total, 4), 256, 0, st>>>(e->d_vols, e->c, e->miss, e->order,
                                                                  e->pool_cap, e->stamp, e->owner);
    eng_commit_kernel<<<1, 1, 0, st>>>(e->c, e->pool_cap);
    // decode the misses of each volume into their slots
    for (uint32_t vi = 0; vi < e->vols.size(); ++vi) {
        if (!(p.used_mask >> vi & 1))
            continue;
        EngVolume &v = e->vols[vi];
        launch_decode_to_cache(ctx, v.stream, v.nwords, v.maxbits, v.miss_list, &e->c->vol_miss[vi],
                               v.n, v.slot_map, e->pool, e->pool_cap);
    }
    return eng_launch_count(e, nq, p.total);
}

extern "C" {

BCMC_API int bcmc_engine_create(bcmc_ctx *ctx, uint64_t pool_bytes, bcmc_engine **out) {
    BCMC_CHECK_ARG(ctx && out, "null argument");
    BCMC_CHECK_ARG(pool_bytes >= 256 * 64, "decoded-block pool must hold at least 64 blocks");
    *out = nullptr;
    bcmc_engine *e = new bcmc_engine;
    e->ctx = ctx;
    e->budget = pool_bytes;
    e->pool_cap = (uint32_t)std::min<uint64_t>(pool_bytes / 256, 0x7FFFFFC0ull);
    // one extra zeroed slot past the end: absent blocks read it
    int rc = dev_alloc((void **)&e->pool, (size_t)e->pool_cap * 256 + 256);
    if (rc == BCMC_OK && hipMemsetAsync((char *)e->pool + (size_t)e->pool_cap * 256, 0, 256, ctx->stream) != hipSuccess)
        rc = BCMC_ERR_HIP;
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->stamp, (size_t)e->pool_cap * 4);
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->owner, (size_t)e->pool_cap * 8);
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->cand, (size_t)e->pool_cap * 4);
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->cand_keys, (size_t)e->pool_cap * 4);
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->order, (size_t)e->pool_cap * 4);
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->d_vols, kMaxVols * sizeof(EngVol));
    if (rc == BCMC_OK)
        rc = dev_alloc((void **)&e->c, sizeof(EngCounters));
    if (rc == BCMC_OK)
        rc = sort_scratch_alloc(&e->sort, e->pool_cap);
    if (rc == BCMC_OK)
        rc = scan_site_alloc(&e->site_cand, e->pool_cap);
    if (rc == BCMC_OK && (hipHostMalloc(&e->ch, sizeof(EngCounters), hipHostMallocCoherent) != hipSuccess ||
                          hipHostMalloc(&e->h_emit, 4, hipHostMallocCoherent) != hipSuccess))
        rc = BCMC_ERR_NOMEM;
    if (rc == BCMC_OK && (hipHostGetDevicePointer((void **)&e->ch_d, e->ch, 0) != hipSuccess ||
                          hipHostGetDevicePointer((void **)&e->h_emit_d, e->h_emit, 0) != hipSuccess))
        rc = BCMC_ERR_HIP;
    if (rc != BCMC_OK) {
        bcmc_engine_destroy(e);
        return rc;
    }
    *e->ch = EngCounters{};
    *e->h_emit = 0;
    BCMC_HIP_CHECK(hipMemsetAsync(e->c, 0, sizeof(EngCounters), ctx->stream));
    BCMC_HIP_CHECK(hipMemsetAsync(e->stamp, 0, (size_t)e->pool_cap * 4, ctx->stream));
    BCMC_HIP_CHECK(hipMemsetAsync(e->owner, 0xFF, (size_t)e->pool_cap * 8, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = e;
    return BCMC_OK;
}

BCMC_API int bcmc_engine_destroy(bcmc_engine *e) {
    if (!e)
        return BCMC_OK;
    (void)hipStreamSynchronize(e->ctx->stream);
    eng_drop_graphs(e);
    for (auto &v : e->vols) {
        dev_free(v.stream);
        dev_free(v.ranges);
        dev_free(v.seg_range);
        dev_free(v.dilated);
        dev_free(v.tile);
        dev_free(v.slot_map);
        dev_free(v.miss_list);
    }
    dev_free(e->d_vols);
    dev_free(e->pool);
    dev_free(e->stamp);
    dev_free(e->owner);
    dev_free(e->cand);
    dev_free(e->cand_keys);
    dev_free(e->order);
    dev_free(e->miss);
    sort_scratch_free(&e->sort);
    dev_free(e->c);
    dev_free(e->flags);
    dev_free(e->act_q);
    dev_free(e->act_b);
    dev_free(e->cand_rows);
    scan_site_free(&e->site_rows);
    dev_free(e->counts);
    dev_free(e->chunk_verts);
    dev_free(e->chunk_verts_pre);
    dev_free(e->d_q);
    dev_free(e->q_act_start);
    dev_free(e->q_pair_end);
    dev_free(e->q_chunk_base);
    dev_free(e->same_list);
    dev_free(e->tile_q);
    for (void *h : {(void *)e->h_q, (void *)e->h_same, (void *)e->h_act, (void *)e->h_vfirst,
                    (void *)e->h_tile_q, (void *)e->h_chunk_base})
        if (h)
            (void)hipHostFree(h);
    dev_free(e->verts);
    scan_site_free(&e->site_active);
    scan_site_free(&e->site_cand);
    scan_site_free(&e->site_chunks);
    if (e->ch)
        (void)hipHostFree(e->ch);
    if (e->h_emit)
        (void)hipHostFree(e->h_emit);
    delete e;
    return BCMC_OK;
}

// add_volume with the ranges from the range pass (ranges == nullptr) or
// persisted ones.
static int eng_add_volume(bcmc_engine *e, const void *stream, uint64_t bytes, double rate, const uint32_t dims[3],
                          const void *ranges, bcmc_ranges_format fmt, uint64_t ranges_bytes, uint32_t *volume_id) {
    BCMC_CHECK_ARG(e && stream && dims && volume_id, "null argument");
    BCMC_CHECK_ARG(e->vols.size() < kMaxVols, "too many volumes (max 64)");
    const double mb = 64.0 * rate;
    BCMC_CHECK_ARG(mb >= 9.0 && mb <= 4096.0 && mb == std::floor(mb),
                   "64*compressionRate must be an integer in [9, 4096]");
    EngVolume v;
    for (int i = 0; i < 3; ++i) {
        BCMC_CHECK_ARG(dims[i] > 0, "volume dims must be positive");
        v.dims[i] = dims[i];
        v.nb[i] = (dims[i] + 3) / 4;
    }
    const uint64_t n = (uint64_t)v.nb[0] * v.nb[1] * v.nb[2];
    BCMC_CHECK_ARG(n < (1ull << 31), "volume has too many blocks");
    v.n = (uint32_t)n;
    v.maxbits = (uint32_t)mb;
    BCMC_CHECK_ARG(bytes >= (n * v.maxbits + 7) / 8, "compressed stream shorter than the volume");
    if (ranges) {
        BCMC_CHECK_ARG(fmt == BCMC_RANGES_BLOCK || fmt == BCMC_RANGES_BLOCK_TILE, "bad ranges format");
        BCMC_CHECK_ARG(ranges_bytes == n * (fmt == BCMC_RANGES_BLOCK ? 8u : 16u),
                       "ranges must hold 8 (block) or 16 (block and tile) bytes per block");
    }
    bcmc_ctx *ctx = e->ctx;
    float2 *parts = nullptr;  // 8 part ranges per block, transient
    // frees the volume's arrays on an error return; disarmed once it is registered
    struct VolGuard {
        EngVolume *v;
        float2 **parts;
        bcmc_ctx *ctx;
        bool armed = true;
        ~VolGuard() {
            if (!armed)
                return;
            (void)hipStreamSynchronize(ctx->stream);
            for (void *p : {(void *)v->stream, (void *)v->ranges, (void *)v->seg_range,
                            (void *)v->dilated, (void *)v->tile, (void *)v->slot_map,
                            (void *)v->miss_list, (void *)*parts})
                dev_free(p);
        }
    } guard{&v, &parts, ctx};
    v.nwords = (bytes + 7) / 8 + 2;
    BCMC_TRY(dev_alloc((void **)&v.stream, v.nwords * 8));
    BCMC_HIP_CHECK(hipMemsetAsync(v.stream, 0, v.nwords * 8, ctx->stream));
    BCMC_HIP_CHECK(hipMemcpyAsync(v.stream, stream, bytes, hipMemcpyDefault, ctx->stream));
    BCMC_TRY(dev_alloc((void **)&v.ranges, n * 8));
    BCMC_TRY(dev_alloc((void **)&v.slot_map, n * 4));
    BCMC_TRY(dev_alloc((void **)&v.miss_list, n * 4));
    BCMC_HIP_CHECK(hipMemsetAsync(v.slot_map, 0xFF, n * 4, ctx->stream));
    BCMC_TRY(dev_alloc((void **)&v.tile, n * 8));
    if (ranges) {
        BCMC_HIP_CHECK(hipMemcpyAsync(v.ranges, ranges, n * 8, hipMemcpyDefault, ctx->stream));
        if (fmt == BCMC_RANGES_BLOCK_TILE)
            BCMC_HIP_CHECK(hipMemcpyAsync(v.tile, (const uint8_t *)ranges + n * 8, n * 8, hipMemcpyDefault,
                                          ctx->stream));
        else
            launch_tile_bound_ranges(ctx, v.ranges, v.dims, v.tile);
    } else {
        BCMC_TRY(dev_alloc((void **)&parts, n * 64));
        launch_block_ranges(ctx, v.stream, v.nwords, v.maxbits, v.dims, v.ranges, parts);
        launch_tile_ranges(ctx, parts, v.nb, v.tile);
    }
    const uint32_t nsegx = div_up(v.nb[0], kSegBlocks), nsegs = nsegx * v.nb[1] * v.nb[2];
    BCMC_TRY(dev_alloc((void **)&v.dilated, n * 8));
    BCMC_TRY(dev_alloc((void **)&v.seg_range, (size_t)nsegs * 8));
    launch_dilate_ranges(ctx, v.ranges, v.nb, v.dilated);
    eng_seg_range_kernel<<<eng_grid(ctx, nsegs), 256, 0, ctx->stream>>>(v.dilated, v.nb[0], nsegx, nsegs,
                                                                         v.seg_range);
    BCMC_HIP_CHECK(hipGetLastError());
    {
        // range bins: edges spread evenly over the volume's finite dilated
        // ranges (the row segments' ranges span them); any ascending edges
        // give exact bounds, even ones make them tight
        std::vector<float2> sr(nsegs);
        BCMC_HIP_CHECK(hipMemcpyAsync(sr.data(), v.seg_range, (size_t)nsegs * 8, hipMemcpyDeviceToHost,
                                      ctx->stream));
        BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        float gmin = INFINITY, gmax = -INFINITY;
        for (const float2 &r : sr) {
            if (std::isfinite(r.x))
                gmin = std::min(gmin, r.x);
            if (std::isfinite(r.y))
                gmax = std::max(gmax, r.y);
        }
        if (!std::isfinite(gmin))
            gmin = 0.0f;
        if (!(gmax > gmin))
            gmax = gmin + 1.0f;
        v.edges.resize(kRangeBins - 1);
        for (uint32_t i = 0; i + 1 < kRangeBins; ++i)
            v.edges[i] = (float)((double)gmin + ((double)gmax - (double)gmin) * (i + 1) / kRangeBins);
        for (uint32_t i = 1; i + 1 < kRangeBins; ++i)
            v.edges[i] = std::max(v.edges[i], v.edges[i - 1]);
        float *d_edges = nullptr;
        uint32_t *d_hist = nullptr;
        const size_t hbytes = (size_t)kRangeBins * kRangeBins * 4;
        int rc = dev_alloc((void **)&d_edges, (kRangeBins - 1) * 4);
        if (rc == BCMC_OK)
            rc = dev_alloc((void **)&d_hist, hbytes);
        std::vector<uint32_t> hist;
        if (rc == BCMC_OK) {
            hist.resize((size_t)kRangeBins * kRangeBins);
            if (hipMemcpyAsync(d_edges, v.edges.data(), (kRangeBins - 1) * 4, hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess ||
                hipMemsetAsync(d_hist, 0, hbytes, ctx->stream) != hipSuccess) {
                rc = BCMC_ERR_HIP;
            } else {
                eng_range_hist_kernel<<<eng_grid(ctx, n), 256, 0, ctx->stream>>>(v.dilated, n, d_edges, d_hist);
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(hist.data(), d_hist, hbytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                    hipStreamSynchronize(ctx->stream) != hipSuccess)
                    rc = BCMC_ERR_HIP;
            }
        }
        dev_free(d_edges);
        dev_free(d_hist);
        if (rc != BCMC_OK) {
            set_error("range histogram of the volume failed");
            return rc;
        }
        // below[a][b] = blocks with bin(lo) >= a and bin(hi) < b
        constexpr uint32_t W = kRangeBins + 1;
        v.below.assign((size_t)W * W, 0u);
        for (int a = (int)kRangeBins - 1; a >= 0; --a) {
            uint32_t row = 0;
            for (uint32_t b = 0; b < kRangeBins; ++b) {
                v.below[(size_t)a * W + b] = v.below[(size_t)(a + 1) * W + b] + row;
                row += hist[(size_t)a * kRangeBins + b];
            }
            v.below[(size_t)a * W + kRangeBins] = v.below[(size_t)(a + 1) * W + kRangeBins] + row;
        }
    }
    dev_free(v.ranges);  // only the dilated and tile ranges are used from here on
    dev_free(parts);
    parts = nullptr;
    v.ranges = nullptr;
    EngVol dv;
    for (int i = 0; i < 3; ++i)
        dv.nb[i] = v.nb[i];
    dv.n = v.n;
    dv.div_x = make_fast_div(v.nb[0]);
    dv.div_y = make_fast_div(v.nb[1]);
    dv.dilated = v.dilated;
    dv.seg_range = v.seg_range;
    dv.tile = v.tile;
    dv.slot_map = v.slot_map;
    dv.miss_list = v.miss_list;
    const uint32_t id = (uint32_t)e->vols.size();
    BCMC_HIP_CHECK(hipMemcpyAsync(e->d_vols + id, &dv, sizeof(dv), hipMemcpyHostToDevice, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    e->vols.push_back(v);
    guard.armed = false;
    // miss list capacity: every block of every volume
    uint64_t all = 0;
    for (auto &w : e->vols)
        all += w.n;
    eng_changed(e);
    e->miss_cap = 0;
    eng_release(e->miss);
    BCMC_TRY(dev_alloc((void **)&e->miss, all * 8));
    e->miss_cap = (uint32_t)std::min<uint64_t>(all, 0xFFFFFFFFu);
    *volume_id = id;
    return BCMC_OK;
}

BCMC_API int bcmc_engine_add_volume(bcmc_engine *e, const void *stream, uint64_t bytes, double rate,
                                    const uint32_t dims[3], uint32_t *volume_id) {
    return eng_add_volume(e, stream, bytes, rate, dims, nullptr, BCMC_RANGES_BLOCK, 0, volume_id);
}

BCMC_API int bcmc_engine_add_volume_ranges(bcmc_engine *e, const void *stream, uint64_t bytes, double rate,
                                           const uint32_t dims[3], const void *ranges, bcmc_ranges_format fmt,
                                           uint64_t ranges_bytes, uint32_t *volume_id) {
    BCMC_CHECK_ARG(ranges, "null ranges");
    return eng_add_volume(e, stream, bytes, rate, dims, ranges, fmt, ranges_bytes, volume_id);
}

BCMC_API int bcmc_engine_set_vertex_format(bcmc_engine *e, bcmc_vertex_format fmt) {
    BCMC_CHECK_ARG(e, "null argument");
    BCMC_CHECK_ARG(fmt >= BCMC_VERTS_LOCAL_F32 && fmt <= BCMC_VERTS_PACKED_U32X2, "bad format");
    if (fmt != e->fmt) {
        (void)hipStreamSynchronize(e->ctx->stream);
        eng_changed(e);
        eng_release(e->verts);
        e->vert_cap = 0;
        e->fmt = fmt;
    }
    return BCMC_OK;
}

}  // extern "C"

// Replays the graph captured for `key`, capturing it the second time the key
// is seen among the last kEngSeen parts, or runs enqueue() directly.
template <class F>
static int eng_run_part(bcmc_engine *e, const EngKey &key, F &&enqueue, bool *replayed) {
    hipStream_t st = e->ctx->stream;
    EngGraph *hit = nullptr;
    if (e->ctx->graph_replay && e->use_graph) {
        for (EngGraph &g : e->graphs)
            hit = g.key == key ? &g : hit;
        const bool seen = std::find(e->seen.begin(), e->seen.end(), key) != e->seen.end();
        if (!hit && seen) {
            // the stream holds only the staging copies of this step: nothing
            // was reallocated (gen is in the key)
            BCMC_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue();
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc == BCMC_OK && ec == hipSuccess && g && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess) {
                if (e->graphs.size() == kEngGraphs) {
                    auto lru = std::min_element(e->graphs.begin(), e->graphs.end(),
                                                [](const EngGraph &u, const EngGraph &v) { return u.last_use < v.last_use; });
                    eng_drop_graph(*lru);
                    e->graphs.erase(lru);
                }
                e->graphs.push_back(EngGraph{key, g, x, 0});
                hit = &e->graphs.back();
            } else {
                if (g)
                    (void)hipGraphDestroy(g);
                eng_drop_graphs(e);
                e->use_graph = false;  // replay is an optimisation: fall back to direct launches
                (void)hipGetLastError();
                if (rc != BCMC_OK)
                    return rc;
            }
        } else if (!hit) {
            if (e->seen.size() < kEngSeen)
                e->seen.push_back(key);
            else
                e->seen[e->seen_next] = key;
            e->seen_next = (e->seen_next + 1) % kEngSeen;
        }
        if (hit)
            hit->last_use = ++e->clock;
    }
    *replayed = hit != nullptr;
    if (hit) {
        BCMC_HIP_CHECK(hipGraphLaunch(hit->exec, st));
        return BCMC_OK;
    }
    return enqueue();
}

// An empty step: advances the epoch, so that the blocks the last step stamped
// may be evicted by the next one.
static int eng_idle_step(bcmc_engine *e) {
    eng_begin_kernel<<<1, 256, 0, e->ctx->stream>>>(e->c, e->q_act_start, 0);
    BCMC_HIP_CHECK(hipGetLastError());
    BCMC_HIP_CHECK(hipMemcpyAsync(e->ch, e->c, sizeof(EngCounters), hipMemcpyDeviceToHost, e->ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
    return BCMC_OK;
}

// One step over nq queries.  The chain runs up to the count pass; from the
// exact vertex counts the host admits the leading queries whose vertices fit
// `budget` bytes (all of them if they fit), sizes the vertex buffer for
// exactly those, and launches the emit for their chunks.  *n_done receives
// the number admitted (0: the first query alone exceeds the budget); their
// results go to results[0 .. *n_done).
static int eng_run(bcmc_engine *e, const uint32_t *volume_ids, const float *isovalues, uint32_t nq,
                   uint64_t budget, uint32_t *n_done, bcmc_query_result *results) {
    BCMC_CHECK_ARG(e && (nq == 0 || (volume_ids && isovalues)), "null argument");
    *n_done = 0;
    if (nq == 0)
        return BCMC_OK;
    if (!e->miss) {
        set_error("engine has no miss list (an earlier add_volume failed to allocate it)");
        return BCMC_ERR_STATE;
    }
    bcmc_ctx *ctx = e->ctx;
    hipStream_t st = ctx->stream;
    std::vector<EngQuery> qs(nq);
    std::vector<bool> used(e->vols.size(), false);
    std::vector<std::vector<uint32_t>> earlier(e->vols.size());
    std::vector<uint32_t> same;
    uint64_t total = 0, rows = 0;
    // every scan tile of the row segments lies in one query
    const uint32_t tile = scan_tile(ctx);
    for (uint32_t i = 0; i < nq; ++i) {
        BCMC_CHECK_ARG(volume_ids[i] < e->vols.size(), "unknown volume id");
        const EngVolume &v = e->vols[volume_ids[i]];
        EngQuery &q = qs[i];
        q = EngQuery{};
        q.vol = volume_ids[i];
        q.iso = isovalues[i];
        q.flag_base = (uint32_t)total;
        q.n = v.n;
        q.row_base = (uint32_t)rows;
        q.nb0 = v.nb[0];
        q.nsegx = div_up(v.nb[0], kSegBlocks);
        q.nrows = q.nsegx * v.nb[1] * v.nb[2];
        q.dilated = v.dilated;
        q.seg_range = v.seg_range;
        q.slot_map = v.slot_map;
        q.same_off = (uint32_t)same.size();
        q.same_cnt = (uint32_t)earlier[q.vol].size();
        same.insert(same.end(), earlier[q.vol].begin(), earlier[q.vol].end());
        earlier[q.vol].push_back(q.flag_base);
        used[volume_ids[i]] = true;
        total += v.n;
        rows += (uint64_t)div_up(q.nrows, tile) * tile;  // tile-aligned
    }
    BCMC_CHECK_ARG(total <= kMaxStepBlocks, "too many blocks in one step");
    const uint64_t row_elems = rows * kSegBlocks;
    BCMC_CHECK_ARG(row_elems < (1ull << 32) - kScanTileMax, "too many row segments in one step");
    BCMC_TRY(eng_ensure_blocks(e, total));
    BCMC_TRY(eng_ensure_rows(e, rows, row_elems));
    BCMC_TRY(eng_ensure_queries(e, nq));
    if (same.size() > e->same_cap || !e->same_list) {
        BCMC_HIP_CHECK(hipStreamSynchronize(st));
        eng_changed(e);
        const uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(same.size(), 1024), e->same_cap * 3 / 2);
        e->same_cap = 0;
        eng_release(e->same_list);
        eng_release_host(e->h_same);
        BCMC_TRY(dev_alloc((void **)&e->same_list, cap * 4));
        BCMC_HIP_CHECK(hipHostMalloc((void **)&e->h_same, cap * 4, hipHostMallocDefault));
        e->same_cap = cap;
    }
    // query of every scan tile of the row segments (+ a last entry)
    const uint64_t ntiles = div_up(rows, tile) + 1;
    if (ntiles > e->tile_q_cap) {
        BCMC_HIP_CHECK(hipStreamSynchronize(st));
        eng_changed(e);
        const uint64_t cap = std::max<uint64_t>(ntiles, e->tile_q_cap * 2);
        e->tile_q_cap = 0;
        eng_release(e->tile_q);
        eng_release_host(e->h_tile_q);
        BCMC_TRY(dev_alloc((void **)&e->tile_q, cap * 4));
        BCMC_HIP_CHECK(hipHostMalloc((void **)&e->h_tile_q, cap * 4, hipHostMallocDefault));
        e->tile_q_cap = cap;
    }
    for (uint64_t t = 0, q = 0; t < ntiles; ++t) {
        const uint64_t r = std::min<uint64_t>(t * tile, rows ? rows - 1 : 0);
        while (q + 1 < nq && qs[q + 1].row_base <= r)
            ++q;
        e->h_tile_q[t] = (uint32_t)q;
    }
    // eviction order (only does work when the misses exceed the free room).
    // A step misses at most every block of each volume it touches; when the
    // free room after the last step covers that, eviction cannot happen and
    // the chain is not launched (assign then never reads the order).
    EngPlan p{nq, rows, row_elems, total, 0, 0};
    uint64_t most_misses = 0;
    for (uint32_t vi = 0; vi < e->vols.size(); ++vi) {
        most_misses += used[vi] ? e->vols[vi].n : 0u;
        p.used_mask |= used[vi] ? 1ull << vi : 0;
    }
    if ((uint64_t)e->pool_cap - e->ch->used < most_misses) {
        // keys are stamps, at most this step's epoch
        uint32_t bits = 0;
        while (bits < 32 && ((uint64_t)(e->ch->epoch + 1) >> bits) != 0)
            ++bits;
        p.sort_passes = std::max<uint32_t>(1, (bits + 7) / 8);
    }
    BCMC_TRY(eng_ensure_chunks(e, eng_max_chunks(total, nq)));
    // the chain reads the step through the uploaded tables; everything the
    // launches bake in (sizes, volumes touched, sort passes, buffers, vertex
    // format, tuned grids) is in the keys
    const EngKey front_key = {nq | (uint64_t)p.sort_passes << 32, rows, total, p.used_mask,
                              (uint64_t)ctx->tune_gen << 8, e->gen};
    // pinned staging: the copies run asynchronously, ahead of the chain on the
    // same stream (the step ends with a stream synchronisation before the
    // staging is written again)
    BCMC_HIP_CHECK(hipMemcpyAsync(e->tile_q, e->h_tile_q, ntiles * 4, hipMemcpyHostToDevice, st));
    std::memcpy(e->h_q, qs.data(), nq * sizeof(EngQuery));
    BCMC_HIP_CHECK(hipMemcpyAsync(e->d_q, e->h_q, nq * sizeof(EngQuery), hipMemcpyHostToDevice, st));
    if (!same.empty()) {
        std::memcpy(e->h_same, same.data(), same.size() * 4);
        BCMC_HIP_CHECK(hipMemcpyAsync(e->same_list, e->h_same, same.size() * 4, hipMemcpyHostToDevice, st));
    }
    bool replayed = false;
    BCMC_TRY(eng_run_part(e, front_key, [&]() { return eng_enqueue_step(e, p); }, &replayed));
    ++e->steps;
    e->graph_steps += replayed ? 1 : 0;
    // the chain's last launch wrote the counters and the per-query results
    // into host-coherent memory
    const unsigned long long *qv = e->h_vfirst;
    const uint32_t *qa = e->h_act;
    BCMC_HIP_CHECK(hipStreamSynchronize(st));
    if (e->ch->short_slots) {
        set_error("decoded-block pool too small for this step: %u blocks found no slot",
                  e->ch->short_slots);
        return BCMC_ERR_NOMEM;
    }
    // each query's vertex range: from its first vertex to the next non-empty
    // query's (or the step total)
    const uint64_t nv = e->ch->num_verts;
    std::vector<uint64_t> vstart(nq + 1);
    vstart[nq] = nv;
    for (uint32_t i = nq; i-- > 0;)
        vstart[i] = qa[i] == kNoPairs ? vstart[i + 1] : qv[i];
    // admission by the output budget: the longest prefix of queries whose
    // vertices fit
    const uint32_t vb = vbytes(e->fmt);
    uint32_t nadm = nq;
    while (nadm > 0 && vstart[nadm] * vb > budget)
        --nadm;
    if (nadm == 0) {
        set_error("output budget of %llu bytes is smaller than one query's %llu vertices",
                  (unsigned long long)budget, (unsigned long long)vstart[1]);
        return BCMC_ERR_NOMEM;
    }
    const uint64_t nv_adm = vstart[nadm];
    // the vertex buffer, sized from the exact counts
    if (nv_adm > e->vert_cap) {
        const uint64_t want = std::max<uint64_t>(nv_adm, e->vert_cap + e->vert_cap / 2);
        eng_changed(e);
        eng_release(e->verts);
        e->vert_cap = 0;
        BCMC_TRY(dev_alloc(&e->verts, want * vb));
        e->vert_cap = want;
    }
    // read by the emit launch from host-coherent memory (the previous step's
    // emit has finished: every step ends with a stream synchronisation)
    *e->h_emit = nadm == nq ? e->ch->num_chunks : e->h_chunk_base[nadm];
    const EngKey back_key = {nq | 1ull << 40, total, 0, 0, (uint64_t)e->fmt | ctx->tune_gen << 8, e->gen};
    BCMC_TRY(eng_run_part(e, back_key, [&]() { return eng_launch_emit(e, nq, total); }, &replayed));
    BCMC_HIP_CHECK(hipStreamSynchronize(st));
    e->last.resize(nadm);
    const uint32_t na = e->ch->num_active;
    for (uint32_t i = 0; i < nadm; ++i) {
        bcmc_query_result r;
        // active pairs run from the query's first to the next non-empty query's
        uint32_t next_a = na;
        for (uint32_t j = i + 1; j < nq; ++j)
            if (qa[j] != kNoPairs) {
                next_a = qa[j];
                break;
            }
        r.num_active_blocks = qa[i] == kNoPairs ? 0u : next_a - qa[i];
        r.vertex_offset = vstart[i];
        r.num_vertices = vstart[i + 1] - vstart[i];
        r.step_decoded_blocks = e->ch->num_miss;
        e->last[i] = r;
        if (results)
            results[i] = r;
    }
    e->vert_view.ctx = ctx;
    e->vert_view.ptr = e->verts;
    e->vert_view.bytes = nv_adm * vb;
    e->vert_view.owned = false;
    *n_done = nadm;
    return BCMC_OK;
}

// A step over the distinct (volume, isovalue bits) queries of a request list:
// request i is answered by the result of the first request equal to it, so
// equal requests share one vertex range.  Admission stays a prefix of the
// requests: the distinct queries are in order of first occurrence, and the
// requests before the first occurrence of an unadmitted query only use
// admitted ones.
static int eng_run_coalesced(bcmc_engine *e, const uint32_t *volume_ids, const float *isovalues,
                             uint32_t nq, uint64_t budget, uint32_t *n_done,
                             bcmc_query_result *results) {
    std::vector<uint32_t> map(nq), first, uid;
    std::vector<float> uiso;
    std::unordered_map<uint64_t, uint32_t> seen;
    seen.reserve(nq * 2u);
    for (uint32_t i = 0; i < nq; ++i) {
        uint32_t bits;
        std::memcpy(&bits, &isovalues[i], 4);
        const uint64_t key = (uint64_t)volume_ids[i] << 32 | bits;
        const auto ins = seen.emplace(key, (uint32_t)uid.size());
        if (ins.second) {
            first.push_back(i);
            uid.push_back(volume_ids[i]);
            uiso.push_back(isovalues[i]);
        }
        map[i] = ins.first->second;
    }
    const uint32_t nu = (uint32_t)uid.size();
    std::vector<bcmc_query_result> ures(nu);
    uint32_t udone = 0;
    BCMC_TRY(eng_run(e, uid.data(), uiso.data(), nu, budget, &udone, ures.data()));
    const uint32_t nadm = udone == nu ? nq : first[udone];
    for (uint32_t i = 0; i < nadm; ++i)
        results[i] = ures[map[i]];
    *n_done = nadm;
    return BCMC_OK;
}

extern "C" {

BCMC_API int bcmc_engine_run(bcmc_engine *e, const uint32_t *volume_ids, const float *isovalues,
                             uint32_t nq, bcmc_query_result *results) {
    uint32_t done = 0;
    BCMC_TRY(eng_run(e, volume_ids, isovalues, nq, ~0ull, &done, results));
    if (done != nq) {
        set_error("step output exceeds the address space");
        return BCMC_ERR_NOMEM;
    }
    return BCMC_OK;
}

BCMC_API bcmc_buf *bcmc_engine_vertex_buffer(bcmc_engine *e) { return e ? &e->vert_view : nullptr; }

BCMC_API int bcmc_engine_stats(bcmc_engine *e, bcmc_engine_info *info) {
    BCMC_CHECK_ARG(e && info, "null argument");
    BCMC_HIP_CHECK(hipMemcpyAsync(e->ch, e->c, sizeof(EngCounters), hipMemcpyDeviceToHost, e->ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
    info->num_volumes = (uint32_t)e->vols.size();
    info->pool_slots = e->pool_cap;
    info->pool_used = e->ch->used;
    info->epoch = e->ch->epoch;
    info->pool_bytes = (uint64_t)e->pool_cap * 256;
    uint64_t dev = info->pool_bytes + (uint64_t)e->pool_cap * 20 + e->blocks_cap * 13 + e->vert_cap * vbytes(e->fmt);
    for (auto &v : e->vols)
        dev += v.nwords * 8 + (uint64_t)v.n * 16;
    info->device_bytes = dev;
    info->steps = e->steps;
    info->graph_steps = e->graph_steps;
    return BCMC_OK;
}

BCMC_API int bcmc_engine_block_bound(bcmc_engine *e, uint32_t volume_id, const float *isovalues, uint32_t n,
                                     uint64_t *blocks) {
    BCMC_CHECK_ARG(e && blocks && (n == 0 || isovalues), "null argument");
    BCMC_CHECK_ARG(volume_id < e->vols.size(), "unknown volume id");
    const EngVolume &v = e->vols[volume_id];
    EngUnion u;
    uint64_t touched = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const int b = (int)v.bin(isovalues[i]);
        uint64_t in;
        touched = u.with(v, b, &in);
        u.add(b, in);
    }
    *blocks = touched;
    return BCMC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Served form: a request loop over an engine (continuous batching)
// ---------------------------------------------------------------------------
// Requests arrive at given times (seconds from the start of the call).  The
// loop keeps one step in flight: every request that has arrived by the time a
// step starts joins it (FIFO, up to max_batch, the pool's free slots and the
// output-vertex budget); requests arriving while a step runs, or left out by
// a bound, wait for the next one.  Latency is
// completion time - arrival time.  The decoded-block pool (the engine's LRU)
// and the compressed volumes stay resident across requests.
extern "C" BCMC_API int bcmc_serve_trace(bcmc_engine *e, const double *arrival_s,
                                         const uint32_t *volume_ids, const float *isovalues,
                                         uint32_t n, const bcmc_serve_opts *opts,
                                         bcmc_serve_result *out) {
    BCMC_CHECK_ARG(e && arrival_s && volume_ids && isovalues && out, "null argument");
    for (uint32_t i = 1; i < n; ++i)
        BCMC_CHECK_ARG(arrival_s[i] >= arrival_s[i - 1], "arrivals must be sorted");
    const uint32_t max_batch = opts && opts->max_batch ? opts->max_batch : 256;
    const uint64_t vert_budget = opts && opts->vertex_budget_bytes ? opts->vertex_budget_bytes
                                                                    : (uint64_t)64 << 30;
    const bool coalesce = !(opts && opts->no_coalesce);
    // capacity planning before the clock starts: a step takes consecutive
    // requests (FIFO), at most max_batch of them, so the largest window of
    // max_batch consecutive requests bounds every step's block and row counts
    {
        uint64_t blocks = 0, rows = 0, best_b = 0, best_r = 0;
        auto vol_rows = [&](uint32_t r) -> uint64_t {
            const EngVolume &v = e->vols[volume_ids[r]];
            const uint64_t nrows = (uint64_t)div_up(v.nb[0], kSegBlocks) * v.nb[1] * v.nb[2];
            const uint32_t tile = scan_tile(e->ctx);
            return (uint64_t)div_up(nrows, tile) * tile;  // tile-aligned, as in a step
        };
        for (uint32_t i = 0; i < n; ++i) {
            BCMC_CHECK_ARG(volume_ids[i] < e->vols.size(), "unknown volume id");
            blocks += e->vols[volume_ids[i]].n;
            rows += vol_rows(i);
            if (i >= max_batch) {
                blocks -= e->vols[volume_ids[i - max_batch]].n;
                rows -= vol_rows(i - max_batch);
            }
            best_b = std::max(best_b, std::min<uint64_t>(blocks, kMaxStepBlocks));
            best_r = std::max(best_r, rows);
        }
        const uint32_t nq = std::min<uint32_t>(n, max_batch);
        BCMC_TRY(eng_ensure_blocks(e, best_b));
        if (best_r * kSegBlocks < (1ull << 32) - kScanTileMax)
            BCMC_TRY(eng_ensure_rows(e, best_r, best_r * kSegBlocks));
        BCMC_TRY(eng_ensure_queries(e, nq));
        BCMC_TRY(eng_ensure_chunks(e, eng_max_chunks(best_b, nq)));
    }
    // the slots stamped by the engine's last step (an earlier serve call or
    // bcmc_engine_run) may not be evicted by the next one; an empty step
    // before the clock makes them evictable, so the first step of this call
    // has no predecessor's blocks to keep
    BCMC_TRY(eng_idle_step(e));
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    // Pool admission: a step's requests may touch at most the pool's slots
    // minus those the previous step stamped (which its successor may not
    // evict), counted per volume by the union bound of the step's isovalues
    // on it (EngUnion: every block but those whose dilated range lies between
    // two neighbouring isovalues or outside all of them).  Output admission:
    // the step's run admits the longest prefix whose exact vertex counts fit
    // the budget; the rest wait at the head of the queue for the next step.
    std::vector<uint32_t> rbin(n);
    std::vector<uint64_t> alone(n);  // a request's bound on its own
    for (uint32_t i = 0; i < n; ++i) {
        const EngVolume &v = e->vols[volume_ids[i]];
        rbin[i] = v.bin(isovalues[i]);
        uint64_t in;
        alone[i] = EngUnion{}.with(v, (int)rbin[i], &in);
    }
    const uint64_t pool = e->pool_cap;
    uint64_t prev_blocks = 0;
    uint32_t next = 0, done = 0, steps = 0;
    std::vector<uint32_t> pending;
    size_t head = 0;
    std::vector<uint32_t> ids;
    std::vector<float> isos;
    std::vector<uint32_t> batch;
    std::vector<bcmc_query_result> res;
    std::vector<EngUnion> uni(e->vols.size());
    std::unordered_map<uint64_t, uint32_t> in_step;  // (volume, isovalue bits) of the step's requests
    while (done < n) {
        const double t = now_s();
        while (next < n && arrival_s[next] <= t)
            pending.push_back(next++);
        if (head == pending.size()) {
            // idle: wait for the next arrival
            const double wait = arrival_s[next] - t;
            if (wait > 0)
                std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            continue;
        }
        if (alone[pending[head]] + prev_blocks > pool && prev_blocks > 0) {
            // the head request needs the previous step's slots: one empty
            // step makes them evictable
            BCMC_TRY(eng_idle_step(e));
            prev_blocks = 0;
        }
        ids.clear();
        isos.clear();
        batch.clear();
        for (auto &u : uni)
            u = EngUnion{};
        in_step.clear();
        uint64_t blocks = 0, step_blocks = 0;
        for (size_t h = head; h < pending.size() && batch.size() < max_batch; ++h) {
            const uint32_t r = pending[h], v = volume_ids[r];
            uint32_t bits;
            std::memcpy(&bits, &isovalues[r], 4);
            // a request equal to one already in the step adds no query
            const bool dup = coalesce && in_step.count((uint64_t)v << 32 | bits) != 0;
            const uint64_t nblk = dup ? 0 : e->vols[v].n;
            const uint64_t before = uni[v].empty() ? 0 : e->vols[v].n - std::min<uint64_t>(e->vols[v].n, uni[v].inactive);
            uint64_t in_after;
            const uint64_t after = uni[v].with(e->vols[v], (int)rbin[r], &in_after);
            const uint64_t grown = blocks - before + after;
            if (!batch.empty() && (grown + prev_blocks > pool || step_blocks + nblk > kMaxStepBlocks))
                break;
            blocks = grown;
            uni[v].add((int)rbin[r], in_after);
            step_blocks += nblk;
            if (coalesce)
                in_step.emplace((uint64_t)v << 32 | bits, r);
            batch.push_back(r);
            ids.push_back(volume_ids[r]);
            isos.push_back(isovalues[r]);
        }
        res.resize(batch.size());
        const double start = now_s();
        uint32_t ndone = 0;
        if (coalesce)
            BCMC_TRY(eng_run_coalesced(e, ids.data(), isos.data(), (uint32_t)batch.size(), vert_budget,
                                       &ndone, res.data()));
        else
            BCMC_TRY(eng_run(e, ids.data(), isos.data(), (uint32_t)batch.size(), vert_budget, &ndone,
                             res.data()));
        const double end = now_s();
        // every request of the batch stamped its blocks, the deferred ones too
        prev_blocks = blocks;
        for (uint32_t k = 0; k < ndone; ++k) {
            const uint32_t r = batch[k];
            out[r].arrival_s = arrival_s[r];
            out[r].start_s = start;
            out[r].finish_s = end;
            out[r].num_vertices = res[k].num_vertices;
            out[r].batch_size = ndone;
            out[r].step = steps;
        }
        head += ndone;
        done += ndone;
        ++steps;
    }
    return BCMC_OK;
}
