/* CPU oracle for the block-compressed marching-cubes path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference method used
 * as the checker by tests/, the smoke check and the benchmark's CPU-baseline
 * leg.  The product path never links or calls this file.
 *
 * Reference: Twinklebear/webgpu-bcmc.  Each function cites the file:line range
 * it restates.  Deliberate divergences from the reference (all documented at the
 * function):
 *   - 64-bit stream bit offsets (zfp_decompress.comp:199 multiplies in u32);
 *   - total block count nbx*nby*nbz (zfp_compute_block_range.comp:22 computes
 *     (px*py)/64*pz, which truncates when px*py is not a multiple of 64);
 *   - LRU `needed` entries past N read as 0 (lru_cache_mark_new_items.comp:32,42
 *     reads past the end of the flags buffer);
 *   - vertex totals are 64-bit (exclusive_scan.js:291 reads back a u32).
 */
#include "bcmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "tri_cases.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static uint32_t align_to(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

/* ======================================================================== */
/* ZFP fixed-rate decode                                                     */
/* ======================================================================== */

/* Sequency order of the 64 coefficients, zfp_decompress.comp:100-185: entry i
 * is the raster index x + 4y + 16z of the i-th coefficient, written here as
 * "xyz" digit triples. */
static const char PERM3D_XYZ[] =
    "000"
    "100010001"
    "011101110200020002"
    "111210201021120102012300030003"
    "211121112022202220310301031130103013"
    "122212221311131113320302032230203023"
    "222321312132231213123033303330"
    "322232223133313331"
    "233323332"
    "333";

static int perm3d(int i)
{
    const char *t = PERM3D_XYZ + 3 * i;
    return (t[0] - '0') + 4 * (t[1] - '0') + 16 * (t[2] - '0');
}

/* LSB-first bit reader over little-endian u64 words, starting at an absolute
 * bit offset.  Restates BlockReader/create_block_reader/read_bit/read_bits
 * (zfp_decompress.comp:188-257): the reference keeps a one-word window shifted
 * right as bits are consumed, which is equivalent to reading consecutive
 * stream bits LSB-first. */
typedef struct {
    const uint64_t *w;
    size_t nwords;
    uint64_t pos;
} reader_t;

static uint32_t rd_bit(reader_t *r)
{
    uint64_t wi = r->pos >> 6;
    uint32_t b = wi < r->nwords ? (uint32_t)((r->w[wi] >> (r->pos & 63)) & 1u) : 0u;
    r->pos++;
    return b;
}

static uint64_t rd_bits(reader_t *r, uint32_t n)
{
    uint64_t v = 0;
    for (uint32_t i = 0; i < n; ++i)
        v |= (uint64_t)rd_bit(r) << i;
    return v;
}

/* Embedded bit-plane decoder, zfp_decompress.comp:265-294. */
static void decode_ints(reader_t *r, uint32_t bits, uint32_t c[64])
{
    memset(c, 0, 64 * sizeof(uint32_t));
    uint32_t n = 0;
    for (uint32_t k = 32; bits != 0 && k-- > 0;) {
        uint32_t m = n < bits ? n : bits;
        bits -= m;
        uint64_t x = rd_bits(r, m);
        for (; n < 64 && bits != 0 && (bits--, rd_bit(r) != 0); x |= (uint64_t)1 << n++) {
            for (; n < 63 && bits != 0 && (bits--, rd_bit(r) == 0); ++n)
                ;
        }
        for (uint32_t i = 0; i < 64; ++i)
            c[i] += (uint32_t)((x >> i) & 1u) << k;
    }
}

static inline uint32_t asr1(uint32_t v) { return (uint32_t)((int32_t)v >> 1); }

/* Inverse decorrelating lift, zfp_decompress.comp:296-332 (int32 wrapping
 * arithmetic, arithmetic right shifts). */
static void inv_lift(uint32_t *b, int s, int idx)
{
    uint32_t x = b[idx], y = b[idx + s], z = b[idx + 2 * s], w = b[idx + 3 * s];
    y += asr1(w);
    w -= asr1(y);
    y += w;
    w <<= 1;
    w -= y;
    z += x;
    x <<= 1;
    x -= z;
    y += z;
    z <<= 1;
    z -= y;
    w += x;
    x <<= 1;
    x -= w;
    b[idx] = x;
    b[idx + s] = y;
    b[idx + 2 * s] = z;
    b[idx + 3 * s] = w;
}

/* zfp_decompress.comp:334-354: along z, then y, then x. */
static void inv_transform(uint32_t *b)
{
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x)
            inv_lift(b, 16, x + 4 * y);
    for (int x = 0; x < 4; ++x)
        for (int z = 0; z < 4; ++z)
            inv_lift(b, 4, 16 * z + x);
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y)
            inv_lift(b, 1, 4 * y + 16 * z);
}

/* decompress_block, zfp_decompress.comp:356-384, with the zero block written
 * explicitly (zfp_decompress_block.comp:33-34 relies on zero-initialised
 * locals).  Block start = block*maxbits computed in 64 bits. */
void orc_decode_block(const uint64_t *words, size_t nwords, uint32_t maxbits,
                      uint64_t block, float out[64])
{
    reader_t r = {words, nwords, block * (uint64_t)maxbits};
    if (rd_bit(&r) == 0) {
        for (int i = 0; i < 64; ++i)
            out[i] = 0.0f;
        return;
    }
    int emax = (int)rd_bits(&r, 8) - 127;
    uint32_t c[64], q[64];
    decode_ints(&r, maxbits - 9, c);
    for (int i = 0; i < 64; ++i)
        q[perm3d(i)] = (c[i] ^ 0xaaaaaaaau) - 0xaaaaaaaau;
    inv_transform(q);
    const float inv_w = ldexpf(1.0f, emax - 30);
    for (int i = 0; i < 64; ++i)
        out[i] = inv_w * (float)(int32_t)q[i];
}

static void pad_dims(const uint32_t dims[3], uint32_t p[3])
{
    for (int i = 0; i < 3; ++i)
        p[i] = align_to(dims[i], 4);
}

void orc_decode_volume(const uint64_t *words, size_t nwords, uint32_t maxbits,
                       const uint32_t dims[3], float *out, int threads)
{
    uint32_t p[3];
    pad_dims(dims, p);
    const uint32_t nb[3] = {p[0] / 4, p[1] / 4, p[2] / 4};
    const int64_t nblocks = (int64_t)nb[0] * nb[1] * nb[2];
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int64_t b = 0; b < nblocks; ++b) {
        float blk[64];
        orc_decode_block(words, nwords, maxbits, (uint64_t)b, blk);
        const uint64_t bx = b % nb[0], by = (b / nb[0]) % nb[1], bz = b / ((uint64_t)nb[0] * nb[1]);
        for (int z = 0; z < 4; ++z)
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x)
                    out[((bz * 4 + z) * p[1] + by * 4 + y) * (uint64_t)p[0] + bx * 4 + x] =
                        blk[16 * z + 4 * y + x];
    }
}

/* ======================================================================== */
/* Block ranges and the active-block test                                    */
/* ======================================================================== */

/* zfp_compute_block_range.comp:19-61: min/max over the voxels of the block that
 * lie inside the (unpadded) volume, starting from (1e20, -1e20). */
void orc_block_ranges(const uint64_t *words, size_t nwords, uint32_t maxbits,
                      const uint32_t dims[3], float *ranges, int threads)
{
    uint32_t p[3];
    pad_dims(dims, p);
    const uint32_t nb[3] = {p[0] / 4, p[1] / 4, p[2] / 4};
    const int64_t nblocks = (int64_t)nb[0] * nb[1] * nb[2];
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int64_t b = 0; b < nblocks; ++b) {
        float blk[64];
        orc_decode_block(words, nwords, maxbits, (uint64_t)b, blk);
        const uint32_t bx = (uint32_t)(b % nb[0]) * 4, by = (uint32_t)((b / nb[0]) % nb[1]) * 4,
                       bz = (uint32_t)(b / ((int64_t)nb[0] * nb[1])) * 4;
        const uint32_t sx = bx + 4 > dims[0] ? dims[0] - bx : 4;
        const uint32_t sy = by + 4 > dims[1] ? dims[1] - by : 4;
        const uint32_t sz = bz + 4 > dims[2] ? dims[2] - bz : 4;
        float lo = 1e20f, hi = -1e20f;
        for (uint32_t z = 0; z < sz; ++z)
            for (uint32_t y = 0; y < sy; ++y)
                for (uint32_t x = 0; x < sx; ++x) {
                    const float v = blk[16 * z + 4 * y + x];
                    lo = v < lo ? v : lo;
                    hi = v > hi ? v : hi;
                }
        ranges[2 * b] = lo;
        ranges[2 * b + 1] = hi;
    }
}

/* compute_block_active.comp:24-61. */
void orc_block_active(const float *ranges, const uint32_t dims[3], float iso,
                      uint32_t *flags, int threads)
{
    uint32_t p[3];
    pad_dims(dims, p);
    const int nb[3] = {(int)p[0] / 4, (int)p[1] / 4, (int)p[2] / 4};
    const int64_t nblocks = (int64_t)nb[0] * nb[1] * nb[2];
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int64_t b = 0; b < nblocks; ++b) {
        const int bx = (int)(b % nb[0]), by = (int)((b / nb[0]) % nb[1]),
                  bz = (int)(b / ((int64_t)nb[0] * nb[1]));
        const float lo = ranges[2 * b], hi = ranges[2 * b + 1];
        uint32_t active = 0;
        if (iso >= lo && iso <= hi) {
            active = 1;
        } else {
            for (int k = -1; k < 2 && !active; ++k)
                for (int j = -1; j < 2 && !active; ++j)
                    for (int i = -1; i < 2 && !active; ++i) {
                        const int cx = bx + i, cy = by + j, cz = bz + k;
                        if ((i == 0 && j == 0 && k == 0) || cx < 0 || cy < 0 || cz < 0 ||
                            cx >= nb[0] || cy >= nb[1] || cz >= nb[2])
                            continue;
                        const int64_t nid = cx + (int64_t)nb[0] * (cy + (int64_t)nb[1] * cz);
                        const float nlo = ranges[2 * nid], nhi = ranges[2 * nid + 1];
                        const float ulo = nlo < lo ? nlo : lo, uhi = nhi > hi ? nhi : hi;
                        if (iso >= ulo && iso <= uhi)
                            active = 1;
                    }
        }
        flags[b] = active;
    }
}

/* ======================================================================== */
/* Scan, compact, sort                                                       */
/* ======================================================================== */

/* ExclusiveScanner.scan (exclusive_scan.js:178-295; prefix_sum.comp,
 * block_prefix_sum.comp, add_block_sums.comp): in-place exclusive u32 scan
 * with wraparound; returns the u32 total of the first n elements. */
uint32_t orc_exclusive_scan(uint32_t *data, uint32_t n)
{
    uint32_t sum = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = data[i];
        data[i] = sum;
        sum += v;
    }
    return sum;
}

/* stream_compact.comp:26-32: out[offsets[i]] = i for every flags[i] != 0. */
void orc_compact(const uint32_t *flags, const uint32_t *offsets, uint32_t n, uint32_t *out)
{
    for (uint32_t i = 0; i < n; ++i)
        if (flags[i] != 0)
            out[offsets[i]] = i;
}

static uint32_t next_pow2(uint32_t x)
{
    x = x - 1;
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    return x + 1;
}

/* RadixSorter.getAlignedSize, radix_sort_by_key.js:155-158 (chunk = 64). */
uint32_t orc_sort_aligned_size(uint32_t n)
{
    return next_pow2((n + 63) / 64) * 64;
}

/* radix_sort_chunk.comp:28-101: 32 one-bit split passes (zeros first, both
 * halves keep their order) over one 64-element chunk. */
static void sort_chunk(uint32_t *k, uint32_t *v)
{
    uint32_t tk[64], tv[64];
    for (uint32_t bit = 0; bit < 32; ++bit) {
        const uint32_t mask = 1u << bit;
        uint32_t total_false = 0;
        for (int i = 0; i < 64; ++i)
            total_false += (k[i] & mask) ? 0 : 1;
        uint32_t f = 0;
        for (int i = 0; i < 64; ++i) {
            if (k[i] & mask) {
                const uint32_t t = (uint32_t)i - f + total_false;
                tk[t] = k[i];
                tv[t] = v[i];
            } else {
                tk[f] = k[i];
                tv[f] = v[i];
                ++f;
            }
        }
        memcpy(k, tk, sizeof(tk));
        memcpy(v, tv, sizeof(tv));
    }
}

/* upper_bound / lower_bound of merge_sorted_chunks.comp:35-61. */
static uint32_t ub(const uint32_t *keys, uint32_t start, uint32_t count, uint32_t e)
{
    while (count > 0) {
        const uint32_t i = start + count / 2;
        if (e >= keys[i]) {
            start = i + 1;
            count -= count / 2 + 1;
        } else {
            count = count / 2;
        }
    }
    return start;
}

static uint32_t lb(const uint32_t *keys, uint32_t start, uint32_t count, uint32_t e)
{
    while (count > 0) {
        const uint32_t i = start + count / 2;
        if (keys[i] < e) {
            start = i + 1;
            count -= count / 2 + 1;
        } else {
            count = count / 2;
        }
    }
    return start;
}

/* RadixSorter.sort, radix_sort_by_key.js:161-400: pad to the aligned size
 * with UINT_MAX (radix_sort_chunk.comp:32-39), sort 64-element chunks, merge
 * pairs of runs log2(chunks) times (merge_sorted_chunks.comp:63-94: an A
 * element lands after every B element with key <= its own, a B element after
 * every A element with key < its own, so equal keys from B precede those from
 * A), optionally reverse the values over the aligned size
 * (reverse_buffer.comp:15-29), then copy n values from offset
 * aligned-n (reverse) or 0 (radix_sort_by_key.js:382-391).  The key buffer's
 * final contents are unspecified by the reference; here it receives the
 * ascending keys. */
void orc_sort_by_key(uint32_t *keys, uint32_t *vals, uint32_t n, int reverse)
{
    const uint32_t chunks = next_pow2((n + 63) / 64);
    const uint32_t aligned = chunks * 64;
    uint32_t *k = malloc(aligned * sizeof(uint32_t)), *v = malloc(aligned * sizeof(uint32_t));
    uint32_t *k2 = malloc(aligned * sizeof(uint32_t)), *v2 = malloc(aligned * sizeof(uint32_t));
    for (uint32_t i = 0; i < aligned; ++i) {
        k[i] = i < n ? keys[i] : 0xffffffffu;
        v[i] = i < n ? vals[i] : 0xffffffffu;
    }
    for (uint32_t c = 0; c < chunks; ++c)
        sort_chunk(k + 64 * c, v + 64 * c);
    for (uint32_t out_size = 128; out_size <= aligned; out_size *= 2) {
        const uint32_t half = out_size / 2;
        for (uint32_t offs = 0; offs < aligned; offs += out_size) {
            for (uint32_t j = 0; j < half; ++j) {
                const uint32_t a_in = offs + j, b_in = offs + half + j;
                const uint32_t a_loc = j + ub(k, offs + half, half, k[a_in]) - half;
                const uint32_t b_loc = j + lb(k, offs, half, k[b_in]);
                k2[a_loc] = k[a_in];
                v2[a_loc] = v[a_in];
                k2[b_loc] = k[b_in];
                v2[b_loc] = v[b_in];
            }
        }
        uint32_t *t = k;
        k = k2;
        k2 = t;
        t = v;
        v = v2;
        v2 = t;
    }
    if (reverse) {
        for (uint32_t i = 0; i < aligned / 2; ++i) {
            const uint32_t t = v[i];
            v[i] = v[aligned - 1 - i];
            v[aligned - 1 - i] = t;
        }
    }
    const uint32_t off = reverse ? aligned - n : 0;
    memcpy(vals, v + off, n * sizeof(uint32_t));
    memcpy(keys, k, aligned * sizeof(uint32_t));
    free(k);
    free(v);
    free(k2);
    free(v2);
}

/* ======================================================================== */
/* LRU cache                                                                 */
/* ======================================================================== */

struct orc_lru {
    uint32_t total, aligned_total, size, elem_bytes;
    int32_t *item_slot;  /* [aligned_total], cachedItemSlots */
    uint32_t *age, *avail;
    int32_t *item;       /* slot AoS of lru_cache_inputs.comp:4-11, held as SoA */
    uint8_t *data;       /* cache contents, size*elem_bytes */
    uint32_t *needs;
};

/* LRUCache constructor, lru_cache.js:4-95: cacheSize = alignTo(initial, 32),
 * slots {age 2+i, available 1, item -1}, every item slot -1. */
orc_lru *orc_lru_create(uint32_t initial_size, uint32_t elem_bytes, uint32_t total)
{
    orc_lru *c = calloc(1, sizeof(orc_lru));
    c->total = total;
    c->aligned_total = align_to(total, 32);
    c->size = align_to(initial_size, 32);
    c->elem_bytes = elem_bytes;
    c->item_slot = malloc(c->aligned_total * sizeof(int32_t));
    for (uint32_t i = 0; i < c->aligned_total; ++i)
        c->item_slot[i] = -1;
    c->needs = calloc(c->aligned_total, sizeof(uint32_t));
    c->age = malloc((size_t)c->size * sizeof(uint32_t) + 4);
    c->avail = malloc((size_t)c->size * sizeof(uint32_t) + 4);
    c->item = malloc((size_t)c->size * sizeof(int32_t) + 4);
    c->data = calloc((size_t)c->size * elem_bytes + 1, 1);
    for (uint32_t s = 0; s < c->size; ++s) {
        c->age[s] = 2 + s;
        c->avail[s] = 1;
        c->item[s] = -1;
    }
    return c;
}

void orc_lru_destroy(orc_lru *c)
{
    if (!c)
        return;
    free(c->item_slot);
    free(c->needs);
    free(c->age);
    free(c->avail);
    free(c->item);
    free(c->data);
    free(c);
}

/* LRUCache.update, lru_cache.js:312-843. */
uint32_t orc_lru_update(orc_lru *c, const uint32_t *needed, uint32_t *new_items)
{
    /* lru_cache_age_slots.comp:14-21 */
    for (uint32_t s = 0; s < c->size; ++s)
        c->age[s] += 1;
    /* lru_cache_mark_new_items.comp:26-44 over alignTo(N,32) items; entries
     * past N count as not needed. */
    for (uint32_t i = 0; i < c->aligned_total; ++i) {
        const uint32_t nd = i < c->total ? needed[i] : 0;
        const int32_t s = c->item_slot[i];
        if (s >= 0) {
            c->needs[i] = 0;
            if (nd == 1) {
                c->age[s] = 0;
                c->avail[s] = 0;
            } else {
                c->avail[s] = 1;
            }
        } else {
            c->needs[i] = nd;
        }
    }
    /* scan + early exit, lru_cache.js:463-481 */
    uint32_t num_new = 0;
    for (uint32_t i = 0; i < c->total; ++i)
        num_new += c->needs[i];
    if (num_new == 0)
        return 0;
    /* compact new item ids, lru_cache.js:485-492 */
    {
        uint32_t off = 0;
        for (uint32_t i = 0; i < c->total; ++i)
            if (c->needs[i] != 0)
                new_items[off++] = i;
    }
    /* scan available slots, lru_cache.js:501 */
    uint32_t num_avail = 0;
    for (uint32_t s = 0; s < c->size; ++s)
        num_avail += c->avail[s];
    /* grow, lru_cache.js:510-672: new slots {100000, 1, -1}, contents kept */
    if (num_avail < num_new) {
        const uint64_t want = (uint64_t)c->size + ((uint64_t)(num_new - num_avail) * 3 + 1) / 2;
        uint32_t new_size = (uint32_t)(want < c->total ? want : c->total);
        new_size = align_to(new_size, 32);
        c->age = realloc(c->age, (size_t)new_size * sizeof(uint32_t) + 4);
        c->avail = realloc(c->avail, (size_t)new_size * sizeof(uint32_t) + 4);
        c->item = realloc(c->item, (size_t)new_size * sizeof(int32_t) + 4);
        c->data = realloc(c->data, (size_t)new_size * c->elem_bytes + 1);
        memset(c->data + (size_t)c->size * c->elem_bytes, 0,
               (size_t)(new_size - c->size) * c->elem_bytes);
        for (uint32_t s = c->size; s < new_size; ++s) {
            c->age[s] = 100000;
            c->avail[s] = 1;
            c->item[s] = -1;
        }
        c->size = new_size;
        num_avail = 0;
        for (uint32_t s = 0; s < c->size; ++s)
            num_avail += c->avail[s];
    }
    /* available slot ids, ages, sort descending, lru_cache.js:684-769 */
    const uint32_t aligned = orc_sort_aligned_size(num_avail);
    uint32_t *keys = malloc(aligned * sizeof(uint32_t));
    uint32_t *ids = malloc(aligned * sizeof(uint32_t));
    {
        uint32_t off = 0;
        for (uint32_t s = 0; s < c->size; ++s)
            if (c->avail[s] != 0)
                ids[off++] = s;
        for (uint32_t j = 0; j < num_avail; ++j)
            keys[j] = c->age[ids[j]];
    }
    orc_sort_by_key(keys, ids, num_avail, 1);
    /* lru_cache_update.comp:19-37 */
    for (uint32_t w = 0; w < num_new; ++w) {
        const uint32_t it = new_items[w], s = ids[w];
        const int32_t prev = c->item[s];
        if (prev != -1)
            c->item_slot[prev] = -1;
        c->age[s] = 0;
        c->item[s] = (int32_t)it;
        c->avail[s] = 0;
        c->item_slot[it] = (int32_t)s;
    }
    free(keys);
    free(ids);
    return num_new;
}

/* LRUCache.reset, lru_cache.js:846-903 with lru_cache_init.comp:19-30. */
void orc_lru_reset(orc_lru *c)
{
    for (uint32_t s = 0; s < c->size; ++s) {
        c->age[s] = 100000;
        c->avail[s] = 1;
        c->item[s] = -1;
    }
    for (uint32_t i = 0; i < c->total; ++i)
        c->item_slot[i] = -1;
}

uint32_t orc_lru_size(const orc_lru *c) { return c->size; }
uint8_t *orc_lru_data(orc_lru *c) { return c->data; }

void orc_lru_state(const orc_lru *c, uint32_t *age, uint32_t *avail, int32_t *item,
                   int32_t *item_slot)
{
    if (age)
        memcpy(age, c->age, c->size * sizeof(uint32_t));
    if (avail)
        memcpy(avail, c->avail, c->size * sizeof(uint32_t));
    if (item)
        memcpy(item, c->item, c->size * sizeof(int32_t));
    if (item_slot)
        memcpy(item_slot, c->item_slot, c->total * sizeof(int32_t));
}

/* ======================================================================== */
/* Marching cubes                                                            */
/* ======================================================================== */

int orc_tri_count(int c) { return (int)strlen(ORC_TRI_CASES[c]); }

int orc_tri_edge(int c, int k)
{
    const char ch = ORC_TRI_CASES[c][k];
    return ch <= '9' ? ch - '0' : ch - 'a' + 10;
}

/* index_to_vertex, compute_vertex_values.comp:48-57 */
static const int I2V[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
/* edge_vertices, compute_block_vertices.comp:10-23 */
static const int EDGE_V[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {6, 5},
                                  {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct orc_mc {
    uint64_t *words;
    size_t nwords;
    uint32_t maxbits, dims[3], pdims[3], nb[3], total;
    int threads;
    float *ranges;
    orc_lru *lru;
    uint32_t *flags, *active, *new_items, *has, *with_verts, *offsets, *counts;
    uint32_t num_active, num_new, num_with;
    uint64_t num_verts;
    float *vf32;
    double *vf64;
    size_t vcap;
};

/* CompressedMarchingCubes.setCompressedVolume (compressed_marching_cubes.js:
 * 432-520) + computeBlockRanges (:522-578).  initial_cache = the LRU's initial
 * slot count (the reference hard-codes ceil(0.05*N), :443-448). */
orc_mc *orc_mc_create(const uint8_t *stream, size_t bytes, uint32_t maxbits,
                      const uint32_t dims[3], uint32_t initial_cache, int threads)
{
    orc_mc *m = calloc(1, sizeof(orc_mc));
    m->nwords = (bytes + 7) / 8;
    m->words = calloc(m->nwords + 1, 8);
    memcpy(m->words, stream, bytes);
    m->maxbits = maxbits;
    m->threads = threads > 0 ? threads : 1;
    for (int i = 0; i < 3; ++i) {
        m->dims[i] = dims[i];
        m->pdims[i] = align_to(dims[i], 4);
        m->nb[i] = m->pdims[i] / 4;
    }
    m->total = m->nb[0] * m->nb[1] * m->nb[2];
    m->ranges = malloc((size_t)m->total * 2 * sizeof(float));
    orc_block_ranges(m->words, m->nwords, maxbits, dims, m->ranges, m->threads);
    m->lru = orc_lru_create(initial_cache, 256, m->total);
    const size_t n = m->total;
    m->flags = calloc(n + 1, 4);
    m->active = calloc(n + 1, 4);
    m->new_items = calloc(n + 1, 4);
    m->has = calloc(n + 1, 4);
    m->with_verts = calloc(n + 1, 4);
    m->offsets = calloc(n + 1, 4);
    m->counts = calloc(n + 1, 4);
    return m;
}

void orc_mc_destroy(orc_mc *m)
{
    if (!m)
        return;
    free(m->words);
    free(m->ranges);
    orc_lru_destroy(m->lru);
    free(m->flags);
    free(m->active);
    free(m->new_items);
    free(m->has);
    free(m->with_verts);
    free(m->offsets);
    free(m->counts);
    free(m->vf32);
    free(m->vf64);
    free(m);
}

static uint32_t block_id(const orc_mc *m, uint32_t x, uint32_t y, uint32_t z)
{
    return x + m->nb[0] * (y + m->nb[1] * z);
}

/* compute_block_dims_with_ghost (compute_vertex_values.comp:96-156) followed by
 * load_block (:178-252): a (dims.x, dims.y, dims.z) <= 5^3 tile of the block
 * and the ghost voxels of its +x/+y/+z face, edge and corner neighbours, read
 * from the decoded-block cache through the item->slot map.  Returns dims. */
static void load_tile(const orc_mc *m, uint32_t bid, float tile[125], uint32_t d[3])
{
    const uint32_t bx = bid % m->nb[0], by = (bid / m->nb[0]) % m->nb[1],
                   bz = bid / (m->nb[0] * m->nb[1]);
    const uint32_t *act = m->flags;
    d[0] = bx + 1 < m->nb[0] ? 5 : 4;
    d[1] = by + 1 < m->nb[1] ? 5 : 4;
    d[2] = bz + 1 < m->nb[2] ? 5 : 4;
    if (d[0] == 5 && d[1] == 5 && d[2] == 5 && act[block_id(m, bx + 1, by + 1, bz + 1)] == 0)
        d[0] = d[1] = d[2] = 4;
    if (d[0] == 5 && d[1] == 5 && act[block_id(m, bx + 1, by + 1, bz)] == 0)
        d[0] = d[1] = 4;
    if (d[0] == 5 && d[2] == 5 && act[block_id(m, bx + 1, by, bz + 1)] == 0)
        d[0] = d[2] = 4;
    if (d[1] == 5 && d[2] == 5 && act[block_id(m, bx, by + 1, bz + 1)] == 0)
        d[1] = d[2] = 4;
    if (d[0] == 5 && act[block_id(m, bx + 1, by, bz)] == 0)
        d[0] = 4;
    if (d[1] == 5 && act[block_id(m, bx, by + 1, bz)] == 0)
        d[1] = 4;
    if (d[2] == 5 && act[block_id(m, bx, by, bz + 1)] == 0)
        d[2] = 4;
    for (int i = 0; i < 125; ++i)
        tile[i] = 0.0f;
    const float *cache = (const float *)orc_lru_data(m->lru);
    const int32_t *slot = m->lru->item_slot;
    /* Every tile voxel (x,y,z) < d comes from the block whose position is
     * (bx + x/4, by + y/4, bz + z/4) at voxel (x%4, y%4, z%4); the reference
     * loads exactly these (own block, faces, edges, corner). */
    for (uint32_t z = 0; z < d[2]; ++z)
        for (uint32_t y = 0; y < d[1]; ++y)
            for (uint32_t x = 0; x < d[0]; ++x) {
                const uint32_t nid = block_id(m, bx + x / 4, by + y / 4, bz + z / 4);
                const int32_t s = slot[nid];
                tile[x + d[0] * (y + d[1] * z)] =
                    cache[(size_t)s * 64 + (x % 4) + 4 * ((y % 4) + 4 * (z % 4))];
            }
}

/* Case index of cell (x,y,z): bit i set iff corner i <= iso
 * (compute_block_has_vertices.comp:30-37). */
static int cell_case(const float *tile, const uint32_t d[3], uint32_t x, uint32_t y,
                     uint32_t z, float iso, float vals[8])
{
    int c = 0;
    for (int i = 0; i < 8; ++i) {
        const uint32_t vx = x + I2V[i][0], vy = y + I2V[i][1], vz = z + I2V[i][2];
        vals[i] = tile[(vz * d[1] + vy) * d[0] + vx];
        if (vals[i] <= iso)
            c |= 1 << i;
    }
    return c;
}

/* Per-block vertex total: compute_block_has_vertices.comp:22-66 and
 * compute_block_voxel_num_verts.comp:25-77 (cells with pos+1 < dims). */
static uint32_t block_count(const orc_mc *m, uint32_t bid, float iso)
{
    float tile[125], vals[8];
    uint32_t d[3], total = 0;
    load_tile(m, bid, tile, d);
    for (uint32_t z = 0; z + 1 < d[2] && z < 4; ++z)
        for (uint32_t y = 0; y + 1 < d[1] && y < 4; ++y)
            for (uint32_t x = 0; x + 1 < d[0] && x < 4; ++x)
                total += (uint32_t)orc_tri_count(cell_case(tile, d, x, y, z, iso, vals));
    return total;
}

/* compute_block_vertices.comp:76-174: cells in ascending x+4y+16z, vertices
 * in table order; position = lerp(corner a, corner b) + voxel + 0.5 in FP32
 * (lerp_verts :62-74), plus the same position with t interpolated in FP64
 * (the |fa-fb| < 0.001 branch is evaluated in FP32 as in the reference). */
static void block_emit(const orc_mc *m, uint32_t bid, float iso, float *o32, double *o64)
{
    float tile[125], vals[8];
    uint32_t d[3];
    uint64_t k = 0;
    load_tile(m, bid, tile, d);
    union {
        uint32_t u;
        float f;
    } bidf;
    bidf.u = bid;
    for (uint32_t z = 0; z < 4; ++z)
        for (uint32_t y = 0; y < 4; ++y)
            for (uint32_t x = 0; x < 4; ++x) {
                if (!(x + 1 < d[0] && y + 1 < d[1] && z + 1 < d[2]))
                    continue;
                const int c = cell_case(tile, d, x, y, z, iso, vals);
                const int nv = orc_tri_count(c);
                for (int j = 0; j < nv; ++j, ++k) {
                    const int e = orc_tri_edge(c, j);
                    const int a = EDGE_V[e][0], b = EDGE_V[e][1];
                    const float fa = vals[a], fb = vals[b];
                    float t = 0.0f;
                    double t64 = 0.0;
                    if (!(fabsf(fa - fb) < 0.001f)) {
                        t = (iso - fa) / (fb - fa);
                        t64 = ((double)iso - (double)fa) / ((double)fb - (double)fa);
                    }
                    const uint32_t vp[3] = {x, y, z};
                    for (int ax = 0; ax < 3; ++ax) {
                        const float va = (float)I2V[a][ax], vb = (float)I2V[b][ax];
                        const float lerp = va + t * (vb - va);
                        o32[4 * k + ax] = (lerp + (float)vp[ax]) + 0.5f;
                        o64[3 * k + ax] = (double)I2V[a][ax] +
                                          t64 * (double)(I2V[b][ax] - I2V[a][ax]) +
                                          (double)vp[ax] + 0.5;
                    }
                    o32[4 * k + 3] = bidf.f;
                }
            }
}

/* CompressedMarchingCubes.computeSurface, compressed_marching_cubes.js:580-809. */
int orc_mc_compute_surface(orc_mc *m, float iso, orc_mc_stats *st)
{
    orc_mc_stats s;
    memset(&s, 0, sizeof(s));
    double t0 = now_s();
    /* computeActiveBlocks (:811-829) */
    orc_block_active(m->ranges, m->dims, iso, m->flags, m->threads);
    uint32_t a = 0;
    for (uint32_t i = 0; i < m->total; ++i)
        a += m->flags[i];
    m->num_active = a;
    s.num_active = a;
    double t1 = now_s();
    s.t_active = t1 - t0;
    m->num_new = 0;
    m->num_with = 0;
    m->num_verts = 0;
    if (a == 0) { /* :630-641: no cache update when nothing is active */
        s.cache_size = orc_lru_size(m->lru);
        if (st)
            *st = s;
        return 0;
    }
    /* LRUCache.update (:648-649) */
    m->num_new = orc_lru_update(m->lru, m->flags, m->new_items);
    s.num_decoded = m->num_new;
    s.cache_size = orc_lru_size(m->lru);
    double t2 = now_s();
    s.t_lru = t2 - t1;
    /* decompressBlocks (:831-913; zfp_decompress_block.comp:29-39) */
    {
        float *cache = (float *)orc_lru_data(m->lru);
        const int32_t *slot = m->lru->item_slot;
        const uint32_t d = m->num_new;
#pragma omp parallel for schedule(static) num_threads(m->threads)
        for (int64_t w = 0; w < (int64_t)d; ++w) {
            const uint32_t it = m->new_items[w];
            orc_decode_block(m->words, m->nwords, m->maxbits, it, cache + (size_t)slot[it] * 64);
        }
    }
    double t3 = now_s();
    s.t_decode = t3 - t2;
    /* compactActiveIDs (:704-707) */
    {
        uint32_t off = 0;
        for (uint32_t i = 0; i < m->total; ++i)
            if (m->flags[i])
                m->active[off++] = i;
    }
    /* computeBlockHasVertices + scan + compact (:715-755) and
     * computeBlockVertexCounts + scan (:993-1042) */
#pragma omp parallel for schedule(dynamic, 256) num_threads(m->threads)
    for (int64_t k = 0; k < (int64_t)a; ++k)
        m->counts[k] = block_count(m, m->active[k], iso);
    uint64_t v = 0;
    uint32_t h = 0;
    for (uint32_t k = 0; k < a; ++k) {
        if (m->counts[k] > 0) {
            m->with_verts[h] = k;
            m->offsets[h] = (uint32_t)v;
            v += m->counts[k];
            ++h;
        }
    }
    m->num_with = h;
    m->num_verts = v;
    s.num_with_verts = h;
    s.num_verts = v;
    double t4 = now_s();
    s.t_count = t4 - t3;
    /* computeVertices (:1044-1122) */
    if (v > m->vcap) {
        free(m->vf32);
        free(m->vf64);
        m->vcap = v;
        m->vf32 = malloc(v * 4 * sizeof(float));
        m->vf64 = malloc(v * 3 * sizeof(double));
    }
    {
        /* block h starts at the 64-bit running offset; recompute it so that
         * totals past 2^32 stay exact */
        uint64_t *off64 = malloc(((size_t)h + 1) * sizeof(uint64_t));
        uint64_t run = 0;
        for (uint32_t i = 0; i < h; ++i) {
            off64[i] = run;
            run += m->counts[m->with_verts[i]];
        }
#pragma omp parallel for schedule(dynamic, 256) num_threads(m->threads)
        for (int64_t i = 0; i < (int64_t)h; ++i) {
            const uint32_t bid = m->active[m->with_verts[i]];
            block_emit(m, bid, iso, m->vf32 + 4 * off64[i], m->vf64 + 3 * off64[i]);
        }
        free(off64);
    }
    s.t_emit = now_s() - t4;
    if (st)
        *st = s;
    return 0;
}

const uint32_t *orc_mc_flags(const orc_mc *m) { return m->flags; }
const uint32_t *orc_mc_active_ids(const orc_mc *m) { return m->active; }
const uint32_t *orc_mc_new_items(const orc_mc *m) { return m->new_items; }
const uint32_t *orc_mc_with_verts(const orc_mc *m) { return m->with_verts; }
const uint32_t *orc_mc_vert_offsets(const orc_mc *m) { return m->offsets; }
const float *orc_mc_verts_f32(const orc_mc *m) { return m->vf32; }
const double *orc_mc_verts_f64(const orc_mc *m) { return m->vf64; }
const float *orc_mc_ranges(const orc_mc *m) { return m->ranges; }
orc_lru *orc_mc_lru(orc_mc *m) { return m->lru; }
uint32_t orc_mc_total_blocks(const orc_mc *m) { return m->total; }

/* Brute force: every cell of the padded volume with the same <= iso rule and
 * triangle table.  Independent of the block/ghost machinery. */
uint64_t orc_mc_bruteforce_count(const float *vol, const uint32_t p[3], float iso)
{
    uint64_t total = 0;
    for (uint32_t z = 0; z + 1 < p[2]; ++z)
        for (uint32_t y = 0; y + 1 < p[1]; ++y)
            for (uint32_t x = 0; x + 1 < p[0]; ++x) {
                int c = 0;
                for (int i = 0; i < 8; ++i) {
                    const size_t idx = ((size_t)(z + I2V[i][2]) * p[1] + y + I2V[i][1]) * p[0] +
                                       x + I2V[i][0];
                    if (vol[idx] <= iso)
                        c |= 1 << i;
                }
                total += (uint64_t)orc_tri_count(c);
            }
    return total;
}
