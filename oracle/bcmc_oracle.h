/* CPU oracle for the block-compressed marching-cubes path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * method (Twinklebear/webgpu-bcmc: the compute shaders and pipeline JS), used by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark as the
 * checker.  The product path never links, loads or calls it.
 *
 * Every function cites the reference file:line range it restates.  Parity of
 * each function is pinned by the tests (libzfp 0.5.5 fixtures, closed forms,
 * brute force); see tests/test_oracle.py.
 */
#ifndef BCMC_ORACLE_H
#define BCMC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ZFP fixed-rate decode (shaders/zfp_decompress.comp) ---------------- */
/* Decode block `block` of a headerless fixed-rate float32 3-D zfp stream into
 * out[64] (x fastest).  maxbits = 64*rate.  nwords = stream length in u64
 * words; bits past the end read as 0. */
void orc_decode_block(const uint64_t *words, size_t nwords, uint32_t maxbits,
                      uint64_t block, float out[64]);
/* Decode the whole padded volume: out[(pz)(py)(px)] with padded dims. */
void orc_decode_volume(const uint64_t *words, size_t nwords, uint32_t maxbits,
                       const uint32_t dims[3], float *out, int threads);

/* ---- block ranges and active test -------------------------------------- */
void orc_block_ranges(const uint64_t *words, size_t nwords, uint32_t maxbits,
                      const uint32_t dims[3], float *ranges /* 2N */, int threads);
void orc_block_active(const float *ranges, const uint32_t dims[3], float iso,
                      uint32_t *flags /* N */, int threads);

/* ---- scan / compact / sort --------------------------------------------- */
uint32_t orc_exclusive_scan(uint32_t *data, uint32_t n); /* in place, returns total */
void orc_compact(const uint32_t *flags, const uint32_t *offsets, uint32_t n,
                 uint32_t *out);
uint32_t orc_sort_aligned_size(uint32_t n);
/* Literal restatement of RadixSorter.sort: keys/vals have aligned size. */
void orc_sort_by_key(uint32_t *keys, uint32_t *vals, uint32_t n, int reverse);

/* ---- LRU cache (js/lru_cache.js) ---------------------------------------- */
typedef struct orc_lru orc_lru;
orc_lru *orc_lru_create(uint32_t initial_size, uint32_t elem_bytes, uint32_t total);
void orc_lru_destroy(orc_lru *);
/* needed[total] (0/1).  Returns numNew; new_items receives the new ids
 * (capacity >= total). */
uint32_t orc_lru_update(orc_lru *, const uint32_t *needed, uint32_t *new_items);
void orc_lru_reset(orc_lru *);
uint32_t orc_lru_size(const orc_lru *);
/* Copies out slot state: age[C], avail[C], item[C] and item_slot[total]. */
void orc_lru_state(const orc_lru *, uint32_t *age, uint32_t *avail, int32_t *item,
                   int32_t *item_slot);
uint8_t *orc_lru_data(orc_lru *);

/* ---- marching cubes pipeline (js/compressed_marching_cubes.js) ---------- */
typedef struct orc_mc orc_mc;
orc_mc *orc_mc_create(const uint8_t *stream, size_t bytes, uint32_t maxbits,
                      const uint32_t dims[3], uint32_t initial_cache, int threads);
void orc_mc_destroy(orc_mc *);

typedef struct {
    uint32_t num_active;      /* A */
    uint32_t num_decoded;     /* D (LRU misses decoded this query) */
    uint32_t num_with_verts;  /* H */
    uint64_t num_verts;       /* V */
    uint32_t cache_size;      /* LRU slots after the query */
    double t_active, t_lru, t_decode, t_count, t_emit; /* seconds */
} orc_mc_stats;

/* Run one query.  Results stay inside the object until the next call. */
int orc_mc_compute_surface(orc_mc *, float iso, orc_mc_stats *stats);
/* Accessors for the last query's arrays. */
const uint32_t *orc_mc_flags(const orc_mc *);          /* N */
const uint32_t *orc_mc_active_ids(const orc_mc *);     /* A */
const uint32_t *orc_mc_new_items(const orc_mc *);      /* D */
const uint32_t *orc_mc_with_verts(const orc_mc *);     /* H: indices into active */
const uint32_t *orc_mc_vert_offsets(const orc_mc *);   /* H: exclusive scan */
const float *orc_mc_verts_f32(const orc_mc *);         /* V*4: x,y,z,bits(block) */
const double *orc_mc_verts_f64(const orc_mc *);        /* V*3 block-local */
const float *orc_mc_ranges(const orc_mc *);            /* 2N */
orc_lru *orc_mc_lru(orc_mc *);
uint32_t orc_mc_total_blocks(const orc_mc *);

/* Brute-force MC vertex count over all cells of a dense padded volume. */
uint64_t orc_mc_bruteforce_count(const float *vol, const uint32_t pdims[3], float iso);
/* Per-case vertex count from the triangle table. */
int orc_tri_count(int c);
int orc_tri_edge(int c, int k);

#ifdef __cplusplus
}
#endif
#endif
