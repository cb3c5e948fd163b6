// Marching-cubes device code shared by the per-query pipeline (mc.hip) and
// the batched/served engine (engine.hip).
//
// Work layout: a wave handles 4 blocks at a time, one per 16-lane DPP row.
// Lane l = y + 4z of a row owns the 4 cells (x = 0..3, y, z) of its block and
// holds the voxel row (x = 0..4, y, z) in registers; the rows at y + 1 and
// z + 1 come from lanes l + 1, l + 4 and l + 5 through DPP row shifts, and the
// lanes on the block's upper faces also load the ghost rows from the +y/+z
// neighbours.  The corner tests (v <= iso) of the 4 rows a lane's cells touch
// are interleaved into one word W with row r's bit at voxel x in bit 4x + r,
// so the 8 corners of cell x are the byte (W >> 4x) & 0xFF; the triangle
// tables are re-indexed by that byte once per workgroup.
//
// Chunks of kChunk active blocks go to a workgroup; a count sweep stores each
// block's cache slots and tile dims in LDS, so the emit sweep only reloads
// voxels.
#pragma once

#include "common.h"
#include "tri_table.h"

namespace bcmc {

static __constant__ unsigned long long kTriEdges[256] = BCMC_TRI_EDGES_INIT;
static __constant__ uint8_t kTriCount[256] = BCMC_TRI_COUNT_INIT;
// edge_vertices (compute_block_vertices.comp:10-23): edge e joins corners
// a = byte e & 15 and b = byte e >> 4 (bytes from the low end; 12 edges)
constexpr unsigned long long kEdgePairsLo = 0x4776565403322110ull;  // edges 0..7
constexpr unsigned long long kEdgePairsHi = 0x73625140ull;          // edges 8..11

// index_to_vertex (compute_vertex_values.comp:48-57): corner i at
// (x, y, z) = (((i+1)>>1)&1, (i>>1)&1, i>>2)
__host__ __device__ constexpr uint32_t corner_x(uint32_t i) { return ((i + 1) >> 1) & 1u; }
__host__ __device__ constexpr uint32_t corner_y(uint32_t i) { return (i >> 1) & 1u; }
__host__ __device__ constexpr uint32_t corner_z(uint32_t i) { return i >> 2; }

constexpr int kMcThreads = 256;          // 4 waves per workgroup
constexpr int kRowBlocks = 4;            // blocks per wave (one per 16-lane row)
constexpr uint32_t kChunk = kMcThreads;  // active blocks per chunk (one per thread in its scan)
// Block tile in LDS: 5^3 voxels at x + 8y + 40z (rows 32-byte aligned).
constexpr uint32_t kTileY = 8, kTileZ = 40, kTileFloats = 200;

// Corner byte B of a cell (bit 4 dx + r, r = dy + 2 dz over the rows
// (y, z), (y+1, z), (y, z+1), (y+1, z+1)) -> reference case index (bit i iff
// corner i <= iso).  Corner i sits at B bit kCornerBit[i].
__host__ __device__ constexpr uint32_t corner_byte_bit(uint32_t i) {
    return 4 * corner_x(i) + corner_y(i) + 2 * corner_z(i);
}
__host__ __device__ constexpr uint32_t case_of_corner_byte(uint32_t byte) {
    uint32_t c = 0;
    for (uint32_t i = 0; i < 8; ++i)
        c |= ((byte >> corner_byte_bit(i)) & 1u) << i;
    return c;
}

// Interpolation data of edge e: corner a's position, b - a, and both corners'
// offsets in the tile.
struct alignas(16) EdgeInfo {
    float a[3];
    float off_a;
    float d[3];
    float off_b;
};

// Tables staged in LDS once per workgroup (the lookups are divergent, so
// they would otherwise be vector loads through the L1).  The count pass needs
// only the vertex counts and the tile-dims rule.
struct CountTables {
    uint8_t count[256];  // by corner byte: vertices of the cell
    uint8_t dims[1024];  // tile-dims rule, see tile_dims_rule
};

struct McTables {
    unsigned long long edges[256];  // by corner byte
    uint8_t count[256];             // by corner byte: vertices of the cell
    uint8_t dims[1024];             // tile-dims rule, see tile_dims_rule
    EdgeInfo edge[12];
};

// compute_block_dims_with_ghost (compute_vertex_values.comp:96-156): the
// ghost planes are part of the tile only where the +x/+y/+z neighbours (and,
// for the shared edges and corner, the diagonal neighbours) are active.
// idx bits 0..6: neighbour +x, +y, +z, +xy, +xz, +yz, +xyz active; bits 7..9:
// the block is not last in x, y, z.  Returns dx5 | dy5 << 1 | dz5 << 2.
BCMC_DEV uint32_t tile_dims_rule(uint32_t idx) {
    const uint32_t a = idx & 0x7Fu;
    bool dx = (idx >> 7) & 1u, dy = (idx >> 8) & 1u, dz = (idx >> 9) & 1u;
    if (dx && dy && dz && !(a & 64))
        dx = dy = dz = false;
    if (dx && dy && !(a & 8))
        dx = dy = false;
    if (dx && dz && !(a & 16))
        dx = dz = false;
    if (dy && dz && !(a & 32))
        dy = dz = false;
    if (dx && !(a & 1))
        dx = false;
    if (dy && !(a & 2))
        dy = false;
    if (dz && !(a & 4))
        dz = false;
    return (dx ? 1u : 0u) | (dy ? 2u : 0u) | (dz ? 4u : 0u);
}

BCMC_DEV void load_count_tables(CountTables *t) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
        t->count[i] = kTriCount[case_of_corner_byte(i)];
    for (uint32_t i = threadIdx.x; i < 1024; i += blockDim.x)
        t->dims[i] = (uint8_t)tile_dims_rule(i);
    __syncthreads();
}

BCMC_DEV void load_mc_tables(McTables *t) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        const uint32_t c = case_of_corner_byte(i);
        t->edges[i] = kTriEdges[c];
        t->count[i] = kTriCount[c];
    }
    for (uint32_t i = threadIdx.x; i < 1024; i += blockDim.x)
        t->dims[i] = (uint8_t)tile_dims_rule(i);
    if (threadIdx.x < 12) {
        const uint32_t e = threadIdx.x;
        const uint32_t pair = (uint32_t)((e < 8 ? kEdgePairsLo >> (8 * e) : kEdgePairsHi >> (8 * (e - 8))) & 0xFFu);
        const uint32_t ca = pair & 15u, cb = pair >> 4;
        EdgeInfo &g = t->edge[e];
        g.a[0] = (float)corner_x(ca);
        g.a[1] = (float)corner_y(ca);
        g.a[2] = (float)corner_z(ca);
        g.d[0] = (float)corner_x(cb) - g.a[0];
        g.d[1] = (float)corner_y(cb) - g.a[1];
        g.d[2] = (float)corner_z(cb) - g.a[2];
        g.off_a = __uint_as_float(corner_x(ca) + kTileY * corner_y(ca) + kTileZ * corner_z(ca));
        g.off_b = __uint_as_float(corner_x(cb) + kTileY * corner_y(cb) + kTileZ * corner_z(cb));
    }
    __syncthreads();
}

BCMC_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of lane l + N of the same 16-lane row (0 past the row's end).
template <int N>
BCMC_DEV uint32_t row_from_next(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + N, 0xF, 0xF, false);  // row_shl:N
}

// Inclusive prefix sum within each 16-lane row.
BCMC_DEV uint32_t row_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    return v;
}

// Inclusive running maximum across the wave (same DPP pattern as the sum).
BCMC_DEV uint32_t wave_incl_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));
    return v;
}

// n / d for n < 2^31 as a multiply-high: with c = ceil(log2 d), k = 31 + c and
// m = ceil(2^k / d), floor(n m / 2^k) = floor(n / d) because the rounding error
// n (m d - 2^k) / (d 2^k) stays below 2^(31-k) <= 1/d.
struct FastDiv {
    uint32_t m, k;
};

inline FastDiv make_fast_div(uint32_t d) {
    uint32_t c = 0;
    while ((1ull << c) < d)
        ++c;
    const uint32_t k = 31 + c;
    return FastDiv{(uint32_t)(((1ull << k) + d - 1) / d), k};
}

BCMC_DEV uint32_t fast_div(uint32_t n, FastDiv f) { return (uint32_t)(((uint64_t)n * f.m) >> f.k); }

// The block a row works on, as seen by each of its lanes.  In the per-query
// pipeline every field but bid/live is uniform; in the batched engine rows
// may belong to different volumes and isovalues.
struct RowBlock {
    const float *cache;        // decoded blocks, 64 floats per slot; slot `cap` reads as zeros
    const int32_t *item_slot;  // block -> cache slot (< 0: not cached)
    const uint32_t *flags;     // active flags of the query
    uint32_t nb[3];
    FastDiv div_x, div_y;      // by nb[0], nb[1]
    uint32_t cap;              // cache slots
    uint32_t tag;              // a neighbour is active iff its flag word equals tag
    float iso;
    uint32_t bid;
    bool live;
};

BCMC_DEV void block_coords(const RowBlock &b, uint32_t &bx, uint32_t &by, uint32_t &bz) {
    const uint32_t q1 = fast_div(b.bid, b.div_x), q2 = fast_div(q1, b.div_y);
    bx = b.bid - q1 * b.nb[0];
    by = q1 - q2 * b.nb[1];
    bz = q2;
}

// Per-block words kept in LDS between the count and emit sweeps of a chunk:
// the cache slots of the 8 sub-blocks (s = sx + 2sy + 4sz, -1 when absent)
// and the tile dims (dx5 | dy5 << 1 | dz5 << 2).
constexpr uint32_t kMetaWords = 9;

// One load per lane: lanes 0..7 of the row fetch the cache slot of sub-block
// l, lanes 8..14 the active flag of neighbour +x, +y, +z, +xy, +xz, +yz,
// +xyz (the reference's check order).
struct RowMeta {
    int32_t v;
    uint32_t grid;  // block not last in x | y << 1 | z << 2
};

BCMC_DEV RowMeta row_meta_load(const RowBlock &b) {
    const uint32_t l = lane_id() & 15u;
    uint32_t bx, by, bz;
    block_coords(b, bx, by, bz);
    RowMeta m;
    m.grid = (bx + 1 < b.nb[0] ? 1u : 0u) | (by + 1 < b.nb[1] ? 2u : 0u) | (bz + 1 < b.nb[2] ? 4u : 0u);
    // lanes 8..14: neighbour code (x | y << 1 | z << 2) 1, 2, 4, 3, 5, 6, 7,
    // one nibble each of 0x7653421 (a select, not a branch tree)
    const uint32_t code = l < 8 ? l : ((0x7653421u >> (4u * (l - 8u))) & 15u);
    const uint32_t x = bx + (code & 1u), y = by + ((code >> 1) & 1u), z = bz + (code >> 2);
    m.v = l < 8 ? -1 : 0;
    if (b.live && l < 15 && x < b.nb[0] && y < b.nb[1] && z < b.nb[2]) {
        const uint32_t idx = x + b.nb[0] * (y + b.nb[1] * z);
        // raw words: testing the flag here would wait for every load in flight
        m.v = l < 8 ? b.item_slot[idx] : (int32_t)b.flags[idx];
    }
    return m;
}

// Stores the row's slots and tile dims into its block's meta words.  Must be
// called by all 64 lanes.
BCMC_DEV void row_meta_store(const RowMeta &m, uint32_t tag, const uint8_t *dims, uint32_t *meta) {
    const uint32_t lane = lane_id(), l = lane & 15u;
    const uint64_t ab = __ballot(l >= 8 && l < 15 && (uint32_t)m.v == tag);
    const uint32_t a = (uint32_t)(ab >> ((lane & 48u) + 8u)) & 0x7Fu;
    if (l < 8)
        meta[l] = (uint32_t)m.v;
    if (l == 8)
        meta[8] = dims[a | (m.grid << 7)];
}

// Voxels held by a lane: its row and, on the upper faces, the ghost rows
// (y+1 | z+1 | y+1,z+1) from the neighbouring blocks.
struct RowTile {
    float r[5], gy[5], gz[5], gyz[5];
    uint32_t dims;  // dx5 | dy5 << 1 | dz5 << 2
    bool live;
};

BCMC_DEV void row_fetch(const RowBlock &b, int32_t s, int32_t sx, uint32_t off, float (&r)[5]) {
    // a slot that is absent (negative: huge as unsigned) or past the cache
    // reads the zero slot `cap`
    const uint32_t si = min((uint32_t)s, b.cap);
    const uint32_t sj = min((uint32_t)sx, b.cap);
    const float4 v = *reinterpret_cast<const float4 *>(b.cache + (uint64_t)si * 64 + off);
    r[0] = v.x;
    r[1] = v.y;
    r[2] = v.z;
    r[3] = v.w;
    r[4] = b.cache[(uint64_t)sj * 64 + off];
}

// load_block (compute_vertex_values.comp:171-252) for this lane's rows, from
// the block's meta words: own row (x = 0..3 of the block, x = 4 from +x), then
// the ghost rows of the upper faces.  Absent blocks read the zero slot.
BCMC_DEV void row_voxels(const RowBlock &b, const uint32_t *meta, RowTile &t) {
    const uint32_t l = lane_id() & 15u, y = l & 3u, z = l >> 2;
    t.dims = meta[8];
    t.live = b.live;
    // all 8 words are read (LDS; selects, not branches); rows without a
    // block and lanes off the upper faces take -1 (absent)
    const bool gy = b.live && y == 3, gz = b.live && z == 3;
    const int32_t s0 = b.live ? (int32_t)meta[0] : -1, s1 = b.live ? (int32_t)meta[1] : -1;
    const int32_t s2 = gy ? (int32_t)meta[2] : -1, s3 = gy ? (int32_t)meta[3] : -1;
    const int32_t s4 = gz ? (int32_t)meta[4] : -1, s5 = gz ? (int32_t)meta[5] : -1;
    const int32_t s6 = gy && gz ? (int32_t)meta[6] : -1, s7 = gy && gz ? (int32_t)meta[7] : -1;
    row_fetch(b, s0, s1, 4 * y + 16 * z, t.r);
    row_fetch(b, s2, s3, 16 * z, t.gy);
    row_fetch(b, s4, s5, 4 * y, t.gz);
    row_fetch(b, s6, s7, 0, t.gyz);
}

// Corner tests of a row: bit 4x iff r[x] <= iso.
BCMC_DEV uint32_t row_mask(const float (&r)[5], float iso) {
    uint32_t m = 0;
#pragma unroll
    for (int x = 0; x < 5; ++x)
        m |= (r[x] <= iso) ? (1u << (4 * x)) : 0u;
    return m;
}

// Corner word of the lane's 4 cells (compute_block_has_vertices.comp:28-37):
// cell x's corner byte is (w >> 4x) & 0xFF.  Cells outside the tile (and all
// cells of rows without a block) get byte 0, whose cell has no vertices: rows
// outside give w = 0, and x3_ok is false when the cell at x = 3 is outside.
// Must be called by all 64 lanes.
struct CornerWord {
    uint32_t w;
    bool x3_ok;
};

BCMC_DEV CornerWord row_corner_word(const RowTile &t, float iso) {
    const uint32_t l = lane_id() & 15u, y = l & 3u, z = l >> 2;
    const uint32_t m00 = row_mask(t.r, iso), my = row_mask(t.gy, iso), mz = row_mask(t.gz, iso),
                   myz = row_mask(t.gyz, iso);
    const uint32_t n1 = row_from_next<1>(m00), n4 = row_from_next<4>(m00), n5 = row_from_next<5>(m00);
    const uint32_t y4 = row_from_next<4>(my), z1 = row_from_next<1>(mz);
    const uint32_t m10 = y < 3 ? n1 : my;
    const uint32_t m01 = z < 3 ? n4 : mz;
    const uint32_t m11 = y < 3 ? (z < 3 ? n5 : z1) : (z < 3 ? y4 : myz);
    // cells need y + 1 < dy and z + 1 < dz; the cell at x = 3 needs dx = 5
    const bool row_ok = t.live && (y < 3 || (t.dims & 2u)) && (z < 3 || (t.dims & 4u));
    CornerWord cw;
    cw.w = row_ok ? (m00 | (m10 << 1) | (m01 << 2) | (m11 << 3)) : 0u;
    cw.x3_ok = (t.dims & 1u) != 0;
    return cw;
}

template <uint32_t X>
BCMC_DEV uint32_t cell_byte(const CornerWord &cw) {
    const uint32_t b = (cw.w >> (4 * X)) & 0xFFu;
    return (X == 3 && !cw.x3_ok) ? 0u : b;
}

// Vertex count of the row's block, valid in lane 15 of the row.
BCMC_DEV uint32_t row_block_count(const RowTile &t, float iso, const uint8_t *count) {
    const CornerWord cw = row_corner_word(t, iso);
    const uint32_t n = (uint32_t)count[cell_byte<0>(cw)] + count[cell_byte<1>(cw)] +
                       count[cell_byte<2>(cw)] + count[cell_byte<3>(cw)];
    return row_incl_scan(n);
}

// Per-wave LDS scratch of the emission of 4 consecutive blocks.
constexpr uint32_t kMaxGroupTris = kRowBlocks * 64 * 5;
struct alignas(16) EmitScratch {
    float tile[kRowBlocks][kTileFloats];
    uint16_t first[kRowBlocks * 64];  // first vertex of each cell, from the group start
    uint8_t corner[kRowBlocks * 64];  // corner byte of each cell
    alignas(16) uint8_t start[kMaxGroupTris];  // at a cell's first triangle: the cell (0 elsewhere)
    uint32_t bid[kRowBlocks], bpos[kRowBlocks][3];
    float iso[kRowBlocks];
    uint32_t meta[kRowBlocks][kMetaWords];  // meta words of the group's blocks
};

typedef float V4f __attribute__((ext_vector_type(4)));
typedef uint32_t V4u __attribute__((ext_vector_type(4)));
typedef uint32_t V2u __attribute__((ext_vector_type(2)));
// cache-policy operand of the raw buffer store builtins: non-temporal
constexpr int kStoreNontemporal = 2;

BCMC_DEV void row_store(float *dst, const float (&r)[5]) {
    *reinterpret_cast<float4 *>(dst) = make_float4(r[0], r[1], r[2], r[3]);
    dst[4] = r[4];
}

// Vertices of 4 blocks that are consecutive in the output (rows 0..3 of the
// wave; rows with live = false have none), written from out0 on
// (compute_block_vertices.comp:76-174).  Vertex order is the reference's:
// blocks in row order, cells in x-fastest order, within a cell its
// triangle-table row.  Triangle-parallel: every cell with vertices marks its
// first triangle in LDS with its index g = 64 row + cell (cells ascend with
// their triangles), so the owner of triangle t is the running maximum of the
// marks up to t.  Vertices at or past `cap` are not written.  FMT is the
// bcmc_vertex_format.  Must be called by all 64 lanes.
template <int FMT>
BCMC_DEV void emit_row_blocks(const RowBlock &b, const RowTile &t, const McTables &tab, EmitScratch &es,
                              unsigned long long out0, unsigned long long cap, void *verts) {
    const uint32_t lane = lane_id(), row = lane >> 4, l = lane & 15u, y = l & 3u, z = l >> 2;
    const CornerWord cw = row_corner_word(t, b.iso);
    const uint32_t cb[4] = {cell_byte<0>(cw), cell_byte<1>(cw), cell_byte<2>(cw), cell_byte<3>(cw)};
    const uint32_t nv[4] = {tab.count[cb[0]], tab.count[cb[1]], tab.count[cb[2]], tab.count[cb[3]]};
    // the tile, for the interpolation
    float *tile = es.tile[row];
    row_store(tile + kTileY * y + kTileZ * z, t.r);
    if (y == 3)
        row_store(tile + kTileY * 4 + kTileZ * z, t.gy);
    if (z == 3)
        row_store(tile + kTileY * y + kTileZ * 4, t.gz);
    if (y == 3 && z == 3)
        row_store(tile + kTileY * 4 + kTileZ * 4, t.gyz);
    // first vertex of each cell: group prefix of the block totals + row scan
    const uint32_t lane_total = nv[0] + nv[1] + nv[2] + nv[3];
    const uint32_t incl = row_incl_scan(lane_total);
    const uint32_t t0 = wave_read_lane(incl, 15), t1 = wave_read_lane(incl, 31),
                   t2 = wave_read_lane(incl, 47), t3 = wave_read_lane(incl, 63);
    const uint32_t total = t0 + t1 + t2 + t3;
    const uint32_t base = row == 0 ? 0u : row == 1 ? t0 : row == 2 ? t0 + t1 : t0 + t1 + t2;
    // vertices this call writes: those below the buffer capacity
    const uint32_t limit = out0 >= cap ? 0u : (uint32_t)min((unsigned long long)total, cap - out0);
    // clear the start marks of this group's triangles (vertex counts are
    // multiples of 3)
    for (uint32_t i = 16 * lane; i < total / 3; i += 1024)
        *reinterpret_cast<uint4 *>(&es.start[i]) = make_uint4(0, 0, 0, 0);
    if (l == 0) {
        uint32_t bx, by, bz;
        block_coords(b, bx, by, bz);
        es.bid[row] = b.bid;
        es.iso[row] = b.iso;
        es.bpos[row][0] = bx;
        es.bpos[row][1] = by;
        es.bpos[row][2] = bz;
    }
    wave_lds_sync();
    uint32_t first = base + incl - lane_total;
#pragma unroll
    for (uint32_t x = 0; x < 4; ++x) {
        const uint32_t g = 64 * row + 4 * l + x;
        es.first[g] = (uint16_t)first;
        es.corner[g] = (uint8_t)cb[x];
        if (nv[x])
            es.start[first / 3] = (uint8_t)g;
        first += nv[x];
    }
    wave_lds_sync();
    // Rounds of 64 triangles: lane t looks up the cell of triangle t0 + t and
    // its 3 edges; then 3 sub-rounds of 64 vertices each, where lane l handles
    // vertex 3 t0 + 64 r + l and fetches its triangle's lookup from lane
    // (64 r + l) / 3 (ds_bpermute), so the stores stay one contiguous run.
    const uint32_t ntri = (limit + 2) / 3;
    uint32_t carry = 0;
    // block-local format: stores through a buffer resource over the group's
    // first `limit` vertices, which drops those past it, so the three
    // sub-rounds of a round are straight-line code (gfx9 raw buffer
    // descriptor, word 3 = 0x00020000: 32-bit data format, no swizzle)
    // (out0 and limit are the same in every lane: the descriptor is built in
    // scalar registers; readfirstlane returns int, so each word goes through
    // uint32_t before it is widened -- a low word with bit 31 set would
    // otherwise sign-extend into the high word of the address)
    const unsigned long long out0_s =
        ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(out0 >> 32)) << 32) |
        (unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)out0);
    const uint32_t limit_s = (uint32_t)__builtin_amdgcn_readfirstlane(limit);
    [[maybe_unused]] const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
        reinterpret_cast<V4f *>(verts) + out0_s, 0, (int)(limit_s * 16u), 0x00020000);
    for (uint32_t t0 = 0; t0 < ntri; t0 += 64) {
        const uint32_t t = t0 + lane;
        const uint32_t g = max(carry, wave_incl_max(t < ntri ? (uint32_t)es.start[t] : 0u));
        carry = wave_read_lane(g, 63);
        uint32_t info = 0;  // edges (12 bits) | g << 12
        if (t < ntri) {
            const uint32_t k = 3 * t - es.first[g];  // first vertex of the triangle in its cell
            info = ((uint32_t)(tab.edges[es.corner[g]] >> (4 * k)) & 0xFFFu) | (g << 12);
        }
#pragma unroll
        for (uint32_t r = 0; r < 3; ++r) {
            const uint32_t lv = 64 * r + lane;          // vertex of this round
            const uint32_t src = (lv * 0xAAABu) >> 17;  // lv / 3
            const uint32_t inf = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)info);
            const uint32_t v = 3 * t0 + lv;
            // lanes past the limit compute from valid LDS entries (info 0 is
            // cell 0, edge 0 of row 0) and do not store
            const uint32_t gg = inf >> 12, j = gg >> 6, cell = gg & 63u;
            const uint32_t m = lv - 3 * src;
            // the edge's two table rows, the block's isovalue and id in one
            // batch of LDS reads, then the two corner values
            const EdgeInfo &ei = tab.edge[(inf >> (4 * m)) & 15u];
            const float4 ea = *reinterpret_cast<const float4 *>(ei.a);  // a[0..2], off_a
            const float4 ed = *reinterpret_cast<const float4 *>(ei.d);  // d[0..2], off_b
            const float iso = es.iso[j];
            const uint32_t bid = es.bid[j];
            const uint32_t cx = cell & 3u, cy = (cell >> 2) & 3u, cz = cell >> 4;
            const float *tc = es.tile[j] + cx + kTileY * cy + kTileZ * cz;
            const float fa = tc[__float_as_uint(ea.w)];
            const float fb = tc[__float_as_uint(ed.w)];
            // lerp_verts (compute_block_vertices.comp:62-74), then + voxel +
            // 0.5; the quotient is formed unconditionally (no branch) and
            // replaced by 0 for a flat edge
            const float q = (iso - fa) / (fb - fa);
            const float tt = fabsf(fa - fb) < 0.001f ? 0.0f : q;
            const float px = (ea.x + tt * ed.x + (float)cx) + 0.5f;
            const float py = (ea.y + tt * ed.y + (float)cy) + 0.5f;
            const float pz = (ea.z + tt * ed.z + (float)cz) + 0.5f;
            const unsigned long long o = out0 + v;
            // vertices are written once and not read back by the pass:
            // non-temporal stores keep them from displacing the block data
            // in the caches
            if constexpr (FMT == BCMC_VERTS_LOCAL_F32) {
                const V4u vv = {__float_as_uint(px), __float_as_uint(py), __float_as_uint(pz), bid};
                __builtin_amdgcn_raw_buffer_store_b128(vv, vrs, (int)(v * 16u), 0, kStoreNontemporal);
                continue;
            }
            if (v >= limit)
                continue;
            if constexpr (FMT == BCMC_VERTS_GLOBAL_F32) {
                float *gp = reinterpret_cast<float *>(verts) + 3 * o;
                __builtin_nontemporal_store((float)(4 * es.bpos[j][0]) + px, gp);
                __builtin_nontemporal_store((float)(4 * es.bpos[j][1]) + py, gp + 1);
                __builtin_nontemporal_store((float)(4 * es.bpos[j][2]) + pz, gp + 2);
            } else {
                // compress_position (block_vertex_position.comp:9-17)
                const uint32_t qx = (uint32_t)(((px - 0.5f) * 0.25f) * 1023.0f);
                const uint32_t qy = (uint32_t)(((py - 0.5f) * 0.25f) * 1023.0f);
                const uint32_t qz = (uint32_t)(((pz - 0.5f) * 0.25f) * 1023.0f);
                const uint32_t packed = ((qx & 0x3ffu) << 20) | ((qy & 0x3ffu) << 10) | (qz & 0x3ffu);
                const V2u vv = {packed, bid};
                __builtin_nontemporal_store(vv, reinterpret_cast<V2u *>(verts) + o);
            }
        }
    }
    wave_lds_sync();  // the scratch is reused by the next group
}

// The marching-cubes pass runs as three launches over chunks of kChunk
// consecutive active blocks: a count pass (per-block vertex counts and
// per-chunk totals), a scan of the chunk totals, and an emit pass that builds
// each chunk's list of blocks with vertices and writes their vertices.

// LDS of the count pass over one chunk: block ids and a caller word per
// entry, per-block vertex counts, the candidate entries, and per wave the
// meta words of the blocks its rows hold.  Entry kChunk is a scratch entry
// for rows without a block.
struct CountShared {
    uint32_t bid[kChunk + 1];
    uint32_t aux[kChunk + 1];
    uint32_t cnt[kChunk];
    uint32_t list[kChunk];
    uint32_t stage[kMcThreads / 64][kRowBlocks][kMetaWords];
};

// LDS of the emit pass over one chunk: block ids and caller words, and the
// list of entries with vertices with their vertex offsets in the chunk.
struct EmitShared {
    uint32_t bid[kChunk + 1];
    uint32_t aux[kChunk + 1];
    uint32_t off[kChunk];
    uint32_t list[kChunk];
};

// Chunk prologue of the count pass (all threads): entry t of the chunk gets
// its block id and caller word; it is a count candidate when the isovalue
// lies in [lo, hi) of its tile range (launch_tile_ranges): with iso < lo
// every corner test fails and with iso >= hi every one passes, so the block's
// cells have case 0 or 255 and no vertices.  Returns the number of candidates
// (listed in cs.list, ascending).
BCMC_DEV uint32_t chunk_prologue(CountShared &cs, bool in, uint32_t bid, uint32_t aux, float2 tile, float iso,
                                 uint32_t *lds_waves) {
    const uint32_t t = threadIdx.x;
    cs.bid[t] = bid;
    cs.aux[t] = aux;
    cs.cnt[t] = 0;
    if (t == 0) {
        cs.bid[kChunk] = 0;
        cs.aux[kChunk] = 0;
    }
    const bool cand = in && iso >= tile.x && iso < tile.y;
    uint32_t ncand;
    const uint32_t pos = block_excl_scan<kMcThreads>(cand ? 1u : 0u, lds_waves, &ncand);
    if (cand)
        cs.list[pos] = t;
    __syncthreads();
    return ncand;
}

// Count sweep over the candidate entries: wave w takes candidates
// 4w + 16q + row (q = 0, 1, ...); the next round's slot/flag loads are issued
// before this round's voxels are used.  setup(e, b) fills the RowBlock of
// entry e (e = kChunk: no block).
template <class Setup>
BCMC_DEV void count_sweep(CountShared &cs, uint32_t ncand, uint32_t wave, const CountTables &tab,
                          Setup &&setup) {
    constexpr uint32_t kRound = (kMcThreads / 64) * kRowBlocks;
    const uint32_t lane = lane_id(), row = lane >> 4;
    const uint32_t first = wave * kRowBlocks;
    if (first >= ncand)
        return;
    uint32_t *stage = cs.stage[wave][row];
    auto setup_c = [&](uint32_t c, uint32_t &e, RowBlock &b) {
        e = c < ncand ? cs.list[c] : kChunk;
        setup(e, b);
        b.live = b.live && e < kChunk;
    };
    RowBlock b;
    uint32_t e;
    setup_c(first + row, e, b);
    RowMeta m = row_meta_load(b);
    for (uint32_t base = first; base < ncand; base += kRound) {
        const RowBlock bq = b;
        const uint32_t eq = e;
        wave_lds_sync();  // the previous round's reads of the stage are done
        row_meta_store(m, bq.tag, tab.dims, stage);
        wave_lds_sync();
        RowTile t;
        row_voxels(bq, stage, t);
        if (base + kRound < ncand) {
            setup_c(base + kRound + row, e, b);
            m = row_meta_load(b);
        }
        const uint32_t c = row_block_count(t, bq.iso, tab.count);
        if ((lane & 15u) == 15u && eq < kChunk)
            cs.cnt[eq] = c;
    }
}

// Emit sweep: wave w takes groups of 4 consecutive entries of the list of
// blocks with vertices (whose outputs are contiguous) and writes their
// vertices from out0 + the group's offset in the chunk.  The next group's
// slot/flag loads are issued before this group's vertices are written.
template <int FMT, class Setup>
BCMC_DEV void emit_sweep(EmitShared &cs, uint32_t wave, uint32_t num_with, unsigned long long out0,
                         const McTables &tab, EmitScratch &es, unsigned long long cap, void *verts,
                         Setup &&setup) {
    constexpr uint32_t kRound = (kMcThreads / 64) * kRowBlocks;
    const uint32_t row = lane_id() >> 4;
    uint32_t hh = wave * kRowBlocks;
    if (hh >= num_with)
        return;
    auto block_of = [&](uint32_t h, RowBlock &b) {
        const bool live = h < num_with;
        setup(live ? cs.list[h] : 0u, b);
        b.live = live;
    };
    RowBlock b;
    block_of(hh + row, b);
    RowMeta m = row_meta_load(b);
    for (; hh < num_with; hh += kRound) {
        const RowBlock bq = b;
        wave_lds_sync();  // the previous group's reads of the meta words are done
        row_meta_store(m, bq.tag, tab.dims, es.meta[row]);
        wave_lds_sync();
        RowTile t;
        row_voxels(bq, es.meta[row], t);
        if (hh + kRound < num_with) {
            block_of(hh + kRound + row, b);
            m = row_meta_load(b);
        }
        emit_row_blocks<FMT>(bq, t, tab, es, out0 + cs.off[hh], cap, verts);
    }
}

// Emit-pass chunk prologue (all threads): the chunk's list of entries with
// vertices (cnt[t] > 0) with their vertex offsets in the chunk.  Returns the
// number of entries with vertices; *ex_h / *ex_v receive this thread's
// exclusive positions and *agg_v the chunk's vertex total.
BCMC_DEV uint32_t emit_prologue(EmitShared &cs, uint32_t c, uint32_t bid, uint32_t aux, uint32_t *lds_waves,
                                uint32_t *ex_h, uint32_t *ex_v, uint32_t *agg_v) {
    const uint32_t t = threadIdx.x;
    cs.bid[t] = bid;
    cs.aux[t] = aux;
    if (t == 0) {
        cs.bid[kChunk] = 0;
        cs.aux[kChunk] = 0;
    }
    uint32_t agg_h;
    *ex_h = block_excl_scan<kMcThreads>(c ? 1u : 0u, lds_waves, &agg_h);
    *ex_v = block_excl_scan<kMcThreads>(c, lds_waves, agg_v);
    if (c) {
        cs.list[*ex_h] = t;
        cs.off[*ex_h] = *ex_v;
    }
    __syncthreads();
    return agg_h;
}

// ---------------------------------------------------------------------------
// Active test (compute_block_active.comp:30-60).  The reference marks block b
// active when iso lies in [lo_b, hi_b], or in [min(lo_b, lo_n), max(hi_b, hi_n)]
// for some in-bounds neighbour n of its 3x3x3 neighbourhood.  Each of these
// intervals contains [lo_b, hi_b], so their union is the single interval
// [min lo, max hi] over the neighbourhood (b included), and the test is one
// pair of comparisons against that dilated range, computed once per volume
// (launch_dilate_ranges).
// min and max select one of their operands, so the comparisons see the same
// float values as the reference's.
// ---------------------------------------------------------------------------
BCMC_DEV bool block_active(float2 dilated, float iso) { return iso >= dilated.x && iso <= dilated.y; }

}  // namespace bcmc
