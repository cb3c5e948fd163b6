// Marching-cubes device code shared by the per-query pipeline (mc.hip) and
// the batched/served engine (engine.hip).
//
// Work layout: a wave handles 4 blocks at a time, one per 16-lane DPP row.
// Lane l = y + 4z of a row owns the 4 cells (x = 0..3, y, z) of its block and
// holds the voxel row (x = 0..4, y, z) in registers; the rows at y + 1 and
// z + 1 come from lanes l + 1, l + 4 and l + 5 through DPP row shifts, and the
// lanes on the block's upper faces also load the ghost rows from the +y/+z
// neighbours.  Corner tests become 5-bit row masks, so a cell's case is a
// fixed gather of 8 bits from 4 masks.
#pragma once

#include "common.h"
#include "tri_table.h"

namespace bcmc {

static __constant__ unsigned long long kTriEdges[256] = BCMC_TRI_EDGES_INIT;
static __constant__ uint8_t kTriCount[256] = BCMC_TRI_COUNT_INIT;
// edge_vertices (compute_block_vertices.comp:10-23) as two nibbles per edge
constexpr uint8_t kEdgeVertsHost[12] = {0x10, 0x21, 0x32, 0x03, 0x54, 0x56,
                                        0x76, 0x47, 0x40, 0x51, 0x62, 0x73};

// index_to_vertex (compute_vertex_values.comp:48-57): corner i at
// (x, y, z) = (((i+1)>>1)&1, (i>>1)&1, i>>2)
__host__ __device__ constexpr uint32_t corner_x(uint32_t i) { return ((i + 1) >> 1) & 1u; }
__host__ __device__ constexpr uint32_t corner_y(uint32_t i) { return (i >> 1) & 1u; }
__host__ __device__ constexpr uint32_t corner_z(uint32_t i) { return i >> 2; }

constexpr int kMcThreads = 256;   // 4 waves per workgroup
constexpr int kRowBlocks = 4;     // blocks per wave (one per 16-lane row)
// Block tile in LDS: 5^3 voxels at x + 8y + 40z (rows 32-byte aligned).
constexpr uint32_t kTileY = 8, kTileZ = 40, kTileFloats = 200;

// Triangle tables staged in LDS once per workgroup (the per-cell lookups are
// divergent, so they would otherwise be vector loads through the L1).
struct McTables {
    unsigned long long edges[256];
    uint8_t count[256];
};

BCMC_DEV void load_mc_tables(McTables *t) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        t->edges[i] = kTriEdges[i];
        t->count[i] = kTriCount[i];
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

// The block a row works on, as seen by each of its lanes.  In the per-query
// pipeline every field but bid/live is uniform; in the batched engine rows
// may belong to different volumes and isovalues.
struct RowBlock {
    const float *cache;        // decoded blocks, 64 floats per slot
    const int32_t *item_slot;  // block -> cache slot (< 0: not cached)
    const uint32_t *flags;     // active flags of the query
    uint32_t nb[3];
    uint32_t cap;              // cache slots (reads outside are skipped)
    uint32_t bid;
    float iso;
    bool live;
};

// Voxels held by a lane: its row and, on the upper faces, the ghost rows
// (y+1 | z+1 | y+1,z+1) from the neighbouring blocks; tile dims (4 or 5 each).
struct RowTile {
    float r[5], gy[5], gz[5], gyz[5];
    uint32_t dx, dy, dz;
    uint32_t bx, by, bz;
    bool live;
};

BCMC_DEV int32_t row_slot(const RowBlock &b, uint32_t bx, uint32_t by, uint32_t bz, uint32_t s) {
    const uint32_t x = bx + (s & 1u), y = by + ((s >> 1) & 1u), z = bz + (s >> 2);
    if (!b.live || x >= b.nb[0] || y >= b.nb[1] || z >= b.nb[2])
        return -1;
    return b.item_slot[x + b.nb[0] * (y + b.nb[1] * z)];
}

BCMC_DEV void row_fetch(const RowBlock &b, int32_t s, int32_t sx, uint32_t off, float (&r)[5]) {
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float w = 0.0f;
    if (s >= 0 && (uint32_t)s < b.cap)
        v = *reinterpret_cast<const float4 *>(b.cache + (uint64_t)(uint32_t)s * 64 + off);
    if (sx >= 0 && (uint32_t)sx < b.cap)
        w = b.cache[(uint64_t)(uint32_t)sx * 64 + off];
    r[0] = v.x;
    r[1] = v.y;
    r[2] = v.z;
    r[3] = v.w;
    r[4] = w;
}

// compute_block_dims_with_ghost + load_block (compute_vertex_values.comp:96-252):
// the ghost planes are part of the tile only where the +x/+y/+z neighbours
// (and, for the shared edges and corner, the diagonal neighbours) are active.
// Must be called by all 64 lanes.
BCMC_DEV void row_load(const RowBlock &b, RowTile &t) {
    const uint32_t lane = lane_id(), l = lane & 15u, y = l & 3u, z = l >> 2;
    t.live = b.live;
    t.bx = b.bid % b.nb[0];
    t.by = (b.bid / b.nb[0]) % b.nb[1];
    t.bz = b.bid / (b.nb[0] * b.nb[1]);
    // neighbour activity, in the reference's check order: lanes 0..6 of the
    // row test +x, +y, +z, +xy, +xz, +yz, +xyz
    bool act = false;
    if (b.live && l < 7) {
        const uint32_t code = l < 3 ? (1u << l) : (l == 3 ? 3u : l == 4 ? 5u : l == 5 ? 6u : 7u);
        const uint32_t nx = t.bx + (code & 1u), ny = t.by + ((code >> 1) & 1u), nz = t.bz + (code >> 2);
        if (nx < b.nb[0] && ny < b.nb[1] && nz < b.nb[2])
            act = b.flags[nx + b.nb[0] * (ny + b.nb[1] * nz)] != 0;
    }
    // cache slots of the sub-blocks this lane reads: s = sx + 2sy + 4sz
    const int32_t s0 = row_slot(b, t.bx, t.by, t.bz, 0), s1 = row_slot(b, t.bx, t.by, t.bz, 1);
    int32_t s2 = -1, s3 = -1, s4 = -1, s5 = -1, s6 = -1, s7 = -1;
    if (y == 3) {
        s2 = row_slot(b, t.bx, t.by, t.bz, 2);
        s3 = row_slot(b, t.bx, t.by, t.bz, 3);
    }
    if (z == 3) {
        s4 = row_slot(b, t.bx, t.by, t.bz, 4);
        s5 = row_slot(b, t.bx, t.by, t.bz, 5);
    }
    if (y == 3 && z == 3) {
        s6 = row_slot(b, t.bx, t.by, t.bz, 6);
        s7 = row_slot(b, t.bx, t.by, t.bz, 7);
    }
    const uint64_t ab = __ballot(act);
    const uint32_t a = (uint32_t)(ab >> (lane & 48u)) & 0x7Fu;
    uint32_t dx = t.bx + 1 < b.nb[0] ? 5 : 4, dy = t.by + 1 < b.nb[1] ? 5 : 4,
             dz = t.bz + 1 < b.nb[2] ? 5 : 4;
    if (dx == 5 && dy == 5 && dz == 5 && !(a & 64))
        dx = dy = dz = 4;
    if (dx == 5 && dy == 5 && !(a & 8))
        dx = dy = 4;
    if (dx == 5 && dz == 5 && !(a & 16))
        dx = dz = 4;
    if (dy == 5 && dz == 5 && !(a & 32))
        dy = dz = 4;
    if (dx == 5 && !(a & 1))
        dx = 4;
    if (dy == 5 && !(a & 2))
        dy = 4;
    if (dz == 5 && !(a & 4))
        dz = 4;
    t.dx = dx;
    t.dy = dy;
    t.dz = dz;
    // voxel rows: own (x = 0..3 of this block, x = 4 from +x), then the ghosts
    row_fetch(b, s0, s1, 4 * y + 16 * z, t.r);
    row_fetch(b, s2, s3, 16 * z, t.gy);
    row_fetch(b, s4, s5, 4 * y, t.gz);
    row_fetch(b, s6, s7, 0, t.gyz);
}

BCMC_DEV uint32_t row_mask(const float (&r)[5], float iso) {
    uint32_t m = 0;
#pragma unroll
    for (int x = 0; x < 5; ++x)
        m |= (r[x] <= iso) ? (1u << x) : 0u;
    return m;
}

// Cases of the lane's 4 cells (bit i iff corner i <= iso,
// compute_block_has_vertices.comp:28-37), packed 8 bits per cell, and the
// per-cell vertex counts (0 for cells outside the tile and for rows without
// a block), packed 8 bits per cell.  Must be called by all 64 lanes.
BCMC_DEV void row_cases(const RowTile &t, float iso, const McTables &tab, uint32_t &cases,
                        uint32_t &counts) {
    const uint32_t l = lane_id() & 15u, y = l & 3u, z = l >> 2;
    const uint32_t m00 = row_mask(t.r, iso), my = row_mask(t.gy, iso), mz = row_mask(t.gz, iso),
                   myz = row_mask(t.gyz, iso);
    const uint32_t n1 = row_from_next<1>(m00), n4 = row_from_next<4>(m00), n5 = row_from_next<5>(m00);
    const uint32_t y4 = row_from_next<4>(my), z1 = row_from_next<1>(mz);
    const uint32_t m10 = y < 3 ? n1 : my;
    const uint32_t m01 = z < 3 ? n4 : mz;
    const uint32_t m11 = y < 3 ? (z < 3 ? n5 : z1) : (z < 3 ? y4 : myz);
    const uint32_t w = m00 | (m10 << 5) | (m01 << 10) | (m11 << 15);
    const bool row_ok = t.live && y + 1 < t.dy && z + 1 < t.dz;
    cases = 0;
    counts = 0;
#pragma unroll
    for (uint32_t x = 0; x < 4; ++x) {
        // corners 0..7 sit at bits x + {0, 1, 6, 5, 10, 11, 16, 15} of w
        const uint32_t s = w >> x;
        const uint32_t c = (s & 3u) | ((s >> 4) & 4u) | ((s >> 2) & 8u) | ((s >> 6) & 0x30u) |
                           ((s >> 10) & 0x40u) | ((s >> 8) & 0x80u);
        const uint32_t nv = (row_ok && x + 1 < t.dx) ? tab.count[c] : 0u;
        cases |= c << (8 * x);
        counts |= nv << (8 * x);
    }
}

BCMC_DEV uint32_t sum_bytes(uint32_t v) { return (v & 0xFF) + ((v >> 8) & 0xFF) + ((v >> 16) & 0xFF) + (v >> 24); }

// Vertex count of the row's block, valid in lane 15 of the row.
BCMC_DEV uint32_t row_block_count(const RowTile &t, float iso, const McTables &tab) {
    uint32_t cases, counts;
    row_cases(t, iso, tab, cases, counts);
    return row_incl_scan(sum_bytes(counts));
}

// Per-wave LDS scratch of the emission of 4 consecutive blocks.
struct EmitScratch {
    float tile[kRowBlocks][kTileFloats];
    uint16_t first[kRowBlocks][64];  // first vertex of each cell, from the group start
    uint8_t cell_case[kRowBlocks][64];
    uint32_t bid[kRowBlocks], bpos[kRowBlocks][3];
    float iso[kRowBlocks];
};

BCMC_DEV void row_store(float *dst, const float (&r)[5]) {
    *reinterpret_cast<float4 *>(dst) = make_float4(r[0], r[1], r[2], r[3]);
    dst[4] = r[4];
}

// Corners a and b of edge e as 4-bit fields, and the tile offset of corner i.
constexpr unsigned long long edge_corner_bits(int which) {
    unsigned long long r = 0;
    for (int e = 0; e < 12; ++e)
        r |= (unsigned long long)((kEdgeVertsHost[e] >> (4 * which)) & 15u) << (4 * e);
    return r;
}
constexpr unsigned long long kEdgeCornerA = edge_corner_bits(0), kEdgeCornerB = edge_corner_bits(1);
BCMC_DEV uint32_t corner_offset(uint32_t i) { return corner_x(i) + kTileY * corner_y(i) + kTileZ * corner_z(i); }

// Vertices of 4 blocks that are consecutive in the output (rows 0..3 of the
// wave; rows with live = false have none), written from out0 on
// (compute_block_vertices.comp:76-174).  Vertex order is the reference's:
// blocks in row order, cells in x-fastest order, within a cell its
// triangle-table row.  Vertex-parallel: lane v of each round finds its cell by
// a binary search of the cells' first-vertex offsets and interpolates, so a
// round writes one contiguous run.  Vertices at or past `cap` are not written.
// Must be called by all 64 lanes.
BCMC_DEV void emit_row_blocks(const RowBlock &b, const RowTile &t, const McTables &tab, EmitScratch &es,
                              unsigned long long out0, unsigned long long cap, void *verts, int fmt) {
    const uint32_t lane = lane_id(), row = lane >> 4, l = lane & 15u, y = l & 3u, z = l >> 2;
    uint32_t cases, counts;
    row_cases(t, b.iso, tab, cases, counts);
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
    const uint32_t lane_total = sum_bytes(counts);
    const uint32_t incl = row_incl_scan(lane_total);
    const uint32_t t0 = wave_read_lane(incl, 15), t1 = wave_read_lane(incl, 31),
                   t2 = wave_read_lane(incl, 47), t3 = wave_read_lane(incl, 63);
    const uint32_t s1 = t0, s2 = t0 + t1, s3 = t0 + t1 + t2, total = s3 + t3;
    uint32_t first = (row == 0 ? 0u : row == 1 ? s1 : row == 2 ? s2 : s3) + incl - lane_total;
#pragma unroll
    for (uint32_t x = 0; x < 4; ++x) {
        es.first[row][4 * l + x] = (uint16_t)first;
        es.cell_case[row][4 * l + x] = (uint8_t)(cases >> (8 * x));
        first += (counts >> (8 * x)) & 0xFFu;
    }
    if (l == 0) {
        es.bid[row] = b.bid;
        es.iso[row] = b.iso;
        es.bpos[row][0] = t.bx;
        es.bpos[row][1] = t.by;
        es.bpos[row][2] = t.bz;
    }
    wave_lds_sync();
    for (uint32_t v = lane; v < total; v += 64) {
        const uint32_t j = (v >= s1 ? 1u : 0u) + (v >= s2 ? 1u : 0u) + (v >= s3 ? 1u : 0u);
        // last cell whose first vertex is <= v (cells after the owner start past v)
        const uint16_t *fj = es.first[j];
        uint32_t g = 0;
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1)
            g = (fj[g + step] <= v) ? g + step : g;
        const uint32_t k = v - fj[g];
        const uint32_t e = (uint32_t)(tab.edges[es.cell_case[j][g]] >> (4 * k)) & 15u;
        const uint32_t ca = (uint32_t)(kEdgeCornerA >> (4 * e)) & 15u,
                       cb = (uint32_t)(kEdgeCornerB >> (4 * e)) & 15u;
        const uint32_t cx = g & 3u, cy = (g >> 2) & 3u, cz = g >> 4;
        const float *tj = es.tile[j] + cx + kTileY * cy + kTileZ * cz;
        const float fa = tj[corner_offset(ca)];
        const float fb = tj[corner_offset(cb)];
        const float iso = es.iso[j];
        // lerp_verts (compute_block_vertices.comp:62-74), then + voxel + 0.5
        float tt = 0.0f;
        if (!(fabsf(fa - fb) < 0.001f))
            tt = (iso - fa) / (fb - fa);
        const float ax = (float)corner_x(ca), ay = (float)corner_y(ca), az = (float)corner_z(ca);
        const float px = (ax + tt * ((float)corner_x(cb) - ax) + (float)cx) + 0.5f;
        const float py = (ay + tt * ((float)corner_y(cb) - ay) + (float)cy) + 0.5f;
        const float pz = (az + tt * ((float)corner_z(cb) - az) + (float)cz) + 0.5f;
        const unsigned long long o = out0 + v;
        if (o >= cap)
            break;
        const uint32_t bid = es.bid[j];
        if (fmt == BCMC_VERTS_LOCAL_F32) {
            reinterpret_cast<float4 *>(verts)[o] = make_float4(px, py, pz, __uint_as_float(bid));
        } else if (fmt == BCMC_VERTS_GLOBAL_F32) {
            float *gp = reinterpret_cast<float *>(verts) + 3 * o;
            gp[0] = (float)(4 * es.bpos[j][0]) + px;
            gp[1] = (float)(4 * es.bpos[j][1]) + py;
            gp[2] = (float)(4 * es.bpos[j][2]) + pz;
        } else {
            // compress_position (block_vertex_position.comp:9-17)
            const uint32_t qx = (uint32_t)(((px - 0.5f) * 0.25f) * 1023.0f);
            const uint32_t qy = (uint32_t)(((py - 0.5f) * 0.25f) * 1023.0f);
            const uint32_t qz = (uint32_t)(((pz - 0.5f) * 0.25f) * 1023.0f);
            const uint32_t packed = ((qx & 0x3ffu) << 20) | ((qy & 0x3ffu) << 10) | (qz & 0x3ffu);
            reinterpret_cast<uint2 *>(verts)[o] = make_uint2(packed, bid);
        }
    }
    wave_lds_sync();  // the scratch is reused by the next group
}

// ---------------------------------------------------------------------------
// Active test (compute_block_active.comp:127-164) restated as a dilation.
// The reference marks a block active when iso lies in its own range, or in
// the union of its range with any in-bounds neighbour's.  With above = iso < lo
// and below = iso > hi, the union with neighbour n misses iso exactly when
// (above && above_n) || (below && below_n), so
//   active = contains || (above && OR_27(!above_n)) || (below && OR_27(!below_n))
// where OR_27 runs over the in-bounds 3x3x3 neighbourhood (the block's own
// bit never contributes: !above is false when above).  The OR is separable:
// the code byte holds the x-direction OR; the y/z part reads 9 code bytes.
// ---------------------------------------------------------------------------
constexpr uint32_t kActContains = 1, kActAbove = 2, kActBelow = 4, kActXNotAbove = 8,
                   kActXNotBelow = 16;

BCMC_DEV uint32_t active_code(const float2 *ranges, uint32_t b, uint32_t x, uint32_t nbx, float iso) {
    const float2 r = ranges[b];
    const bool above = iso < r.x, below = iso > r.y;
    bool xna = !above, xnb = !below;
    if (x > 0) {
        const float2 l = ranges[b - 1];
        xna |= !(iso < l.x);
        xnb |= !(iso > l.y);
    }
    if (x + 1 < nbx) {
        const float2 h = ranges[b + 1];
        xna |= !(iso < h.x);
        xnb |= !(iso > h.y);
    }
    const bool contains = iso >= r.x && iso <= r.y;
    return (contains ? kActContains : 0u) | (above ? kActAbove : 0u) | (below ? kActBelow : 0u) |
           (xna ? kActXNotAbove : 0u) | (xnb ? kActXNotBelow : 0u);
}

BCMC_DEV uint32_t active_from_codes(const uint8_t *code, uint32_t b, const uint32_t (&nb)[3]) {
    const uint32_t c = code[b];
    if (c & kActContains)
        return 1u;
    const uint32_t need = (c & kActAbove) ? kActXNotAbove : (c & kActBelow) ? kActXNotBelow : 0u;
    if (!need)
        return 0u;
    const uint32_t plane = nb[0] * nb[1];
    const uint32_t y = (b / nb[0]) % nb[1], z = b / plane;
    uint32_t acc = 0;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = (int)y + dy, zz = (int)z + dz;
            if (yy < 0 || zz < 0 || yy >= (int)nb[1] || zz >= (int)nb[2])
                continue;
            acc |= code[(int64_t)b + (int64_t)dy * nb[0] + (int64_t)dz * plane];
        }
    }
    return (acc & need) ? 1u : 0u;
}

}  // namespace bcmc
