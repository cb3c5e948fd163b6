// Marching-cubes device code shared by the per-query pipeline (mc.hip) and
// the batched/served engine (engine.hip).
#pragma once

#include "common.h"
#include "tri_table.h"

namespace bcmc {

static __constant__ unsigned long long kTriEdges[256] = BCMC_TRI_EDGES_INIT;
static __constant__ uint8_t kTriCount[256] = BCMC_TRI_COUNT_INIT;
// edge_vertices (compute_block_vertices.comp:10-23) as two nibbles per edge
static __constant__ uint8_t kEdgeVerts[12] = {0x10, 0x21, 0x32, 0x03, 0x54, 0x56,
                                       0x76, 0x47, 0x40, 0x51, 0x62, 0x73};

// index_to_vertex (compute_vertex_values.comp:48-57): corner i at
// (x, y, z) = (((i+1)>>1)&1, (i>>1)&1, i>>2)
BCMC_DEV uint32_t corner_x(uint32_t i) { return ((i + 1) >> 1) & 1u; }
BCMC_DEV uint32_t corner_y(uint32_t i) { return (i >> 1) & 1u; }
BCMC_DEV uint32_t corner_z(uint32_t i) { return i >> 2; }

struct McView {
    const float *cache;
    const int32_t *item_slot;
    const uint32_t *flags;
    uint32_t nb[3];
    uint32_t capacity;  // cache slots (reads outside are skipped)
};

constexpr int kMcThreads = 256;  // 4 waves, one block each
constexpr int kTileStride = 128;

// compute_block_dims_with_ghost + load_block (compute_vertex_values.comp:96-252).
// Tile layout x + 5y + 25z.  Returns the dims through d[3] (wave uniform).
BCMC_DEV void load_tile(const McView &m, uint32_t bid, float *tile, int32_t *slots, uint32_t (&d)[3]) {
    const uint32_t lane = lane_id();
    const uint32_t bx = bid % m.nb[0], by = (bid / m.nb[0]) % m.nb[1], bz = bid / (m.nb[0] * m.nb[1]);
    // neighbour activity: +x, +y, +z, +xy, +xz, +yz, +xyz
    bool act = false;
    if (lane < 7) {
        const uint32_t code = lane < 3 ? (1u << lane) : (lane == 3 ? 3u : lane == 4 ? 5u : lane == 5 ? 6u : 7u);
        const uint32_t nx = bx + (code & 1u), ny = by + ((code >> 1) & 1u), nz = bz + (code >> 2);
        if (nx < m.nb[0] && ny < m.nb[1] && nz < m.nb[2])
            act = m.flags[nx + m.nb[0] * (ny + m.nb[1] * nz)] != 0;
    }
    const uint64_t a = __ballot(act);
    uint32_t dx = bx + 1 < m.nb[0] ? 5 : 4, dy = by + 1 < m.nb[1] ? 5 : 4, dz = bz + 1 < m.nb[2] ? 5 : 4;
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
    d[0] = dx;
    d[1] = dy;
    d[2] = dz;
    // cache slots of the (up to 8) blocks the tile touches
    if (lane < 8) {
        const uint32_t sx = lane & 1u, sy = (lane >> 1) & 1u, sz = lane >> 2;
        const bool need = (!sx || dx == 5) && (!sy || dy == 5) && (!sz || dz == 5);
        slots[lane] = need ? m.item_slot[(bx + sx) + m.nb[0] * ((by + sy) + m.nb[1] * (bz + sz))] : -1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t t = lane; t < 125; t += 64) {
        const uint32_t x = t % 5, y = (t / 5) % 5, z = t / 25;
        float v = 0.0f;
        if (x < dx && y < dy && z < dz) {
            const int32_t s = slots[(x >> 2) + 2 * (y >> 2) + 4 * (z >> 2)];
            if (s >= 0 && (uint32_t)s < m.capacity)
                v = m.cache[(uint64_t)(uint32_t)s * 64 + (x & 3) + 4 * (y & 3) + 16 * (z & 3)];
        }
        tile[t] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cell of this lane: case index (bit i iff corner i <= iso) or -1 if the cell
// is outside the tile (compute_block_has_vertices.comp:28-37).
BCMC_DEV int cell_case(const float *tile, const uint32_t (&d)[3], float iso, uint32_t &base) {
    const uint32_t lane = lane_id();
    const uint32_t x = lane & 3u, y = (lane >> 2) & 3u, z = lane >> 4;
    base = x + 5 * y + 25 * z;
    if (!(x + 1 < d[0] && y + 1 < d[1] && z + 1 < d[2]))
        return -1;
    int c = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        const float v = tile[base + corner_x(i) + 5 * corner_y(i) + 25 * corner_z(i)];
        c |= (v <= iso) ? (1 << i) : 0;
    }
    return c;
}


// Vertices of one block (compute_block_vertices.comp:76-174): a wave prefix
// sum of the per-cell counts, then each lane interpolates and writes the
// vertices of its cell starting at out0 + its prefix.
BCMC_DEV void emit_block_vertices(const float *tile, const uint32_t (&d)[3], float iso, uint32_t bid,
                                  const uint32_t (&nb)[3], unsigned long long base_offset,
                                  void *verts, int fmt) {
    uint32_t base;
    const int c = cell_case(tile, d, iso, base);
    const uint32_t nv = c >= 0 ? kTriCount[c] : 0u;
    const uint32_t excl = wave_incl_scan(nv) - nv;
    const unsigned long long out0 = base_offset + excl;
    const uint32_t lane = lane_id();
    const float vx = (float)(lane & 3u), vy = (float)((lane >> 2) & 3u), vz = (float)(lane >> 4);
    const unsigned long long edges = c >= 0 ? kTriEdges[c] : 0ull;
    const uint32_t bx = bid % nb[0], by = (bid / nb[0]) % nb[1],
                   bz = bid / (nb[0] * nb[1]);
    for (uint32_t j = 0; j < nv; ++j) {
        const uint32_t e = (uint32_t)(edges >> (4 * j)) & 15u;
        const uint32_t ev = kEdgeVerts[e], ca = ev & 15u, cb = ev >> 4;
        const float fa = tile[base + corner_x(ca) + 5 * corner_y(ca) + 25 * corner_z(ca)];
        const float fb = tile[base + corner_x(cb) + 5 * corner_y(cb) + 25 * corner_z(cb)];
        // lerp_verts (compute_block_vertices.comp:62-74), then + voxel + 0.5
        float t = 0.0f;
        if (!(fabsf(fa - fb) < 0.001f))
            t = (iso - fa) / (fb - fa);
        const float ax = (float)corner_x(ca), ay = (float)corner_y(ca), az = (float)corner_z(ca);
        const float px = (ax + t * ((float)corner_x(cb) - ax) + vx) + 0.5f;
        const float py = (ay + t * ((float)corner_y(cb) - ay) + vy) + 0.5f;
        const float pz = (az + t * ((float)corner_z(cb) - az) + vz) + 0.5f;
        const unsigned long long o = out0 + j;
        if (fmt == BCMC_VERTS_LOCAL_F32) {
            reinterpret_cast<float4 *>(verts)[o] = make_float4(px, py, pz, __uint_as_float(bid));
        } else if (fmt == BCMC_VERTS_GLOBAL_F32) {
            float *g = reinterpret_cast<float *>(verts) + 3 * o;
            g[0] = (float)(4 * bx) + px;
            g[1] = (float)(4 * by) + py;
            g[2] = (float)(4 * bz) + pz;
        } else {
            // compress_position (block_vertex_position.comp:9-17)
            const uint32_t qx = (uint32_t)(((px - 0.5f) * 0.25f) * 1023.0f);
            const uint32_t qy = (uint32_t)(((py - 0.5f) * 0.25f) * 1023.0f);
            const uint32_t qz = (uint32_t)(((pz - 0.5f) * 0.25f) * 1023.0f);
            const uint32_t packed = ((qx & 0x3ffu) << 20) | ((qy & 0x3ffu) << 10) | (qz & 0x3ffu);
            reinterpret_cast<uint2 *>(verts)[o] = make_uint2(packed, bid);
        }
    }
}

BCMC_DEV uint32_t block_vertex_count(const float *tile, const uint32_t (&d)[3], float iso) {
    uint32_t base;
    const int c = cell_case(tile, d, iso, base);
    return wave_sum(c >= 0 ? kTriCount[c] : 0u);
}

}  // namespace bcmc
