// Marching-cubes device code shared by the per-query pipeline (mc.hip) and
// the batched/served engine (engine.hip).
#pragma once

#include "common.h"
#include "tri_table.h"

#include <type_traits>

namespace bcmc {

// compile-time unrolled loop: f(std::integral_constant<int, j>) for j < N
template <int N, class F, int J = 0>
BCMC_DEV void unroll_for(F &&f) {
    if constexpr (J < N) {
        f(std::integral_constant<int, J>{});
        unroll_for<N, F, J + 1>(static_cast<F &&>(f));
    }
}

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

struct McView {
    const float *cache;
    const int32_t *item_slot;
    const uint32_t *flags;
    uint32_t nb[3];
    uint32_t capacity;  // cache slots (reads outside are skipped)
};

constexpr int kMcThreads = 256;  // 4 waves, one block each
constexpr int kTileStride = 128;

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

// Per-wave LDS scratch of the vertex-parallel emission: the case and first
// vertex of each of the 64 cells, and the owning cell of each vertex.
struct EmitScratch {
    uint8_t cell_case[64];
    uint16_t cell_first[64];
    uint8_t vert_cell[64 * 15];
};

// Corners a and b of edge e as 4-bit fields (edge_vertices,
// compute_block_vertices.comp:10-23), and the tile offset of corner i.
constexpr unsigned long long edge_corner_bits(int which) {
    unsigned long long r = 0;
    for (int e = 0; e < 12; ++e)
        r |= (unsigned long long)((kEdgeVertsHost[e] >> (4 * which)) & 15u) << (4 * e);
    return r;
}
constexpr unsigned long long kEdgeCornerA = edge_corner_bits(0), kEdgeCornerB = edge_corner_bits(1);
BCMC_DEV uint32_t corner_offset(uint32_t i) { return corner_x(i) + 5 * corner_y(i) + 25 * corner_z(i); }

BCMC_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cell of this lane: case index (bit i iff corner i <= iso) or -1 if the cell
// is outside the tile (compute_block_has_vertices.comp:28-37).
BCMC_DEV int cell_case(const float *tile, const uint32_t (&d)[3], float iso) {
    const uint32_t lane = lane_id();
    const uint32_t x = lane & 3u, y = (lane >> 2) & 3u, z = lane >> 4;
    const uint32_t base = x + 5 * y + 25 * z;
    if (!(x + 1 < d[0] && y + 1 < d[1] && z + 1 < d[2]))
        return -1;
    int c = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        const float v = tile[base + corner_offset(i)];
        c |= (v <= iso) ? (1 << i) : 0;
    }
    return c;
}

// Vertices of one block (compute_block_vertices.comp:76-174), vertex-parallel:
// the cells' counts are prefix-summed across the wave, each cell marks its
// vertices' owner in LDS, then lane v of each round interpolates vertex v, so
// the 64 lanes of a round write one contiguous run of the vertex buffer.
// Vertex order is the reference's: cells in x-fastest order, within a cell the
// order of its triangle-table row.  Vertices at or past `cap` are not written.
BCMC_DEV void emit_block_vertices(const float *tile, const uint32_t (&d)[3], float iso, uint32_t bid,
                                  const uint32_t (&nb)[3], unsigned long long out0,
                                  unsigned long long cap, void *verts, int fmt, const McTables &tab,
                                  EmitScratch &es) {
    const uint32_t lane = lane_id();
    const int c = cell_case(tile, d, iso);
    const uint32_t nv = c >= 0 ? tab.count[c] : 0u;
    const uint32_t incl = wave_incl_scan(nv);
    const uint32_t first = incl - nv, total = wave_read_lane(incl, 63);
    es.cell_case[lane] = (uint8_t)c;
    es.cell_first[lane] = (uint16_t)first;
    for (uint32_t k = 0; k < nv; ++k)
        es.vert_cell[first + k] = (uint8_t)lane;
    wave_lds_sync();
    const uint32_t bx = bid % nb[0], by = (bid / nb[0]) % nb[1], bz = bid / (nb[0] * nb[1]);
    for (uint32_t v = lane; v < total; v += 64) {
        const uint32_t cell = es.vert_cell[v];
        const uint32_t k = v - es.cell_first[cell];
        const uint32_t e = (uint32_t)(tab.edges[es.cell_case[cell]] >> (4 * k)) & 15u;
        const uint32_t ca = (uint32_t)(kEdgeCornerA >> (4 * e)) & 15u,
                       cb = (uint32_t)(kEdgeCornerB >> (4 * e)) & 15u;
        const uint32_t cx = cell & 3u, cy = (cell >> 2) & 3u, cz = cell >> 4;
        const uint32_t cbase = cx + 5 * cy + 25 * cz;
        const float fa = tile[cbase + corner_offset(ca)];
        const float fb = tile[cbase + corner_offset(cb)];
        // lerp_verts (compute_block_vertices.comp:62-74), then + voxel + 0.5
        float t = 0.0f;
        if (!(fabsf(fa - fb) < 0.001f))
            t = (iso - fa) / (fb - fa);
        const float ax = (float)corner_x(ca), ay = (float)corner_y(ca), az = (float)corner_z(ca);
        const float px = (ax + t * ((float)corner_x(cb) - ax) + (float)cx) + 0.5f;
        const float py = (ay + t * ((float)corner_y(cb) - ay) + (float)cy) + 0.5f;
        const float pz = (az + t * ((float)corner_z(cb) - az) + (float)cz) + 0.5f;
        const unsigned long long o = out0 + v;
        if (o >= cap)
            break;
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
    wave_lds_sync();  // the scratch is reused by the next block
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

// ---------------------------------------------------------------------------
// Block tiles with their ghost layer (compute_block_dims_with_ghost and
// load_block, compute_vertex_values.comp:96-252): 5^3 voxels at x + 5y + 25z,
// the ghost planes taken from the +x/+y/+z neighbours only when those are
// active.  NB tiles per wave with the dependent load rounds shared by all NB blocks:
// (1) neighbour activity flags and all 8 candidate cache slots, (2) the tile
// voxels.  bid[], live[] and the views are wave-uniform; block j may belong to
// a different volume/query than block j' (the batched engine).
// ---------------------------------------------------------------------------
// Views: `Views::get<j>()` returns the McView of block j (compile-time j).
template <int NB>
struct SameView {
    McView m;
    template <int J>
    BCMC_DEV const McView &get() const { return m; }
};
template <int NB>
struct PerBlockViews {
    McView m[NB];
    template <int J>
    BCMC_DEV const McView &get() const { return m[J]; }
};

template <int NB, class Views>
BCMC_DEV void load_tiles(const Views &mv, const uint32_t (&bid)[NB], const bool (&live)[NB],
                         float *tiles /* NB*128 */, int32_t *slots /* NB*8 */, uint32_t (&d)[NB][3]) {
    const uint32_t lane = lane_id();
    uint32_t bx[NB], by[NB], bz[NB], nbx[NB], nby[NB], nbz[NB];
    const uint32_t *flg[NB];
    const int32_t *ism[NB];
    const float *cache[NB];
    uint32_t cap[NB];
    unroll_for<NB>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const McView &m = mv.template get<j>();
        nbx[j] = m.nb[0];
        nby[j] = m.nb[1];
        nbz[j] = m.nb[2];
        flg[j] = m.flags;
        ism[j] = m.item_slot;
        cache[j] = m.cache;
        cap[j] = m.capacity;
        bx[j] = bid[j] % nbx[j];
        by[j] = (bid[j] / nbx[j]) % nby[j];
        bz[j] = bid[j] / (nbx[j] * nby[j]);
    });
    // lane j*8 + s: flag (s < 7: +x,+y,+z,+xy,+xz,+yz,+xyz) and slot of sub-block s
    bool act = false;
    if (lane < 8u * NB) {
        const uint32_t j = lane / 8, s = lane % 8;
        uint32_t jx = bx[0], jy = by[0], jz = bz[0], nx = nbx[0], ny = nby[0], nz = nbz[0];
        bool jl = live[0];
        const uint32_t *f = flg[0];
        const int32_t *sm = ism[0];
#pragma unroll
        for (int q = 1; q < NB; ++q) {
            const bool pick = j == (uint32_t)q;
            jx = pick ? bx[q] : jx;
            jy = pick ? by[q] : jy;
            jz = pick ? bz[q] : jz;
            nx = pick ? nbx[q] : nx;
            ny = pick ? nby[q] : ny;
            nz = pick ? nbz[q] : nz;
            jl = pick ? live[q] : jl;
            f = pick ? flg[q] : f;
            sm = pick ? ism[q] : sm;
        }
        // flag codes in the reference's check order; slot sub-block s = sx + 2sy + 4sz
        const uint32_t fcode = s < 3 ? (1u << s) : (s == 3 ? 3u : s == 4 ? 5u : s == 5 ? 6u : 7u);
        const uint32_t fx = jx + (fcode & 1u), fy = jy + ((fcode >> 1) & 1u), fz = jz + (fcode >> 2);
        if (jl && s < 7 && fx < nx && fy < ny && fz < nz)
            act = f[fx + nx * (fy + ny * fz)] != 0;
        const uint32_t sx = jx + (s & 1u), sy = jy + ((s >> 1) & 1u), sz = jz + (s >> 2);
        int32_t slot = -1;
        if (jl && sx < nx && sy < ny && sz < nz)
            slot = sm[sx + nx * (sy + ny * sz)];
        slots[lane] = slot;
    }
    const uint64_t ab = __ballot(act);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const uint32_t a = (uint32_t)(ab >> (8 * j)) & 0x7Fu;
        uint32_t dx = bx[j] + 1 < nbx[j] ? 5 : 4, dy = by[j] + 1 < nby[j] ? 5 : 4,
                 dz = bz[j] + 1 < nbz[j] ? 5 : 4;
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
        d[j][0] = dx;
        d[j][1] = dy;
        d[j][2] = dz;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // voxel loads of all NB tiles issued together (2 per lane per tile)
    float vv[NB][2];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t t = lane + 64u * h;
            const uint32_t x = t % 5, y = (t / 5) % 5, z = t / 25;
            float v = 0.0f;
            if (t < 125 && live[j] && x < d[j][0] && y < d[j][1] && z < d[j][2]) {
                const int32_t s = slots[8 * j + (x >> 2) + 2 * (y >> 2) + 4 * (z >> 2)];
                if (s >= 0 && (uint32_t)s < cap[j])
                    v = cache[j][(uint64_t)(uint32_t)s * 64 + (x & 3) + 4 * (y & 3) + 16 * (z & 3)];
            }
            vv[j][h] = v;
        }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        tiles[128 * j + lane] = vv[j][0];
        if (lane + 64 < 125)
            tiles[128 * j + lane + 64] = vv[j][1];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

BCMC_DEV uint32_t block_vertex_count(const float *tile, const uint32_t (&d)[3], float iso,
                                    const McTables &tab) {
    const int c = cell_case(tile, d, iso);
    return wave_sum(c >= 0 ? tab.count[c] : 0u);
}

}  // namespace bcmc
