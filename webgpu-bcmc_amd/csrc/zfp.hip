// ZFP fixed-rate float32 3-D block decoding on gfx950.
//
// Restates shaders/zfp_decompress.comp (bit reader :188-257, embedded
// bit-plane decoder :265-294, negabinary + sequency permutation :101-185,
// 259-263, 373-375, inverse lift :296-354, dequantisation :356-384) as one
// lane per block.  The fixed-rate words of the 64 blocks of a wave are first
// copied into LDS with coalesced loads; each lane then decodes its block from
// LDS.  Group-tested runs of zero bits are consumed with one count-trailing-
// zeros per newly significant coefficient instead of one bit at a time.
// Bit offsets are 64-bit (the reference multiplies block_index*max_bits in
// u32, zfp_decompress.comp:199).
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace bcmc {

// zfp's sequency order (zfp_decompress.comp:101-185): raster index of the
// i-th decoded coefficient, from "xyz" digit triples.
static constexpr char kPermXYZ[] =
    "000100010001011101110200020002111210201021120102012300030003"
    "211121112022202220310301031130103013122212221311131113320302"
    "032230203023222321312132231213123033303330322232223133313331"
    "233323332333";

__host__ __device__ constexpr int perm3d(int i) {
    return (kPermXYZ[3 * i] - '0') + 4 * (kPermXYZ[3 * i + 1] - '0') +
           16 * (kPermXYZ[3 * i + 2] - '0');
}

// LSB-first window over u64 words (LDS or global): the 64 bits from bit
// position p of the block, read from words i and i + 1 of its first word
// (bits past the block are junk and are masked off by the caller).  No
// reader state: every read is one pair of word loads and a funnel shift, so
// the decoder loops carry only the bit position.
// kGuard: words at or past nw read as 0 (a block read straight from a stream
// buffer of nw words; LDS-staged words carry their own zero padding).
template <bool kGuard>
struct BitWindow {
    const uint64_t *w;
    uint32_t p0;  // the block's first bit within w[0]
    uint64_t nw;  // words readable from w (kGuard)
    BCMC_DEV uint64_t at(uint32_t p) const {
        const uint32_t a = p0 + p, i = a >> 6, o = a & 63u;
        uint64_t lo, hi;
        if constexpr (kGuard) {
            lo = i < nw ? w[i] : 0ull;
            hi = i + 1 < nw ? w[i + 1] : 0ull;
        } else {
            lo = w[i];
            hi = w[i + 1];
        }
        return o ? ((lo >> o) | (hi << (64u - o))) : lo;
    }
};

// Decode one block into v[64] (raster order); an all-zero block (flag bit 0)
// decodes to +0.
template <bool kGuard>
BCMC_DEV void decode_block(const BitWindow<kGuard> &bw, uint32_t maxbits, float (&v)[64]) {
    const uint64_t head = bw.at(0);
    if ((head & 1u) == 0) {
#pragma unroll
        for (int i = 0; i < 64; ++i)
            v[i] = 0.0f;
        return;
    }
    const int emax = (int)((head >> 1) & 0xFFu) - 127;
    uint32_t c[64];
#pragma unroll
    for (int i = 0; i < 64; ++i)
        c[i] = 0;
    uint32_t p = 9, bits = maxbits - 9, n = 0;
    for (uint32_t k = 32; bits != 0 && k-- > 0;) {
        // the n coefficients already significant: one bit each, verbatim
        const uint32_t m = n < bits ? n : bits;
        bits -= m;
        uint64_t x = 0;
        if (m) {
            x = bw.at(p) & (m == 64 ? ~0ull : ((1ull << m) - 1ull));
            p += m;
        }
        // group tests + unary runs (zfp_decompress.comp:282-287): a group
        // bit 1 is followed by zeros up to the next significant coefficient
        // (one count-trailing-zeros over the window), or by 63 - n zeros when
        // the run reaches the last coefficient, which is then implied
        while (n < 64 && bits != 0) {
            const uint64_t win = bw.at(p);
            --bits;
            ++p;
            if ((win & 1u) == 0)
                break;
            const uint32_t lim = min(63u - n, bits);
            const uint64_t pk = (win >> 1) & ((1ull << lim) - 1ull);
            const uint32_t q = pk ? (uint32_t)__builtin_ctzll(pk) : lim;
            const uint32_t adv = pk ? q + 1 : lim;
            p += adv;
            bits -= adv;
            n += q;
            x |= 1ull << n;
            ++n;
        }
        // deposit bit plane k (zfp_decompress.comp:290-292); bits >= n are 0
        if (x) {
            const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
            for (int i = 0; i < 16; ++i)
                c[i] += ((lo >> i) & 1u) << k;
            if (n > 16) {
#pragma unroll
                for (int i = 16; i < 32; ++i)
                    c[i] += ((lo >> i) & 1u) << k;
            }
            if (n > 32) {
#pragma unroll
                for (int i = 32; i < 48; ++i)
                    c[i] += ((hi >> (i - 32)) & 1u) << k;
            }
            if (n > 48) {
#pragma unroll
                for (int i = 48; i < 64; ++i)
                    c[i] += ((hi >> (i - 32)) & 1u) << k;
            }
        }
    }
    // negabinary -> two's complement, sequency -> raster (register renaming)
    int32_t q[64];
#pragma unroll
    for (int i = 0; i < 64; ++i)
        q[perm3d(i)] = (int32_t)((c[i] ^ 0xaaaaaaaau) - 0xaaaaaaaau);
    // inverse lifts along z, y, x (zfp_decompress.comp:296-354), wrapping int32
    auto lift = [&](int i0, int s) {
        uint32_t x = (uint32_t)q[i0], y = (uint32_t)q[i0 + s], z = (uint32_t)q[i0 + 2 * s],
                 w = (uint32_t)q[i0 + 3 * s];
        y += (uint32_t)((int32_t)w >> 1);
        w -= (uint32_t)((int32_t)y >> 1);
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
        q[i0] = (int32_t)x;
        q[i0 + s] = (int32_t)y;
        q[i0 + 2 * s] = (int32_t)z;
        q[i0 + 3 * s] = (int32_t)w;
    };
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x)
            lift(x + 4 * y, 16);
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int z = 0; z < 4; ++z)
            lift(16 * z + x, 4);
#pragma unroll
    for (int z = 0; z < 4; ++z)
#pragma unroll
        for (int y = 0; y < 4; ++y)
            lift(4 * y + 16 * z, 1);
    const float inv_w = __builtin_ldexpf(1.0f, emax - 30);
#pragma unroll
    for (int i = 0; i < 64; ++i)
        v[i] = inv_w * (float)q[i];
}

// ---------------------------------------------------------------------------
// Decoder for blocks staged in LDS (64 | maxbits: W = maxbits / 64 words).
//
// A wave stages its 64 blocks' fixed-rate bits as dwords, block b at
// [b * kStride, b * kStride + 2W), followed by two all-ones dwords.  The ones
// stand for the budget's end: a run of zeros that reaches it stops there, as
// the reference's bit counter stops it (zfp_decompress.comp:282-287), so the
// run needs no budget test.  Every read is the 64-bit window at bit p: three
// dword loads and two v_alignbit.
//
// The bit planes (k = 31 .. 0, the same k in every lane at each step) are
// kept as 64-bit words X[k] in registers -- the verbatim bits of the n
// coefficients already significant, plus a 1 at each coefficient that becomes
// significant in the plane -- and turned into the 64 coefficients at the end
// by two 32x32 bit-matrix transposes, instead of depositing every plane's
// bits into 64 registers.
// ---------------------------------------------------------------------------
template <int W>
struct Staged {
    static constexpr int kStride = 2 * W + 3;  // odd: lanes' windows spread over banks
    static constexpr uint32_t kEnd = 64u * W;  // maxbits
};

BCMC_DEV void window64(const uint32_t *s, uint32_t p, uint32_t &lo, uint32_t &hi) {
    const uint32_t *q = s + __builtin_amdgcn_ubfe(p, 5u, 27u);
    const uint32_t a = q[0], b = q[1], c = q[2];
    lo = __builtin_amdgcn_alignbit(b, a, p);  // the shift uses p mod 32
    hi = __builtin_amdgcn_alignbit(c, b, p);
}

// In-place transpose of a 32x32 bit matrix: on entry bit i of a[k] is element
// (k, i), on exit bit k of a[i] is.  Quadrant swaps at distance 16 and 8 are
// byte permutes, those at 4, 2, 1 two shifts and two bit-field inserts.
BCMC_DEV void transpose32(uint32_t (&a)[32]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t x = a[k], y = a[k + 16];
        a[k] = __builtin_amdgcn_perm(y, x, 0x05040100u);
        a[k + 16] = __builtin_amdgcn_perm(y, x, 0x07060302u);
    }
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        if (k & 8)
            continue;
        const uint32_t x = a[k], y = a[k + 8];
        a[k] = __builtin_amdgcn_perm(y, x, 0x06020400u);
        a[k + 8] = __builtin_amdgcn_perm(y, x, 0x07030501u);
    }
    constexpr uint32_t kMask[3] = {0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
#pragma unroll
    for (int st = 0; st < 3; ++st) {
        const int j = 4 >> st;
        const uint32_t m = kMask[st];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            if (k & j)
                continue;
            const uint32_t x = a[k], y = a[k + j];
            a[k] = (x & m) | ((y << j) & ~m);
            a[k + j] = ((x >> j) & m) | (y & ~m);
        }
    }
}

template <int W>
BCMC_DEV void decode_block_staged(const uint32_t *s, float (&v)[64]) {
    constexpr uint32_t E = Staged<W>::kEnd;
    const uint32_t head = s[0];
    if ((head & 1u) == 0) {
#pragma unroll
        for (int i = 0; i < 64; ++i)
            v[i] = 0.0f;
        return;
    }
    const int emax = (int)((head >> 1) & 0xFFu) - 127;
    uint32_t xl[32], xh[32];  // bit plane k: coefficients 0..31 in xl[k], 32..63 in xh[k]
    uint32_t p = 9, n = 0;
#pragma unroll
    for (int k = 31; k >= 0; --k) {
        uint64_t x = 0;
        if (p < E) {
            // the n significant coefficients' bits, verbatim
            const uint32_t m = min(n, E - p);
            uint32_t lo, hi;
            window64(s, p, lo, hi);
            const uint64_t w0 = (uint64_t)hi << 32 | lo;
            x = w0 & (m >= 64 ? ~0ull : (1ull << m) - 1ull);
            p += m;
            // group tests (zfp_decompress.comp:282-287): a group bit 1, then
            // zeros up to the next significant coefficient and its 1; a run
            // reaching coefficient 63 implies it without a terminating bit.
            // Each group bit is read from the window that ends the step
            // before it (here: m < 64 whenever n < 64), so a step runs only
            // for a group bit 1, and a 0 ends the plane without a step.
            bool more = n < 64 && p < E;
            if (more && ((w0 >> m) & 1u) == 0) {
                ++p;
                more = false;
            }
            while (more) {
                uint32_t wl, wh;
                window64(s, p, wl, wh);
                // zeros after the group bit: ctz of the window shifted by one
                // (bit 63 forced to 1: a window of zeros is a run of >= 63)
                const uint64_t t = ((uint64_t)__builtin_amdgcn_alignbit(1u, wh, 1u) << 32) |
                                   __builtin_amdgcn_alignbit(wh, wl, 1u);
                const uint32_t r = (uint32_t)__builtin_ctzll(t);
                const uint32_t adv = min(r + 2u, 64u - n);
                n = min(n + r + 1u, 64u);  // one past the coefficient found
                x |= 1ull << (n - 1u);
                // the next group bit: bit adv of this window (adv == 64 only
                // after a run of 62 zeros from n = 0: read it)
                const uint64_t w = (uint64_t)wh << 32 | wl;
                uint32_t g = (uint32_t)(w >> (adv & 63u)) & 1u;
                p += adv;
                more = n < 64 && p < E;
                if (more && adv == 64u)
                    g = (s[p >> 5] >> (p & 31u)) & 1u;
                if (more && g == 0) {
                    ++p;
                    more = false;
                }
            }
        }
        // negabinary: the planes of 0xaaaaaaaa (the odd ones) are inverted
        xl[k] = (k & 1) ? ~(uint32_t)x : (uint32_t)x;
        xh[k] = (k & 1) ? ~(uint32_t)(x >> 32) : (uint32_t)(x >> 32);
    }
    transpose32(xl);
    transpose32(xh);
    int32_t q[64];
#pragma unroll
    for (int i = 0; i < 64; ++i)
        q[perm3d(i)] = (int32_t)((i < 32 ? xl[i] : xh[i - 32]) - 0xaaaaaaaau);
    // inverse lifts along z, y, x (zfp_decompress.comp:296-354), wrapping
    // int32; each step pair such as y += w; w <<= 1; w -= y is formed as
    // (y + w, w - y), which is the same value modulo 2^32
    auto lift = [&](int i0, int st) {
        uint32_t x = (uint32_t)q[i0], y = (uint32_t)q[i0 + st], z = (uint32_t)q[i0 + 2 * st],
                 w = (uint32_t)q[i0 + 3 * st];
        y += (uint32_t)((int32_t)w >> 1);
        w -= (uint32_t)((int32_t)y >> 1);
        uint32_t t = y;
        y = t + w;
        w = w - t;
        t = z;
        z = t + x;
        x = x - t;
        t = y;
        y = t + z;
        z = z - t;
        t = w;
        w = t + x;
        x = x - t;
        q[i0] = (int32_t)x;
        q[i0 + st] = (int32_t)y;
        q[i0 + 2 * st] = (int32_t)z;
        q[i0 + 3 * st] = (int32_t)w;
    };
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x)
            lift(x + 4 * y, 16);
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int z = 0; z < 4; ++z)
            lift(16 * z + x, 4);
#pragma unroll
    for (int z = 0; z < 4; ++z)
#pragma unroll
        for (int y = 0; y < 4; ++y)
            lift(4 * y + 16 * z, 1);
    const float inv_w = __builtin_ldexpf(1.0f, emax - 30);
#pragma unroll
    for (int i = 0; i < 64; ++i)
        v[i] = inv_w * (float)q[i];
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// Workgroups of 64, 128 or 256 threads (one block per lane; the autotuner
// picks the size, which changes nothing but speed); each wave stages its
// chunk's words in its own part of the dynamic LDS.
constexpr int kZfpThreads = 256;  // largest workgroup

enum ZfpMode { kModeRanges = 0, kModeCache = 1, kModeDense = 2, kModeLruAssign = 3 };

struct ZfpArgs {
    const uint64_t *stream;
    uint64_t nwords;
    uint32_t maxbits;
    const uint32_t *ids;       // block ids (cache/dense); null: identity
    const uint32_t *n_dev;     // device count (nullable)
    uint32_t n_host;
    const int32_t *item_slot;  // cache mode: destination slot per block id
    uint32_t capacity;         // cache mode: slots in `out` (others are skipped)
    float *out;                // cache / dense output
    float2 *ranges;            // ranges mode: range over the in-volume voxels
    float2 *part_ranges;       // ranges mode (nullable): 8 part ranges per block (all 64 voxels, faces, edges, voxel 0)
    uint32_t dims[3], nb[3];
    // LRU-assign mode: new item w takes slot order[w] (lru_cache_update.comp)
    const uint32_t *order;
    uint32_t *slot_age, *slot_avail;
    int32_t *slot_item, *item_slot_w;
    uint32_t total;            // totalElements
    QState *q;
};

// W = u64 words per block when 64 | maxbits (LDS-staged), 0 = generic.
// Dwords of LDS staging per wave.
__host__ __device__ constexpr int zfp_stage_dwords(int w) { return w > 0 ? 64 * (2 * w + 3) : 1; }

template <int W, int MODE>
__global__ __launch_bounds__(kZfpThreads) void zfp_kernel(ZfpArgs a) {
    constexpr int kStride = Staged<W>::kStride;
    constexpr int kStage = zfp_stage_dwords(W);
    static_assert(W == 0 || kStage == 64 * kStride, "staging stride");
    extern __shared__ uint32_t zfp_stage[];
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64, wpb = blockDim.x / 64;
    uint32_t n = a.n_dev ? *a.n_dev : a.n_host;
    if constexpr (MODE == kModeLruAssign) {
        // no assignment without available slots (lru_cache.js:717-769)
        n = a.q->num_avail_total ? a.q->num_new : 0u;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.q->avail_display = a.q->num_avail_total - n;
            a.q->cache_size = a.q->cache_size_next;
        }
        // the grown cache has more slots than the storage holds: assign
        // nothing (the host enlarges the storage and launches this again)
        if (a.q->cache_size_next > a.capacity) {
            if (blockIdx.x == 0 && threadIdx.x == 0)
                a.q->cache_short = 1;
            n = 0;
        }
    }
    const uint32_t waves_total = gridDim.x * wpb;
    for (uint32_t chunk = blockIdx.x * wpb + wave; (uint64_t)chunk * 64 < n; chunk += waves_total) {
        const uint32_t i = chunk * 64 + lane;
        const bool live = i < n;
        const uint32_t blk = live ? (a.ids ? a.ids[i] : i) : 0u;
        // destination lookups, issued before the decode so that their
        // latency hides behind it
        int32_t slot = -1, prev = -1;
        uint32_t s_new = 0xFFFFFFFFu;
        if constexpr (MODE == kModeLruAssign) {
            if (live)
                s_new = a.order[i];
        }
        float v[64];
        if constexpr (W > 0) {
            uint32_t *sw = zfp_stage + wave * kStage;
            // copy of the wave's fixed-rate words into LDS, block by block at
            // a stride of kStride dwords, each followed by two all-ones dwords
            // (the budget's end).  Word j of the wave is word j % W of the
            // chunk's block j / W, whose id lane j / W holds; all W loads of a
            // lane are issued before any is stored.
            uint32_t bsrc[W];
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const uint32_t src = ((uint32_t)t * 64u + lane) / W;
                bsrc[t] = a.ids ? (uint32_t)__shfl((int)blk, (int)src, 64) : chunk * 64 + src;
            }
            uint64_t words[W];
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const uint32_t j = (uint32_t)t * 64u + lane;
                const uint64_t wi = (uint64_t)bsrc[t] * W + j % W;
                words[t] = (chunk * 64 + j / W < n && wi < a.nwords) ? a.stream[wi] : 0ull;
            }
            if constexpr (MODE == kModeCache) {
                if (live)
                    slot = a.item_slot[blk];
            }
            if constexpr (MODE == kModeLruAssign) {
                if (s_new < a.capacity)
                    prev = a.slot_item[s_new];
            }
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const uint32_t j = (uint32_t)t * 64u + lane;
                uint32_t *d = sw + (j / W) * kStride + 2 * (j % W);
                d[0] = (uint32_t)words[t];
                d[1] = (uint32_t)(words[t] >> 32);
            }
            sw[lane * kStride + 2 * W] = ~0u;
            sw[lane * kStride + 2 * W + 1] = ~0u;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (live)
                decode_block_staged<W>(sw + lane * kStride, v);
            __builtin_amdgcn_wave_barrier();
        } else {
            if constexpr (MODE == kModeCache) {
                if (live)
                    slot = a.item_slot[blk];
            }
            if constexpr (MODE == kModeLruAssign) {
                if (s_new < a.capacity)
                    prev = a.slot_item[s_new];
            }
            const uint64_t bit0 = (uint64_t)blk * a.maxbits, w0 = bit0 >> 6;
            if (live)
                decode_block(BitWindow<true>{a.stream + w0, (uint32_t)(bit0 & 63), a.nwords > w0 ? a.nwords - w0 : 0ull},
                             a.maxbits, v);
        }
        if (!live)
            continue;
        if constexpr (MODE == kModeRanges) {
            // zfp_compute_block_range.comp:42-60
            const uint32_t bx = (blk % a.nb[0]) * 4, by = ((blk / a.nb[0]) % a.nb[1]) * 4,
                           bz = (blk / (a.nb[0] * a.nb[1])) * 4;
            const uint32_t sx = bx + 4 > a.dims[0] ? a.dims[0] - bx : 4;
            const uint32_t sy = by + 4 > a.dims[1] ? a.dims[1] - by : 4;
            const uint32_t sz = bz + 4 > a.dims[2] ? a.dims[2] - bz : 4;
            float lo = 1e20f, hi = -1e20f;
            // part ranges over all 64 decoded voxels (padding included): the
            // block, its x = 0 / y = 0 / z = 0 faces, the x = y = 0, x = z = 0
            // and y = z = 0 edges and voxel 0 -- the parts a block's -x / -y /
            // -z / ... neighbours read as ghost voxels (launch_tile_ranges)
            float plo[8], phi[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                plo[k] = v[0];
                phi[k] = v[0];
            }
#pragma unroll
            for (int j = 0; j < 64; ++j) {
                const bool in = (uint32_t)(j & 3) < sx && (uint32_t)((j >> 2) & 3) < sy &&
                                (uint32_t)(j >> 4) < sz;
                const float x = v[j];
                lo = (in && x < lo) ? x : lo;
                hi = (in && x > hi) ? x : hi;
                const bool x0 = (j & 3) == 0, y0 = ((j >> 2) & 3) == 0, z0 = (j >> 4) == 0;
                const bool part[8] = {true, x0, y0, z0, x0 && y0, x0 && z0, y0 && z0, j == 0};
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (part[k]) {
                        plo[k] = fminf(plo[k], x);
                        phi[k] = fmaxf(phi[k], x);
                    }
                }
            }
            a.ranges[blk] = make_float2(lo, hi);
            if (a.part_ranges) {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    a.part_ranges[(uint64_t)blk * 8 + k] = make_float2(plo[k], phi[k]);
            }
        } else {
            uint64_t dst = (uint64_t)i * 64;
            if constexpr (MODE == kModeCache) {
                if (slot < 0 || (uint32_t)slot >= a.capacity)
                    continue;
                dst = (uint64_t)(uint32_t)slot * 64;
            }
            if constexpr (MODE == kModeLruAssign) {
                // lru_cache_update.comp:25-36: evict the slot's previous item,
                // install this one (new items are never cached ones, so no two
                // threads write the same item entry)
                const uint32_t s = s_new;
                if (s >= a.capacity || blk >= a.total)
                    continue;  // cannot happen for a consistent state
                if (prev >= 0 && (uint32_t)prev < a.total)
                    a.item_slot_w[prev] = -1;
                a.slot_age[s] = 0;
                a.slot_item[s] = (int32_t)blk;
                a.slot_avail[s] = 0;
                a.item_slot_w[blk] = (int32_t)s;
                dst = (uint64_t)s * 64;
            }
            float4 *o = reinterpret_cast<float4 *>(a.out + dst);
#pragma unroll
            for (int j = 0; j < 16; ++j)
                o[j] = make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
        }
    }
}

// `threads` per workgroup: 64, 128 or 256.
template <int MODE>
static void launch_mode(const ZfpArgs &a, uint32_t grid, uint32_t threads, hipStream_t s) {
    const uint32_t W = a.maxbits % 64 == 0 ? a.maxbits / 64 : 0;
    const size_t lds = (size_t)(threads / 64) * zfp_stage_dwords((int)W) * 4;
    switch (W) {
    case 1: zfp_kernel<1, MODE><<<grid, threads, lds, s>>>(a); break;
    case 2: zfp_kernel<2, MODE><<<grid, threads, lds, s>>>(a); break;
    case 4: zfp_kernel<4, MODE><<<grid, threads, lds, s>>>(a); break;
    case 8: zfp_kernel<8, MODE><<<grid, threads, lds, s>>>(a); break;
    case 16: zfp_kernel<16, MODE><<<grid, threads, lds, s>>>(a); break;
    default: zfp_kernel<0, MODE><<<grid, threads, lds, s>>>(a); break;
    }
}

static ZfpArgs base_args(const uint64_t *stream, uint64_t nwords, uint32_t maxbits) {
    ZfpArgs a{};
    a.stream = stream;
    a.nwords = nwords;
    a.maxbits = maxbits;
    return a;
}

void launch_block_ranges(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords, uint32_t maxbits,
                         const uint32_t dims[3], float2 *ranges, float2 *part_ranges) {
    ZfpArgs a = base_args(stream, nwords, maxbits);
    for (int i = 0; i < 3; ++i) {
        a.dims[i] = dims[i];
        a.nb[i] = (dims[i] + 3) / 4;
    }
    a.n_host = a.nb[0] * a.nb[1] * a.nb[2];
    a.ranges = ranges;
    a.part_ranges = part_ranges;
    const uint32_t grid = div_up(a.n_host, kZfpThreads);
    if (grid)
        launch_mode<kModeRanges>(a, grid, kZfpThreads, ctx->stream);
}

// Per block: [min lo, max hi] over the in-grid blocks at offsets LO..1 on
// each axis: LO = -1 is the 3x3x3 neighbourhood (the active test's interval),
// LO = 0 the block and its +x/+y/+z, edge and corner neighbours (a bound on the
// tile range from block ranges alone).  Block ranges cover voxels inside the
// volume only, while the cells of a block last on an axis whose size is not a
// multiple of 4 also read its padding voxels: bit k of `padded` set marks axis
// k so, and such blocks get (-inf, inf).
template <int LO>
__global__ void union_ranges_kernel(const float2 *ranges, uint32_t nbx, uint32_t nby, uint32_t nbz,
                                    uint32_t padded, float2 *out) {
    const uint64_t n = (uint64_t)nbx * nby * nbz;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(b % nbx), y = (uint32_t)((b / nbx) % nby), z = (uint32_t)(b / ((uint64_t)nbx * nby));
        if (((padded & 1u) && x + 1 == nbx) || ((padded & 2u) && y + 1 == nby) || ((padded & 4u) && z + 1 == nbz)) {
            out[b] = make_float2(-INFINITY, INFINITY);
            continue;
        }
        float lo = ranges[b].x, hi = ranges[b].y;
        for (int dz = LO; dz <= 1; ++dz)
            for (int dy = LO; dy <= 1; ++dy) {
                const int yy = (int)y + dy, zz = (int)z + dz;
                if (yy < 0 || zz < 0 || yy >= (int)nby || zz >= (int)nbz)
                    continue;
                const float2 *row = ranges + (uint64_t)nbx * ((uint32_t)yy + (uint64_t)nby * (uint32_t)zz);
                for (int dx = LO; dx <= 1; ++dx) {
                    const int xx = (int)x + dx;
                    if (xx < 0 || xx >= (int)nbx)
                        continue;
                    const float2 v = row[xx];
                    lo = fminf(lo, v.x);
                    hi = fmaxf(hi, v.y);
                }
            }
        out[b] = make_float2(lo, hi);
    }
}

// Per block: [min, max] over its whole 5^3 tile, i.e. every decoded voxel its
// cells can read: its own 64 (padding included) and the ghost voxels of its
// in-grid +x / +y / +z / +xy / +xz / +yz / +xyz neighbours, which are their
// x = 0 face, y = 0 face, z = 0 face, x = y = 0 edge, x = z = 0 edge,
// y = z = 0 edge and voxel 0 (part ranges 1..7 of the range pass).
__global__ void tile_ranges_kernel(const float2 *parts, uint32_t nbx, uint32_t nby, uint32_t nbz,
                                   float2 *out) {
    const uint64_t n = (uint64_t)nbx * nby * nbz;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(b % nbx), y = (uint32_t)((b / nbx) % nby),
                       z = (uint32_t)(b / ((uint64_t)nbx * nby));
        float lo = parts[b * 8].x, hi = parts[b * 8].y;
        for (uint32_t s = 1; s < 8; ++s) {
            const uint32_t xx = x + (s & 1u), yy = y + ((s >> 1) & 1u), zz = z + (s >> 2);
            if (xx >= nbx || yy >= nby || zz >= nbz)
                continue;
            // neighbour s: its part k reads with x0 for +x, y0 for +y, z0 for +z
            const uint32_t k = s == 1 ? 1u : s == 2 ? 2u : s == 4 ? 3u : s == 3 ? 4u : s == 5 ? 5u : s == 6 ? 6u : 7u;
            const float2 v = parts[(xx + (uint64_t)nbx * (yy + (uint64_t)nby * zz)) * 8 + k];
            lo = fminf(lo, v.x);
            hi = fmaxf(hi, v.y);
        }
        out[b] = make_float2(lo, hi);
    }
}

void launch_tile_ranges(bcmc_ctx *ctx, const float2 *part_ranges, const uint32_t nb[3], float2 *out) {
    const uint64_t n = (uint64_t)nb[0] * nb[1] * nb[2];
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, 256), persistent_grid(ctx, 8)));
    tile_ranges_kernel<<<grid, 256, 0, ctx->stream>>>(part_ranges, nb[0], nb[1], nb[2], out);
}

void launch_dilate_ranges(bcmc_ctx *ctx, const float2 *ranges, const uint32_t nb[3], float2 *out) {
    const uint64_t n = (uint64_t)nb[0] * nb[1] * nb[2];
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, 256), persistent_grid(ctx, 8)));
    union_ranges_kernel<-1><<<grid, 256, 0, ctx->stream>>>(ranges, nb[0], nb[1], nb[2], 0u, out);
}

void launch_tile_bound_ranges(bcmc_ctx *ctx, const float2 *ranges, const uint32_t dims[3], float2 *out) {
    const uint32_t nb[3] = {(dims[0] + 3) / 4, (dims[1] + 3) / 4, (dims[2] + 3) / 4};
    const uint32_t padded = (dims[0] % 4 ? 1u : 0u) | (dims[1] % 4 ? 2u : 0u) | (dims[2] % 4 ? 4u : 0u);
    const uint64_t n = (uint64_t)nb[0] * nb[1] * nb[2];
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, 256), persistent_grid(ctx, 8)));
    union_ranges_kernel<0><<<grid, 256, 0, ctx->stream>>>(ranges, nb[0], nb[1], nb[2], padded, out);
}

void launch_decode_to_cache(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords,
                            uint32_t maxbits, const uint32_t *ids, const uint32_t *n_dev,
                            uint32_t n_max, const int32_t *item_slot, float *cache,
                            uint32_t capacity) {
    ZfpArgs a = base_args(stream, nwords, maxbits);
    a.ids = ids;
    a.n_dev = n_dev;
    a.n_host = n_max;
    a.item_slot = item_slot;
    a.capacity = capacity;
    a.out = cache;
    const uint32_t threads = ctx->tune.decode_threads;
    const uint32_t grid = std::min(div_up(n_max, threads), persistent_grid(ctx, ctx->tune.decode));
    if (grid)
        launch_mode<kModeCache>(a, grid, threads, ctx->stream);
}

void launch_decode_assign(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords, uint32_t maxbits,
                          const uint32_t *new_items, uint32_t n_max, const uint32_t *order,
                          uint32_t *slot_age, uint32_t *slot_avail, int32_t *slot_item,
                          int32_t *item_slot, uint32_t total, float *cache, uint32_t capacity,
                          QState *q) {
    ZfpArgs a = base_args(stream, nwords, maxbits);
    a.ids = new_items;
    a.n_host = n_max;
    a.order = order;
    a.slot_age = slot_age;
    a.slot_avail = slot_avail;
    a.slot_item = slot_item;
    a.item_slot_w = item_slot;
    a.total = total;
    a.capacity = capacity;
    a.out = cache;
    a.q = q;
    const uint32_t threads = ctx->tune.decode_threads;
    const uint32_t grid = std::max(1u, std::min(div_up(n_max, threads), persistent_grid(ctx, ctx->tune.decode)));
    launch_mode<kModeLruAssign>(a, grid, threads, ctx->stream);
}

void launch_decode_dense(bcmc_ctx *ctx, const uint64_t *stream, uint64_t nwords, uint32_t maxbits,
                         const uint32_t *ids, uint32_t n, float *out) {
    ZfpArgs a = base_args(stream, nwords, maxbits);
    a.ids = ids;
    a.n_host = n;
    a.out = out;
    const uint32_t grid = div_up(n, kZfpThreads);
    if (grid)
        launch_mode<kModeDense>(a, grid, kZfpThreads, ctx->stream);
}

// ---------------------------------------------------------------------------
// Fixed-rate encoder (zfp 0.5.5 compress, float32 3-D, fixed-rate mode): the
// inverse of the decoder above, byte-identical to libzfp's zfp_compress.  One
// lane per block: gather with zfp's partial-block padding, block exponent,
// quantise to int32, forward decorrelating lifts (x, then y, then z),
// sequency permutation + negabinary, embedded bit-plane coding with group
// tests, zero padding up to exactly max_bits bits.  The stream words must be
// zero on entry; a block's first and last words may be shared with its
// neighbours (max_bits not a multiple of 64) and are merged with atomic ORs,
// every other word is owned by one block and stored plainly.
// ---------------------------------------------------------------------------
constexpr int kEncThreads = 256;

// zfp's pad_block for one line of 4 values with n valid (n <= 4)
BCMC_DEV void pad_line(float &p0, float &p1, float &p2, float &p3, uint32_t n) {
    p0 = n == 0 ? 0.0f : p0;
    p1 = n < 2 ? p0 : p1;
    p2 = n < 3 ? p1 : p2;
    p3 = n < 4 ? p0 : p3;
}

BCMC_DEV uint32_t asr1(uint32_t v) { return (uint32_t)((int32_t)v >> 1); }

// forward lift of 4 values (int32, wrapping; zfp fwd_lift)
BCMC_DEV void fwd_lift4(uint32_t &x, uint32_t &y, uint32_t &z, uint32_t &w) {
    x += w; x = asr1(x); w -= x;
    z += y; z = asr1(z); y -= z;
    x += z; x = asr1(x); z -= x;
    w += y; w = asr1(w); y -= w;
    w += asr1(y); y -= asr1(w);
}

// float -> int32 as x86's truncating conversion (what libzfp's (int32)(s*f)
// compiles to): values outside the int32 range, infinities and NaN give
// INT32_MIN.  In range it is the plain truncation.
BCMC_DEV uint32_t quantise(float v) {
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (uint32_t)(int32_t)v : 0x80000000u;
}

// LSB-first writer of one block's bits into zeroed stream words.
struct BitSink {
    uint64_t *out;
    uint64_t word;               // index of the word being filled
    uint64_t first, last;        // the block's first and last words
    bool share_first, share_last;
    uint64_t acc = 0;
    uint32_t fill;               // bits of acc in use
    BCMC_DEV void flush() {
        const bool shared = (word == first && share_first) || (word == last && share_last);
        if (shared)
            atomicOr(reinterpret_cast<unsigned long long *>(out + word), (unsigned long long)acc);
        else
            out[word] = acc;
        ++word;
    }
    // the low n bits of v (bits above n are zero), n <= 64
    BCMC_DEV void put(uint64_t v, uint32_t n) {
        acc |= fill < 64 ? v << fill : 0ull;
        if (fill + n >= 64) {
            const uint32_t used = 64 - fill;  // bits of v that went into acc
            flush();
            acc = used < 64 ? v >> used : 0ull;
            fill = fill + n - 64;
        } else {
            fill += n;
        }
    }
    BCMC_DEV void finish() {
        if (fill)
            flush();
    }
};

BCMC_DEV uint64_t low_bits(uint64_t x, uint32_t m) { return m >= 64 ? x : x & ((1ull << m) - 1ull); }

__global__ __launch_bounds__(kEncThreads) void zfp_encode_kernel(const float *vol, uint32_t d0, uint32_t d1,
                                                                  uint32_t d2, uint32_t nbx, uint32_t nby,
                                                                  uint64_t nblocks, uint32_t maxbits,
                                                                  uint64_t *out) {
    const uint64_t stride = (uint64_t)gridDim.x * kEncThreads;
    for (uint64_t b = (uint64_t)blockIdx.x * kEncThreads + threadIdx.x; b < nblocks; b += stride) {
        const uint32_t bx = (uint32_t)(b % nbx), by = (uint32_t)((b / nbx) % nby),
                       bz = (uint32_t)(b / ((uint64_t)nbx * nby));
        const uint32_t x0 = 4 * bx, y0 = 4 * by, z0 = 4 * bz;
        const uint32_t nx = min(4u, d0 - x0), ny = min(4u, d1 - y0), nz = min(4u, d2 - z0);
        // gather with clamped coordinates, then pad along x, y, z (the
        // clamped values are all overwritten by the padding)
        float f[64];
#pragma unroll
        for (int z = 0; z < 4; ++z)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const uint64_t row = ((uint64_t)min(z0 + z, d2 - 1) * d1 + min(y0 + y, d1 - 1)) * d0;
#pragma unroll
                for (int x = 0; x < 4; ++x)
                    f[16 * z + 4 * y + x] = vol[row + min(x0 + x, d0 - 1)];
            }
#pragma unroll
        for (int r = 0; r < 16; ++r)
            pad_line(f[4 * r], f[4 * r + 1], f[4 * r + 2], f[4 * r + 3], nx);
#pragma unroll
        for (int z = 0; z < 4; ++z)
#pragma unroll
            for (int x = 0; x < 4; ++x)
                pad_line(f[16 * z + x], f[16 * z + 4 + x], f[16 * z + 8 + x], f[16 * z + 12 + x], ny);
#pragma unroll
        for (int r = 0; r < 16; ++r)
            pad_line(f[r], f[16 + r], f[32 + r], f[48 + r], nz);
        float amax = 0.0f;
#pragma unroll
        for (int i = 0; i < 64; ++i)
            amax = fmaxf(amax, fabsf(f[i]));
        // zfp's exponent: frexp exponent of the largest magnitude, at least
        // 1 - 127 (subnormals), -127 for an all-zero block
        const uint32_t eb = (__float_as_uint(amax) >> 23) & 0xFFu;
        const int emax = amax > 0.0f ? (eb == 0 ? -126 : (int)eb - 126) : -127;
        const uint32_t e = (uint32_t)(emax + 127);
        const uint64_t start = b * maxbits;
        BitSink w;
        w.out = out;
        w.first = start >> 6;
        w.last = (start + maxbits - 1) >> 6;
        w.word = w.first;
        w.share_first = (start & 63u) != 0;
        w.share_last = ((start + maxbits) & 63u) != 0;
        w.fill = (uint32_t)(start & 63u);
        if (e == 0)
            continue;  // flag bit 0 and zero padding: the words are already zero
        w.put(2ull * e + 1ull, 9);
        const float s = ldexpf(1.0f, 30 - emax);
        uint32_t q[64];
#pragma unroll
        for (int i = 0; i < 64; ++i)
            q[i] = quantise(s * f[i]);
#pragma unroll
        for (int r = 0; r < 16; ++r)
            fwd_lift4(q[4 * r], q[4 * r + 1], q[4 * r + 2], q[4 * r + 3]);
#pragma unroll
        for (int z = 0; z < 4; ++z)
#pragma unroll
            for (int x = 0; x < 4; ++x)
                fwd_lift4(q[16 * z + x], q[16 * z + 4 + x], q[16 * z + 8 + x], q[16 * z + 12 + x]);
#pragma unroll
        for (int r = 0; r < 16; ++r)
            fwd_lift4(q[r], q[16 + r], q[32 + r], q[48 + r]);
        uint32_t u[64];
#pragma unroll
        for (int i = 0; i < 64; ++i)
            u[i] = (q[perm3d(i)] + 0xaaaaaaaau) ^ 0xaaaaaaaau;
        // embedded coding: per bit plane, the first n coefficients' bits
        // verbatim, then group tests: a 1 says another coefficient becomes
        // significant, followed by the run of zeros up to it and its 1 bit
        uint32_t bits = maxbits - 9, n = 0;
        for (int k = 31; k >= 0 && bits; --k) {
            uint64_t x = 0;
#pragma unroll
            for (int i = 0; i < 64; ++i)
                x |= (uint64_t)((u[i] >> k) & 1u) << i;
            const uint32_t m = min(n, bits);
            bits -= m;
            w.put(low_bits(x, m), m);
            x = m >= 64 ? 0ull : x >> m;
            while (n < 64 && bits) {
                --bits;
                w.put(x != 0, 1);
                if (!x)
                    break;
                // zeros up to the next 1, capped by the coefficient count
                // and the budget
                const uint32_t tz = (uint32_t)__builtin_ctzll(x), cap = min(63u - n, bits);
                if (tz < cap) {
                    w.put(1ull << tz, tz + 1);
                    bits -= tz + 1;
                    x >>= tz;
                    n += tz;
                } else {
                    w.put(0, cap);
                    bits -= cap;
                    x = cap >= 64 ? 0ull : x >> cap;
                    n += cap;
                }
                x >>= 1;
                ++n;
            }
        }
        w.finish();
    }
}

void launch_encode(bcmc_ctx *ctx, const float *vol, const uint32_t d[3], uint32_t maxbits, uint64_t *out) {
    const uint32_t nbx = div_up(d[0], 4), nby = div_up(d[1], 4), nbz = div_up(d[2], 4);
    const uint64_t n = (uint64_t)nbx * nby * nbz;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(div_up(n, kEncThreads), persistent_grid(ctx, 8)));
    zfp_encode_kernel<<<grid, kEncThreads, 0, ctx->stream>>>(vol, d[0], d[1], d[2], nbx, nby, n, maxbits, out);
}

}  // namespace bcmc

using namespace bcmc;

// With timing enabled, the device time of the launch just issued is kept in
// ctx->last_kernel_ms (HIP events on the context stream; waits for it).
static int finish_timing(bcmc_ctx *ctx) {
    if (!ctx->timing)
        return BCMC_OK;
    BCMC_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
    BCMC_HIP_CHECK(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    BCMC_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return BCMC_OK;
}

extern "C" BCMC_API int bcmc_zfp_decode_blocks(bcmc_ctx *ctx, bcmc_buf *stream, uint32_t max_bits,
                                               bcmc_buf *ids, uint32_t n, bcmc_buf *out) {
    BCMC_CHECK_ARG(ctx && stream && out, "null argument");
    BCMC_CHECK_ARG(max_bits >= 9, "max_bits must be >= 9");
    BCMC_CHECK_ARG(out->bytes >= (uint64_t)n * 256, "output buffer too small");
    BCMC_CHECK_ARG(!ids || ids->bytes >= (uint64_t)n * 4, "ids buffer too small");
    // every block read must lie inside the stream (plus the 8-byte prefetch pad)
    if (!ids)
        BCMC_CHECK_ARG((uint64_t)n * max_bits <= stream->bytes * 8, "stream too short for n blocks");
    if (ctx->timing)
        BCMC_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
    launch_decode_dense(ctx, (const uint64_t *)stream->ptr, stream->bytes / 8, max_bits,
                        ids ? (const uint32_t *)ids->ptr : nullptr, n, (float *)out->ptr);
    BCMC_HIP_CHECK(hipGetLastError());
    return finish_timing(ctx);
}

extern "C" BCMC_API int bcmc_zfp_block_ranges(bcmc_ctx *ctx, bcmc_buf *stream, uint32_t max_bits,
                                              const uint32_t dims[3], bcmc_buf *ranges) {
    BCMC_CHECK_ARG(ctx && stream && ranges && dims, "null argument");
    BCMC_CHECK_ARG(max_bits >= 9, "max_bits must be >= 9");
    const uint64_t nb = (uint64_t)((dims[0] + 3) / 4) * ((dims[1] + 3) / 4) * ((dims[2] + 3) / 4);
    BCMC_CHECK_ARG(nb > 0 && nb < (1ull << 32), "bad dims");
    BCMC_CHECK_ARG(ranges->bytes >= nb * 8, "ranges buffer too small");
    BCMC_CHECK_ARG(nb * max_bits <= stream->bytes * 8, "stream too short for the volume");
    if (ctx->timing)
        BCMC_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
    launch_block_ranges(ctx, (const uint64_t *)stream->ptr, stream->bytes / 8, max_bits, dims,
                        (float2 *)ranges->ptr);
    BCMC_HIP_CHECK(hipGetLastError());
    return finish_timing(ctx);
}

extern "C" BCMC_API int bcmc_zfp_encode(bcmc_ctx *ctx, bcmc_buf *field, const uint32_t dims[3],
                                        uint32_t max_bits, bcmc_buf *stream) {
    BCMC_CHECK_ARG(ctx && field && stream && dims, "null argument");
    BCMC_CHECK_ARG(max_bits >= 9 && max_bits <= 4096, "max_bits must be in [9, 4096]");
    BCMC_CHECK_ARG(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    const uint64_t nb = (uint64_t)((dims[0] + 3) / 4) * ((dims[1] + 3) / 4) * ((dims[2] + 3) / 4);
    const uint64_t voxels = (uint64_t)dims[0] * dims[1] * dims[2];
    BCMC_CHECK_ARG(field->bytes >= voxels * 4, "field buffer smaller than dims");
    const uint64_t words = (nb * max_bits + 63) / 64;
    BCMC_CHECK_ARG(stream->bytes >= words * 8, "stream buffer smaller than the encoded volume");
    BCMC_CHECK_ARG(nb < (1ull << 40), "volume has too many blocks");
    if (ctx->timing)
        BCMC_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
    // the kernel ORs bits into shared words: the whole stream buffer starts at zero
    BCMC_HIP_CHECK(hipMemsetAsync(stream->ptr, 0, stream->bytes, ctx->stream));
    launch_encode(ctx, (const float *)field->ptr, dims, max_bits, (uint64_t *)stream->ptr);
    BCMC_HIP_CHECK(hipGetLastError());
    return finish_timing(ctx);
}
