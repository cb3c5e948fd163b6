/* Synthetic inputs for the block-compressed marching-cubes path: seeded scalar
 * fields and a zfp-compatible fixed-rate float32 3-D encoder.
 *
 * The reference consumes streams written by LLNL zfp (js/bcmc.js:27-38 fetches
 * `<name>.zfp`, a headerless fixed-rate stream; js/volumes.js gives the rates).
 * No such file is available offline, so inputs are synthesised here.  The
 * encoder restates the published zfp 0.5.5 fixed-rate algorithm (float32, 3-D,
 * 4^3 blocks, 64-bit stream words, no header): block-floating-point cast,
 * forward decorrelating lift, sequency reorder with negabinary mapping,
 * embedded bit-plane coding with group testing, zero padding to maxbits.  The
 * tests pin its output byte-for-byte against streams written by libzfp 0.5.5.
 *
 * Random numbers: SplitMix64 (Steele/Lea/Flood 2014), counter form
 *   z = seed + (i+1)*0x9E3779B97F4A7C15, mixed with the 30/27/31 xor-shift
 *   multiplies; u = (z >> 11) * 2^-53 in [0,1).
 * Fields are evaluated in FP64, normalised to [0,1] over the volume, and
 * rounded to FP32.  All loops have a fixed summation order, so results do not
 * depend on the thread count; build without -ffast-math and with
 * -ffp-contract=off so every host produces the same bytes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------------------
 * SplitMix64
 * ------------------------------------------------------------------------ */
static uint64_t sm64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double u01_at(uint64_t seed, uint64_t i) { return (double)(sm64_at(seed, i) >> 11) * 0x1.0p-53; }

typedef struct {
    uint64_t seed, i;
} rng_t;

static double rng_u01(rng_t *r) { return u01_at(r->seed, r->i++); }
static double rng_range(rng_t *r, double lo, double hi) { return lo + (hi - lo) * rng_u01(r); }

/* ---------------------------------------------------------------------------
 * Fields.  kind: 0 = Gaussians, 1 = Fourier modes, 2 = blobs (+ optional noise)
 * ------------------------------------------------------------------------ */
enum { FIELD_GAUSS = 0, FIELD_FOURIER = 1, FIELD_BLOBS = 2 };

typedef struct {
    int kind;
    uint64_t seed;
    uint32_t dims[3];
    uint32_t count;            /* number of Gaussians / modes / blobs */
    double p0, p1, p2, p3;     /* kind-specific ranges, see make_params */
    double noise;              /* blobs: noise amplitude (0 = none) */
    /* derived per-item parameters */
    double *item;              /* count * 8 */
    /* per-axis tables */
    double *tab[3];            /* gauss/blobs: count*dim; fourier: count*dim*2 (re, im) */
    double fmin, fmax;         /* smooth-part normalisation */
    double gmin, gmax;         /* after-noise normalisation */
} field_t;

/* Gaussians (kind 0): count items; per item draw cx, cy, cz ~ U[0,dim),
 * sigma ~ U[p0,p1], amplitude ~ U[p2,p3] in that order.  Blobs (kind 2) use
 * the same draw order.  Fourier (kind 1): per mode draw |k| ~ U[p0,p1],
 * cos(theta) ~ U[-1,1], phi ~ U[0,2pi), phase ~ U[0,2pi); amplitude |k|^-5/6;
 * mode = a cos(2 pi (kx x/Lx + ky y/Ly + kz z/Lz) + phase). */
static void field_init(field_t *f)
{
    rng_t r = {f->seed, 0};
    f->item = calloc((size_t)f->count * 8, sizeof(double));
    for (int a = 0; a < 3; ++a)
        f->tab[a] = calloc((size_t)f->count * f->dims[a] * 2, sizeof(double));
    for (uint32_t m = 0; m < f->count; ++m) {
        double *it = f->item + 8 * m;
        if (f->kind == FIELD_FOURIER) {
            const double kmag = rng_range(&r, f->p0, f->p1);
            const double ct = rng_range(&r, -1.0, 1.0);
            const double ph = rng_range(&r, 0.0, 2.0 * M_PI);
            const double phase = rng_range(&r, 0.0, 2.0 * M_PI);
            const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
            it[0] = kmag * st * cos(ph);
            it[1] = kmag * st * sin(ph);
            it[2] = kmag * ct;
            it[3] = pow(kmag, -5.0 / 6.0);
            it[4] = phase;
            for (int a = 0; a < 3; ++a) {
                double *t = f->tab[a] + (size_t)m * f->dims[a] * 2;
                for (uint32_t x = 0; x < f->dims[a]; ++x) {
                    const double ang = 2.0 * M_PI * it[a] * (double)x / (double)f->dims[a];
                    t[2 * x] = cos(ang);
                    t[2 * x + 1] = sin(ang);
                }
            }
        } else {
            it[0] = rng_range(&r, 0.0, f->dims[0]);
            it[1] = rng_range(&r, 0.0, f->dims[1]);
            it[2] = rng_range(&r, 0.0, f->dims[2]);
            it[3] = rng_range(&r, f->p0, f->p1);
            it[4] = rng_range(&r, f->p2, f->p3);
            const double inv = 1.0 / (2.0 * it[3] * it[3]);
            for (int a = 0; a < 3; ++a) {
                double *t = f->tab[a] + (size_t)m * f->dims[a];
                for (uint32_t x = 0; x < f->dims[a]; ++x) {
                    const double d = (double)x - it[a];
                    t[x] = exp(-d * d * inv);
                }
            }
        }
    }
}

static void field_free(field_t *f)
{
    free(f->item);
    for (int a = 0; a < 3; ++a)
        free(f->tab[a]);
}

/* Smooth part of the field over one row (all x) at (y, z). */
static void field_row(const field_t *f, uint32_t y, uint32_t z, double *row)
{
    const uint32_t nx = f->dims[0];
    for (uint32_t x = 0; x < nx; ++x)
        row[x] = 0.0;
    for (uint32_t m = 0; m < f->count; ++m) {
        const double *it = f->item + 8 * m;
        if (f->kind == FIELD_FOURIER) {
            const double *tx = f->tab[0] + (size_t)m * nx * 2;
            const double *ty = f->tab[1] + (size_t)m * f->dims[1] * 2 + 2 * y;
            const double *tz = f->tab[2] + (size_t)m * f->dims[2] * 2 + 2 * z;
            /* b = a e^{i phase} e^{i ky y} e^{i kz z} */
            const double cr = cos(it[4]) * it[3], ci = sin(it[4]) * it[3];
            const double yzr = ty[0] * tz[0] - ty[1] * tz[1], yzi = ty[0] * tz[1] + ty[1] * tz[0];
            const double br = cr * yzr - ci * yzi, bi = cr * yzi + ci * yzr;
            for (uint32_t x = 0; x < nx; ++x)
                row[x] += br * tx[2 * x] - bi * tx[2 * x + 1];
        } else {
            const double *tx = f->tab[0] + (size_t)m * nx;
            const double w = it[4] * f->tab[1][(size_t)m * f->dims[1] + y] *
                             f->tab[2][(size_t)m * f->dims[2] + z];
            for (uint32_t x = 0; x < nx; ++x)
                row[x] += w * tx[x];
        }
    }
}

/* Standard normal for voxel i (Box-Muller on two counter-based uniforms). */
static double noise_at(uint64_t seed, uint64_t i)
{
    const uint64_t ns = seed ^ 0xD1B54A32D192ED03ull;
    const double u1 = 1.0 - u01_at(ns, 2 * i), u2 = u01_at(ns, 2 * i + 1);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Final normalised value (FP64) at (x, y, z) given the smooth row value. */
static double field_final(const field_t *f, double smooth, uint64_t vi)
{
    double v = (smooth - f->fmin) / (f->fmax - f->fmin);
    if (f->noise > 0.0) {
        v += f->noise * noise_at(f->seed, vi);
        v = (v - f->gmin) / (f->gmax - f->gmin);
    }
    return v;
}

static int nthreads(int t) { return t > 0 ? t : 1; }

/* Pass 1 (and 2 with noise): global min/max.  When `store` is non-null the
 * smooth field is kept there (x fastest) so later passes need not recompute it. */
static void field_minmax(field_t *f, int threads, double *store)
{
    const uint32_t nx = f->dims[0], ny = f->dims[1], nz = f->dims[2];
    double lo = INFINITY, hi = -INFINITY;
#pragma omp parallel num_threads(nthreads(threads))
    {
        double *tmp = malloc(nx * sizeof(double));
        double l = INFINITY, h = -INFINITY;
#pragma omp for schedule(static)
        for (int64_t zy = 0; zy < (int64_t)ny * nz; ++zy) {
            double *row = store ? store + (size_t)zy * nx : tmp;
            field_row(f, (uint32_t)(zy % ny), (uint32_t)(zy / ny), row);
            for (uint32_t x = 0; x < nx; ++x) {
                l = row[x] < l ? row[x] : l;
                h = row[x] > h ? row[x] : h;
            }
        }
#pragma omp critical
        {
            lo = l < lo ? l : lo;
            hi = h > hi ? h : hi;
        }
        free(tmp);
    }
    f->fmin = lo;
    f->fmax = hi;
    if (f->noise > 0.0) {
        lo = INFINITY;
        hi = -INFINITY;
#pragma omp parallel num_threads(nthreads(threads))
        {
            double *tmp = malloc(nx * sizeof(double));
            double l = INFINITY, h = -INFINITY;
#pragma omp for schedule(static)
            for (int64_t zy = 0; zy < (int64_t)ny * nz; ++zy) {
                double *row = store ? store + (size_t)zy * nx : tmp;
                if (!store)
                    field_row(f, (uint32_t)(zy % ny), (uint32_t)(zy / ny), row);
                for (uint32_t x = 0; x < nx; ++x) {
                    const double v = (row[x] - f->fmin) / (f->fmax - f->fmin) +
                                     f->noise * noise_at(f->seed, (uint64_t)zy * nx + x);
                    l = v < l ? v : l;
                    h = v > h ? v : h;
                }
            }
#pragma omp critical
            {
                lo = l < lo ? l : lo;
                hi = h > hi ? h : hi;
            }
            free(tmp);
        }
        f->gmin = lo;
        f->gmax = hi;
    }
}

/* ---------------------------------------------------------------------------
 * zfp 0.5.5 fixed-rate float32 3-D block encoder
 * ------------------------------------------------------------------------ */

/* Sequency order (the zfp perm_3 table): entry i is x + 4y + 16z of the i-th
 * coefficient, as "xyz" digit triples. */
static const char PERM_XYZ[] =
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

static int perm_idx(int i)
{
    const char *t = PERM_XYZ + 3 * i;
    return (t[0] - '0') + 4 * (t[1] - '0') + 16 * (t[2] - '0');
}

typedef struct {
    uint64_t *words; /* caller-zeroed */
    uint64_t pos;    /* absolute bit position */
} bitwriter_t;

static void put_bit(bitwriter_t *w, uint32_t b)
{
    if (b)
        w->words[w->pos >> 6] |= 1ull << (w->pos & 63);
    w->pos++;
}

/* writes the low n bits of x LSB-first, returns x >> n */
static uint64_t put_bits(bitwriter_t *w, uint64_t x, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i) {
        put_bit(w, (uint32_t)(x & 1u));
        x >>= 1;
    }
    return x;
}

/* pad partial block of width n along stride s (zfp pad_block) */
static void pad_block(float *p, uint32_t n, uint32_t s)
{
    switch (n) {
    case 0:
        p[0 * s] = 0.0f; /* fall through */
    case 1:
        p[1 * s] = p[0 * s]; /* fall through */
    case 2:
        p[2 * s] = p[1 * s]; /* fall through */
    case 3:
        p[3 * s] = p[0 * s]; /* fall through */
    default:
        break;
    }
}

static inline uint32_t asr(uint32_t v, int s) { return (uint32_t)((int32_t)v >> s); }

/* forward decorrelating lift on 4 values at stride s (int32, wrapping) */
static void fwd_lift(uint32_t *p, int s)
{
    uint32_t x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
    x += w; x = asr(x, 1); w -= x;
    z += y; z = asr(z, 1); y -= z;
    x += z; x = asr(x, 1); z -= x;
    w += y; w = asr(w, 1); y -= w;
    w += asr(y, 1); y -= asr(w, 1);
    p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

static int exponent_of(float x)
{
    if (x > 0.0f) {
        int e;
        frexpf(x, &e);
        return e > 1 - 127 ? e : 1 - 127;
    }
    return -127;
}

/* Encode one 4^3 block at the writer's position using exactly maxbits bits. */
static void encode_block(bitwriter_t *w, const float *fb, uint32_t maxbits)
{
    const uint64_t start = w->pos;
    float amax = 0.0f;
    for (int i = 0; i < 64; ++i) {
        const float a = fabsf(fb[i]);
        if (amax < a)
            amax = a;
    }
    const int emax = exponent_of(amax);
    const uint32_t e = (uint32_t)(emax + 127);
    if (e) {
        put_bits(w, 2ull * e + 1, 9);
        uint32_t ib[64], ub[64];
        const float s = ldexpf(1.0f, 30 - emax);
        for (int i = 0; i < 64; ++i)
            ib[i] = (uint32_t)(int32_t)(s * fb[i]);
        for (int z = 0; z < 4; ++z)
            for (int y = 0; y < 4; ++y)
                fwd_lift(ib + 4 * y + 16 * z, 1);
        for (int x = 0; x < 4; ++x)
            for (int z = 0; z < 4; ++z)
                fwd_lift(ib + 16 * z + x, 4);
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x)
                fwd_lift(ib + x + 4 * y, 16);
        for (int i = 0; i < 64; ++i)
            ub[i] = (ib[perm_idx(i)] + 0xaaaaaaaau) ^ 0xaaaaaaaau;
        /* embedded bit-plane coder with group testing */
        uint32_t bits = maxbits - 9, n = 0;
        for (uint32_t k = 32; bits && k-- > 0;) {
            uint64_t x = 0;
            for (uint32_t i = 0; i < 64; ++i)
                x += (uint64_t)((ub[i] >> k) & 1u) << i;
            const uint32_t m = n < bits ? n : bits;
            bits -= m;
            x = put_bits(w, x, m);
            for (; n < 64 && bits && (bits--, put_bit(w, x != 0), x != 0); x >>= 1, n++)
                for (; n < 63 && bits && (bits--, put_bit(w, (uint32_t)(x & 1u)), !(x & 1u));
                     x >>= 1, n++)
                    ;
        }
    } else {
        put_bit(w, 0);
    }
    w->pos = start + maxbits; /* zero padding up to the fixed rate */
}

/* Gather block (bx,by,bz) of an nx*ny*nz float volume with zfp's partial-block
 * padding (pad along x per row, then y per column, then z). */
static void gather_block(const float *vol, const uint32_t d[3], uint32_t bx, uint32_t by,
                         uint32_t bz, float q[64])
{
    const uint32_t x0 = bx * 4, y0 = by * 4, z0 = bz * 4;
    const uint32_t nx = d[0] - x0 < 4 ? d[0] - x0 : 4, ny = d[1] - y0 < 4 ? d[1] - y0 : 4,
                   nz = d[2] - z0 < 4 ? d[2] - z0 : 4;
    if (nx == 4 && ny == 4 && nz == 4) {
        for (uint32_t z = 0; z < 4; ++z)
            for (uint32_t y = 0; y < 4; ++y)
                for (uint32_t x = 0; x < 4; ++x)
                    q[16 * z + 4 * y + x] =
                        vol[((size_t)(z0 + z) * d[1] + y0 + y) * d[0] + x0 + x];
        return;
    }
    for (uint32_t z = 0; z < nz; ++z) {
        for (uint32_t y = 0; y < ny; ++y) {
            for (uint32_t x = 0; x < nx; ++x)
                q[16 * z + 4 * y + x] = vol[((size_t)(z0 + z) * d[1] + y0 + y) * d[0] + x0 + x];
            pad_block(q + 16 * z + 4 * y, nx, 1);
        }
        for (uint32_t x = 0; x < 4; ++x)
            pad_block(q + 16 * z + x, ny, 4);
    }
    for (uint32_t y = 0; y < 4; ++y)
        for (uint32_t x = 0; x < 4; ++x)
            pad_block(q + 4 * y + x, nz, 16);
}

static uint64_t nblocks_of(const uint32_t d[3])
{
    return (uint64_t)((d[0] + 3) / 4) * ((d[1] + 3) / 4) * ((d[2] + 3) / 4);
}

/* Size in bytes of a fixed-rate stream for dims at maxbits (u64 words). */
EXPORT uint64_t bcmc_synth_stream_bytes(const uint32_t dims[3], uint32_t maxbits)
{
    return (nblocks_of(dims) * maxbits + 63) / 64 * 8;
}

/* Encode blocks [b0, b1) of vol into out (bit offset b*maxbits).  When maxbits
 * is a multiple of 64 every block owns whole words and blocks encode in
 * parallel; otherwise blocks share words and encode in order. */
static void encode_range(const float *vol, const uint32_t d[3], uint32_t maxbits, uint64_t *out,
                         uint64_t b0, uint64_t b1, uint64_t zoff, int threads)
{
    const uint32_t nbx = (d[0] + 3) / 4, nby = (d[1] + 3) / 4;
    if (maxbits % 64 == 0) {
#pragma omp parallel for schedule(static) num_threads(nthreads(threads))
        for (int64_t b = (int64_t)b0; b < (int64_t)b1; ++b) {
            float q[64];
            const uint64_t bl = (uint64_t)b - zoff;
            gather_block(vol, d, (uint32_t)(b % nbx), (uint32_t)((b / nbx) % nby),
                         (uint32_t)(bl / ((uint64_t)nbx * nby)), q);
            bitwriter_t w = {out, (uint64_t)b * maxbits};
            encode_block(&w, q, maxbits);
        }
    } else {
        for (uint64_t b = b0; b < b1; ++b) {
            float q[64];
            const uint64_t bl = b - zoff;
            gather_block(vol, d, (uint32_t)(b % nbx), (uint32_t)((b / nbx) % nby),
                         (uint32_t)(bl / ((uint64_t)nbx * nby)), q);
            bitwriter_t w = {out, b * maxbits};
            encode_block(&w, q, maxbits);
        }
    }
}

/* Encode a whole float32 volume (x fastest).  out must hold
 * bcmc_synth_stream_bytes(dims, maxbits) bytes; it is zeroed here. */
EXPORT int bcmc_synth_encode(const float *vol, const uint32_t dims[3], uint32_t maxbits,
                             uint8_t *out, int threads)
{
    if (maxbits < 9)
        return -1;
    const uint64_t bytes = bcmc_synth_stream_bytes(dims, maxbits);
    memset(out, 0, bytes);
    encode_range(vol, dims, maxbits, (uint64_t *)out, 0, nblocks_of(dims), 0, threads);
    return 0;
}

static int field_setup(field_t *f, int kind, uint64_t seed, const uint32_t dims[3], uint32_t count,
                       const double params[4], double noise)
{
    memset(f, 0, sizeof(*f));
    if (kind < 0 || kind > 2 || dims[0] == 0 || dims[1] == 0 || dims[2] == 0)
        return -1;
    f->kind = kind;
    f->seed = seed;
    memcpy(f->dims, dims, sizeof(f->dims));
    f->count = count;
    f->p0 = params[0];
    f->p1 = params[1];
    f->p2 = params[2];
    f->p3 = params[3];
    f->noise = kind == FIELD_BLOBS ? noise : 0.0;
    field_init(f);
    return 0;
}

/* Write the normalised float32 field (x fastest). */
EXPORT int bcmc_synth_field(int kind, uint64_t seed, const uint32_t dims[3], uint32_t count,
                            const double params[4], double noise, float *out, int threads)
{
    field_t f;
    if (field_setup(&f, kind, seed, dims, count, params, noise))
        return -1;
    const uint32_t nx = dims[0], ny = dims[1], nz = dims[2];
    double *store = malloc((size_t)nx * ny * nz * sizeof(double));
    field_minmax(&f, threads, store);
#pragma omp parallel num_threads(nthreads(threads))
    {
        double *tmp = malloc(nx * sizeof(double));
#pragma omp for schedule(static)
        for (int64_t zy = 0; zy < (int64_t)ny * nz; ++zy) {
            double *row = store ? store + (size_t)zy * nx : tmp;
            if (!store)
                field_row(&f, (uint32_t)(zy % ny), (uint32_t)(zy / ny), row);
            for (uint32_t x = 0; x < nx; ++x)
                out[(size_t)zy * nx + x] = (float)field_final(&f, row[x], (uint64_t)zy * nx + x);
        }
        free(tmp);
    }
    free(store);
    field_free(&f);
    return 0;
}

/* Generate and encode a field one 4-slice slab at a time.  The FP64 field is
 * kept in host memory for volumes up to 2^31 voxels (16 GiB) and recomputed
 * per pass beyond that.  Produces the same bytes as bcmc_synth_field followed
 * by bcmc_synth_encode. */
EXPORT int bcmc_synth_stream(int kind, uint64_t seed, const uint32_t dims[3], uint32_t count,
                             const double params[4], double noise, uint32_t maxbits, uint8_t *out,
                             int threads)
{
    const int store_ok = (uint64_t)dims[0] * dims[1] * dims[2] <= (1ull << 31);
    field_t f;
    if (maxbits < 9 || field_setup(&f, kind, seed, dims, count, params, noise))
        return -1;
    const uint32_t nx = dims[0], ny = dims[1], nz = dims[2];
    /* keep the FP64 field when it fits (saves recomputing it per pass) */
    double *store = store_ok ? malloc((size_t)nx * ny * nz * sizeof(double)) : NULL;
    field_minmax(&f, threads, store);
    const uint32_t nbx = (nx + 3) / 4, nby = (ny + 3) / 4, nbz = (nz + 3) / 4;
    const uint64_t bytes = bcmc_synth_stream_bytes(dims, maxbits);
    memset(out, 0, bytes);
    float *slab = malloc((size_t)nx * ny * 4 * sizeof(float));
    for (uint32_t bz = 0; bz < nbz; ++bz) {
        const uint32_t z0 = bz * 4, sz = nz - z0 < 4 ? nz - z0 : 4;
#pragma omp parallel num_threads(nthreads(threads))
        {
            double *tmp = malloc(nx * sizeof(double));
#pragma omp for schedule(static)
            for (int64_t zy = 0; zy < (int64_t)ny * sz; ++zy) {
                const uint32_t y = (uint32_t)(zy % ny), z = z0 + (uint32_t)(zy / ny);
                const uint64_t vi = ((uint64_t)z * ny + y) * nx;
                const double *row = store ? store + vi : tmp;
                if (!store)
                    field_row(&f, y, z, tmp);
                for (uint32_t x = 0; x < nx; ++x)
                    slab[(size_t)zy * nx + x] = (float)field_final(&f, row[x], vi + x);
            }
            free(tmp);
        }
        const uint32_t sd[3] = {nx, ny, sz};
        const uint64_t b0 = (uint64_t)bz * nbx * nby;
        encode_range(slab, sd, maxbits, (uint64_t *)out, b0, b0 + (uint64_t)nbx * nby, b0, threads);
    }
    free(slab);
    free(store);
    field_free(&f);
    return 0;
}
