// Context, buffers, errors and the scan / compact / sort primitives of the
// C ABI (include/bcmc.h).
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

namespace bcmc {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_err; }

// Injected allocation failure for the failure-handling tests:
// BCMC_FAIL_ALLOC_AFTER=k makes the k-th device allocation (1-based) of the
// process fail with BCMC_ERR_NOMEM.
static int fail_after() {
    static int v = [] {
        const char *e = getenv("BCMC_FAIL_ALLOC_AFTER");
        return e ? atoi(e) : 0;
    }();
    return v;
}
static int g_alloc_count = 0;

int dev_alloc(void **p, size_t bytes) {
    *p = nullptr;
    ++g_alloc_count;
    if (fail_after() > 0 && g_alloc_count == fail_after()) {
        set_error("device allocation of %zu bytes failed (injected by BCMC_FAIL_ALLOC_AFTER)", bytes);
        return BCMC_ERR_NOMEM;
    }
    if (bytes == 0)
        bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return BCMC_ERR_NOMEM;
    }
    return BCMC_OK;
}

void dev_free(void *p) {
    if (p)
        (void)hipFree(p);
}

int make_buf(bcmc_ctx *ctx, uint64_t bytes, bcmc_buf **out) {
    bcmc_buf *b = new (std::nothrow) bcmc_buf;
    if (!b) {
        set_error("out of host memory");
        return BCMC_ERR_NOMEM;
    }
    b->ctx = ctx;
    b->bytes = bytes;
    int rc = dev_alloc(&b->ptr, bytes);
    if (rc != BCMC_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return BCMC_OK;
}

int scan_site_alloc(ScanSite *s, uint64_t max_elems) {
    // one status word per tile of the smallest shape
    s->max_tiles = std::max<uint32_t>(1, div_up(max_elems, kScanTileMin));
    BCMC_TRY(dev_alloc((void **)&s->ctrl, 16));
    BCMC_TRY(dev_alloc((void **)&s->status, (size_t)s->max_tiles * 8));
    BCMC_HIP_CHECK(hipMemset(s->ctrl, 0, 16));
    BCMC_HIP_CHECK(hipMemset(s->status, 0, (size_t)s->max_tiles * 8));
    // the memsets run on the null stream, which the library's non-blocking
    // streams do not wait for: finish them before any kernel reads the words
    BCMC_HIP_CHECK(hipDeviceSynchronize());
    return BCMC_OK;
}

void scan_site_free(ScanSite *s) {
    dev_free(s->ctrl);
    dev_free(s->status);
    s->ctrl = nullptr;
    s->status = nullptr;
}

// ---- primitive kernels -----------------------------------------------------
struct InplaceScanOp {
    uint32_t *data;
    uint32_t n;
    unsigned long long *total_out;
    BCMC_DEV uint32_t size() const { return n; }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[1], Item &) const { v[0] = data[i]; }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        data[i] = (uint32_t)e[0];
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { *total_out = t[0]; }
};

struct FlagCompactOp {
    const uint32_t *flags;
    uint32_t *out;
    uint32_t n;
    unsigned long long *total_out;
    BCMC_DEV uint32_t size() const { return n; }
    BCMC_DEV void begin() const {}
    using Tile = NoTile;
    using Item = NoItem;
    BCMC_DEV Tile tile(uint32_t) const { return {}; }
    BCMC_DEV void value(const Tile &, uint32_t i, uint32_t (&v)[1], Item &) const {
        v[0] = flags[i] != 0 ? 1u : 0u;
    }
    BCMC_DEV void emit(const Tile &, uint32_t i, const uint32_t (&v)[1], const Item &,
                       const unsigned long long (&e)[1]) const {
        if (v[0])
            out[e[0]] = i;
    }
    BCMC_DEV void total(const unsigned long long (&t)[1]) const { *total_out = t[0]; }
};

__global__ void write_total_kernel(uint32_t *data, uint32_t n, const unsigned long long *total) {
    data[n] = (uint32_t)*total;
}

// stream_compact.comp:26-32
// (writes whose offset falls outside `out` are dropped instead of faulting)
__global__ void compact_ids_kernel(uint32_t n, const uint32_t *flags, const uint32_t *offsets,
                                   uint32_t *out, uint64_t out_cap) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (flags[i] != 0 && offsets[i] < out_cap)
            out[offsets[i]] = i;
}

// 16-byte vector copy (n16 = bytes / 16), then a byte tail
__global__ void copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n16,
                            uint8_t *__restrict__ dtail, const uint8_t *__restrict__ stail,
                            uint32_t ntail) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < ntail)
        dtail[threadIdx.x] = stail[threadIdx.x];
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Scratch owned by the context for the primitive entry points.
struct CtxScratch {
    ScanSite site;
    uint64_t site_elems = 0;
    unsigned long long *total = nullptr;
    SortScratch sort;
    uint32_t sort_elems = 0;
};

static CtxScratch *scratch_of(bcmc_ctx *ctx);

// World-space positions of block-local vertices, as the reference's vertex
// shader places them (mc_isosurface.vert:51-55): scale * ((4 * block_pos +
// local) / dims - 0.5), each step a separately rounded float operation.
// The local positions are the exact float ones (the shader reads them back
// from the 10-bit packed form).
__global__ void world_positions_kernel(const float4 *__restrict__ v, uint64_t n, uint32_t nbx, uint32_t nby,
                                       float3 dims, float3 scale, float *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 p = v[i];
        const uint32_t bid = __float_as_uint(p.w);
        const uint32_t bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / nbx / nby;
        const float gx = (float)(4 * bx) + p.x, gy = (float)(4 * by) + p.y, gz = (float)(4 * bz) + p.z;
        out[3 * i] = scale.x * (gx / dims.x - 0.5f);
        out[3 * i + 1] = scale.y * (gy / dims.y - 0.5f);
        out[3 * i + 2] = scale.z * (gz / dims.z - 0.5f);
    }
}

}  // namespace bcmc

// ---- look-back state of the scan that allocates nothing ---------------------
// bcmc_scan_exclusive_u32_noalloc keeps its scan state here instead of in
// memory the context allocates: bytes 0..11 the tile counter, the
// finished-workgroup counter and the epoch (bcmc::ScanSite::ctrl), bytes
// 16..23 the scan total, and from byte 256 one status word per 8192-element
// tile.  Every device the library runs on holds its own copy for the life of
// the process.  In host code `scratch` names a host-side stand-in, never the
// device copy: host code uses the address hipGetSymbolAddress returns for the
// current device (noalloc_setup), kept per device.
constexpr uint32_t kNoAllocTiles = 131072;
constexpr int kNoAllocItems = 16;  // its tile shape, whatever the context's tuning
constexpr uint64_t kNoAllocMaxElems = (uint64_t)kNoAllocTiles * bcmc::kScanThreads * kNoAllocItems;  // 2^30
constexpr size_t kNoAllocTotalOff = 16, kNoAllocStatusOff = 256;
constexpr size_t kNoAllocBytes = kNoAllocStatusOff + (size_t)kNoAllocTiles * 8;  // 1 MiB + 256 B
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoAllocMaxDevices];

// Once per device, with that device current (bcmc_ctx_create has just made it
// so): look the device's copy of `scratch` up and zero it on the context's
// stream, so the first scan sees clear counters, epoch 0 and unpublished
// tiles.  Later scans leave the counters at zero themselves (the last
// workgroup resets them) and tell stale status words by their epoch.
static int noalloc_setup(bcmc_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
        return BCMC_OK;  // the no-allocation scan then reports BCMC_ERR_STATE
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[ctx->device])
        return BCMC_OK;
    void *p = nullptr;
    BCMC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BCMC_CHECK_ARG(p != nullptr && ((uintptr_t)p & 255) == 0, "scratch symbol address");
    const uint64_t words = kNoAllocBytes / 4;
    bcmc::fill_u32_kernel<<<(uint32_t)bcmc::div_up(words, 256), 256, 0, ctx->stream>>>(
        (uint32_t *)p, 0u, words);
    BCMC_HIP_CHECK(hipGetLastError());
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    g_noalloc_ptr[ctx->device] = p;
    return BCMC_OK;
}

static void *noalloc_ptr(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

using namespace bcmc;

// Context-owned scratch, kept in a side table so bcmc_ctx stays a plain struct.
static std::mutex g_scratch_mu;
static std::map<bcmc_ctx *, CtxScratch *> g_scratch;

CtxScratch *bcmc::scratch_of(bcmc_ctx *ctx) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(ctx);
    if (it != g_scratch.end())
        return it->second;
    CtxScratch *s = new CtxScratch;
    g_scratch[ctx] = s;
    return s;
}

static int ensure_site(bcmc_ctx *ctx, CtxScratch *s, uint64_t elems) {
    if (!s->total)
        BCMC_TRY(dev_alloc((void **)&s->total, 8));
    if (s->site.ctrl && s->site_elems >= elems)
        return BCMC_OK;
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    scan_site_free(&s->site);
    BCMC_TRY(scan_site_alloc(&s->site, std::max<uint64_t>(elems, kScanTileMax)));
    s->site_elems = std::max<uint64_t>(elems, kScanTileMax);
    return BCMC_OK;
}

extern "C" {

BCMC_API const char *bcmc_last_error(void) { return bcmc::last_error(); }

BCMC_API const char *bcmc_version(void) { return "bcmc-mi355x 0.1 (gfx950)"; }

BCMC_API int bcmc_ctx_create(int device, bcmc_ctx **out) {
    BCMC_CHECK_ARG(out, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        set_error("no HIP device available (%s)", hipGetErrorString(e));
        return BCMC_ERR_NO_DEVICE;
    }
    BCMC_CHECK_ARG(device >= 0 && device < count, "device index out of range");
    hipDeviceProp_t prop;
    BCMC_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 only", device,
                  prop.gcnArchName);
        return BCMC_ERR_NO_DEVICE;
    }
    BCMC_HIP_CHECK(hipSetDevice(device));
    bcmc_ctx *c = new bcmc_ctx;
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        set_error("HIP stream/event/pinned allocation failed");
        delete c;
        return BCMC_ERR_HIP;
    }
    c->pinned_bytes = 4096;
    const int rc = noalloc_setup(c);
    if (rc != BCMC_OK) {
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_destroy(bcmc_ctx *ctx) {
    if (!ctx)
        return BCMC_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch.find(ctx);
        if (it != g_scratch.end()) {
            scan_site_free(&it->second->site);
            dev_free(it->second->total);
            sort_scratch_free(&it->second->sort);
            delete it->second;
            g_scratch.erase(it);
        }
    }
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipHostFree(ctx->pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_sync(bcmc_ctx *ctx) {
    BCMC_CHECK_ARG(ctx, "ctx is null");
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_info(bcmc_ctx *ctx, int *num_cus, char *name, size_t name_len) {
    BCMC_CHECK_ARG(ctx, "ctx is null");
    if (num_cus)
        *num_cus = ctx->num_cus;
    if (name && name_len)
        snprintf(name, name_len, "%s", ctx->name);
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_enable_timing(bcmc_ctx *ctx, int enable) {
    BCMC_CHECK_ARG(ctx, "ctx is null");
    ctx->timing = enable != 0;
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_set_graph_replay(bcmc_ctx *ctx, int enable) {
    BCMC_CHECK_ARG(ctx, "null context");
    ctx->graph_replay = enable != 0;
    return BCMC_OK;
}

BCMC_API int bcmc_ctx_get_graph_replay(bcmc_ctx *ctx, int *enabled) {
    BCMC_CHECK_ARG(ctx && enabled, "null argument");
    *enabled = ctx->graph_replay ? 1 : 0;
    return BCMC_OK;
}

BCMC_API double bcmc_ctx_last_kernel_ms(bcmc_ctx *ctx) { return ctx ? ctx->last_kernel_ms : 0.0; }

BCMC_API int bcmc_buf_alloc(bcmc_ctx *ctx, uint64_t bytes, bcmc_buf **out) {
    BCMC_CHECK_ARG(ctx && out, "null argument");
    return make_buf(ctx, bytes, out);
}

BCMC_API int bcmc_buf_wrap(bcmc_ctx *ctx, void *device_ptr, uint64_t bytes, bcmc_buf **out) {
    BCMC_CHECK_ARG(ctx && out && device_ptr, "null argument");
    bcmc_buf *b = new bcmc_buf;
    b->ctx = ctx;
    b->ptr = device_ptr;
    b->bytes = bytes;
    b->owned = false;
    *out = b;
    return BCMC_OK;
}

BCMC_API int bcmc_buf_free(bcmc_buf *buf) {
    if (!buf)
        return BCMC_OK;
    if (buf->owned && buf->ptr) {
        (void)hipStreamSynchronize(buf->ctx->stream);
        dev_free(buf->ptr);
    }
    delete buf;
    return BCMC_OK;
}

BCMC_API int bcmc_buf_upload(bcmc_buf *buf, uint64_t offset, const void *src, uint64_t bytes) {
    BCMC_CHECK_ARG(buf && (src || bytes == 0), "null argument");
    BCMC_CHECK_ARG(offset <= buf->bytes && bytes <= buf->bytes - offset, "upload out of range");
    if (!bytes)
        return BCMC_OK;
    BCMC_HIP_CHECK(hipMemcpyAsync((char *)buf->ptr + offset, src, bytes, hipMemcpyHostToDevice,
                                  buf->ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(buf->ctx->stream));
    return BCMC_OK;
}

BCMC_API int bcmc_buf_download(bcmc_buf *buf, uint64_t offset, void *dst, uint64_t bytes) {
    BCMC_CHECK_ARG(buf && (dst || bytes == 0), "null argument");
    BCMC_CHECK_ARG(offset <= buf->bytes && bytes <= buf->bytes - offset, "download out of range");
    if (!bytes)
        return BCMC_OK;
    BCMC_HIP_CHECK(hipMemcpyAsync(dst, (const char *)buf->ptr + offset, bytes,
                                  hipMemcpyDeviceToHost, buf->ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(buf->ctx->stream));
    return BCMC_OK;
}

BCMC_API int bcmc_buf_fill_u32(bcmc_buf *buf, uint64_t offset, uint32_t value, uint64_t count) {
    BCMC_CHECK_ARG(buf, "null buffer");
    BCMC_CHECK_ARG(offset % 4 == 0 && offset <= buf->bytes && count * 4 <= buf->bytes - offset,
                   "fill out of range");
    if (!count)
        return BCMC_OK;
    const uint32_t grid = std::min<uint64_t>(div_up(count, 256), 4096);
    fill_u32_kernel<<<grid, 256, 0, buf->ctx->stream>>>((uint32_t *)((char *)buf->ptr + offset),
                                                         value, count);
    BCMC_HIP_CHECK(hipGetLastError());
    return BCMC_OK;
}

BCMC_API int bcmc_buf_copy(bcmc_buf *dst, uint64_t dst_offset, bcmc_buf *src, uint64_t src_offset,
                           uint64_t bytes) {
    BCMC_CHECK_ARG(dst && src, "null buffer");
    BCMC_CHECK_ARG(dst_offset <= dst->bytes && bytes <= dst->bytes - dst_offset &&
                       src_offset <= src->bytes && bytes <= src->bytes - src_offset,
                   "copy out of range");
    BCMC_CHECK_ARG(dst_offset % 16 == 0 && src_offset % 16 == 0, "offsets must be 16-byte aligned");
    if (!bytes)
        return BCMC_OK;
    bcmc_ctx *ctx = dst->ctx;
    const uint64_t n16 = bytes / 16;
    const uint32_t ntail = (uint32_t)(bytes % 16);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(n16, 256),
                                                                             persistent_grid(ctx, 16)));
    char *d = (char *)dst->ptr + dst_offset;
    const char *sp = (const char *)src->ptr + src_offset;
    if (ctx->timing)
        BCMC_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
    copy_kernel<<<grid, 256, 0, ctx->stream>>>((uint4 *)d, (const uint4 *)sp, n16,
                                               (uint8_t *)d + n16 * 16,
                                               (const uint8_t *)sp + n16 * 16, ntail);
    BCMC_HIP_CHECK(hipGetLastError());
    if (ctx->timing) {
        BCMC_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
        BCMC_HIP_CHECK(hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        BCMC_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_kernel_ms = ms;
    }
    return BCMC_OK;
}

BCMC_API uint64_t bcmc_buf_size(const bcmc_buf *buf) { return buf ? buf->bytes : 0; }
BCMC_API void *bcmc_buf_device_ptr(const bcmc_buf *buf) { return buf ? buf->ptr : nullptr; }

BCMC_API uint32_t bcmc_scan_aligned_size(uint32_t n) { return (n + 511) / 512 * 512; }

BCMC_API int bcmc_scan_exclusive_u32(bcmc_ctx *ctx, bcmc_buf *data, uint32_t aligned_n, uint32_t n,
                                     uint32_t *total) {
    BCMC_CHECK_ARG(ctx && data, "null argument");
    BCMC_CHECK_ARG(bcmc_scan_aligned_size(aligned_n) == aligned_n,
                   "GPU input must be aligned to getAlignedSize");
    BCMC_CHECK_ARG(n <= aligned_n && (uint64_t)aligned_n * 4 <= data->bytes,
                   "scan size exceeds the buffer");
    CtxScratch *s = scratch_of(ctx);
    BCMC_TRY(ensure_site(ctx, s, n));
    InplaceScanOp op{(uint32_t *)data->ptr, n, s->total};
    launch_scan<1>(ctx, s->site, s->site, op, n);
    if (n < aligned_n)
        write_total_kernel<<<1, 1, 0, ctx->stream>>>((uint32_t *)data->ptr, n, s->total);
    BCMC_HIP_CHECK(hipGetLastError());
    unsigned long long t = 0;
    BCMC_HIP_CHECK(hipMemcpyAsync(&t, s->total, 8, hipMemcpyDeviceToHost, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (total)
        *total = (uint32_t)t;
    return BCMC_OK;
}

BCMC_API int bcmc_scan_exclusive_u32_noalloc(bcmc_ctx *ctx, bcmc_buf *data, uint32_t aligned_n,
                                             uint32_t n, uint32_t *total) {
    BCMC_CHECK_ARG(ctx && data, "null argument");
    BCMC_CHECK_ARG(bcmc_scan_aligned_size(aligned_n) == aligned_n,
                   "GPU input must be aligned to getAlignedSize");
    if ((uint64_t)n > kNoAllocMaxElems) {
        set_error("invalid argument: n = %u exceeds the no-allocation scan's maximum of %llu "
                  "elements", n, (unsigned long long)kNoAllocMaxElems);
        return BCMC_ERR_INVALID;
    }
    BCMC_CHECK_ARG(n <= aligned_n && (uint64_t)aligned_n * 4 <= data->bytes,
                   "scan size exceeds the buffer");
    void *p = noalloc_ptr(ctx->device);
    if (!p) {
        set_error("no scan scratch on device %d", ctx->device);
        return BCMC_ERR_STATE;
    }
    ScanSite site;
    site.ctrl = (uint32_t *)p;
    site.status = (unsigned long long *)((char *)p + kNoAllocStatusOff);
    site.max_tiles = kNoAllocTiles;
    unsigned long long *tot = (unsigned long long *)((char *)p + kNoAllocTotalOff);
    InplaceScanOp op{(uint32_t *)data->ptr, n, tot};
    launch_scan_items<1, kNoAllocItems>(ctx, site, site, op, n);
    if (n < aligned_n)
        write_total_kernel<<<1, 1, 0, ctx->stream>>>((uint32_t *)data->ptr, n, tot);
    BCMC_HIP_CHECK(hipGetLastError());
    // the total comes back through the context's pinned staging words
    BCMC_HIP_CHECK(hipMemcpyAsync(ctx->pinned, tot, 8, hipMemcpyDeviceToHost, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (total)
        *total = (uint32_t)*(const unsigned long long *)ctx->pinned;
    return BCMC_OK;
}

BCMC_API int bcmc_compact_ids(bcmc_ctx *ctx, uint32_t n, bcmc_buf *flags, bcmc_buf *offsets,
                              bcmc_buf *out) {
    BCMC_CHECK_ARG(ctx && flags && offsets && out, "null argument");
    BCMC_CHECK_ARG((uint64_t)n * 4 <= flags->bytes && (uint64_t)n * 4 <= offsets->bytes,
                   "n exceeds the input buffers");
    if (!n)
        return BCMC_OK;
    const uint32_t grid = std::min<uint32_t>(div_up(n, 256), 8192);
    compact_ids_kernel<<<grid, 256, 0, ctx->stream>>>(n, (const uint32_t *)flags->ptr,
                                                      (const uint32_t *)offsets->ptr,
                                                      (uint32_t *)out->ptr, out->bytes / 4);
    BCMC_HIP_CHECK(hipGetLastError());
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return BCMC_OK;
}

BCMC_API int bcmc_compact_flags(bcmc_ctx *ctx, uint32_t n, bcmc_buf *flags, bcmc_buf *out,
                                uint32_t *count) {
    BCMC_CHECK_ARG(ctx && flags && out, "null argument");
    BCMC_CHECK_ARG((uint64_t)n * 4 <= flags->bytes && (uint64_t)n * 4 <= out->bytes,
                   "n exceeds the buffers");
    CtxScratch *s = scratch_of(ctx);
    BCMC_TRY(ensure_site(ctx, s, n));
    FlagCompactOp op{(const uint32_t *)flags->ptr, (uint32_t *)out->ptr, n, s->total};
    launch_scan<1>(ctx, s->site, s->site, op, n);
    BCMC_HIP_CHECK(hipGetLastError());
    unsigned long long t = 0;
    BCMC_HIP_CHECK(hipMemcpyAsync(&t, s->total, 8, hipMemcpyDeviceToHost, ctx->stream));
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (count)
        *count = (uint32_t)t;
    return BCMC_OK;
}

BCMC_API uint32_t bcmc_sort_aligned_size(uint32_t n) {
    uint32_t chunks = 1;
    while (chunks < (n + 63) / 64)
        chunks *= 2;
    return chunks * 64;
}

BCMC_API int bcmc_sort_u32_by_key(bcmc_ctx *ctx, bcmc_buf *keys, bcmc_buf *vals, uint32_t n,
                                  int reverse) {
    BCMC_CHECK_ARG(ctx && keys && vals, "null argument");
    const uint32_t aligned = bcmc_sort_aligned_size(n);
    BCMC_CHECK_ARG((uint64_t)aligned * 4 <= keys->bytes && (uint64_t)aligned * 4 <= vals->bytes,
                   "keys/values must hold getAlignedSize(n) entries");
    if (n == 0)
        return BCMC_OK;
    CtxScratch *s = scratch_of(ctx);
    if (s->sort_elems < aligned) {
        BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        sort_scratch_free(&s->sort);
        BCMC_TRY(sort_scratch_alloc(&s->sort, aligned));
        s->sort_elems = aligned;
    }
    launch_sort_by_key(ctx, &s->sort, (uint32_t *)keys->ptr, (uint32_t *)vals->ptr, n, reverse);
    // values: the kept slice; keys: the sorted keys (unspecified by the reference)
    BCMC_HIP_CHECK(hipMemcpyAsync(vals->ptr, s->sort.vals[1], (size_t)n * 4,
                                  hipMemcpyDeviceToDevice, ctx->stream));
    BCMC_HIP_CHECK(hipMemcpyAsync(keys->ptr, s->sort.keys[1], (size_t)n * 4,
                                  hipMemcpyDeviceToDevice, ctx->stream));
    BCMC_HIP_CHECK(hipGetLastError());
    BCMC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return BCMC_OK;
}

BCMC_API int bcmc_vertices_world(bcmc_ctx *ctx, bcmc_buf *verts, uint64_t first, uint64_t n,
                                 const uint32_t dims[3], const float scale[3], bcmc_buf *out) {
    BCMC_CHECK_ARG(ctx && verts && dims && scale && out, "null argument");
    BCMC_CHECK_ARG(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    BCMC_CHECK_ARG(first <= verts->bytes / 16 && n <= verts->bytes / 16 - first,
                   "vertex range past the end of the buffer");
    BCMC_CHECK_ARG(out->bytes / 12 >= n, "output buffer smaller than n float xyz");
    if (n == 0)
        return BCMC_OK;
    const uint32_t nbx = (dims[0] + 3) / 4, nby = (dims[1] + 3) / 4;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(n, 256), persistent_grid(ctx, 8)));
    world_positions_kernel<<<grid, 256, 0, ctx->stream>>>(
        reinterpret_cast<const float4 *>(verts->ptr) + first, n, nbx, nby,
        make_float3((float)dims[0], (float)dims[1], (float)dims[2]), make_float3(scale[0], scale[1], scale[2]),
        (float *)out->ptr);
    BCMC_HIP_CHECK(hipGetLastError());
    return BCMC_OK;
}

}  // extern "C"
