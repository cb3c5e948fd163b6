"""Python host side of the MI355X block-compressed marching-cubes library.

Mirrors the reference's JS classes (Twinklebear/webgpu-bcmc js/*.js) over the
C ABI of lib/libbcmc_amd.so (include/bcmc.h) with ctypes:

  Device                   navigator.gpu device + queue      (js/bcmc.js:1-20)
  Buffer                   GPUBuffer
  ExclusiveScanPipeline    js/exclusive_scan.js:14-122
  ExclusiveScanner         js/exclusive_scan.js:124-295
  StreamCompact            js/stream_compact.js (alias StreamCompactor)
  RadixSorter              js/radix_sort_by_key.js
  LRUCache                 js/lru_cache.js
  CompressedMarchingCubes  js/compressed_marching_cubes.js

Every call goes to the HIP library; there is no CPU fallback.  Importing the
module without the built library raises, and creating a Device without a
gfx950 GPU raises BcmcError.
"""
import ctypes
import os
import re
import time

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_ROOT, "lib", "libbcmc_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_ROOT), "include", "bcmc.h")

VERTS_LOCAL_F32, VERTS_GLOBAL_F32, VERTS_PACKED_U32X2 = 0, 1, 2
ARRAY_RANGES, ARRAY_ACTIVE_FLAGS, ARRAY_ACTIVE_IDS, ARRAY_NEW_ITEMS, ARRAY_WITH_VERTS, \
    ARRAY_VERT_OFFSETS = range(6)
RANGES_BLOCK, RANGES_BLOCK_TILE = 0, 1
STATUS = {0: "OK", -1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "OVERFLOW", -5: "STATE",
          -6: "NO_DEVICE"}


class BcmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bcmc error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class _Perf(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in (
        "computeActiveBlocks", "cacheUpdate", "decompression", "compactActiveIDs",
        "computeBlockHasVertices", "compactBlocksWithVertices", "computeVertices", "total_ms")] + [
        ("numActiveBlocks", ctypes.c_uint32), ("numBlocksDecompressed", ctypes.c_uint32),
        ("numBlocksWithVertices", ctypes.c_uint32), ("cacheSize", ctypes.c_uint32),
        ("numVertices", ctypes.c_uint64)]


class _LruInfo(ctypes.Structure):
    _fields_ = [("cache_size", ctypes.c_uint32), ("total_elements", ctypes.c_uint32),
                ("element_size", ctypes.c_uint32), ("slots_available", ctypes.c_uint32),
                ("reserved_bytes", ctypes.c_uint64), ("logical_bytes", ctypes.c_uint64)]


class _McInfo(ctypes.Structure):
    _fields_ = [("total_blocks", ctypes.c_uint32), ("num_active_blocks", ctypes.c_uint32),
                ("num_blocks_with_vertices", ctypes.c_uint32), ("new_decompressed", ctypes.c_uint32),
                ("num_vertices", ctypes.c_uint64), ("vertex_capacity", ctypes.c_uint64),
                ("dims", ctypes.c_uint32 * 3), ("padded_dims", ctypes.c_uint32 * 3),
                ("max_bits", ctypes.c_uint32), ("scale", ctypes.c_float * 3),
                ("format", ctypes.c_int)]


class _QueryResult(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_uint64), ("vertex_offset", ctypes.c_uint64),
                ("num_active_blocks", ctypes.c_uint32), ("step_decoded_blocks", ctypes.c_uint32)]


class _EngineInfo(ctypes.Structure):
    _fields_ = [("num_volumes", ctypes.c_uint32), ("pool_slots", ctypes.c_uint32),
                ("pool_used", ctypes.c_uint32), ("epoch", ctypes.c_uint32),
                ("pool_bytes", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("steps", ctypes.c_uint32), ("graph_steps", ctypes.c_uint32)]


LRU_PERF_KEYS = ("lruMarkNewItems", "lruScanNeededSlots", "lruNumNewItems", "lruCompactNewItems",
                 "lruScanSlotsAvailable", "lruNumSlotsAvailable", "lruGrowCache",
                 "lruCompactAvailableSlots", "lruPrepKeyValue", "lruTotalSortTime", "lruWriteNewItems")


class _LruPerf(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in LRU_PERF_KEYS]


class _Tuning(ctypes.Structure):
    _fields_ = [("scan_wg_per_cu", ctypes.c_uint32), ("sort_wg_per_cu", ctypes.c_uint32),
                ("decode_wg_per_cu", ctypes.c_uint32), ("mc_wg_per_cu", ctypes.c_uint32),
                ("count_wg_per_cu", ctypes.c_uint32),
                ("scan_ms", ctypes.c_float), ("sort_ms", ctypes.c_float),
                ("decode_ms", ctypes.c_float), ("mc_ms", ctypes.c_float),
                ("scan_items", ctypes.c_uint32), ("sort_items", ctypes.c_uint32),
                ("decode_threads", ctypes.c_uint32)]


class _ServeOpts(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("vertex_budget_bytes", ctypes.c_uint64),
                ("no_coalesce", ctypes.c_uint32)]


class _ServeResult(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("start_s", ctypes.c_double),
                ("finish_s", ctypes.c_double), ("num_vertices", ctypes.c_uint64),
                ("batch_size", ctypes.c_uint32), ("step", ctypes.c_uint32)]


_lib = None


def exported_symbols():
    """Names declared BCMC_API in include/bcmc.h."""
    import re
    txt = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"BCMC_API\s+[\w\s\*]+?\b(bcmc_\w+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "bcmc_last_error": (ctypes.c_char_p, []),
        "bcmc_version": (ctypes.c_char_p, []),
        "bcmc_ctx_create": (i32, [i32, P(vp)]),
        "bcmc_ctx_destroy": (i32, [vp]),
        "bcmc_ctx_sync": (i32, [vp]),
        "bcmc_ctx_info": (i32, [vp, P(i32), ctypes.c_char_p, ctypes.c_size_t]),
        "bcmc_ctx_enable_timing": (i32, [vp, i32]),
        "bcmc_ctx_set_graph_replay": (i32, [vp, i32]),
        "bcmc_ctx_get_graph_replay": (i32, [vp, P(i32)]),
        "bcmc_ctx_last_kernel_ms": (ctypes.c_double, [vp]),
        "bcmc_ctx_get_tuning": (i32, [vp, P(_Tuning)]),
        "bcmc_ctx_set_tuning": (i32, [vp, P(_Tuning)]),
        "bcmc_ctx_autotune": (i32, [vp, P(_Tuning)]),
        "bcmc_buf_alloc": (i32, [vp, u64, P(vp)]),
        "bcmc_buf_wrap": (i32, [vp, vp, u64, P(vp)]),
        "bcmc_buf_free": (i32, [vp]),
        "bcmc_buf_upload": (i32, [vp, u64, vp, u64]),
        "bcmc_buf_download": (i32, [vp, u64, vp, u64]),
        "bcmc_buf_fill_u32": (i32, [vp, u64, u32, u64]),
        "bcmc_buf_copy": (i32, [vp, u64, vp, u64, u64]),
        "bcmc_buf_size": (u64, [vp]),
        "bcmc_buf_device_ptr": (vp, [vp]),
        "bcmc_scan_aligned_size": (u32, [u32]),
        "bcmc_scan_exclusive_u32": (i32, [vp, vp, u32, u32, P(u32)]),
        "bcmc_scan_exclusive_u32_noalloc": (i32, [vp, vp, u32, u32, P(u32)]),
        "bcmc_compact_ids": (i32, [vp, u32, vp, vp, vp]),
        "bcmc_compact_flags": (i32, [vp, u32, vp, vp, P(u32)]),
        "bcmc_sort_aligned_size": (u32, [u32]),
        "bcmc_sort_u32_by_key": (i32, [vp, vp, vp, u32, i32]),
        "bcmc_lru_create": (i32, [vp, u32, u32, u32, P(vp)]),
        "bcmc_lru_destroy": (i32, [vp]),
        "bcmc_lru_update": (i32, [vp, vp, P(u32), P(vp)]),
        "bcmc_lru_update_timed": (i32, [vp, vp, P(u32), P(vp), P(_LruPerf)]),
        "bcmc_lru_reset": (i32, [vp]),
        "bcmc_lru_get_info": (i32, [vp, P(_LruInfo)]),
        "bcmc_lru_read_state": (i32, [vp, vp, vp, vp, vp]),
        "bcmc_lru_cache_buffer": (vp, [vp]),
        "bcmc_mc_create": (i32, [vp, P(vp)]),
        "bcmc_mc_destroy": (i32, [vp]),
        "bcmc_mc_set_compressed_volume": (i32, [vp, vp, u64, ctypes.c_double, P(u32), P(ctypes.c_float), u32]),
        "bcmc_mc_set_compressed_volume_ranges": (i32, [vp, vp, u64, ctypes.c_double, P(u32), P(ctypes.c_float),
                                                       u32, vp, i32, u64]),
        "bcmc_mc_get_block_ranges": (i32, [vp, i32, vp, u64]),
        "bcmc_mc_set_vertex_format": (i32, [vp, i32]),
        "bcmc_mc_compute_surface": (i32, [vp, ctypes.c_float, P(_Perf), P(u64)]),
        "bcmc_mc_get_info": (i32, [vp, P(_McInfo)]),
        "bcmc_mc_vertex_buffer": (vp, [vp]),
        "bcmc_mc_lru": (vp, [vp]),
        "bcmc_mc_array_buffer": (vp, [vp, i32]),
        "bcmc_mc_memory_use": (i32, [vp, P(u64), P(u64)]),
        "bcmc_engine_create": (i32, [vp, u64, P(vp)]),
        "bcmc_engine_destroy": (i32, [vp]),
        "bcmc_engine_add_volume": (i32, [vp, vp, u64, ctypes.c_double, P(u32), P(u32)]),
        "bcmc_engine_add_volume_ranges": (i32, [vp, vp, u64, ctypes.c_double, P(u32), vp, i32, u64, P(u32)]),
        "bcmc_engine_set_vertex_format": (i32, [vp, i32]),
        "bcmc_engine_run": (i32, [vp, vp, vp, u32, vp]),
        "bcmc_engine_vertex_buffer": (vp, [vp]),
        "bcmc_engine_stats": (i32, [vp, P(_EngineInfo)]),
        "bcmc_engine_block_bound": (i32, [vp, u32, vp, u32, P(u64)]),
        "bcmc_serve_trace": (i32, [vp, vp, vp, vp, u32, P(_ServeOpts), vp]),
        "bcmc_zfp_decode_blocks": (i32, [vp, vp, u32, vp, u32, vp]),
        "bcmc_zfp_block_ranges": (i32, [vp, vp, u32, P(u32), vp]),
        "bcmc_zfp_encode": (i32, [vp, vp, P(u32), u32, vp]),
        "bcmc_vertices_world": (i32, [vp, vp, u64, u64, P(u32), P(ctypes.c_float), vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise BcmcError(rc, lib().bcmc_last_error().decode(errors="replace"))


def _source(data):
    """(pointer, bytes, keep-alive) of host bytes or a device Buffer."""
    if isinstance(data, Buffer):
        return lib().bcmc_buf_device_ptr(data.h), data.nbytes, data
    a = np.ascontiguousarray(np.frombuffer(memoryview(data), np.uint8))
    return a.ctypes.data, a.nbytes, a


def _ranges_source(ranges, dims):
    """(pointer, format, bytes, keep-alive) of persisted block ranges: 8 bytes
    a block (RANGES_BLOCK) or 16 (RANGES_BLOCK_TILE), host array or Buffer."""
    nb = int(np.prod([(int(d) + 3) // 4 for d in dims]))
    if not isinstance(ranges, Buffer):
        ranges = np.ascontiguousarray(ranges, np.float32)
    ptr, nbytes, keep = _source(ranges)
    if nbytes not in (8 * nb, 16 * nb):
        raise ValueError(f"block ranges of {nbytes} bytes for {nb} blocks: expected {8 * nb} or {16 * nb}")
    return ptr, RANGES_BLOCK if nbytes == 8 * nb else RANGES_BLOCK_TILE, nbytes, keep


def _u32arr(vals):
    return (ctypes.c_uint32 * 3)(*[int(v) for v in vals])


class Device:
    """One GPU and its in-order queue (a bcmc_ctx)."""

    def __init__(self, index=0):
        h = ctypes.c_void_p()
        _check(lib().bcmc_ctx_create(index, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().bcmc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sync(self):
        _check(lib().bcmc_ctx_sync(self.h))

    def info(self):
        n = ctypes.c_int()
        name = ctypes.create_string_buffer(128)
        _check(lib().bcmc_ctx_info(self.h, ctypes.byref(n), name, 128))
        return {"num_cus": n.value, "name": name.value.decode()}

    _TUNE_KEYS = ("scan", "sort", "decode", "mc", "count")
    _TILE_KEYS = ("scan_items", "sort_items", "decode_threads")

    @staticmethod
    def _tuning_dict(t):
        d = {k: int(getattr(t, f"{k}_wg_per_cu")) for k in Device._TUNE_KEYS}
        d["tiles"] = {k: int(getattr(t, k)) for k in Device._TILE_KEYS}
        d["ms"] = {k: float(getattr(t, f"{k}_ms")) for k in Device._TUNE_KEYS[:3]}
        return d

    def tuning(self):
        """Workgroups per CU of the scan / sort / decode / mc kernel families
        and the tile shapes (scan and sort elements per thread, decoder
        blocks per workgroup)."""
        t = _Tuning()
        _check(lib().bcmc_ctx_get_tuning(self.h, ctypes.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, **shapes):
        """Keys: scan, sort, decode, mc, count (workgroups per CU) and
        scan_items, sort_items (8, 16, 32), decode_threads (64, 128, 256)."""
        t = _Tuning()
        _check(lib().bcmc_ctx_get_tuning(self.h, ctypes.byref(t)))
        for k, v in shapes.items():
            if k in self._TUNE_KEYS:
                setattr(t, f"{k}_wg_per_cu", int(v))
            elif k in self._TILE_KEYS:
                setattr(t, k, int(v))
            else:
                raise ValueError(f"unknown tuning key {k!r}")
        _check(lib().bcmc_ctx_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self):
        """Times the launch-shape candidates on synthetic data and keeps the
        fastest per kernel family; returns the chosen shapes and their times."""
        t = _Tuning()
        _check(lib().bcmc_ctx_autotune(self.h, ctypes.byref(t)))
        return self._tuning_dict(t)

    def set_graph_replay(self, enable):
        """Launch the query chains of pipelines and engines on this device
        directly (False, the default) or replay them from captured HIP graphs
        (True); results are identical either way."""
        _check(lib().bcmc_ctx_set_graph_replay(self.h, 1 if enable else 0))

    @property
    def graph_replay(self):
        v = ctypes.c_int()
        _check(lib().bcmc_ctx_get_graph_replay(self.h, ctypes.byref(v)))
        return bool(v.value)

    def create_buffer(self, nbytes):
        return Buffer.alloc(self, nbytes)

    def buffer_from(self, array):
        a = np.ascontiguousarray(array)
        b = Buffer.alloc(self, max(a.nbytes, 4))
        b.upload(a)
        return b


class Buffer:
    """A device allocation (GPUBuffer)."""

    def __init__(self, device, handle, owned=True):
        self.device, self.h, self.owned = device, handle, owned

    @classmethod
    def alloc(cls, device, nbytes):
        h = ctypes.c_void_p()
        _check(lib().bcmc_buf_alloc(device.h, int(nbytes), ctypes.byref(h)))
        return cls(device, h)

    @property
    def nbytes(self):
        return int(lib().bcmc_buf_size(self.h))

    def upload(self, array, offset=0):
        a = np.ascontiguousarray(array)
        _check(lib().bcmc_buf_upload(self.h, offset, a.ctypes.data, a.nbytes))

    def download(self, dtype=np.uint8, count=None, offset=0):
        dt = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dt.itemsize
        out = np.empty(int(count), dt)
        _check(lib().bcmc_buf_download(self.h, offset, out.ctypes.data, out.nbytes))
        return out

    def copy_from(self, src, nbytes, offset=0, src_offset=0):
        """copyBufferToBuffer(src, src_offset, self, offset, nbytes)."""
        _check(lib().bcmc_buf_copy(self.h, offset, src.h, src_offset, nbytes))

    def fill_u32(self, value, count, offset=0):
        _check(lib().bcmc_buf_fill_u32(self.h, offset, value, count))

    def destroy(self):
        if self.owned and getattr(self, "h", None):
            lib().bcmc_buf_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _view(device, handle):
    return Buffer(device, ctypes.c_void_p(handle), owned=False) if handle else None


# ---------------------------------------------------------------------------
# Primitives
# ---------------------------------------------------------------------------
class ExclusiveScanPipeline:
    """js/exclusive_scan.js:14-122."""

    def __init__(self, device):
        self.device = device

    def getAlignedSize(self, size):
        return int(lib().bcmc_scan_aligned_size(size))

    def prepareInput(self, cpu_array):
        a = np.asarray(cpu_array, np.uint32)
        aligned = self.getAlignedSize(a.size)
        buf = Buffer.alloc(self.device, max(aligned, 1) * 4)
        padded = np.zeros(aligned, np.uint32)
        padded[:a.size] = a
        buf.upload(padded)
        return ExclusiveScanner(self, buf, aligned)

    def prepareGPUInput(self, gpu_buffer, aligned_size, data_size=None):
        if self.getAlignedSize(aligned_size) != aligned_size:
            raise BcmcError(-1, "GPU input must be aligned to getAlignedSize")
        return ExclusiveScanner(self, gpu_buffer, aligned_size)


class ExclusiveScanner:
    """js/exclusive_scan.js:124-295: scan(n) -> total."""

    def __init__(self, pipeline, gpu_buffer, aligned_size):
        self.pipeline, self.inputBuf, self.inputSize = pipeline, gpu_buffer, aligned_size

    def scan(self, data_size):
        total = ctypes.c_uint32()
        _check(lib().bcmc_scan_exclusive_u32(self.pipeline.device.h, self.inputBuf.h,
                                             self.inputSize, data_size, ctypes.byref(total)))
        return total.value


class StreamCompact:
    """js/stream_compact.js:1-165."""

    def __init__(self, device):
        self.device = device

    def compactActiveIDs(self, num_elements, is_active, offsets, output):
        _check(lib().bcmc_compact_ids(self.device.h, num_elements, is_active.h, offsets.h, output.h))

    def compactFlags(self, num_elements, flags, output):
        """Fused single-pass scan + compaction; returns the count."""
        n = ctypes.c_uint32()
        _check(lib().bcmc_compact_flags(self.device.h, num_elements, flags.h, output.h,
                                        ctypes.byref(n)))
        return n.value


StreamCompactor = StreamCompact


class RadixSorter:
    """js/radix_sort_by_key.js."""

    def __init__(self, device):
        self.device = device

    def getAlignedSize(self, size):
        return int(lib().bcmc_sort_aligned_size(size))

    def sort(self, keys, values, size, reverse=False):
        _check(lib().bcmc_sort_u32_by_key(self.device.h, keys.h, values.h, size, 1 if reverse else 0))


class LRUCache:
    """js/lru_cache.js: LRUCache(device, scan, compact, initialSize, elementSize, totalElements)."""

    def __init__(self, device, scan_pipeline, stream_compact, initial_size, element_size,
                 total_elements, _handle=None):
        self.device = device
        if _handle is not None:
            self.h, self._owned = _handle, False
        else:
            h = ctypes.c_void_p()
            _check(lib().bcmc_lru_create(device.h, initial_size, element_size, total_elements,
                                         ctypes.byref(h)))
            self.h, self._owned = h, True

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "h", None):
            lib().bcmc_lru_destroy(self.h)
            self.h = None

    def update(self, item_needed, perf_tracker=None):
        """LRUCache.update; with a perf_tracker dict, each of the reference's
        11 lru* keys gets one entry (stage device ms or count)."""
        n = ctypes.c_uint32()
        ids = ctypes.c_void_p()
        perf = _LruPerf() if perf_tracker is not None else None
        _check(lib().bcmc_lru_update_timed(self.h, item_needed.h, ctypes.byref(n), ctypes.byref(ids),
                                           ctypes.byref(perf) if perf is not None else None))
        if perf is not None:
            for k in LRU_PERF_KEYS:
                perf_tracker.setdefault(k, []).append(getattr(perf, k))
        return n.value, (Buffer(self.device, ids) if ids.value else None)

    def reset(self):
        _check(lib().bcmc_lru_reset(self.h))

    def info(self):
        i = _LruInfo()
        _check(lib().bcmc_lru_get_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _LruInfo._fields_}

    @property
    def cacheSize(self):
        return self.info()["cache_size"]

    def state(self):
        info = self.info()
        c, t = info["cache_size"], info["total_elements"]
        age = np.zeros(max(c, 1), np.uint32)
        avail = np.zeros(max(c, 1), np.uint32)
        item = np.zeros(max(c, 1), np.int32)
        slot = np.zeros(max(t, 1), np.int32)
        _check(lib().bcmc_lru_read_state(self.h, age.ctypes.data, avail.ctypes.data,
                                         item.ctypes.data, slot.ctypes.data))
        return age[:c], avail[:c], item[:c], slot[:t]

    @property
    def cache(self):
        return _view(self.device, lib().bcmc_lru_cache_buffer(self.h))


class CompressedMarchingCubes:
    """js/compressed_marching_cubes.js."""

    PERF_KEYS = ("computeActiveBlocks", "numActiveBlocks", "cacheUpdate", "numBlocksDecompressed",
                 "decompression", "compactActiveIDs", "computeBlockHasVertices",
                 "numBlocksWithVertices", "compactBlocksWithVertices", "numVertices",
                 "computeVertices")

    def __init__(self, device):
        self.device = device
        h = ctypes.c_void_p()
        _check(lib().bcmc_mc_create(device.h, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().bcmc_mc_destroy(self.h)
            self.h = None

    def setCompressedVolume(self, volume, compression_rate, volume_dims, volume_scale=(1, 1, 1),
                            initial_cache_slots=0, block_ranges=None):
        """volume: host bytes or a device Buffer.  block_ranges (optional):
        persisted ranges from blockRanges() -- float32, 2 (min, max) or 4
        (and the tile range) values per block, host array or Buffer -- in
        place of the full-volume decode."""
        ptr, nbytes, keep = _source(volume)
        scale = (ctypes.c_float * 3)(*[float(s) for s in volume_scale])
        if block_ranges is None:
            _check(lib().bcmc_mc_set_compressed_volume(self.h, ptr, nbytes, float(compression_rate),
                                                       _u32arr(volume_dims), scale, int(initial_cache_slots)))
        else:
            rp, fmt, rbytes, rkeep = _ranges_source(block_ranges, volume_dims)
            _check(lib().bcmc_mc_set_compressed_volume_ranges(self.h, ptr, nbytes, float(compression_rate),
                                                              _u32arr(volume_dims), scale,
                                                              int(initial_cache_slots), rp, fmt, rbytes))
        del keep

    def blockRanges(self, with_tile=True):
        """The loaded volume's block ranges for persisting: float32 (N, 2)
        (min, max) per block, followed by the N tile ranges when with_tile."""
        n = self.info()["total_blocks"]
        out = np.empty((2 * n if with_tile else n, 2), np.float32)
        _check(lib().bcmc_mc_get_block_ranges(self.h, RANGES_BLOCK_TILE if with_tile else RANGES_BLOCK,
                                              out.ctypes.data, out.nbytes))
        return out

    def setVertexFormat(self, fmt):
        _check(lib().bcmc_mc_set_vertex_format(self.h, fmt))

    def computeSurface(self, isovalue, perf_tracker=None, timed=None):
        """Returns the vertex count; fills perf_tracker (dict of lists) like the reference.

        timed (default: whether perf_tracker is given) records the per-stage
        device times in lastPerf; untimed queries run in the device's launch
        mode (direct launches, or the replayed graph of the query chain:
        Device.set_graph_replay) and lastPerf holds the counts only."""
        if timed is None:
            timed = perf_tracker is not None
        perf = _Perf()
        nv = ctypes.c_uint64()
        _check(lib().bcmc_mc_compute_surface(self.h, float(isovalue),
                                             ctypes.byref(perf) if timed else None, ctypes.byref(nv)))
        if timed:
            self.lastPerf = {f: getattr(perf, f) for f, _ in _Perf._fields_}
        else:
            i = self.info()
            self.lastPerf = {"numActiveBlocks": i["num_active_blocks"],
                             "numBlocksDecompressed": i["new_decompressed"],
                             "numBlocksWithVertices": i["num_blocks_with_vertices"],
                             "numVertices": i["num_vertices"]}
        if perf_tracker is not None:
            if "computeActiveBlocks" not in perf_tracker:
                perf_tracker["totalBlocks"] = self.info()["total_blocks"]
                for k in self.PERF_KEYS:
                    perf_tracker[k] = []
            for k in self.PERF_KEYS:
                perf_tracker[k].append(self.lastPerf[k])
        return int(nv.value)

    def info(self):
        i = _McInfo()
        _check(lib().bcmc_mc_get_info(self.h, ctypes.byref(i)))
        d = {f: getattr(i, f) for f, _ in _McInfo._fields_}
        for k in ("dims", "padded_dims", "scale"):
            d[k] = list(d[k])
        return d

    @property
    def numVertices(self):
        return self.info()["num_vertices"]

    @property
    def numActiveBlocks(self):
        return self.info()["num_active_blocks"]

    @property
    def numBlocksWithVertices(self):
        return self.info()["num_blocks_with_vertices"]

    @property
    def totalBlocks(self):
        return self.info()["total_blocks"]

    @property
    def newDecompressed(self):
        return self.info()["new_decompressed"]

    @property
    def vertexBuffer(self):
        return _view(self.device, lib().bcmc_mc_vertex_buffer(self.h))

    @property
    def lruCache(self):
        h = lib().bcmc_mc_lru(self.h)
        return LRUCache(self.device, None, None, 0, 0, 0, _handle=ctypes.c_void_p(h)) if h else None

    def array(self, which, dtype=np.uint32):
        b = _view(self.device, lib().bcmc_mc_array_buffer(self.h, which))
        return b.download(dtype) if b is not None and b.nbytes else np.zeros(0, dtype)

    def vertices(self):
        info = self.info()
        n, fmt = info["num_vertices"], info["format"]
        b = self.vertexBuffer
        if fmt == VERTS_LOCAL_F32:
            return b.download(np.float32, 4 * n).reshape(-1, 4)
        if fmt == VERTS_GLOBAL_F32:
            return b.download(np.float32, 3 * n).reshape(-1, 3)
        return b.download(np.uint32, 2 * n).reshape(-1, 2)

    def reportMemoryUse(self):
        t, c = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().bcmc_mc_memory_use(self.h, ctypes.byref(t), ctypes.byref(c)))
        return {"total": t.value, "cache": c.value}


def isovalue_benchmark_values(kind, value_range, random=None):
    """The 100 isovalues of the reference app's benchmarks
    (js/run_benchmark.js:1-47): 'random' draws U[lo, hi) from random()
    (default numpy's generator), 'sweepUp' / 'sweepDown' move by
    (hi - lo) / 100 per step from lo / hi, accumulated."""
    lo, hi = value_range
    if kind == "random":
        draw = random or np.random.default_rng().random
        return [draw() * (hi - lo) + lo for _ in range(100)]
    if kind not in ("sweepUp", "sweepDown"):
        raise ValueError(f"unknown benchmark {kind!r} (random, sweepUp or sweepDown)")
    out, cur, step = [], lo if kind == "sweepUp" else hi, (hi - lo) / 100
    for _ in range(100):
        cur = cur + step if kind == "sweepUp" else cur - step
        out.append(cur)
    return out


def run_isovalue_benchmark(mc, kind, value_range, dataset_name="volume", random=None, enable_cache=True):
    """One benchmark on a CompressedMarchingCubes as the reference app drives it
    (js/bcmc.js:319-370): LRU reset, then a computeSurface with a perfTracker
    whenever the isovalue changed (every step without the cache: reset
    first).  Returns (perf results, the file name the app saves them under)."""
    perf = {"isovalue": [], "totalTime": [], "memUse": []}
    mc.lruCache.reset()
    current = None
    for iso in isovalue_benchmark_values(kind, value_range, random):
        if not enable_cache:
            mc.lruCache.reset()
        if iso != current:
            current = iso
            t = time.perf_counter()
            mc.computeSurface(iso, perf)
            perf["totalTime"].append((time.perf_counter() - t) * 1e3)
            perf["isovalue"].append(iso)
            perf["memUse"].append(mc.reportMemoryUse())
    return perf, f"bcmc-explicit-{dataset_name}-{kind}.json"


def request_trace(rng, n, num_volumes, repeat=0.5, window=16):
    """Synthetic served-form requests: volume ids uniform over the volumes,
    isovalues U[0.1, 0.9] of which a fraction `repeat` re-use one of the last
    `window` isovalues of the same volume.  Returns (volume ids, isovalues)."""
    vols = rng.integers(0, num_volumes, n)
    hist = [[] for _ in range(num_volumes)]
    isos = np.empty(n, np.float32)
    for i, v in enumerate(vols):
        h = hist[v]
        if h and rng.random() < repeat:
            iso = h[int(rng.integers(0, len(h)))]
        else:
            iso = np.float32(rng.uniform(0.1, 0.9))
        isos[i] = iso
        h.append(iso)
        del h[:-window]
    return vols.astype(np.uint32), isos


class Engine:
    """Batched / served form: resident volumes + a shared decoded-block pool;
    run() answers many (volume, isovalue) queries in one step."""

    def __init__(self, device, pool_bytes):
        self.device = device
        h = ctypes.c_void_p()
        _check(lib().bcmc_engine_create(device.h, int(pool_bytes), ctypes.byref(h)))
        self.h = h
        self.fmt = VERTS_LOCAL_F32

    def __del__(self):
        if getattr(self, "h", None):
            lib().bcmc_engine_destroy(self.h)
            self.h = None

    def add_volume(self, stream, rate, dims, block_ranges=None):
        """stream: host bytes or a device Buffer; block_ranges as
        CompressedMarchingCubes.setCompressedVolume takes them."""
        ptr, nbytes, keep = _source(stream)
        vid = ctypes.c_uint32()
        if block_ranges is None:
            _check(lib().bcmc_engine_add_volume(self.h, ptr, nbytes, float(rate), _u32arr(dims),
                                                ctypes.byref(vid)))
        else:
            rp, fmt, rbytes, rkeep = _ranges_source(block_ranges, dims)
            _check(lib().bcmc_engine_add_volume_ranges(self.h, ptr, nbytes, float(rate), _u32arr(dims), rp,
                                                       fmt, rbytes, ctypes.byref(vid)))
        del keep
        return vid.value

    def set_vertex_format(self, fmt):
        _check(lib().bcmc_engine_set_vertex_format(self.h, fmt))
        self.fmt = fmt

    def run(self, volume_ids, isovalues):
        ids = np.ascontiguousarray(volume_ids, np.uint32)
        isos = np.ascontiguousarray(isovalues, np.float32)
        res = (_QueryResult * max(len(ids), 1))()
        _check(lib().bcmc_engine_run(self.h, ids.ctypes.data, isos.ctypes.data, len(ids), res))
        return [{f: getattr(r, f) for f, _ in _QueryResult._fields_} for r in res[:len(ids)]]

    def vertices(self, result, first=0, count=None):
        """Vertices first .. first + count (default: all) of a query's range
        in the last step's vertex buffer."""
        b = _view(self.device, lib().bcmc_engine_vertex_buffer(self.h))
        n = result["num_vertices"] - first if count is None else count
        off = result["vertex_offset"] + first
        if self.fmt == VERTS_LOCAL_F32:
            return b.download(np.float32, 4 * n, 16 * off).reshape(-1, 4)
        if self.fmt == VERTS_GLOBAL_F32:
            return b.download(np.float32, 3 * n, 12 * off).reshape(-1, 3)
        return b.download(np.uint32, 2 * n, 8 * off).reshape(-1, 2)

    def serve(self, arrival_s, volume_ids, isovalues, max_batch=0, vertex_budget_bytes=0, coalesce=True):
        """Replay a request trace in real time (continuous batching); returns a
        structured array with arrival/start/finish times, vertices, batch size.
        With coalesce, requests of one step with the same volume and isovalue
        are answered by one query."""
        arr = np.ascontiguousarray(arrival_s, np.float64)
        ids = np.ascontiguousarray(volume_ids, np.uint32)
        isos = np.ascontiguousarray(isovalues, np.float32)
        n = len(arr)
        out = (_ServeResult * max(n, 1))()
        opts = _ServeOpts(max_batch, vertex_budget_bytes, 0 if coalesce else 1)
        _check(lib().bcmc_serve_trace(self.h, arr.ctypes.data, ids.ctypes.data, isos.ctypes.data, n,
                                      ctypes.byref(opts), out))
        return {f: np.array([getattr(r, f) for r in out[:n]]) for f, _ in _ServeResult._fields_}

    def stats(self):
        i = _EngineInfo()
        _check(lib().bcmc_engine_stats(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _EngineInfo._fields_}

    def block_bound(self, volume_id, isovalues):
        """Pool slots one step of queries at these isovalues on the volume may
        need (the served loop's admission bound)."""
        isos = np.ascontiguousarray(isovalues, np.float32)
        out = ctypes.c_uint64()
        _check(lib().bcmc_engine_block_bound(self.h, volume_id, isos.ctypes.data, len(isos),
                                             ctypes.byref(out)))
        return int(out.value)


# ---------------------------------------------------------------------------
# Standalone decode helpers
# ---------------------------------------------------------------------------
def decode_blocks(device, stream_buf, max_bits, n, ids_buf=None):
    out = Buffer.alloc(device, max(n, 1) * 256)
    _check(lib().bcmc_zfp_decode_blocks(device.h, stream_buf.h, max_bits,
                                        ids_buf.h if ids_buf is not None else None, n, out.h))
    return out


def world_positions(device, verts_buf, n, dims, scale=(1, 1, 1), first=0):
    """World-space float xyz (n x 3, on the device) of vertices first..first+n
    of a block-local float vertex buffer, placed as the reference's renderer
    places them: scale * ((4 * block_pos + local) / dims - 0.5)."""
    out = Buffer.alloc(device, max(12 * n, 4))
    sc = (ctypes.c_float * 3)(*[float(v) for v in scale])
    _check(lib().bcmc_vertices_world(device.h, verts_buf.h, first, n, _u32arr(dims), sc, out.h))
    return out


def stream_bytes(dims, max_bits):
    """Bytes of a fixed-rate stream of a volume: whole 64-bit words."""
    nb = int(np.prod([(d + 3) // 4 for d in dims]))
    return (nb * max_bits + 63) // 64 * 8


def encode(device, field_buf, dims, max_bits):
    """Fixed-rate zfp encoding of a float32 device volume (x fastest) on the
    GPU; returns a device stream buffer of stream_bytes(dims, max_bits) bytes
    plus two zero words, so the decoders can read it in place."""
    out = Buffer.alloc(device, stream_bytes(dims, max_bits) + 16)
    _check(lib().bcmc_zfp_encode(device.h, field_buf.h, _u32arr(dims), max_bits, out.h))
    return out


_CRATE = re.compile(r"(\w+)_(\d+)x(\d+)x(\d+)_(\w+)\.*")
_CRATE_RATE = re.compile(r"\.crate(\d+)$")


def crate_info(filename):
    """Dims and compression rate of a zfp volume file from its name, e.g.
    'vertebra_512x512x512_uint16.raw.crate4' -> ((512, 512, 512), 4).  Dims
    follow the reference's name rule (js/volumes.js:147-152); the rate is the
    number after '.crate' (None when the name has none)."""
    base = os.path.basename(filename)
    m = _CRATE.search(base)
    if not m:
        raise ValueError(f"no <name>_<X>x<Y>x<Z>_<type> pattern in {base!r}")
    r = _CRATE_RATE.search(base)
    return (int(m.group(2)), int(m.group(3)), int(m.group(4))), (int(r.group(1)) if r else None)


def load_crate(filename, rate=None):
    """Read a headerless zfp fixed-rate stream file; returns (bytes, rate,
    dims) as setCompressedVolume takes them.  rate overrides the name's."""
    dims, name_rate = crate_info(filename)
    rate = rate if rate is not None else name_rate
    if rate is None:
        raise ValueError(f"{filename!r}: no '.crate<rate>' suffix and no rate given")
    with open(filename, "rb") as f:
        data = np.frombuffer(f.read(), np.uint8)
    nb = int(np.prod([(d + 3) // 4 for d in dims]))
    if data.size * 8 < nb * int(round(64 * rate)):
        raise ValueError(f"{filename!r}: {data.size} bytes are fewer than a {dims} volume "
                         f"at rate {rate} needs")
    return data, rate, dims


def block_ranges(device, stream_buf, max_bits, dims):
    nb = int(np.prod([(d + 3) // 4 for d in dims]))
    out = Buffer.alloc(device, nb * 8)
    _check(lib().bcmc_zfp_block_ranges(device.h, stream_buf.h, max_bits, _u32arr(dims), out.h))
    return out
