"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg as the checker of the HIP path; the product never
imports it.  The C sources cite the reference file:line each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    c = ctypes
    sig = {
        "orc_decode_block": (None, [u64p, c.c_size_t, c.c_uint32, c.c_uint64, f32p]),
        "orc_decode_volume": (None, [u64p, c.c_size_t, c.c_uint32, u32p, f32p, c.c_int]),
        "orc_block_ranges": (None, [u64p, c.c_size_t, c.c_uint32, u32p, f32p, c.c_int]),
        "orc_block_active": (None, [f32p, u32p, c.c_float, u32p, c.c_int]),
        "orc_exclusive_scan": (c.c_uint32, [u32p, c.c_uint32]),
        "orc_compact": (None, [u32p, u32p, c.c_uint32, u32p]),
        "orc_sort_aligned_size": (c.c_uint32, [c.c_uint32]),
        "orc_sort_by_key": (None, [u32p, u32p, c.c_uint32, c.c_int]),
        "orc_lru_create": (c.c_void_p, [c.c_uint32, c.c_uint32, c.c_uint32]),
        "orc_lru_destroy": (None, [c.c_void_p]),
        "orc_lru_update": (c.c_uint32, [c.c_void_p, u32p, u32p]),
        "orc_lru_reset": (None, [c.c_void_p]),
        "orc_lru_size": (c.c_uint32, [c.c_void_p]),
        "orc_lru_state": (None, [c.c_void_p, u32p, u32p, i32p, i32p]),
        "orc_mc_create": (c.c_void_p, [u8p, c.c_size_t, c.c_uint32, u32p, c.c_uint32, c.c_int]),
        "orc_mc_destroy": (None, [c.c_void_p]),
        "orc_mc_compute_surface": (c.c_int, [c.c_void_p, c.c_float, c.c_void_p]),
        "orc_mc_flags": (c.c_void_p, [c.c_void_p]),
        "orc_mc_active_ids": (c.c_void_p, [c.c_void_p]),
        "orc_mc_new_items": (c.c_void_p, [c.c_void_p]),
        "orc_mc_with_verts": (c.c_void_p, [c.c_void_p]),
        "orc_mc_vert_offsets": (c.c_void_p, [c.c_void_p]),
        "orc_mc_verts_f32": (c.c_void_p, [c.c_void_p]),
        "orc_mc_verts_f64": (c.c_void_p, [c.c_void_p]),
        "orc_mc_ranges": (c.c_void_p, [c.c_void_p]),
        "orc_mc_lru": (c.c_void_p, [c.c_void_p]),
        "orc_mc_total_blocks": (c.c_uint32, [c.c_void_p]),
        "orc_mc_bruteforce_count": (c.c_uint64, [f32p, u32p, c.c_float]),
        "orc_tri_count": (c.c_int, [c.c_int]),
        "orc_tri_edge": (c.c_int, [c.c_int, c.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _words(stream):
    b = np.frombuffer(bytes(stream), dtype=np.uint8)
    pad = (-len(b)) % 8
    if pad:
        b = np.concatenate([b, np.zeros(pad, np.uint8)])
    return np.ascontiguousarray(b.view(np.uint64))


def nblocks(dims):
    return int(np.prod([(d + 3) // 4 for d in dims]))


def padded(dims):
    return [(d + 3) // 4 * 4 for d in dims]


def decode_block(stream, maxbits, block):
    w = _words(stream)
    out = np.zeros(64, np.float32)
    lib().orc_decode_block(w, w.size, maxbits, block, out)
    return out


def decode_volume(stream, maxbits, dims, threads=8):
    """Decoded padded volume as a (pz, py, px) float32 array."""
    w = _words(stream)
    p = padded(dims)
    out = np.zeros(p[0] * p[1] * p[2], np.float32)
    lib().orc_decode_volume(w, w.size, maxbits, np.asarray(dims, np.uint32), out, threads)
    return out.reshape(p[2], p[1], p[0])


def block_ranges(stream, maxbits, dims, threads=8):
    w = _words(stream)
    out = np.zeros(2 * nblocks(dims), np.float32)
    lib().orc_block_ranges(w, w.size, maxbits, np.asarray(dims, np.uint32), out, threads)
    return out.reshape(-1, 2)


def block_active(ranges, dims, iso, threads=8):
    r = np.ascontiguousarray(ranges, np.float32).reshape(-1)
    out = np.zeros(r.size // 2, np.uint32)
    lib().orc_block_active(r, np.asarray(dims, np.uint32), np.float32(iso), out, threads)
    return out


def exclusive_scan(data):
    d = np.array(data, np.uint32)
    total = lib().orc_exclusive_scan(d, d.size)
    return d, total


def compact(flags, offsets):
    f = np.ascontiguousarray(flags, np.uint32)
    o = np.ascontiguousarray(offsets, np.uint32)
    n = int(np.count_nonzero(f))
    out = np.zeros(max(int(o.max()) + 1 if n else 0, 1), np.uint32)
    lib().orc_compact(f, o, f.size, out)
    return out[:n]


def sort_aligned_size(n):
    return lib().orc_sort_aligned_size(n)


def sort_by_key(keys, vals, reverse):
    """Returns (vals[:n] as the reference leaves them, ascending keys)."""
    n = len(keys)
    a = sort_aligned_size(n)
    k = np.full(a, 0xFFFFFFFF, np.uint32)
    v = np.full(a, 0xFFFFFFFF, np.uint32)
    k[:n] = keys
    v[:n] = vals
    lib().orc_sort_by_key(k, v, n, 1 if reverse else 0)
    return v[:n].copy(), k


class LRU:
    """LRUCache (js/lru_cache.js) restated."""

    def __init__(self, initial_size, elem_bytes, total):
        self.total = total
        self.h = lib().orc_lru_create(initial_size, elem_bytes, total)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_lru_destroy(self.h)
            self.h = None

    def update(self, needed):
        nd = np.ascontiguousarray(needed, np.uint32)
        out = np.zeros(max(self.total, 1), np.uint32)
        n = lib().orc_lru_update(self.h, nd, out)
        return n, out[:n].copy()

    def reset(self):
        lib().orc_lru_reset(self.h)

    @property
    def size(self):
        return lib().orc_lru_size(self.h)

    def state(self):
        c = self.size
        age = np.zeros(max(c, 1), np.uint32)
        avail = np.zeros(max(c, 1), np.uint32)
        item = np.zeros(max(c, 1), np.int32)
        slot = np.zeros(max(self.total, 1), np.int32)
        lib().orc_lru_state(self.h, age, avail, item, slot)
        return age[:c], avail[:c], item[:c], slot[: self.total]


class _Stats(ctypes.Structure):
    _fields_ = [
        ("num_active", ctypes.c_uint32),
        ("num_decoded", ctypes.c_uint32),
        ("num_with_verts", ctypes.c_uint32),
        ("num_verts", ctypes.c_uint64),
        ("cache_size", ctypes.c_uint32),
        ("t_active", ctypes.c_double),
        ("t_lru", ctypes.c_double),
        ("t_decode", ctypes.c_double),
        ("t_count", ctypes.c_double),
        ("t_emit", ctypes.c_double),
    ]


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    ct = np.ctypeslib.as_ctypes_type(dtype)
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(n,)).copy()


class MC:
    """CompressedMarchingCubes (js/compressed_marching_cubes.js) restated."""

    def __init__(self, stream, maxbits, dims, initial_cache=None, threads=8):
        self.dims = list(dims)
        self.n = nblocks(dims)
        if initial_cache is None:
            initial_cache = int(np.ceil(self.n * 0.05))
        buf = np.frombuffer(bytes(stream), np.uint8).copy()
        self.h = lib().orc_mc_create(buf, buf.size, maxbits, np.asarray(dims, np.uint32),
                                     initial_cache, threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mc_destroy(self.h)
            self.h = None

    def compute_surface(self, iso):
        st = _Stats()
        lib().orc_mc_compute_surface(self.h, np.float32(iso), ctypes.byref(st))
        self.stats = {f: getattr(st, f) for f, _ in _Stats._fields_}
        return int(st.num_verts)

    def results(self):
        L, s = lib(), self.stats
        a, d, h, v = s["num_active"], s["num_decoded"], s["num_with_verts"], s["num_verts"]
        return {
            "flags": _arr(L.orc_mc_flags(self.h), self.n, np.uint32),
            "active_ids": _arr(L.orc_mc_active_ids(self.h), a, np.uint32),
            "new_items": _arr(L.orc_mc_new_items(self.h), d, np.uint32),
            "with_verts": _arr(L.orc_mc_with_verts(self.h), h, np.uint32),
            "vert_offsets": _arr(L.orc_mc_vert_offsets(self.h), h, np.uint32),
            "verts_f32": _arr(L.orc_mc_verts_f32(self.h), 4 * v, np.float32).reshape(-1, 4),
            "verts_f64": _arr(L.orc_mc_verts_f64(self.h), 3 * v, np.float64).reshape(-1, 3),
        }

    def ranges(self):
        return _arr(lib().orc_mc_ranges(self.h), 2 * self.n, np.float32).reshape(-1, 2)

    def lru_reset(self):
        lib().orc_lru_reset(lib().orc_mc_lru(self.h))

    def lru_state(self):
        h = lib().orc_mc_lru(self.h)
        c = lib().orc_lru_size(h)
        age = np.zeros(max(c, 1), np.uint32)
        avail = np.zeros(max(c, 1), np.uint32)
        item = np.zeros(max(c, 1), np.int32)
        slot = np.zeros(max(self.n, 1), np.int32)
        lib().orc_lru_state(h, age, avail, item, slot)
        return age[:c], avail[:c], item[:c], slot[: self.n]


def bruteforce_count(vol, iso):
    v = np.ascontiguousarray(vol, np.float32)
    p = np.asarray([v.shape[2], v.shape[1], v.shape[0]], np.uint32)
    return int(lib().orc_mc_bruteforce_count(v.reshape(-1), p, np.float32(iso)))


def tri_counts():
    return np.array([lib().orc_tri_count(c) for c in range(256)], np.int32)


def tri_table_ints():
    """The 256x16 table in the reference's -1-terminated layout."""
    out = []
    for c in range(256):
        n = lib().orc_tri_count(c)
        row = [lib().orc_tri_edge(c, k) for k in range(n)] + [-1]
        row += [0] * (16 - len(row))
        out.extend(row)
    return out
