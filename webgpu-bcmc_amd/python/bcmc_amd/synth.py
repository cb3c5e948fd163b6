"""Synthetic inputs: seeded fields and the zfp fixed-rate encoder (host C).

Wraps lib/libbcmc_synth.so.  Field kinds: 0 Gaussians, 1 Fourier modes,
2 Gaussian blobs (+ optional white noise).  Parameters and draw order are
documented in csrc/synth/bcmc_synth.c.
"""
import ctypes
import hashlib
import os

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                    "lib", "libbcmc_synth.so")
_lib = None

GAUSS, FOURIER, BLOBS = 0, 1, 2


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(_LIB)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        lib.bcmc_synth_stream_bytes.restype = ctypes.c_uint64
        lib.bcmc_synth_stream_bytes.argtypes = [u32p, ctypes.c_uint32]
        lib.bcmc_synth_encode.restype = ctypes.c_int
        lib.bcmc_synth_encode.argtypes = [np.ctypeslib.ndpointer(np.float32, flags="C"), u32p,
                                          ctypes.c_uint32, np.ctypeslib.ndpointer(np.uint8, flags="C"),
                                          ctypes.c_int]
        lib.bcmc_synth_field.restype = ctypes.c_int
        lib.bcmc_synth_field.argtypes = [ctypes.c_int, ctypes.c_uint64, u32p, ctypes.c_uint32, f64p,
                                         ctypes.c_double, np.ctypeslib.ndpointer(np.float32, flags="C"),
                                         ctypes.c_int]
        lib.bcmc_synth_stream.restype = ctypes.c_int
        lib.bcmc_synth_stream.argtypes = [ctypes.c_int, ctypes.c_uint64, u32p, ctypes.c_uint32, f64p,
                                          ctypes.c_double, ctypes.c_uint32,
                                          np.ctypeslib.ndpointer(np.uint8, flags="C"), ctypes.c_int]
        _lib = lib
    return _lib


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def stream_bytes(dims, maxbits):
    return int(_load().bcmc_synth_stream_bytes(np.asarray(dims, np.uint32), maxbits))


def encode(vol, maxbits, threads=None):
    """Encode a (nz, ny, nx) float32 volume; returns (stream bytes as uint8, dims)."""
    v = np.ascontiguousarray(vol, np.float32)
    dims = np.asarray([v.shape[2], v.shape[1], v.shape[0]], np.uint32)
    out = np.zeros(stream_bytes(dims, maxbits), np.uint8)
    rc = _load().bcmc_synth_encode(v.reshape(-1), dims, maxbits, out, threads or default_threads())
    if rc != 0:
        raise ValueError("bcmc_synth_encode failed")
    return out


def field(kind, seed, dims, count, params, noise=0.0, threads=None):
    """Normalised float32 field as a (nz, ny, nx) array."""
    d = np.asarray(dims, np.uint32)
    out = np.zeros(int(d[0]) * int(d[1]) * int(d[2]), np.float32)
    rc = _load().bcmc_synth_field(kind, seed, d, count, np.asarray(params, np.float64), noise, out,
                                  threads or default_threads())
    if rc != 0:
        raise ValueError("bcmc_synth_field failed")
    return out.reshape(int(d[2]), int(d[1]), int(d[0]))


def stream(kind, seed, dims, count, params, maxbits, noise=0.0, threads=None):
    """Generate and encode slab by slab; returns the compressed stream (uint8)."""
    d = np.asarray(dims, np.uint32)
    out = np.zeros(stream_bytes(d, maxbits), np.uint8)
    rc = _load().bcmc_synth_stream(kind, seed, d, count, np.asarray(params, np.float64), noise,
                                   maxbits, out, threads or default_threads())
    if rc != 0:
        raise ValueError("bcmc_synth_stream failed")
    return out


# Named workloads (dims, field, rate).  Draw parameters per kind:
#   Gaussians/blobs: params = (sigma_lo, sigma_hi, amp_lo, amp_hi)
#   Fourier:         params = (k_lo, k_hi, 0, 0)
WORKLOADS = {
    "gauss256_r8": dict(dims=(256, 256, 256), kind=GAUSS, seed=0, count=16,
                        params=(8.0, 40.0, 0.5, 1.0), rate=8, noise=0.0),
    "fourier512_r4": dict(dims=(512, 512, 512), kind=FOURIER, seed=1, count=512,
                          params=(1.0, 64.0, 0.0, 0.0), rate=4, noise=0.0),
    "blobs1024x1024x1080_r2": dict(dims=(1024, 1024, 1080), kind=BLOBS, seed=2, count=256,
                                   params=(16.0, 96.0, 0.5, 1.0), rate=2, noise=0.05),
}


def blob_workload(dims, seed, rate, count=256):
    return dict(dims=tuple(dims), kind=BLOBS, seed=seed, count=count,
                params=(16.0, 96.0, 0.5, 1.0), rate=rate, noise=0.0)


# The served set: 8 device-resident blob volumes of uneven size, seeds 7-14,
# zfp rates cycling 2/4/8 (the 1024x1024x1080 volume at 2).
SERVED_SIZES = [(128, 128, 128), (192, 192, 192), (256, 256, 256), (320, 320, 320),
                (384, 384, 384), (512, 512, 512), (768, 768, 768), (1024, 1024, 1080)]
SERVED_RATES = [2, 4, 8, 2, 4, 8, 2, 2]


def served_volumes(threads=None):
    """[(stream, rate, dims)] of the served set."""
    out = []
    for i, (dims, rate) in enumerate(zip(SERVED_SIZES, SERVED_RATES)):
        s, _ = make_stream(blob_workload(dims, 7 + i, rate), threads=threads)
        out.append((s, rate, dims))
    return out


def make_stream(w, threads=None):
    """Compressed stream for a workload dict; returns (uint8 stream, sha256 hex)."""
    s = stream(w["kind"], w["seed"], w["dims"], w["count"], w["params"], 64 * w["rate"],
               noise=w.get("noise", 0.0), threads=threads)
    return s, hashlib.sha256(s.tobytes()).hexdigest()
