"""Writes the zfp golden fixtures (tests/golden/zfp_fixtures.npz).

Runs only where LLNL zfp 0.5.5 is installed (/opt/conda/lib/libzfp.so.0.5.5 in
the build container); the committed .npz is what the tests read everywhere
else.  zfp is the third-party codec whose streams the reference decodes
(js/bcmc.js:27-38 loads `<dataset>.zfp`; shaders/zfp_decompress.comp restates
the zfp decoder), so its own compress/decompress pin both our encoder and the
oracle's decoder:
  stream_<name>  : zfp_compress of field_<name> in fixed-rate mode (no header)
  decoded_<name> : zfp_decompress of that stream
Only libzfp's public API is used.

Usage: python tests/golden/make_zfp_fixtures.py
"""
import ctypes
import hashlib
import json
import os

import numpy as np

ZFP = "/opt/conda/lib/libzfp.so.0.5.5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "zfp_fixtures.npz")


def _zfp():
    z = ctypes.CDLL(ZFP)
    vp = ctypes.c_void_p
    z.zfp_field_3d.restype = vp
    z.zfp_field_3d.argtypes = [vp, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
    z.zfp_field_free.argtypes = [vp]
    z.zfp_stream_open.restype = vp
    z.zfp_stream_open.argtypes = [vp]
    z.zfp_stream_close.argtypes = [vp]
    z.zfp_stream_set_rate.restype = ctypes.c_double
    z.zfp_stream_set_rate.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_uint, ctypes.c_int]
    z.zfp_stream_maximum_size.restype = ctypes.c_size_t
    z.zfp_stream_maximum_size.argtypes = [vp, vp]
    z.stream_open.restype = vp
    z.stream_open.argtypes = [vp, ctypes.c_size_t]
    z.stream_close.argtypes = [vp]
    z.zfp_stream_set_bit_stream.argtypes = [vp, vp]
    z.zfp_stream_rewind.argtypes = [vp]
    z.zfp_compress.restype = ctypes.c_size_t
    z.zfp_compress.argtypes = [vp, vp]
    z.zfp_decompress.restype = ctypes.c_size_t
    z.zfp_decompress.argtypes = [vp, vp]
    return z


def zfp_roundtrip(z, vol, rate):
    vol = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = vol.shape
    fld = z.zfp_field_3d(vol.ctypes.data, 3, nx, ny, nz)
    zs = z.zfp_stream_open(None)
    z.zfp_stream_set_rate(zs, rate, 3, 3, 0)
    cap = z.zfp_stream_maximum_size(zs, fld)
    buf = ctypes.create_string_buffer(cap)
    bs = z.stream_open(buf, cap)
    z.zfp_stream_set_bit_stream(zs, bs)
    z.zfp_stream_rewind(zs)
    size = z.zfp_compress(zs, fld)
    assert size > 0
    stream = np.frombuffer(buf.raw[:size], np.uint8).copy()
    out = np.zeros_like(vol)
    fld2 = z.zfp_field_3d(out.ctypes.data, 3, nx, ny, nz)
    z.zfp_stream_rewind(zs)
    assert z.zfp_decompress(zs, fld2) == size
    z.zfp_field_free(fld)
    z.zfp_field_free(fld2)
    z.stream_close(bs)
    z.zfp_stream_close(zs)
    return stream, out


def smooth(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.zeros(shape)
    for _ in range(6):
        c = rng.uniform(0, 1, 3) * [nx, ny, nz]
        s = rng.uniform(2, 8)
        a = rng.uniform(-1, 1)
        f += a * np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) / (2 * s * s))
    return f.astype(np.float32)


def special():
    """8x8x8 volume whose 8 blocks exercise the decoder's corner cases."""
    v = np.zeros((8, 8, 8), np.float32)
    rng = np.random.default_rng(7)
    blocks = [
        np.zeros((4, 4, 4)),                                   # zero block (flag bit 0)
        np.full((4, 4, 4), 0.75),                              # constant block
        rng.uniform(-1e-30, 1e-30, (4, 4, 4)),                 # emax well below -30
        rng.uniform(-3e4, 3e4, (4, 4, 4)),                     # large magnitude, mixed sign
        rng.normal(0, 1, (4, 4, 4)),                           # white noise
        np.indices((4, 4, 4)).sum(0) * 0.125,                  # linear ramp
        np.where(rng.uniform(0, 1, (4, 4, 4)) < 0.1, 1.0, 0.0),  # sparse spikes
        rng.uniform(0.5, 0.5000001, (4, 4, 4)),                # tiny dynamic range
    ]
    for i, b in enumerate(blocks):
        bx, by, bz = i % 2, (i // 2) % 2, i // 4
        v[bz * 4:bz * 4 + 4, by * 4:by * 4 + 4, bx * 4:bx * 4 + 4] = b
    return v


CASES = [
    # name, field, rate
    ("smooth16_r1", lambda: smooth((16, 16, 16), 1), 1.0),
    ("smooth16_r2", lambda: smooth((16, 16, 16), 1), 2.0),
    ("smooth16_r4", lambda: smooth((16, 16, 16), 1), 4.0),
    ("smooth16_r8", lambda: smooth((16, 16, 16), 1), 8.0),
    ("smooth16_r16", lambda: smooth((16, 16, 16), 2), 16.0),
    ("smooth16_r32", lambda: smooth((16, 16, 16), 2), 32.0),
    ("partial_13x10x7_r2", lambda: smooth((7, 10, 13), 3), 2.0),
    ("partial_13x10x7_r4", lambda: smooth((7, 10, 13), 3), 4.0),
    ("partial_5x6x9_r8", lambda: smooth((9, 6, 5), 4), 8.0),
    ("partial_20x12x36_r4", lambda: smooth((36, 12, 20), 5), 4.0),
    ("rate_1p5_12x8x8", lambda: smooth((8, 8, 12), 6), 1.5),
    ("rate_0p5_8x8x8", lambda: smooth((8, 8, 8), 6), 0.5),
    ("special_r2", special, 2.0),
    ("special_r4", special, 4.0),
    ("special_r8", special, 8.0),
    ("noise16_r4", lambda: np.random.default_rng(9).uniform(0, 1, (16, 16, 16)).astype(np.float32), 4.0),
]


def main():
    z = _zfp()
    wb = ctypes.c_uint.in_dll(z, "stream_word_bits").value
    assert wb == 64, wb
    arrays, meta = {}, {}
    for name, fn, rate in CASES:
        vol = fn()
        stream, dec = zfp_roundtrip(z, vol, rate)
        arrays["field_" + name] = vol
        arrays["stream_" + name] = stream
        arrays["decoded_" + name] = dec
        meta[name] = {"dims": [vol.shape[2], vol.shape[1], vol.shape[0]],
                      "maxbits": int(round(64 * rate)),
                      "sha256": hashlib.sha256(stream.tobytes()).hexdigest()}
    arrays["meta"] = np.frombuffer(json.dumps({"zfp": ctypes.c_char_p.in_dll(
        z, "zfp_version_string").value.decode(), "word_bits": wb, "cases": meta}).encode(), np.uint8)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
