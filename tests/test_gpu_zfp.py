"""ZFP decode and block ranges on the GPU vs libzfp fixtures and the oracle."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import FIX, META, dev, say  # noqa: F401

pytestmark = pytest.mark.gpu


def _padded_stream(dev, s):
    # the library buffers hold the stream plus zero padding
    b = np.zeros(len(s) + 16, np.uint8)
    b[:len(s)] = s
    return dev.buffer_from(b)


@pytest.mark.parametrize("name", sorted(META["cases"]))
def test_decode_matches_libzfp_fixture(dev, name):
    say("zfp fixture", name)
    m = META["cases"][name]
    s = FIX["stream_" + name]
    nb = oracle.nblocks(m["dims"])
    out = bcmc_amd.decode_blocks(dev, _padded_stream(dev, s), m["maxbits"], nb)
    blocks = out.download(np.float32, nb * 64).reshape(nb, 64)
    want = np.stack([oracle.decode_block(s, m["maxbits"], b) for b in range(nb)])
    assert np.array_equal(blocks.view(np.uint32), want.view(np.uint32))
    # and against libzfp's own decompression over the in-volume voxels
    dx, dy, dz = m["dims"]
    p = oracle.padded(m["dims"])
    vol = np.zeros((p[2], p[1], p[0]), np.float32)
    nbx, nby = p[0] // 4, p[1] // 4
    for b in range(nb):
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        vol[bz*4:bz*4+4, by*4:by*4+4, bx*4:bx*4+4] = blocks[b].reshape(4, 4, 4)
    assert np.array_equal(vol[:dz, :dy, :dx].view(np.uint32), FIX["decoded_" + name].view(np.uint32))


@pytest.mark.parametrize("name", ["partial_13x10x7_r4", "partial_20x12x36_r4", "special_r8",
                                  "rate_1p5_12x8x8", "smooth16_r2"])
def test_block_ranges(dev, name):
    m = META["cases"][name]
    s = FIX["stream_" + name]
    r = bcmc_amd.block_ranges(dev, _padded_stream(dev, s), m["maxbits"], m["dims"])
    got = r.download(np.float32).reshape(-1, 2)
    assert np.array_equal(got.view(np.uint32), oracle.block_ranges(s, m["maxbits"], m["dims"]).view(np.uint32))


@pytest.mark.parametrize("rate", [1, 2, 4, 8, 16])
def test_decode_synthetic_volume(dev, rate):
    say("zfp synthetic rate", rate)
    f = synth.field(synth.FOURIER, 21, (64, 48, 40), 64, (1, 16, 0, 0))
    s = synth.encode(f, 64 * rate)
    nb = oracle.nblocks((64, 48, 40))
    sb = _padded_stream(dev, s)
    out = bcmc_amd.decode_blocks(dev, sb, 64 * rate, nb).download(np.float32, nb * 64)
    want = np.concatenate([oracle.decode_block(s, 64 * rate, b) for b in range(nb)])
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    # list decode (cache path shape): a shuffled id list
    ids = np.random.default_rng(rate).permutation(nb)[: nb // 3].astype(np.uint32)
    out2 = bcmc_amd.decode_blocks(dev, sb, 64 * rate, len(ids), dev.buffer_from(ids))
    got = out2.download(np.float32, len(ids) * 64).reshape(-1, 64)
    assert np.array_equal(got.view(np.uint32), want.reshape(nb, 64)[ids].view(np.uint32))


def _random_stream(rng, nblocks, maxbits, densities):
    """Streams of random bits (every block's flag bit set) at several densities
    of ones: sparse ones give long zero runs (up to and past the run that
    reaches coefficient 63, and runs cut by the bit budget), dense ones give
    group tests every few bits.  Any bit sequence is a valid fixed-rate stream."""
    per = nblocks // len(densities)
    bits = np.concatenate([(rng.random(per * maxbits) < d).astype(np.uint8) for d in densities])
    bits[::maxbits] = 1
    return np.packbits(bits, bitorder="little"), per * len(densities)


@pytest.mark.parametrize("maxbits", [64, 128, 256, 512, 1024, 100, 37])
def test_decode_random_bitstreams(dev, maxbits):
    say("zfp random bitstreams, maxbits", maxbits)
    rng = np.random.default_rng(maxbits)
    s, nb = _random_stream(rng, 6000, maxbits, (0.005, 0.02, 0.08, 0.3, 0.6, 0.95))
    out = bcmc_amd.decode_blocks(dev, _padded_stream(dev, s), maxbits, nb)
    got = out.download(np.float32, nb * 64).reshape(nb, 64)
    want = np.stack([oracle.decode_block(s, maxbits, b) for b in range(nb)])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_decode_past_2_pow_32_bits(dev):
    """Blocks whose first bit lies at or beyond bit 2^32 of the stream (the
    reference's u32 offset block_index * max_bits wraps there,
    zfp_decompress.comp:199): rate 8, blocks around 2^32 / 512 = 8,388,608."""
    maxbits = 512
    first = (1 << 32) // maxbits - 40
    nb = first + 80
    s = np.zeros(nb * maxbits // 8, np.uint8)
    f = synth.field(synth.GAUSS, 17, (16, 16, 20), 6, (3, 8, 0.5, 1))
    enc = synth.encode(f, maxbits)  # 80 blocks
    s[first * maxbits // 8:] = enc[:80 * maxbits // 8]
    say("zfp decode past 2^32 bits:", len(s), "bytes")
    ids = np.arange(first, nb, dtype=np.uint32)
    out = bcmc_amd.decode_blocks(dev, _padded_stream(dev, s), maxbits, len(ids), dev.buffer_from(ids))
    got = out.download(np.float32, len(ids) * 64).reshape(-1, 64)
    tail = s[first * maxbits // 8:]
    want = np.stack([oracle.decode_block(tail, maxbits, b) for b in range(80)])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.count_nonzero(want) > 0
