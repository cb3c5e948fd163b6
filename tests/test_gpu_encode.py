"""GPU fixed-rate zfp encoder: byte-identical to libzfp 0.5.5 (the committed
fixtures) and to the host encoder (itself pinned to those fixtures) on
partial blocks, fractional rates and larger synthetic volumes; its streams
decode on the GPU to the oracle's values."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import FIX, META, dev, say  # noqa: F401

pytestmark = pytest.mark.gpu


def _gpu_encode(dev, vol, dims, maxbits):
    out = bcmc_amd.encode(dev, dev.buffer_from(np.ascontiguousarray(vol, np.float32)), dims, maxbits)
    n = bcmc_amd.stream_bytes(dims, maxbits)
    got = out.download(np.uint8, n + 16)
    assert not got[n:].any()  # the read-ahead words stay zero
    return got[:n], out


@pytest.mark.parametrize("name", sorted(META["cases"]))
def test_encode_matches_libzfp_fixture(dev, name):
    say("gpu encode fixture", name)
    m = META["cases"][name]
    got, _ = _gpu_encode(dev, FIX["field_" + name], m["dims"], m["maxbits"])
    assert got.tobytes() == FIX["stream_" + name].tobytes()


@pytest.mark.parametrize("dims,maxbits", [((33, 17, 9), 64), ((33, 17, 9), 96), ((21, 22, 23), 37),
                                          ((5, 4, 3), 9), ((64, 48, 40), 256), ((30, 31, 29), 100),
                                          ((16, 16, 16), 4096)])
def test_encode_matches_host_encoder(dev, dims, maxbits):
    say("gpu encode vs host", dims, maxbits)
    f = synth.field(synth.BLOBS, sum(dims) + maxbits, dims, 6, (3, 10, 0.5, 1), noise=0.05)
    # zero, constant and tiny-exponent blocks among the smooth ones
    f[:4, :4, :4] = 0
    f[-4:, :4, -4:] = np.float32(0.25)
    f[:4, -4:, -4:] *= np.float32(1e-31)
    got, _ = _gpu_encode(dev, f, dims, maxbits)
    assert got.tobytes() == synth.encode(f, maxbits).tobytes()


def test_encoded_stream_decodes_on_the_gpu(dev):
    say("gpu encode -> gpu decode")
    dims, maxbits = (40, 36, 44), 128
    f = synth.field(synth.GAUSS, 9, dims, 8, (3, 10, 0.5, 1))
    host, buf = _gpu_encode(dev, f, dims, maxbits)
    nb = oracle.nblocks(dims)
    dec = bcmc_amd.decode_blocks(dev, buf, maxbits, nb).download(np.float32, nb * 64).reshape(nb, 64)
    want = np.stack([oracle.decode_block(host, maxbits, b) for b in range(nb)])
    assert np.array_equal(dec.view(np.uint32), want.view(np.uint32))


def test_encode_full_size_volume(dev):
    """A 256^3 blob volume at rate 4 (262,144 blocks, several blocks per lane
    of the grid) equals the host encoder byte for byte."""
    say("gpu encode 256^3 r4")
    dims = (256, 256, 256)
    f = synth.field(synth.BLOBS, 3, dims, 16, (8, 40, 0.5, 1))
    got, _ = _gpu_encode(dev, f, dims, 256)
    assert got.tobytes() == synth.encode(f, 256).tobytes()


def test_encode_argument_errors(dev):
    f = dev.buffer_from(np.zeros(8 * 8 * 8, np.float32))
    small = bcmc_amd.Buffer.alloc(dev, 8)
    for call in (lambda: bcmc_amd.lib().bcmc_zfp_encode(dev.h, f.h, bcmc_amd._u32arr((8, 8, 8)), 256, small.h),
                 lambda: bcmc_amd.lib().bcmc_zfp_encode(dev.h, f.h, bcmc_amd._u32arr((8, 8, 9)), 256, small.h),
                 lambda: bcmc_amd.lib().bcmc_zfp_encode(dev.h, f.h, bcmc_amd._u32arr((8, 8, 8)), 8, small.h)):
        assert call() == -1
