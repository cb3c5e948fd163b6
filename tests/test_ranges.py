"""Persisted block ranges: a volume loaded from block ranges saved beside its
stream (the server-side computation compressed_marching_cubes.js:523-524
suggests) in place of the full-volume decode at load.

CPU: the rule that bounds a block's tile range from block ranges alone
contains the tile range computed from the oracle's decoded (padded) volume,
and the padding clause of the rule is needed.  GPU: the exported ranges equal
the oracle's; objects and engines loaded from them (block ranges only, with
tile ranges, from host or device memory) give the oracle's results."""
import ctypes

import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth

VOLS = [((37, 42, 29), 2, 3), ((40, 36, 44), 4, 4), ((33, 20, 17), 8, 5), ((64, 61, 48), 2, 6)]


def _stream(dims, rate, seed):
    f = synth.field(synth.BLOBS, seed, dims, 12, (3, 12, 0.5, 1))
    return synth.encode(f, 64 * rate)


def _nb(dims):
    return [(d + 3) // 4 for d in dims]


def exact_tile(stream, maxbits, dims):
    """Per block [min, max] over every decoded voxel its cells can read: the
    4^3 block (padding included) and one more layer towards each in-grid +x /
    +y / +z neighbour, from the oracle's padded volume."""
    vol = oracle.decode_volume(stream, maxbits, dims)
    nb = _nb(dims)
    out = np.empty((nb[2], nb[1], nb[0], 2), np.float32)
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                r = vol[4 * bz:4 * bz + 4 + (bz + 1 < nb[2]), 4 * by:4 * by + 4 + (by + 1 < nb[1]),
                        4 * bx:4 * bx + 4 + (bx + 1 < nb[0])]
                out[bz, by, bx] = r.min(), r.max()
    return out.reshape(-1, 2)


def tile_bound(ranges, dims, padding_rule=True):
    """The bound from block ranges: union over the block and its in-grid +x /
    +y / +z, edge and corner neighbours; (-inf, inf) for blocks last on an
    axis whose size is not a multiple of 4."""
    nb = _nb(dims)
    r = np.asarray(ranges, np.float32).reshape(nb[2], nb[1], nb[0], 2)
    lo, hi = r[..., 0].copy(), r[..., 1].copy()
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                sub = r[dz:, dy:, dx:]
                z, y, x = sub.shape[:3]
                lo[:z, :y, :x] = np.minimum(lo[:z, :y, :x], sub[..., 0])
                hi[:z, :y, :x] = np.maximum(hi[:z, :y, :x], sub[..., 1])
    if padding_rule:
        for axis, d in zip((2, 1, 0), dims):
            if d % 4:
                idx = [slice(None)] * 3
                idx[axis] = -1
                lo[tuple(idx)], hi[tuple(idx)] = -np.inf, np.inf
    return np.stack([lo, hi], -1).reshape(-1, 2)


@pytest.mark.parametrize("dims,rate,seed", VOLS)
def test_tile_bound_contains_the_tile(dims, rate, seed):
    s = _stream(dims, rate, seed)
    t = exact_tile(s, 64 * rate, dims)
    b = tile_bound(oracle.block_ranges(s, 64 * rate, dims), dims)
    assert (b[:, 0] <= t[:, 0]).all() and (b[:, 1] >= t[:, 1]).all()


def test_block_ranges_do_not_bound_padding_voxels():
    """Without the padding clause the union misses decoded padding values:
    block ranges cover voxels inside the volume only."""
    dims, rate, seed = VOLS[0]
    s = _stream(dims, rate, seed)
    t = exact_tile(s, 64 * rate, dims)
    b = tile_bound(oracle.block_ranges(s, 64 * rate, dims), dims, padding_rule=False)
    assert ((b[:, 0] > t[:, 0]) | (b[:, 1] < t[:, 1])).any()


def _check_mc(mc, ref, isos, tag):
    for iso in isos:
        assert mc.computeSurface(iso) == ref.compute_surface(iso), (tag, iso)
        r = ref.results()
        assert np.array_equal(mc.array(bcmc_amd.ARRAY_ACTIVE_IDS), r["active_ids"]), (tag, iso)
        assert np.array_equal(mc.array(bcmc_amd.ARRAY_NEW_ITEMS), r["new_items"]), (tag, iso)
        assert np.array_equal(mc.vertices().view(np.uint32), r["verts_f32"].view(np.uint32)), (tag, iso)
        for a, b in zip(mc.lruCache.state(), ref.lru_state()):
            assert np.array_equal(a, b), (tag, iso)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,rate,seed", VOLS)
def test_persisted_ranges_give_the_oracle_results(dims, rate, seed):
    from gpu_common import device, say
    dev = device()
    say("persisted ranges", dims, rate)
    s = _stream(dims, rate, seed)
    n, mb = oracle.nblocks(dims), 64 * rate
    cache = n // 4
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), cache)
    r8, r16 = mc.blockRanges(False), mc.blockRanges(True)
    want = oracle.block_ranges(s, mb, dims)
    assert np.array_equal(r8.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(r16[:n].view(np.uint32), want.view(np.uint32))
    t = exact_tile(s, mb, dims)
    assert np.array_equal(r16[n:].view(np.uint32), t.view(np.uint32))

    # the same stream and ranges from device memory
    sbuf = dev.buffer_from(np.frombuffer(s.tobytes(), np.uint8))
    rbuf = bcmc_amd.block_ranges(dev, sbuf, mb, dims)
    isos = [float(want[n // 2].mean()), 0.3, 0.55, float(want[-1, 1]), 0.45]
    for tag, stream, ranges in (("block", s, r8), ("tile", s, r16), ("device", sbuf, rbuf)):
        m2 = bcmc_amd.CompressedMarchingCubes(dev)
        m2.setCompressedVolume(stream, rate, dims, (1, 1, 1), cache, block_ranges=ranges)
        got = m2.blockRanges(True)
        assert np.array_equal(got[:n].view(np.uint32), want.view(np.uint32)), tag
        tile = t if tag == "tile" else tile_bound(want, dims)
        assert np.array_equal(got[n:].view(np.uint32), tile.view(np.uint32)), tag
        _check_mc(m2, oracle.MC(s, mb, dims, initial_cache=cache), isos, tag)


@pytest.mark.gpu
def test_engine_with_persisted_ranges():
    from gpu_common import device
    dev = device()
    streams = [(_stream(d, r, seed), r, d) for d, r, seed in VOLS[:3]]
    eng = bcmc_amd.Engine(dev, 6000 * 256)
    ids = []
    for k, (s, r, d) in enumerate(streams):
        rng = oracle.block_ranges(s, 64 * r, d)
        if k == 1:
            rng = np.concatenate([rng, exact_tile(s, 64 * r, d)])
        ids.append(eng.add_volume(s, r, d, block_ranges=rng if k < 2 else None))
    rs = np.random.default_rng(11)
    for step in range(3):
        qv = rs.integers(0, len(ids), 8)
        qi = rs.uniform(0.2, 0.8, 8).astype(np.float32)
        res = eng.run(qv, qi)
        for k in range(len(qv)):
            s, r, d = streams[qv[k]]
            ref = oracle.MC(s, 64 * r, d)
            assert res[k]["num_vertices"] == ref.compute_surface(float(qi[k])), (step, k)
            got = eng.vertices(res[k])
            assert np.array_equal(got.view(np.uint32), ref.results()["verts_f32"].view(np.uint32)), (step, k)


@pytest.mark.gpu
def test_persisted_ranges_errors():
    from gpu_common import device
    dev = device()
    dims, rate, seed = VOLS[1]
    s = _stream(dims, rate, seed)
    n = oracle.nblocks(dims)
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    with pytest.raises(bcmc_amd.BcmcError) as e:
        mc.blockRanges()
    assert e.value.code == -5  # STATE
    with pytest.raises(ValueError):
        mc.setCompressedVolume(s, rate, dims, block_ranges=np.zeros(3 * n, np.float32))
    # the C ABI checks the size against the format and the format itself
    L, r = bcmc_amd.lib(), np.zeros(4 * n, np.float32)
    d3, sc = (ctypes.c_uint32 * 3)(*dims), (ctypes.c_float * 3)(1, 1, 1)
    for fmt, nbytes in ((bcmc_amd.RANGES_BLOCK, 16 * n), (bcmc_amd.RANGES_BLOCK_TILE, 8 * n), (7, 8 * n)):
        assert L.bcmc_mc_set_compressed_volume_ranges(mc.h, s.ctypes.data, s.nbytes, float(rate), d3, sc, 0,
                                                      r.ctypes.data, fmt, nbytes) == -1, fmt
    assert L.bcmc_mc_set_compressed_volume_ranges(mc.h, s.ctypes.data, s.nbytes, float(rate), d3, sc, 0,
                                                  None, 0, 8 * n) == -1
    eng = bcmc_amd.Engine(dev, 1000 * 256)
    with pytest.raises(ValueError):
        eng.add_volume(s, rate, dims, block_ranges=np.zeros(2 * n + 2, np.float32))
