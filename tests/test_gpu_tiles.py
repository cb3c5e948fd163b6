"""Tile shapes (scan and radix-sort elements per thread, zfp decoder blocks
per workgroup) at their extremes, with the smallest and largest persistent
grids: the scan, compaction, sort, LRU update, marching-cubes queries and an
engine step all give the oracle's bit-exact results at every shape."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import say

pytestmark = pytest.mark.gpu

SHAPES = {
    "smallest": dict(scan=1, sort=1, decode=1, mc=1, count=1, scan_items=8, sort_items=8, decode_threads=64),
    "largest": dict(scan=8, sort=16, decode=16, mc=8, count=16, scan_items=32, sort_items=32,
                    decode_threads=256),
    "mixed": dict(scan=2, sort=4, decode=8, mc=5, count=8, scan_items=32, sort_items=8, decode_threads=128),
}


@pytest.fixture(scope="module", params=list(SHAPES))
def tdev(request):
    d = bcmc_amd.Device(0)
    d.set_tuning(**SHAPES[request.param])
    t = d.tuning()
    assert t["tiles"] == {k: SHAPES[request.param][k] for k in ("scan_items", "sort_items", "decode_threads")}
    say("tile shape", request.param, t)
    return d


def test_scan_at_tile_shape(tdev):
    rng = np.random.default_rng(5)
    for n in (1, 4095, 4097, 8193, 16383, 16385, 100003, 1000003):
        data = rng.integers(0, 1000, n).astype(np.uint32)
        pipe = bcmc_amd.ExclusiveScanPipeline(tdev)
        sc = pipe.prepareInput(data)
        total = sc.scan(n)
        want, wt = oracle.exclusive_scan(data)
        assert total == wt, n
        assert np.array_equal(sc.inputBuf.download(np.uint32, n), want), n


def test_compact_at_tile_shape(tdev):
    rng = np.random.default_rng(6)
    n = 70001
    flags = (rng.uniform(0, 1, n) < 0.4).astype(np.uint32)
    offs, total = oracle.exclusive_scan(flags)
    fb = tdev.buffer_from(flags)
    out = tdev.create_buffer(n * 4)
    assert bcmc_amd.StreamCompactor(tdev).compactFlags(n, fb, out) == total
    assert np.array_equal(out.download(np.uint32, total), np.flatnonzero(flags))


@pytest.mark.parametrize("reverse", [False, True])
def test_sort_at_tile_shape(tdev, reverse):
    rng = np.random.default_rng(7 + reverse)
    for n in (65, 2049, 5000, 70000):
        for keys in (rng.integers(0, 5, n).astype(np.uint32),
                     rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)):
            vals = rng.permutation(n).astype(np.uint32)
            sorter = bcmc_amd.RadixSorter(tdev)
            a = sorter.getAlignedSize(n)
            kpad, vpad = np.zeros(a, np.uint32), np.zeros(a, np.uint32)
            kpad[:n], vpad[:n] = keys, vals
            vb = tdev.buffer_from(vpad)
            sorter.sort(tdev.buffer_from(kpad), vb, n, reverse)
            want, _ = oracle.sort_by_key(keys, vals, reverse)
            assert np.array_equal(vb.download(np.uint32, n), want), n


def test_lru_at_tile_shape(tdev):
    total, initial = 20000, 700
    rng = np.random.default_rng(8)
    gpu = bcmc_amd.LRUCache(tdev, None, None, initial, 256, total)
    cpu = oracle.LRU(initial, 256, total)
    for step in range(8):
        needed = np.zeros(total, np.uint32)
        lo = int(rng.integers(0, total))
        needed[lo:lo + int(rng.integers(1, total // 3))] = 1
        needed[rng.uniform(0, 1, total) < 0.05] = 1
        n_g, ids_g = gpu.update(tdev.buffer_from(needed))
        n_c, ids_c = cpu.update(needed)
        assert n_g == n_c, step
        if n_c:
            assert np.array_equal(ids_g.download(np.uint32, n_g), ids_c), step
        for a, b, what in zip(gpu.state(), cpu.state(), ("age", "avail", "item", "item_slot")):
            assert np.array_equal(a, b), (step, what)
        if step == 3:
            gpu.reset()
            cpu.reset()


@pytest.mark.parametrize("rate", [2, 4, 8])
def test_queries_at_tile_shape(tdev, rate):
    """Decoders (LRU-assign and cache modes at the workgroup size), active
    scan, LRU sort and the marching-cubes pass in one query sequence."""
    dims = (80, 72, 60)
    f = synth.field(synth.FOURIER, 3, dims, 128, (1, 24, 0, 0))
    s = synth.encode(f, 64 * rate)
    n = oracle.nblocks(dims)
    mc = bcmc_amd.CompressedMarchingCubes(tdev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), n // 5)
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=n // 5)
    for iso in (0.3, 0.55, 0.45, 0.7):
        assert mc.computeSurface(iso) == ref.compute_surface(iso), iso
        r = ref.results()
        assert np.array_equal(mc.array(bcmc_amd.ARRAY_ACTIVE_IDS), r["active_ids"]), iso
        assert np.array_equal(mc.array(bcmc_amd.ARRAY_NEW_ITEMS), r["new_items"]), iso
        assert np.array_equal(mc.vertices().view(np.uint32), r["verts_f32"].view(np.uint32)), iso
        for a, b in zip(mc.lruCache.state(), ref.lru_state()):
            assert np.array_equal(a, b), iso


def test_engine_step_at_tile_shape(tdev):
    """Row segments aligned to the scan tile of the shape; every query of a
    step over uneven volumes equals the oracle's single query."""
    vols = [((48, 40, 36), 4, 4), ((24, 28, 20), 2, 5), ((70, 64, 66), 8, 9)]
    streams = []
    for dims, rate, seed in vols:
        f = synth.field(synth.BLOBS, seed, dims, 6, (3, 10, 0.5, 1))
        streams.append((synth.encode(f, 64 * rate), rate, dims))
    eng = bcmc_amd.Engine(tdev, 4000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    rng = np.random.default_rng(9)
    for step in range(3):
        qv = rng.integers(0, len(ids), 9)
        qi = rng.uniform(0.2, 0.8, 9).astype(np.float32)
        res = eng.run(qv, qi)
        for k in range(len(qv)):
            s, r, d = streams[qv[k]]
            ref = oracle.MC(s, 64 * r, d)
            assert res[k]["num_vertices"] == ref.compute_surface(float(qi[k])), (step, k)
            got = eng.vertices(res[k])
            assert np.array_equal(got.view(np.uint32), ref.results()["verts_f32"].view(np.uint32)), (step, k)
