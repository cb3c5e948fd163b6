"""CompressedMarchingCubes on the GPU vs the oracle (js/compressed_marching_cubes.js).

Per query: active flags and ids, decoded ids, cache state, blocks with
vertices, vertex offsets and counts are bit-exact; vertex positions are
bit-exact with the oracle's FP32 evaluation and within 1e-5 voxel units of
its FP64 interpolation.
"""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import dev, replay_dev, say  # noqa: F401

pytestmark = pytest.mark.gpu

TOL = 1e-5  # voxel units, block-local FP32 positions vs FP64


def _volume(dims, kind, seed, rate, count, params):
    f = synth.field(kind, seed, dims, count, params)
    return synth.encode(f, 64 * rate)


def _check_query(mc, ref, iso):
    nv = mc.computeSurface(iso)
    nv_ref = ref.compute_surface(iso)
    r = ref.results()
    st = ref.stats
    assert nv == nv_ref
    info = mc.info()
    assert info["num_active_blocks"] == st["num_active"]
    assert info["new_decompressed"] == st["num_decoded"]
    assert info["num_blocks_with_vertices"] == st["num_with_verts"]
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_ACTIVE_FLAGS), r["flags"])
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_ACTIVE_IDS), r["active_ids"])
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_NEW_ITEMS), r["new_items"])
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_WITH_VERTS), r["with_verts"])
    offs = mc.array(bcmc_amd.ARRAY_VERT_OFFSETS, np.uint64)
    assert np.array_equal(offs.astype(np.uint32), r["vert_offsets"])
    for a, b, what in zip(mc.lruCache.state(), ref.lru_state(), ("age", "avail", "item", "slot")):
        assert np.array_equal(a, b), what
    return nv, r


@pytest.mark.parametrize("dims,rate,cache,isos", [
    ((32, 32, 32), 4, 64, [0.3, 0.5, 0.45]),
    ((40, 24, 52), 8, 32, [0.2, 0.6]),
    ((13, 10, 7), 2, 0, [0.4]),
    ((64, 64, 64), 2, 256, [0.25, 0.35, 0.5, 0.65, 0.35]),
])
def test_pipeline_matches_oracle(dev, dims, rate, cache, isos):
    say("mc", dims, rate, cache, isos)
    s = _volume(dims, synth.GAUSS, 31, rate, 10, (3, 12, 0.5, 1))
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), cache)
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=cache if cache else None)
    r_g = mc.array(bcmc_amd.ARRAY_RANGES, np.float32).reshape(-1, 2)
    assert np.array_equal(r_g.view(np.uint32), ref.ranges().view(np.uint32))
    for iso in isos:
        nv, r = _check_query(mc, ref, iso)
        v = mc.vertices()
        assert v.shape == (nv, 4)
        assert np.array_equal(v.view(np.uint32), r["verts_f32"].view(np.uint32))
        if nv:
            assert np.max(np.abs(v[:, :3].astype(np.float64) - r["verts_f64"])) <= TOL


def test_turbulent_sweep_with_eviction(dev):
    dims, rate = (96, 80, 64), 4
    say("mc fourier sweep", dims)
    f = synth.field(synth.FOURIER, 1, dims, 128, (1, 24, 0, 0))
    s = synth.encode(f, 64 * rate)
    n = oracle.nblocks(dims)
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), n // 4)
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=n // 4)
    for iso in np.linspace(0.2, 0.8, 8).astype(np.float32):
        nv, r = _check_query(mc, ref, float(iso))
        v = mc.vertices()
        assert np.array_equal(v.view(np.uint32), r["verts_f32"].view(np.uint32))
    # a reset, then the sweep again: every slot assignment follows the
    # post-reset tie order, checked slot for slot against the oracle's reset
    mc.lruCache.reset()
    ref.lru_reset()
    for iso in (0.5, 0.3, 0.7):
        nv, r = _check_query(mc, ref, iso)
        v = mc.vertices()
        assert np.array_equal(v.view(np.uint32), r["verts_f32"].view(np.uint32))
        for a, b, what in zip(mc.lruCache.state(), ref.lru_state(), ("age", "avail", "item", "slot")):
            assert np.array_equal(a, b), (iso, what)


def test_graph_replay_matches_the_oracle(replay_dev):
    """With graph replay on, untimed queries replay a captured graph of the
    query chain (captured again after the vertex buffer or the cache storage
    grows); a sweep with eviction, a reset and the sweep again equal the
    oracle query for query, and so do timed queries mixed in."""
    dims, rate = (64, 48, 40), 4
    say("mc graph replay", dims)
    f = synth.field(synth.FOURIER, 3, dims, 96, (1, 16, 0, 0))
    s = synth.encode(f, 64 * rate)
    n = oracle.nblocks(dims)
    mc = bcmc_amd.CompressedMarchingCubes(replay_dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), n // 4)
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=n // 4)
    isos = np.linspace(0.2, 0.8, 6).astype(np.float32)
    for sweep in range(2):
        for k, iso in enumerate(isos):
            if k == 3:
                nv = mc.computeSurface(float(iso), {})
                assert nv == ref.compute_surface(float(iso))
                r = ref.results()
            else:
                nv, r = _check_query(mc, ref, float(iso))
            v = mc.vertices()
            assert np.array_equal(v.view(np.uint32), r["verts_f32"].view(np.uint32)), (sweep, k)
        mc.lruCache.reset()
        ref.lru_reset()


def test_vertex_formats(dev):
    dims, rate, iso = (48, 40, 36), 4, 0.45
    s = _volume(dims, synth.BLOBS, 5, rate, 6, (4, 12, 0.5, 1))
    ref = oracle.MC(s, 64 * rate, dims)
    nv = ref.compute_surface(iso)
    loc = ref.results()["verts_f32"]
    bid = loc[:, 3].view(np.uint32)
    nb = [(d + 3) // 4 for d in dims]
    bpos = np.stack([bid % nb[0], (bid // nb[0]) % nb[1], bid // (nb[0] * nb[1])], 1)
    for fmt in (bcmc_amd.VERTS_GLOBAL_F32, bcmc_amd.VERTS_PACKED_U32X2):
        mc = bcmc_amd.CompressedMarchingCubes(dev)
        mc.setCompressedVolume(s, rate, dims)
        mc.setVertexFormat(fmt)
        assert mc.computeSurface(iso) == nv
        v = mc.vertices()
        if fmt == bcmc_amd.VERTS_GLOBAL_F32:
            want = (4 * bpos).astype(np.float32) + loc[:, :3]
            assert np.array_equal(v.view(np.uint32), want.view(np.uint32))
        else:
            q = (((loc[:, :3] - np.float32(0.5)) * np.float32(0.25)) * np.float32(1023)).astype(np.uint32)
            packed = ((q[:, 0] & 0x3FF) << 20) | ((q[:, 1] & 0x3FF) << 10) | (q[:, 2] & 0x3FF)
            assert np.array_equal(v[:, 0], packed) and np.array_equal(v[:, 1], bid)


def test_empty_and_full_isovalues(dev):
    dims = (24, 24, 24)
    s = _volume(dims, synth.GAUSS, 8, 4, 4, (3, 8, 0.5, 1))
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, 4, dims)
    ref = oracle.MC(s, 256, dims)
    for iso in (2.0, -1.0, 0.5, 2.0):
        _check_query(mc, ref, iso)


def _tie_isovalues(ranges):
    """Isovalues equal to decoded voxel values: a block's minimum and another
    block's maximum (both taken from the block ranges, which are voxel values)."""
    r = np.asarray(ranges, np.float32)
    return [float(r[len(r) // 2, 0]), float(r[len(r) // 3, 1])]


def test_nonfinite_and_tied_isovalues(dev):
    """NaN and +-inf isovalues make every comparison of the active test false
    (no active block, no vertex: compute_block_active.comp:30-60); isovalues
    equal to voxel values exercise the comparisons at equality in the active
    test, the case index and the interpolation.  A ragged volume, so partial
    blocks take part; the cache state carries across the queries."""
    dims, rate = (30, 21, 13), 4
    say("mc non-finite and tied isovalues", dims)
    s = _volume(dims, synth.GAUSS, 12, rate, 6, (3, 8, 0.5, 1))
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims)
    ref = oracle.MC(s, 64 * rate, dims)
    ties = _tie_isovalues(ref.ranges())
    for iso in [float("nan"), float("inf"), -float("inf"), *ties, 0.5, float("nan"), ties[0]]:
        nv, r = _check_query(mc, ref, iso)
        v = mc.vertices()
        assert np.array_equal(v.view(np.uint32), r["verts_f32"].view(np.uint32)), iso
        if not np.isfinite(iso):
            assert nv == 0 and mc.info()["num_active_blocks"] == 0, iso
        else:
            assert nv > 0, iso
            assert np.max(np.abs(v[:, :3].astype(np.float64) - r["verts_f64"])) <= TOL


def test_perf_tracker_keys(dev):
    dims = (32, 32, 32)
    s = _volume(dims, synth.GAUSS, 9, 4, 6, (3, 8, 0.5, 1))
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, 4, dims)
    perf = {}
    mc.computeSurface(0.4, perf)
    mc.computeSurface(0.5, perf)
    for k in bcmc_amd.CompressedMarchingCubes.PERF_KEYS:
        assert len(perf[k]) == 2
    assert perf["totalBlocks"] == 512
    mem = mc.reportMemoryUse()
    assert mem["total"] > mem["cache"] > 0


def test_launch_shapes_leave_results_unchanged(dev):
    """Autotuned and extreme launch shapes give the oracle's bit-exact results."""
    d = bcmc_amd.Device(0)
    t = d.autotune()
    say("mc autotune", t)
    assert all(1 <= t[k] <= 64 for k in ("scan", "sort", "decode", "mc", "count"))
    assert all(v > 0 for v in t["ms"].values())
    with pytest.raises(bcmc_amd.BcmcError):
        d.set_tuning(scan=0)
    dims, rate = (96, 80, 64), 4
    f = synth.field(synth.FOURIER, 2, dims, 128, (1, 24, 0, 0))
    s = synth.encode(f, 64 * rate)
    n = oracle.nblocks(dims)
    assert t["tiles"]["scan_items"] in (8, 16, 32) and t["tiles"]["sort_items"] in (8, 16, 32)
    assert t["tiles"]["decode_threads"] in (64, 128, 256)
    with pytest.raises(bcmc_amd.BcmcError):
        d.set_tuning(scan_items=12)
    with pytest.raises(bcmc_amd.BcmcError):
        d.set_tuning(decode_threads=96)
    for shape in (dict(scan=1, sort=1, decode=1, mc=1, count=1, scan_items=8, sort_items=8, decode_threads=64),
                  dict(scan=16, sort=8, decode=32, mc=8, count=16, scan_items=32, sort_items=32,
                       decode_threads=256)):
        say("mc launch shape", shape)
        d.set_tuning(**shape)
        got = d.tuning()
        assert {k: got[k] if k in got else got["tiles"][k] for k in shape} == shape
        mc = bcmc_amd.CompressedMarchingCubes(d)
        mc.setCompressedVolume(s, rate, dims, (1, 1, 1), n // 4)
        ref = oracle.MC(s, 64 * rate, dims, initial_cache=n // 4)
        for iso in (0.3, 0.5, 0.7, 0.5):
            _check_query(mc, ref, iso)
            v = mc.vertices()
            assert np.array_equal(v.view(np.uint32), ref.results()["verts_f32"].view(np.uint32))
    d.set_tuning(**{k: t[k] for k in ("scan", "sort", "decode", "mc", "count")}, **t["tiles"])


def test_world_positions_follow_the_renderer_formula(dev):
    """World-space expansion of the block-local vertices: the renderer's
    placement (mc_isosurface.vert:51-55) evaluated in float32 with numpy,
    separately rounded steps, bit for bit; and within float rounding of the
    FP64 positions."""
    say("world positions")
    dims, rate, scale = (37, 30, 26), 4, (1.0, 1.0, 26 / 37)
    f = synth.field(synth.GAUSS, 8, dims, 6, (3, 10, 0.5, 1))
    s = synth.encode(f, 64 * rate)
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), 200)
    nv = mc.computeSurface(0.45)
    assert nv > 1000
    local = mc.vertices()
    w = bcmc_amd.world_positions(dev, mc.vertexBuffer, nv, dims, scale).download(np.float32, 3 * nv)
    w = w.reshape(-1, 3)
    bid = local[:, 3].view(np.uint32)
    nbx, nby = (dims[0] + 3) // 4, (dims[1] + 3) // 4
    bpos = np.stack([bid % nbx, (bid // nbx) % nby, bid // nbx // nby], 1)
    g = (4 * bpos).astype(np.float32) + local[:, :3]
    want = np.float32(scale) * (g / np.float32(dims) - np.float32(0.5))
    assert want.dtype == np.float32
    assert np.array_equal(w.view(np.uint32), want.view(np.uint32))
    exact = np.array(scale) * ((4 * bpos + local[:, :3].astype(np.float64)) / np.array(dims) - 0.5)
    assert np.abs(w - exact).max() < 1e-6
    # a sub-range, and the argument checks
    part = bcmc_amd.world_positions(dev, mc.vertexBuffer, 5, dims, scale, first=nv - 5)
    assert np.array_equal(part.download(np.float32, 15).reshape(5, 3).view(np.uint32),
                          want[-5:].view(np.uint32))
    with pytest.raises(bcmc_amd.BcmcError):
        bcmc_amd.world_positions(dev, mc.vertexBuffer, 10, dims, scale, first=nv - 5)


def test_cache_memory_follows_the_logical_size(dev):
    """The LRU's slot contents are stored for its logical size plus room for
    growth, not for every block.  A first query that grows a 64-slot cache
    past its storage assigns nothing, enlarges the storage (contents copied)
    and runs its assignment, decode and MC pass again: counts, cache state and
    vertices still equal the oracle's, and so do the queries after it."""
    dims, rate = (320, 320, 320), 4
    f = synth.field(synth.BLOBS, 4, dims, 64, (8, 40, 0.5, 1))
    s = synth.encode(f, 64 * rate)
    n = oracle.nblocks(dims)
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), 64)
    info0 = mc.lruCache.info()
    say("cache storage before the first query", info0)
    assert info0["logical_bytes"] == 64 * 256 and info0["reserved_bytes"] < 128 * 256 + 256
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=64)
    for iso in (0.5, 0.45, 0.55, 0.52):
        nv, r = _check_query(mc, ref, iso)
        assert np.array_equal(mc.vertices().view(np.uint32), r["verts_f32"].view(np.uint32)), iso
        info = mc.lruCache.info()
        say("  iso", iso, "cache slots", info["cache_size"], "storage", info["reserved_bytes"])
        assert info["logical_bytes"] == info["cache_size"] * 256 < info["reserved_bytes"]
    # the first query grew the cache far past its 64 slots (the re-run
    # path); the storage stays well below all blocks decoded
    assert ref.stats["cache_size"] > 10000
    assert info["reserved_bytes"] < n * 256 // 2
    rep = mc.reportMemoryUse()
    assert rep["cache"] == info["reserved_bytes"] and rep["total"] > rep["cache"]
