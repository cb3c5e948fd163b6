"""The three single-volume workloads at their full sizes, end to end against
the oracle: a 256^3 rate-8 Gaussian field with every block cached (one
isovalue), a 512^3 rate-4 Fourier-mode field with the cache at 25% of the
blocks (a reset, then the 0.2..0.8 sweep of 8 isovalues) and a
1024x1024x1080 rate-2 blob field with 5% white noise and the cache at 10%
(one isovalue, also in the global FP32 xyz vertex format).  Per query:
counts, active and decoded ids, blocks with vertices, vertex offsets and the
cache state bit-exact; vertices bit-exact
with the oracle's FP32 evaluation and within 1e-5 voxel units of its FP64
interpolation.  The volumes come from the same seeded generators as the
benchmark."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import dev, say  # noqa: F401

pytestmark = pytest.mark.gpu

TOL = 1e-5
THREADS = 16


def _check(mc, ref, iso, tag):
    nv = mc.computeSurface(float(iso))
    nv_ref = ref.compute_surface(float(iso))
    assert nv == nv_ref, tag
    st, info = ref.stats, mc.info()
    assert info["num_active_blocks"] == st["num_active"], tag
    assert info["new_decompressed"] == st["num_decoded"], tag
    assert info["num_blocks_with_vertices"] == st["num_with_verts"], tag
    r = ref.results()
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_ACTIVE_IDS), r["active_ids"]), tag
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_NEW_ITEMS), r["new_items"]), tag
    assert np.array_equal(mc.array(bcmc_amd.ARRAY_WITH_VERTS), r["with_verts"]), tag
    offs = mc.array(bcmc_amd.ARRAY_VERT_OFFSETS, np.uint64)
    assert np.array_equal(offs.astype(np.uint32), r["vert_offsets"]), tag
    v = mc.vertices()
    assert np.array_equal(v.view(np.uint32), r["verts_f32"].view(np.uint32)), tag
    if nv:
        assert np.abs(v[:, :3].astype(np.float64) - r["verts_f64"]).max() <= TOL, tag
    for a, b, what in zip(mc.lruCache.state(), ref.lru_state(), ("age", "avail", "item", "slot")):
        assert np.array_equal(a, b), (tag, what)
    say("  iso", iso, "vertices", nv)
    return nv, r


def _setup(dev, name, cache_frac):
    w = synth.WORKLOADS[name]
    s, sha = synth.make_stream(w, threads=THREADS)
    dims, rate = w["dims"], w["rate"]
    n = oracle.nblocks(dims)
    cache = int(np.ceil(n * cache_frac))
    say(name, dims, "rate", rate, "cache", cache, "sha256", sha[:16])
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(s, rate, dims, (1, 1, 1), cache)
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=cache, threads=THREADS)
    return mc, ref


def test_gauss256_rate8_all_cached(dev):
    mc, ref = _setup(dev, "gauss256_r8", 1.0)
    assert _check(mc, ref, 0.3, "gauss256")[0] > 0


def test_fourier512_rate4_sweep_after_reset(dev):
    mc, ref = _setup(dev, "fourier512_r4", 0.25)
    mc.lruCache.reset()
    ref.lru_reset()
    for iso in np.linspace(0.2, 0.8, 8).astype(np.float32):
        _check(mc, ref, iso, ("fourier512", float(iso)))


def test_blobs1024x1024x1080_rate2_one_query(dev):
    mc, ref = _setup(dev, "blobs1024x1024x1080_r2", 0.10)
    nv, r = _check(mc, ref, 0.5, "blobs1024")
    assert nv > 0
    # the cache's storage follows its logical size (the query grows it from
    # 10% of the blocks to what it needs, plus room for the next growth), not
    # the 4.6 GB of all blocks decoded
    info = mc.lruCache.info()
    say("  cache slots", info["cache_size"], "storage", info["reserved_bytes"])
    all_blocks = oracle.nblocks(synth.WORKLOADS["blobs1024x1024x1080_r2"]["dims"]) * 256
    assert info["logical_bytes"] <= info["reserved_bytes"] <= info["logical_bytes"] * 3 // 2 + (64 << 20)
    assert info["reserved_bytes"] < all_blocks
    # the same query with FP32 global xyz output (12 B per vertex): exactly
    # 4 * block_pos + local in FP32, within 1e-5 + one FP32 ulp of |p| of
    # the FP64 position; compared in slices to bound host memory
    dims = synth.WORKLOADS["blobs1024x1024x1080_r2"]["dims"]
    nb = [(d + 3) // 4 for d in dims]
    del mc
    mg = bcmc_amd.CompressedMarchingCubes(dev)
    w = synth.WORKLOADS["blobs1024x1024x1080_r2"]
    s, _ = synth.make_stream(w, threads=THREADS)
    mg.setCompressedVolume(s, w["rate"], dims, (1, 1, 1), int(np.ceil(oracle.nblocks(dims) * 0.10)))
    mg.setVertexFormat(bcmc_amd.VERTS_GLOBAL_F32)
    assert mg.computeSurface(0.5) == nv
    buf = mg.vertexBuffer
    step = 50_000_000
    worst = 0.0
    for a in range(0, nv, step):
        b = min(nv, a + step)
        got = buf.download(np.float32, 3 * (b - a), 12 * a).reshape(-1, 3)
        loc = r["verts_f32"][a:b]
        bid = loc[:, 3].view(np.uint32).astype(np.int64)
        bpos = np.stack([bid % nb[0], (bid // nb[0]) % nb[1], bid // (nb[0] * nb[1])], 1)
        want = (4 * bpos).astype(np.float32) + loc[:, :3]
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), a
        exact = 4.0 * bpos + r["verts_f64"][a:b]
        tol = 1e-5 + np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - exact)
        assert np.all(err <= tol), a
        worst = max(worst, float(err.max()))
    say("  global FP32 vertices", nv, "max |fp32 - fp64|", worst)


def test_vertex_indices_past_2_pow_31(dev):
    """A query with more than 2^31 vertices: a 640x640x600 field of uniform
    noise at rate 8 and the median isovalue (about 9.6 vertices per cell,
    2.35G in all).  The block-local vertices agree with the global FP32
    format at indices on both sides of 2^31 and at the end (4 * block_pos +
    local, evaluated as the emit evaluates it), and the batched engine's
    vertices of the same query equal the plain pipeline's there."""
    dims = (640, 640, 600)
    nb = [(d + 3) // 4 for d in dims]
    n = int(np.prod(nb))
    field = np.random.default_rng(12).random((dims[2], dims[1], dims[0]), dtype=np.float32)
    fbuf = bcmc_amd.Buffer.alloc(dev, field.nbytes)
    fbuf.upload(field)
    del field
    stream = bcmc_amd.encode(dev, fbuf, dims, 64 * 8)
    fbuf.destroy()
    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(stream, 8, dims, (1, 1, 1), n)
    nv = mc.computeSurface(0.5)
    say("  noise 640x640x600 r8 at 0.5:", nv, "vertices")
    assert nv > 2**31 + (1 << 20)
    k = 4096
    starts = [0, 2**31 - k, 2**31 - 64, nv - k]
    local = [mc.vertexBuffer.download(np.float32, 4 * k, 16 * s).reshape(-1, 4) for s in starts]
    mc.setVertexFormat(bcmc_amd.VERTS_GLOBAL_F32)
    assert mc.computeSurface(0.5) == nv
    glob = [mc.vertexBuffer.download(np.float32, 3 * k, 12 * s).reshape(-1, 3) for s in starts]
    for s, lv, gv in zip(starts, local, glob):
        bid = lv[:, 3].view(np.uint32).astype(np.int64)
        pos = np.stack([bid % nb[0], (bid // nb[0]) % nb[1], bid // (nb[0] * nb[1])], 1)
        want = (4 * pos).astype(np.float32) + lv[:, :3]
        assert np.array_equal(gv.view(np.uint32), want.view(np.uint32)), s
    del mc
    eng = bcmc_amd.Engine(dev, n * 256)
    vid = eng.add_volume(stream, 8, dims)
    res = eng.run([vid], [0.5])[0]
    assert res["num_vertices"] == nv and res["vertex_offset"] == 0
    for s, lv in zip(starts, local):
        got = eng.vertices(res, s, k)
        assert np.array_equal(got.view(np.uint32), lv.view(np.uint32)), s
