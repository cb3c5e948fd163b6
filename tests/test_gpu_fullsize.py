"""The three single-volume workloads at their full sizes, end to end against
the oracle: a 256^3 rate-8 Gaussian field with every block cached (one
isovalue), a 512^3 rate-4 Fourier-mode field with the cache at 25% of the
blocks (a reset, then the 0.2..0.8 sweep of 8 isovalues) and a
1024x1024x1080 rate-2 blob field with 10% white noise and the cache at 10%
(one isovalue).  Per query: counts, active and decoded ids, blocks with
vertices, vertex offsets and the cache state bit-exact; vertices bit-exact
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
    return nv


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
    assert _check(mc, ref, 0.3, "gauss256") > 0


def test_fourier512_rate4_sweep_after_reset(dev):
    mc, ref = _setup(dev, "fourier512_r4", 0.25)
    mc.lruCache.reset()
    ref.lru_reset()
    for iso in np.linspace(0.2, 0.8, 8).astype(np.float32):
        _check(mc, ref, iso, ("fourier512", float(iso)))


def test_blobs1024x1024x1080_rate2_one_query(dev):
    mc, ref = _setup(dev, "blobs1024x1024x1080_r2", 0.10)
    assert _check(mc, ref, 0.5, "blobs1024") > 0
