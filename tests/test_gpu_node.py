"""The Node.js facade (the drop-in for the reference's JS classes) on the GPU,
compared with the oracle: scan, compaction, sort, and a computeSurface sweep
with cache slots, counts and vertices bit-exact."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from bcmc_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "webgpu-bcmc_amd", "lib", "bcmc_addon.node")


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_facade_sweep(tmp_path):
    assert os.path.exists(ADDON), "N-API addon not built"
    dims, rate, cache = (40, 32, 36), 4, 96
    f = synth.field(synth.GAUSS, 17, dims, 8, (3, 10, 0.5, 1))
    s = synth.encode(f, 64 * rate)
    rng = np.random.default_rng(5)
    params = {"dims": dims, "rate": rate, "cache": cache, "isos": [0.3, 0.5, 0.42],
              "scan_input": rng.integers(0, 100, 3000).tolist(),
              "flags": (rng.uniform(0, 1, 5000) < 0.3).astype(int).tolist(),
              "sort_keys": rng.integers(0, 6, 300).tolist()}
    (tmp_path / "params.json").write_text(json.dumps(params))
    (tmp_path / "stream.bin").write_bytes(s.tobytes())
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "facade_parity.js"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads((tmp_path / "out.json").read_text())
    want, total = oracle.exclusive_scan(np.array(params["scan_input"], np.uint32))
    assert out["scan_total"] == total and out["scan"] == want.tolist()
    assert out["compact"] == np.flatnonzero(params["flags"]).tolist()
    keys = np.array(params["sort_keys"], np.uint32)
    sv, _ = oracle.sort_by_key(keys, np.arange(len(keys), dtype=np.uint32), True)
    assert out["sorted"] == sv.tolist()
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=cache)
    for i, iso in enumerate(params["isos"]):
        nv = ref.compute_surface(iso)
        c = out["counts"][i]
        assert c["numVertices"] == nv
        assert c["numActiveBlocks"] == ref.stats["num_active"]
        assert c["numBlocksWithVertices"] == ref.stats["num_with_verts"]
        assert c["newDecompressed"] == ref.stats["num_decoded"]
        assert c["cacheSize"] == ref.stats["cache_size"]
        v = np.fromfile(tmp_path / f"verts_{i}.bin", np.float32).reshape(-1, 4)
        assert np.array_equal(v.view(np.uint32), ref.results()["verts_f32"].view(np.uint32))
        slots = np.fromfile(tmp_path / f"slots_{i}.bin", np.int32)
        assert np.array_equal(slots, ref.lru_state()[3])
    assert out["perfLen"] == 3 and "computeVertices" in out["perfKeys"]
    assert out["after_reset_age0"] == 100000
    # blockRanges() persists the oracle's block ranges; objects loaded from
    # them (with and without tile ranges) give the oracle's results
    ranges = np.fromfile(tmp_path / "ranges.bin", np.float32).reshape(-1, 2)
    assert np.array_equal(ranges.view(np.uint32), oracle.block_ranges(s, 64 * rate, dims).view(np.uint32))
    for name in ("block", "tile"):
        ref = oracle.MC(s, 64 * rate, dims, initial_cache=cache)
        for i, iso in enumerate(params["isos"]):
            c = out["persisted"][name][i]
            assert c["numVertices"] == ref.compute_surface(iso), (name, iso)
            assert c["newDecompressed"] == ref.stats["num_decoded"], (name, iso)
            v = np.fromfile(tmp_path / f"pverts_{name}_{i}.bin", np.float32).reshape(-1, 4)
            assert np.array_equal(v.view(np.uint32), ref.results()["verts_f32"].view(np.uint32)), (name, iso)


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_sweep_benchmark_matches_oracle(tmp_path):
    """The facade's sweepUp benchmark (100 isovalues after an LRU reset, on a
    cache too small for the volume) gives the oracle's counts, decoded blocks
    included, at every step."""
    assert os.path.exists(ADDON), "N-API addon not built"
    dims, rate, cache, rng_ = (36, 40, 28), 4, 160, [0.2, 0.8]
    f = synth.field(synth.BLOBS, 23, dims, 6, (3, 10, 0.5, 1))
    s = synth.encode(f, 64 * rate)
    (tmp_path / "params.json").write_text(json.dumps({"dims": dims, "rate": rate, "cache": cache,
                                                      "range": rng_}))
    (tmp_path / "stream.bin").write_bytes(s.tobytes())
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "benchmark_protocol.js"), str(tmp_path),
                        "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads((tmp_path / "out.json").read_text())
    perf = out["perf"]
    assert out["fileName"] == "bcmc-explicit-blob-sweepUp.json"
    assert len(perf["isovalue"]) == 100 and len(perf["numVertices"]) == 100
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=cache)
    ref.compute_surface(0.5)
    ref.lru_reset()
    cur = rng_[0]
    for k in range(100):
        cur += (rng_[1] - rng_[0]) / 100
        assert perf["isovalue"][k] == cur
        nv = ref.compute_surface(cur)
        assert perf["numVertices"][k] == nv, k
        assert perf["numActiveBlocks"][k] == ref.stats["num_active"], k
        assert perf["numBlocksDecompressed"][k] == ref.stats["num_decoded"], k
        assert perf["numBlocksWithVertices"][k] == ref.stats["num_with_verts"], k


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_engine_from_node_matches_oracle(tmp_path):
    """The batched and served engine reached from Node: every query's count
    and block-local vertices equal the oracle's single computeSurface, and
    the served trace gives the same counts."""
    assert os.path.exists(ADDON), "N-API addon not built"
    vols = [((32, 28, 24), 4, 31), ((20, 24, 36), 8, 32), ((13, 10, 7), 2, 33)]
    streams = []
    for i, (dims, rate, seed) in enumerate(vols):
        f = synth.field(synth.BLOBS, seed, dims, 6, (3, 10, 0.5, 1))
        streams.append(synth.encode(f, 64 * rate))
        (tmp_path / f"vol_{i}.bin").write_bytes(streams[-1].tobytes())
    rng = np.random.default_rng(41)
    qv = rng.integers(0, len(vols), 9).tolist()
    qi = rng.uniform(0.2, 0.8, 9).astype(np.float32).tolist()
    arrivals = np.cumsum(rng.exponential(1e-4, 9)).tolist()
    fdims, frate = (21, 18, 11), 1.5
    fld = synth.field(synth.GAUSS, 34, fdims, 4, (3, 10, 0.5, 1))
    (tmp_path / "field.bin").write_bytes(fld.tobytes())
    scale = [1.0, 0.5, 2.0]
    (tmp_path / "params.json").write_text(json.dumps({
        "poolBytes": 50000 * 256, "vols": [{"dims": d, "rate": r} for d, r, _ in vols],
        "qv": qv, "qi": qi, "arrivals": arrivals, "scale": scale, "fieldDims": fdims, "fieldRate": frate}))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "engine_facade.js"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads((tmp_path / "out.json").read_text())
    verts = np.fromfile(tmp_path / "verts.bin", np.float32).reshape(-1, 4)
    off = 0
    for k in range(len(qv)):
        dims, rate, _ = vols[qv[k]]
        ref = oracle.MC(streams[qv[k]], 64 * rate, dims)
        nv = ref.compute_surface(np.float32(qi[k]))
        got = out["res"][k]
        assert got["numVertices"] == nv and got["vertexOffset"] == off, k
        want = ref.results()["verts_f32"]
        assert np.array_equal(verts[off:off + nv].view(np.uint32), want.view(np.uint32)), k
        assert out["served"][k]["numVertices"] == nv, k
        off += nv
    assert out["stats"]["numVolumes"] == 3 and out["stats"]["steps"] >= 2
    # world positions of query 0 (the renderer's placement, float32 steps)
    dims0 = vols[qv[0]][0]
    n0 = out["res"][0]["numVertices"]
    loc = verts[:n0]
    bid = loc[:, 3].view(np.uint32)
    nbx, nby = (dims0[0] + 3) // 4, (dims0[1] + 3) // 4
    bpos = np.stack([bid % nbx, (bid // nbx) % nby, bid // nbx // nby], 1)
    want = np.float32(scale) * (((4 * bpos).astype(np.float32) + loc[:, :3]) / np.float32(dims0)
                                - np.float32(0.5))
    world = np.fromfile(tmp_path / "world.bin", np.float32).reshape(-1, 3)
    assert np.array_equal(world.view(np.uint32), want.view(np.uint32))
    # the GPU encoder from Node: the host encoder's bytes, then two zero words
    enc = np.fromfile(tmp_path / "encoded.bin", np.uint8)
    host = synth.encode(fld, int(64 * frate))
    assert enc[:len(host)].tobytes() == host.tobytes() and not enc[len(host):].any()


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_async_facade_keeps_the_event_loop_running(tmp_path):
    """Engine.serve of requests arriving over a second runs on a worker
    thread: a 5 ms setInterval timer keeps ticking while it is pending.  The
    LRUCache.update perfTracker gets one entry per update under each of the
    reference's 11 lru* keys (counts equal to the oracle's), and operations
    submitted without awaiting complete in call order with the oracle's
    counts."""
    assert os.path.exists(ADDON), "N-API addon not built"
    dims, rate, cache = (48, 40, 36), 4, 200
    f = synth.field(synth.BLOBS, 51, dims, 6, (3, 10, 0.5, 1))
    s = synth.encode(f, 64 * rate)
    (tmp_path / "vol.bin").write_bytes(s.tobytes())
    rng = np.random.default_rng(52)
    n = 40
    arrivals = np.linspace(0.0, 1.0, n).tolist()
    isos = rng.uniform(0.3, 0.7, n).astype(np.float32).tolist()
    lru_total, lru_initial = 3000, 200
    (tmp_path / "params.json").write_text(json.dumps({
        "dims": dims, "rate": rate, "cache": cache, "poolBytes": 20000 * 256, "arrivals": arrivals,
        "isos": isos, "lruTotal": lru_total, "lruInitial": lru_initial}))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "async_timer.js"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads((tmp_path / "out.json").read_text())
    assert out["serveMs"] >= 900
    assert out["serveTicks"] >= 0.3 * out["serveMs"] / 5, out["serveTicks"]
    assert len(out["served"]) == n
    for k in range(n):
        ref = oracle.MC(s, 64 * rate, dims)
        assert out["served"][k]["numVertices"] == ref.compute_surface(np.float32(isos[k])), k
    keys = ["lruMarkNewItems", "lruScanNeededSlots", "lruNumNewItems", "lruCompactNewItems",
            "lruScanSlotsAvailable", "lruNumSlotsAvailable", "lruGrowCache", "lruCompactAvailableSlots",
            "lruPrepKeyValue", "lruTotalSortTime", "lruWriteNewItems"]
    perf = out["perf"]
    assert sorted(perf) == sorted(keys)
    assert all(len(perf[k]) == 2 for k in keys)
    cpu = oracle.LRU(lru_initial, 256, lru_total)
    size0 = len(cpu.state()[0])  # the constructor's slots, all available
    needed = np.zeros(lru_total, np.uint32)
    needed[::3] = 1
    n1, _ = cpu.update(needed)
    n2, _ = cpu.update(needed)
    assert perf["lruNumNewItems"] == [n1, n2] and n1 > 0 and n2 == 0
    assert out["updateNew"] == n1
    assert perf["lruMarkNewItems"][0] > 0 and perf["lruTotalSortTime"][0] > 0
    assert perf["lruGrowCache"][0] > 0  # 1000 new items into 200 slots
    assert perf["lruNumSlotsAvailable"][0] == size0 >= lru_initial
    assert all(perf[k][1] == 0 for k in keys[3:])  # no new items: as the reference
    ref = oracle.MC(s, 64 * rate, dims, initial_cache=cache)
    counts = [ref.compute_surface(np.float32(iso)) for iso in isos[:3]]
    assert out["surfaceVertices"] == counts[0]
    assert out["order"] == [["a", counts[1]], ["b", counts[2]]]
