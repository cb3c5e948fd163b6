"""The batched and served forms at their full sizes against the oracle.

Batched: 64 (volume, isovalue) queries in one engine step over Gaussian-blob
volumes of uneven size (128^3 rate 8, 256^3 rate 4, 384^3 rate 4, 512^3 rate
2; seeds 3-6), isovalues U[0.1, 0.9].  Served: a few hundred Poisson requests
over 8 resident blob volumes from 128^3 to 1024x1024x1080 (seeds 7-14, rates
2/4/8/2/4/8/2/2), half of them repeating a recent isovalue of their volume.
Every query's vertex count (and, batched, its active blocks and vertex offset)
equals the oracle's single computeSurface on that volume
(compressed_marching_cubes.js:580-809); vertices are compared bit for bit for
a seeded sample.  The volumes come from the same seeded generators as the
benchmark."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import dev, replay_dev, say  # noqa: F401

pytestmark = pytest.mark.gpu

THREADS = 16
BATCH = [((128, 128, 128), 8, 3), ((256, 256, 256), 4, 4), ((384, 384, 384), 4, 5),
         ((512, 512, 512), 2, 6)]
SERVED = [(d, r, 7 + i) for i, (d, r) in enumerate(zip(synth.SERVED_SIZES, synth.SERVED_RATES))]


def _volumes(spec):
    out = []
    for dims, rate, seed in spec:
        s, sha = synth.make_stream(synth.blob_workload(dims, seed, rate), threads=THREADS)
        say("volume", dims, "rate", rate, "seed", seed, "sha256", sha[:16])
        out.append((s, rate, dims))
    return out


class _Oracle:
    """One oracle pipeline per volume; results cached per (volume, isovalue)."""

    def __init__(self, vols):
        self.vols, self.mcs, self.memo = vols, {}, {}

    def query(self, v, iso, verts=False):
        key = (int(v), float(np.float32(iso)))
        hit = self.memo.get(key)
        if hit is not None and (hit[2] is not None or not verts):
            return hit
        if v not in self.mcs:
            s, rate, dims = self.vols[v]
            self.mcs[v] = oracle.MC(s, 64 * rate, dims, threads=THREADS)
        mc = self.mcs[v]
        nv = mc.compute_surface(float(np.float32(iso)))
        out = (nv, mc.stats["num_active"], mc.results()["verts_f32"] if verts else None)
        self.memo[key] = out
        return out


def _decoded_bytes(vols):
    return sum(oracle.nblocks(d) for _, _, d in vols) * 256


def test_batched_64_queries_over_uneven_volumes(replay_dev):
    """With graph replay on: the step is captured the second time and
    replayed the third, and every run equals the oracle."""
    dev = replay_dev
    vols = _volumes(BATCH)
    rng = np.random.default_rng(3)
    ids = rng.integers(0, len(vols), 64).astype(np.uint32)
    isos = rng.uniform(0.1, 0.9, 64).astype(np.float32)
    sample = set(np.random.default_rng(33).choice(64, 10, replace=False).tolist())
    eng = bcmc_amd.Engine(dev, 2 * _decoded_bytes(vols))
    for s, r, d in vols:
        eng.add_volume(s, r, d)
    orc = _Oracle(vols)
    for rep in range(3):  # same shape again: captured the second time, replayed the third
        say("batched step", rep)
        res = eng.run(ids, isos)
        off = 0
        for k in range(64):
            nv, na, verts = orc.query(ids[k], isos[k], verts=(rep == 0 and k in sample))
            assert res[k]["num_vertices"] == nv, (rep, k)
            assert res[k]["num_active_blocks"] == na, (rep, k)
            assert res[k]["vertex_offset"] == off, (rep, k)
            off += nv
            if verts is not None and rep == 0:
                got = eng.vertices(res[k])
                assert np.array_equal(got.view(np.uint32), verts.view(np.uint32)), k
        say("  vertices in the step", off)
    assert eng.stats()["graph_steps"] >= 1


def test_served_poisson_requests_over_8_volumes(dev):
    vols = _volumes(SERVED)
    rng = np.random.default_rng(7)
    n = 192
    ids, isos = bcmc_amd.request_trace(rng, n, len(vols))
    arrival = np.cumsum(rng.exponential(1.0 / 1000.0, n))
    eng = bcmc_amd.Engine(dev, 2 * _decoded_bytes(vols))
    for s, r, d in vols:
        eng.add_volume(s, r, d)
    say("serving", n, "requests at 1000 req/s")
    out = eng.serve(arrival, ids, isos, max_batch=64, vertex_budget_bytes=200 << 30)
    assert np.all(out["start_s"] >= out["arrival_s"]) and np.all(out["finish_s"] >= out["start_s"])
    orc = _Oracle(vols)
    for i in np.argsort(ids, kind="stable"):  # volume by volume: one oracle pipeline at a time
        nv, _, _ = orc.query(ids[i], isos[i])
        assert out["num_vertices"][i] == nv, i
    say("  counts equal; mean batch", float(out["batch_size"].mean()))
    # vertices of a sample of the served (volume, isovalue) pairs, answered by
    # the same engine after the trace (pool state churned by it)
    pick = np.random.default_rng(8).choice(n, 6, replace=False)
    pick = [i for i in pick if ids[i] != len(vols) - 1][:4] + \
        [int(np.flatnonzero(ids == len(vols) - 1)[0])]
    res = eng.run(ids[pick], isos[pick])
    for k, i in enumerate(pick):
        nv, _, verts = orc.query(ids[i], isos[i], verts=True)
        assert res[k]["num_vertices"] == nv
        got = eng.vertices(res[k])
        assert np.array_equal(got.view(np.uint32), verts.view(np.uint32)), i
