"""Batched form on the GPU: every query of a multi-volume step equals the
oracle's single-query computeSurface (counts and FP32 vertices bit-exact),
independent of the other queries in the step and of the pool state."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from bcmc_amd import synth
from gpu_common import dev, replay_dev, say  # noqa: F401

pytestmark = pytest.mark.gpu

VOLS = [((32, 32, 32), 8, 3), ((48, 40, 36), 4, 4), ((24, 28, 20), 2, 5), ((13, 10, 7), 4, 6)]


def _streams():
    out = []
    for dims, rate, seed in VOLS:
        f = synth.field(synth.BLOBS, seed, dims, 6, (3, 10, 0.5, 1))
        out.append((synth.encode(f, 64 * rate), rate, dims))
    return out


def _oracle(s, rate, dims, iso):
    ref = oracle.MC(s, 64 * rate, dims)
    nv = ref.compute_surface(iso)
    return nv, ref.results()["verts_f32"]


@pytest.mark.parametrize("pool_blocks", [100000, 3000])
def test_batched_queries_match_single_queries(dev, pool_blocks):
    say("engine pool", pool_blocks)
    streams = _streams()
    eng = bcmc_amd.Engine(dev, pool_blocks * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    rng = np.random.default_rng(pool_blocks)
    cache = {}
    for step in range(4):
        nq = int(rng.integers(1, 12))
        qv = rng.integers(0, len(ids), nq)
        qi = rng.uniform(0.15, 0.85, nq).astype(np.float32)
        res = eng.run(qv, qi)
        for k in range(nq):
            s, r, d = streams[qv[k]]
            key = (int(qv[k]), float(qi[k]))
            if key not in cache:
                cache[key] = _oracle(s, r, d, float(qi[k]))
            nv, verts = cache[key]
            assert res[k]["num_vertices"] == nv, (step, k)
            got = eng.vertices(res[k])
            assert np.array_equal(got.view(np.uint32), verts.view(np.uint32)), (step, k)
    st = eng.stats()
    assert st["pool_used"] <= st["pool_slots"]


def test_pool_too_small_is_an_error(dev):
    s, r, d = _streams()[1]
    eng = bcmc_amd.Engine(dev, 64 * 256)
    vid = eng.add_volume(s, r, d)
    with pytest.raises(bcmc_amd.BcmcError) as e:
        eng.run([vid], [0.5])
    assert e.value.code == -3


def test_served_trace_results_match_single_queries(dev):
    say("serve trace")
    streams = _streams()
    eng = bcmc_amd.Engine(dev, 200000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    rng = np.random.default_rng(11)
    n = 60
    arrival = np.cumsum(rng.exponential(1.0 / 4000.0, n))  # 4000 req/s: requests pile up
    vol = rng.integers(0, len(ids), n)
    iso = rng.uniform(0.15, 0.85, n).astype(np.float32)
    rep = rng.uniform(0, 1, n) < 0.5
    for i in range(1, n):  # half the requests repeat a recent isovalue of their volume
        prev = [j for j in range(max(0, i - 16), i) if vol[j] == vol[i]]
        if rep[i] and prev:
            iso[i] = iso[prev[-1]]
    out = eng.serve(arrival, vol, iso, max_batch=16)
    assert np.all(out["finish_s"] >= out["arrival_s"]) and np.all(out["start_s"] >= out["arrival_s"])
    assert out["batch_size"].max() > 1  # continuous batching actually batched
    cache = {}
    for i in range(n):
        key = (int(vol[i]), float(iso[i]))
        if key not in cache:
            s, r, d = streams[vol[i]]
            cache[key] = _oracle(s, r, d, float(iso[i]))[0]
        assert out["num_vertices"][i] == cache[key], i


@pytest.mark.parametrize("pool_blocks", [200000, 1200])
def test_coalesced_serving_matches_separate_queries(dev, pool_blocks):
    """Requests all waiting at once, most of them exact repeats: served with
    coalescing (equal requests of a step answered by one query) and with
    every request a query of its own, every request gets the oracle's count;
    repeats take no pool room, so coalescing needs no more steps.  A pool of
    1200 blocks holds one step of a few distinct requests only."""
    say("coalesced serving, pool", pool_blocks)
    streams = _streams()
    eng = bcmc_amd.Engine(dev, pool_blocks * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    rng = np.random.default_rng(pool_blocks + 5)
    n = 96
    vol = rng.integers(0, len(ids), n).astype(np.uint32)
    choices = rng.uniform(0.2, 0.8, (len(ids), 5)).astype(np.float32)
    iso = choices[vol, rng.integers(0, 5, n)]
    out_c = eng.serve(np.zeros(n), vol, iso, max_batch=32)
    out_s = eng.serve(np.zeros(n), vol, iso, max_batch=32, coalesce=False)
    cache = {}
    for i in range(n):
        key = (int(vol[i]), float(iso[i]))
        if key not in cache:
            s, r, d = streams[vol[i]]
            cache[key] = _oracle(s, r, d, float(iso[i]))[0]
        assert out_c["num_vertices"][i] == cache[key], ("coalesced", i)
        assert out_s["num_vertices"][i] == cache[key], ("separate", i)
    assert out_c["step"].max() <= out_s["step"].max()


def _piled_trace(streams, n, seed):
    """n requests at 4000 req/s (they pile up behind each step) over every
    volume, isovalues where the volumes have surface."""
    rng = np.random.default_rng(seed)
    arrival = np.cumsum(rng.exponential(1.0 / 4000.0, n))
    vol = rng.integers(0, len(streams), n).astype(np.uint32)
    iso = rng.uniform(0.3, 0.7, n).astype(np.float32)
    return arrival, vol, iso


def _oracle_active(s, rate, dims, iso):
    ref = oracle.MC(s, 64 * rate, dims)
    nv = ref.compute_surface(iso)
    return nv, ref.results()["active_ids"].copy()


def test_admission_bound_covers_active_union(dev):
    """The served loop's pool bound for a set of isovalues on a volume is at
    least the union of their active blocks (the oracle's) and at most the
    volume's blocks; it never shrinks as isovalues are added, repeats add
    nothing and no isovalue means no block."""
    streams = _streams()
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    rng = np.random.default_rng(41)
    for k, (s, r, d) in enumerate(streams):
        say("admission bound, volume", k)
        nb = oracle.nblocks(d)
        isos = rng.uniform(0.05, 0.95, 12).astype(np.float32)
        act = [set(_oracle_active(s, r, d, float(x))[1].tolist()) for x in isos]
        last = 0
        for m in range(1, 13):
            union = set().union(*act[:m])
            b = eng.block_bound(ids[k], isos[:m])
            assert len(union) <= b <= nb, (k, m, len(union), b)
            assert b >= last, (k, m)
            last = b
        assert eng.block_bound(ids[k], np.concatenate([isos[:3], isos[:3]])) == \
            eng.block_bound(ids[k], isos[:3])
        assert eng.block_bound(ids[k], []) == 0


def test_served_trace_waits_for_pool_room(dev):
    """A pool that holds every single request's active blocks but not those
    of a step of piled-up requests: the served loop admits what fits (an
    empty step frees the previous step's slots when the head needs them)
    instead of failing, and every request still equals the oracle."""
    streams = _streams()
    n = 48
    arrival, vol, iso = _piled_trace(streams, n, 31)
    act = {}
    for i in range(n):
        key = (int(vol[i]), float(iso[i]))
        if key not in act:
            s, r, d = streams[vol[i]]
            act[key] = _oracle_active(s, r, d, float(iso[i]))
    one = max(len(a) for _, a in act.values())
    pool = one + 16
    # the union of active blocks of the first 16 requests, per volume
    union = sum(len(set().union(*[set(act[(int(vol[i]), float(iso[i]))][1].tolist())
                                  for i in range(16) if vol[i] == v] or [set()]))
                for v in range(len(streams)))
    say("pool", pool, "largest request", one, "first 16 requests", union)
    assert union > pool
    eng = bcmc_amd.Engine(dev, pool * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    assert ids == list(range(len(streams)))
    with pytest.raises(bcmc_amd.BcmcError) as e:  # as one step it does not fit
        eng.run(vol[:16], iso[:16])
    assert e.value.code == -3
    eng = bcmc_amd.Engine(dev, pool * 256)
    for s, r, d in streams:
        eng.add_volume(s, r, d)
    out = eng.serve(arrival, vol, iso, max_batch=16)
    for i in range(n):
        assert out["num_vertices"][i] == act[(int(vol[i]), float(iso[i]))][0], i
    assert np.all(np.diff(out["step"]) >= 0)  # FIFO
    assert out["batch_size"].max() > 1
    st = eng.stats()
    assert st["pool_used"] <= st["pool_slots"] == pool
    # the vertices of the last step are there and exact
    last = out["step"] == out["step"].max()
    k = int(np.nonzero(last)[0][0])
    s, r, d = streams[vol[k]]
    res = eng.run([vol[k]], [iso[k]])[0]
    nv, verts = _oracle(s, r, d, float(iso[k]))
    assert res["num_vertices"] == nv
    assert np.array_equal(eng.vertices(res).view(np.uint32), verts.view(np.uint32))


def test_served_trace_defers_past_the_vertex_budget(dev):
    """An output budget that holds the largest single request but not a
    step's worth: each step emits the leading requests that fit (exact vertex
    counts), the rest wait at the head of the queue; a request larger than
    the budget alone is an error."""
    streams = _streams()
    n = 40
    arrival, vol, iso = _piled_trace(streams, n, 37)
    nv = {}
    for i in range(n):
        key = (int(vol[i]), float(iso[i]))
        if key not in nv:
            s, r, d = streams[vol[i]]
            nv[key] = _oracle(s, r, d, float(iso[i]))[0]
    counts = np.array([nv[(int(vol[i]), float(iso[i]))] for i in range(n)], np.uint64)
    budget = int(counts.max()) * 16 + 16
    say("budget", budget, "bytes; all requests", int(counts.sum()) * 16)
    assert counts.sum() * 16 > 2 * budget
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    for s, r, d in streams:
        eng.add_volume(s, r, d)
    out = eng.serve(arrival, vol, iso, max_batch=16, vertex_budget_bytes=budget)
    assert np.array_equal(out["num_vertices"].astype(np.uint64), counts)
    steps = out["step"]
    assert np.all(np.diff(steps) >= 0)
    for st_ in np.unique(steps):
        sel = steps == st_
        assert int(counts[sel].sum()) * 16 <= budget, st_
        assert np.all(out["batch_size"][sel] == sel.sum()), st_
    assert out["batch_size"].max() > 1
    # one request larger than the budget alone
    big = int(np.argmax(counts))
    with pytest.raises(bcmc_amd.BcmcError) as e:
        eng.serve([0.0], [vol[big]], [iso[big]], vertex_budget_bytes=int(counts[big]) * 16 - 16)
    assert e.value.code == -3


def test_coalesced_requests_under_the_vertex_budget(dev):
    """Coalesced requests under an output budget: the budget counts each
    distinct query of a step once, admission stays a prefix of the requests
    (a repeat of an admitted query is admitted with it), and every request
    gets the oracle's count."""
    streams = _streams()
    rng = np.random.default_rng(43)
    n = 48
    vol = rng.integers(0, len(streams), n).astype(np.uint32)
    choices = rng.uniform(0.3, 0.7, (len(streams), 4)).astype(np.float32)
    iso = choices[vol, rng.integers(0, 4, n)]
    nv = {}
    for i in range(n):
        key = (int(vol[i]), float(iso[i]))
        if key not in nv:
            s, r, d = streams[vol[i]]
            nv[key] = _oracle(s, r, d, float(iso[i]))[0]
    counts = np.array([nv[(int(vol[i]), float(iso[i]))] for i in range(n)], np.uint64)
    budget = 2 * int(counts.max()) * 16 + 16
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    for s, r, d in streams:
        eng.add_volume(s, r, d)
    out = eng.serve(np.zeros(n), vol, iso, max_batch=16, vertex_budget_bytes=budget)
    assert np.array_equal(out["num_vertices"].astype(np.uint64), counts)
    steps = out["step"]
    assert np.all(np.diff(steps) >= 0)
    shared = False
    for st_ in np.unique(steps):
        sel = np.flatnonzero(steps == st_)
        keys = {(int(vol[i]), float(iso[i])) for i in sel}
        assert sum(nv[k] for k in keys) * 16 <= budget, st_
        shared |= len(keys) < len(sel)
    assert shared  # some step answered a repeat by an earlier request's query
    assert steps.max() + 1 < n  # steps took several requests


def test_result_is_the_same_alone_and_in_the_mix(dev):
    """A query's vertices are bit-identical whether it runs alone or among
    other queries of other volumes (and after the pool has been churned)."""
    streams = _streams()
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    probe_v, probe_iso = ids[1], 0.47
    alone = eng.run([probe_v], [probe_iso])[0]
    want = eng.vertices(alone).copy()
    rng = np.random.default_rng(5)
    for trial in range(3):
        nq = 9
        qv = list(rng.integers(0, len(ids), nq))
        qi = list(rng.uniform(0.15, 0.85, nq).astype(np.float32))
        pos = int(rng.integers(0, nq + 1))
        qv.insert(pos, probe_v)
        qi.insert(pos, np.float32(probe_iso))
        res = eng.run(qv, qi)
        got = eng.vertices(res[pos])
        assert res[pos]["num_vertices"] == alone["num_vertices"], trial
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), trial


def test_many_queries_in_one_step(dev):
    """More queries than one workgroup of the chunk table, with empty queries
    (isovalues outside the field) between them; counts equal the oracle's."""
    streams = _streams()
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    rng = np.random.default_rng(21)
    nq = 300
    qv = rng.integers(0, len(ids), nq)
    qi = rng.uniform(0.15, 0.85, nq).astype(np.float32)
    qi[rng.uniform(0, 1, nq) < 0.2] = 5.0  # no active blocks
    res = eng.run(qv, qi)
    cache = {}
    off = 0
    for k in range(nq):
        key = (int(qv[k]), float(qi[k]))
        if key not in cache:
            s, r, d = streams[qv[k]]
            cache[key] = _oracle(s, r, d, float(qi[k]))
        nv, verts = cache[key]
        assert res[k]["num_vertices"] == nv, k
        assert res[k]["vertex_offset"] == off, k
        off += nv
        if k % 37 == 0:
            got = eng.vertices(res[k])
            assert np.array_equal(got.view(np.uint32), verts.view(np.uint32)), k


@pytest.mark.parametrize("pool_blocks", [100000, 3000])
def test_repeated_step_shapes_replay_a_graph(replay_dev, pool_blocks):
    """A step shape (volumes in query order) seen a second time is captured
    into a HIP graph and replayed whenever it comes again; replayed steps with
    new isovalues, after a step of another shape and (small pool) with the
    eviction chain inside the graph all still equal the oracle's single
    queries."""
    say("engine graph replay, pool", pool_blocks)
    streams = _streams()
    eng = bcmc_amd.Engine(replay_dev, pool_blocks * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    shape_a = np.array([ids[0], ids[1], ids[1], ids[2], ids[3], ids[0]], np.uint32)
    shape_b = shape_a[1:4]
    rng = np.random.default_rng(pool_blocks + 1)
    cache = {}
    plan = [shape_a] * 5 + [shape_b] + [shape_a] * 3
    for step, qv in enumerate(plan):
        qi = rng.uniform(0.15, 0.85, len(qv)).astype(np.float32)
        res = eng.run(qv, qi)
        for k in range(len(qv)):
            s, r, d = streams[qv[k]]
            key = (int(qv[k]), float(qi[k]))
            if key not in cache:
                cache[key] = _oracle(s, r, d, float(qi[k]))
            nv, verts = cache[key]
            assert res[k]["num_vertices"] == nv, (step, k)
            got = eng.vertices(res[k])
            assert np.array_equal(got.view(np.uint32), verts.view(np.uint32)), (step, k)
    st = eng.stats()
    assert st["steps"] == len(plan)
    # vertex-buffer growth re-captures, so not every repeat replays
    assert st["graph_steps"] >= 2, st


def test_steps_launch_directly_by_default(dev):
    """Without graph replay on the device, no step replays a graph."""
    assert not dev.graph_replay
    streams = _streams()
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    for _ in range(3):
        res = eng.run([ids[0], ids[2]], [0.4, 0.6])
        for k, (v, iso) in enumerate(((0, 0.4), (2, 0.6))):
            assert res[k]["num_vertices"] == _oracle(*streams[v], iso)[0]
    st = eng.stats()
    assert st["steps"] == 3 and st["graph_steps"] == 0, st


def test_nonfinite_and_tied_isovalues_in_steps(dev):
    """NaN and +-inf isovalues (every comparison of the active test false: no
    active block, no vertex) and isovalues equal to voxel values (a block's
    minimum, another block's maximum), beside ordinary ones on every volume:
    each query of a batched step, in two orders, equals the oracle's single
    query (counts and vertices bit-exact); served (coalesced and not), every
    request gets the oracle's count; the admission bound of such a set still
    covers the union of its active blocks."""
    say("engine non-finite and tied isovalues")
    streams = _streams()
    eng = bcmc_amd.Engine(dev, 100000 * 256)
    ids = [eng.add_volume(s, r, d) for s, r, d in streams]
    qv, qi, want = [], [], []
    for v, (s, r, d) in zip(ids, streams):
        rg = np.asarray(oracle.MC(s, 64 * r, d).ranges(), np.float32)
        ties = [float(rg[len(rg) // 2, 0]), float(rg[len(rg) // 3, 1])]
        for iso in [float("nan"), float("inf"), -float("inf"), *ties, 0.5]:
            qv.append(v)
            qi.append(iso)
            want.append(_oracle(s, r, d, iso))
            if not np.isfinite(iso):
                assert want[-1][0] == 0
        active = set()
        for iso in ties:
            active |= set(_oracle_active(s, r, d, iso)[1].tolist())
        assert eng.block_bound(v, [float("nan"), float("inf"), -float("inf"), *ties]) >= len(active)
    qv = np.array(qv, np.uint32)
    qi = np.array(qi, np.float32)
    order = np.random.default_rng(17).permutation(len(qv))
    for perm in (np.arange(len(qv)), order):
        res = eng.run(qv[perm], qi[perm])
        for k, j in enumerate(perm):
            nv, verts = want[j]
            assert res[k]["num_vertices"] == nv, (k, qi[j])
            got = eng.vertices(res[k])
            assert np.array_equal(got.view(np.uint32), verts.view(np.uint32)), (k, qi[j])
    for coalesce in (True, False):
        out = eng.serve(np.zeros(len(qv)), qv[order], qi[order], max_batch=8, coalesce=coalesce)
        for k, j in enumerate(order):
            assert out["num_vertices"][k] == want[j][0], (coalesce, k, qi[j])
