#!/usr/bin/env python3
"""Benchmark of the block-compressed marching-cubes path on one MI355X.

Metric: iso-queries/s (computeSurface calls per second) of the plain per-query
pipeline on the 512^3 turbulence-like workload: sum of 512 random Fourier
modes (|k| in [1,64], amplitude |k|^-5/6, seed 1) normalised to [0,1], zfp
fixed rate 4 bits/value, LRU cache initially 25% of the blocks, a sequential
sweep of 8 isovalues linspace(0.2, 0.8).  One step = cache reset + the 8-query
sweep.  The compressed volume is resident in HBM before timing starts.

Also reported: the ZFP decode rate of the cache-decode kernel and of the
load-time range pass, the HBM roofline fraction of the dominant kernel, p50/p99
query latency, and the CPU oracle (the reference method restated in C, OpenMP)
timed on this host on a bounded sample of the same sweep.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "webgpu-bcmc_amd", "python"))

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


STAGE_KERNEL = {"computeVertices": "mc_emit_kernel<0>", "computeBlockHasVertices": "mc_count_kernel",
                "decompression": "zfp_kernel<4,3>", "computeActiveBlocks": "scan_kernel<2,ActiveOp>"}


def profiled_traffic(workload, stage):
    """HBM bytes per launch of the stage's kernel on this workload, from the
    committed FETCH_SIZE/WRITE_SIZE counter passes (copy-calibrated), or None."""
    import re
    kernel = STAGE_KERNEL.get(stage)
    path = os.path.join(ROOT, "profiles", f"{kernel.split('<')[0]}_{workload}_hbm_traffic.txt") \
        if kernel else None
    if not path or not os.path.exists(path):
        return None, None
    m = re.search(rf"^traffic_bytes_per_launch {re.escape(kernel)}: (\d+)$", open(path).read(), re.M)
    return (int(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def host_cpu():
    """The host's CPU model and the threads the CPU baseline uses: the CPUs
    this process may run on (nproc), capped by OMP_NUM_THREADS when set."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(nproc, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else nproc
    return {"model": model, "nproc": nproc, "cpu_count": os.cpu_count(),
            "omp_num_threads": omp, "threads": threads}


def stage_bytes(s, rate, n_blocks):
    """Algorithmic HBM bytes per stage of one query: what each stage must read
    and write at least once (ids, ranges, slot words, block data, ghost voxels,
    vertices), from the query's counts."""
    a, d, h, v = s["numActiveBlocks"], s["numBlocksDecompressed"], s["numBlocksWithVertices"], \
        s["numVertices"]
    return {
        # dilated ranges 8N + item slots 4N + flags 4N + active ids 4A + miss ids 4D
        "computeActiveBlocks": 16 * n_blocks + 4 * a + 4 * d,
        # payload 8r + block id 4 + slot 4 + 64 floats 256, per decoded block
        "decompression": d * (8 * rate + 4 + 4 + 256),
        # count: active id 4 + tile range 8 + count 4 per active block; per
        # candidate its slots/flags 60, own block 256 and ghost voxels <= 244
        # (the candidates, which the pass does not report, include the H blocks
        # with vertices: H is used, a lower bound)
        "computeBlockHasVertices": 16 * a + 560 * h,
        # chunk totals read and prefixes written
        "compactBlocksWithVertices": 20 * ((a + 255) // 256),
        # emit: count + id 8 per active block; per block with vertices its
        # slots/flags 60, own block 256, ghost voxels <= 244, list entry 4 + 8;
        # 16 B per vertex
        "computeVertices": 8 * a + 572 * h + 16 * v,
    }


BATCH_SIZES = [(128, 128, 128), (256, 256, 256), (384, 384, 384), (512, 512, 512)]
BATCH_RATES = [8, 4, 4, 2]


def plain_replay(bcmc_amd, dev, vols, ids, isos, cache_frac=0.25):
    """The plain per-query pipeline on a request sequence: one
    CompressedMarchingCubes per volume (resident, its own LRU cache), queries
    answered one at a time.  Returns per-query service times (s) and vertex counts."""
    mcs = []
    for stream, rate, dims in vols:
        nb = int(np.prod([(d + 3) // 4 for d in dims]))
        mc = bcmc_amd.CompressedMarchingCubes(dev)
        mc.setCompressedVolume(stream, rate, dims, (1, 1, 1), max(32, int(cache_frac * nb)))
        mcs.append(mc)
    # warm: one query per volume (allocations, first growth)
    for mc in mcs:
        mc.computeSurface(0.5)
    dev.sync()
    t_q, nv = np.empty(len(ids)), np.empty(len(ids), np.uint64)
    for k, (v, iso) in enumerate(zip(ids, isos)):
        t = time.perf_counter()
        nv[k] = mcs[v].computeSurface(float(iso))
        t_q[k] = time.perf_counter() - t
    del mcs
    return t_q, nv


def served_bench(bcmc_amd, synth, dev, args):
    t0 = time.time()
    vols = synth.served_volumes()
    log(f"bench: served volumes generated in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(20240607)
    ids, isos = bcmc_amd.request_trace(rng, args.served_queries, len(vols))
    t_plain, nv_plain = plain_replay(bcmc_amd, dev, vols, ids, isos)
    plain_qps = len(ids) / t_plain.sum()
    log(f"bench: plain replay {len(ids)} queries: {plain_qps:.1f} q/s")
    decoded = sum(int(np.prod([(d + 3) // 4 for d in v[2]])) for v in vols) * 256
    eng = bcmc_amd.Engine(dev, min(args.served_pool_gib << 30, 2 * decoded))
    for s, rate, dims in vols:
        eng.add_volume(s, rate, dims)
    # admission by a 200 GB HBM budget for decoded blocks and output
    # vertices: the pool takes its share, the vertex budget the rest
    hbm_budget = 200 * 10**9
    budget = hbm_budget - eng.stats()["pool_bytes"]
    # warm-up: a saturation pass over requests of their own draw (buffer
    # growth, graph capture and first decodes happen here, not in the timed
    # trace)
    wid, wiso = bcmc_amd.request_trace(np.random.default_rng(99), args.served_batch * 4, len(vols))
    eng.serve(np.zeros(len(wid)), wid, wiso, max_batch=args.served_batch, vertex_budget_bytes=budget)
    dev.sync()
    sat = eng.serve(np.zeros(len(ids)), ids, isos, max_batch=args.served_batch,
                    vertex_budget_bytes=budget)
    makespan = float(sat["finish_s"].max())
    served_qps = len(ids) / makespan
    exact = bool(np.array_equal(sat["num_vertices"].astype(np.uint64), nv_plain))
    lat = sat["finish_s"] - sat["arrival_s"]
    # requests answered by an equal request's query (same step, volume and
    # isovalue)
    seen, shared = set(), 0
    for st_, v_, x_ in zip(sat["step"], ids, isos):
        key = (int(st_), int(v_), float(x_))
        shared += key in seen
        seen.add(key)
    log(f"bench: served saturation {served_qps:.1f} q/s in {int(sat['step'].max()) + 1} steps "
        f"({shared} requests coalesced); counts equal to plain: {exact}")
    # the same trace with every request a query of its own
    sep = eng.serve(np.zeros(len(ids)), ids, isos, max_batch=args.served_batch,
                    vertex_budget_bytes=budget, coalesce=False)
    sep_qps = len(ids) / float(sep["finish_s"].max())
    sep_exact = bool(np.array_equal(sep["num_vertices"].astype(np.uint64), nv_plain))
    sep_lat = sep["finish_s"] - sep["arrival_s"]
    log(f"bench: served saturation without coalescing {sep_qps:.1f} q/s; counts equal to plain: "
        f"{sep_exact}")
    out = {"workload": "8 resident blob volumes 128^3..1024x1024x1080 (seeds 7-14, zfp rates "
                       "2/4/8/2/4/8/2/2), isovalues U[0.1,0.9] with 50% repeats of the last 16 of "
                       "the same volume",
           "queries": len(ids), "plain_qps": round(plain_qps, 2),
           "plain_latency_ms": {"p50": round(float(np.percentile(t_plain, 50)) * 1e3, 3),
                                "p99": round(float(np.percentile(t_plain, 99)) * 1e3, 3)},
           "served_qps": round(served_qps, 2), "served_vs_plain": round(served_qps / plain_qps, 3),
           "saturation_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                     "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
           "counts_equal_plain": exact, "steps": int(sat["step"].max()) + 1,
           "coalesced_requests": int(shared),
           "no_coalesce": {"served_qps": round(sep_qps, 2),
                           "served_vs_plain": round(sep_qps / plain_qps, 3),
                           "saturation_latency_ms": {
                               "p50": round(float(np.percentile(sep_lat, 50)) * 1e3, 3),
                               "p99": round(float(np.percentile(sep_lat, 99)) * 1e3, 3)},
                           "counts_equal_plain": sep_exact, "steps": int(sep["step"].max()) + 1},
           "vertices_total": int(nv_plain.sum()),
           "mean_batch": round(float(sat["batch_size"].mean()), 1),
           "hbm_budget_bytes": hbm_budget, "vertex_budget_bytes": budget,
           "pool_bytes": eng.stats()["pool_bytes"], "max_batch": args.served_batch, "poisson": []}
    st = eng.stats()
    out["engine_steps"], out["graph_replayed_steps"] = st["steps"], st["graph_steps"]
    del eng
    # Poisson legs: fresh requests (their own draws, disjoint from the
    # saturation trace), on an engine whose pool holds what the plain
    # replay's caches hold (25% of every volume's blocks), cold at the first
    # leg, so that requests miss and evict as they would in service
    pool25 = sum(max(32, int(0.25 * int(np.prod([(d + 3) // 4 for d in v[2]])))) for v in vols) * 256
    eng = bcmc_amd.Engine(dev, pool25)
    for s, rate, dims in vols:
        eng.add_volume(s, rate, dims)
    out["poisson_pool_bytes"] = eng.stats()["pool_bytes"]
    budget = hbm_budget - out["poisson_pool_bytes"]
    leg_rng = np.random.default_rng(77)
    for rate in args.served_rates:
        n = min(args.served_queries, max(64, int(rate * args.served_seconds)))
        lid, liso = bcmc_amd.request_trace(leg_rng, n, len(vols))
        arr = np.cumsum(leg_rng.exponential(1.0 / rate, n))
        r = eng.serve(arr, lid, liso, max_batch=args.served_batch, vertex_budget_bytes=budget)
        lat = r["finish_s"] - r["arrival_s"]
        span = float(r["finish_s"].max() - arr[0])
        out["poisson"].append({"rate_rps": rate, "queries": n, "qps": round(n / span, 2),
                               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                               "mean_batch": round(float(r["batch_size"].mean()), 2),
                               "steps": int(r["step"].max()) + 1})
        log(f"bench: served poisson {rate} req/s: p50 {out['poisson'][-1]['p50_ms']} ms "
            f"p99 {out['poisson'][-1]['p99_ms']} ms, mean batch {out['poisson'][-1]['mean_batch']}")
    # the saturation trace again, through the pool of the plain replay's size
    sat = eng.serve(np.zeros(len(ids)), ids, isos, max_batch=args.served_batch,
                    vertex_budget_bytes=budget)
    q25 = len(ids) / float(sat["finish_s"].max())
    out["pool25"] = {"served_qps": round(q25, 2), "served_vs_plain": round(q25 / plain_qps, 3),
                     "steps": int(sat["step"].max()) + 1,
                     "counts_equal_plain": bool(np.array_equal(sat["num_vertices"].astype(np.uint64),
                                                               nv_plain))}
    log(f"bench: served saturation, pool of the plain caches' size: {q25:.1f} q/s")
    del eng
    return out


def batched_bench(bcmc_amd, synth, dev, args):
    vols = []
    for i, (dims, rate) in enumerate(zip(BATCH_SIZES, BATCH_RATES)):
        s, _ = synth.make_stream(synth.blob_workload(dims, 3 + i, rate))
        vols.append((s, rate, dims))
    rng = np.random.default_rng(3)
    ids = rng.integers(0, len(vols), 64).astype(np.uint32)
    isos = rng.uniform(0.1, 0.9, 64).astype(np.float32)
    t_plain, nv_plain = plain_replay(bcmc_amd, dev, vols, ids, isos)
    decoded = sum(int(np.prod([(d + 3) // 4 for d in v[2]])) for v in vols) * 256
    eng = bcmc_amd.Engine(dev, 2 * decoded)
    for s, rate, dims in vols:
        eng.add_volume(s, rate, dims)
    eng.run(ids, isos)  # warm (pool fill, buffer sizes)
    dev.sync()
    reps = []
    for _ in range(args.batched_reps):
        t = time.perf_counter()
        res = eng.run(ids, isos)
        reps.append(time.perf_counter() - t)
    nv = np.array([r["num_vertices"] for r in res], np.uint64)
    step = float(np.median(reps))
    out = {"workload": "64 queries over blob volumes 128^3 r8, 256^3 r4, 384^3 r4, 512^3 r2 "
                       "(seeds 3-6), isovalues U[0.1,0.9], one engine step",
           "step_ms": round(step * 1e3, 3), "qps": round(64 / step, 2),
           "plain_qps": round(64 / t_plain.sum(), 2),
           "batched_vs_plain": round(t_plain.sum() / step, 3),
           "counts_equal_plain": bool(np.array_equal(nv, nv_plain)),
           "graph_replayed_steps": eng.stats()["graph_steps"], "timed_steps": len(reps)}
    log(f"bench: batched step {out['step_ms']} ms ({out['qps']} q/s; plain {out['plain_qps']} q/s)")
    del eng
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fourier512_r4")
    ap.add_argument("--cpu-sweeps", type=int, default=8,
                    help="sweeps of the 8 isovalues timed on the CPU oracle (0 = skip)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device copy kernel of this size (HBM counter calibration)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--served-queries", type=int, default=2048)
    ap.add_argument("--served-batch", type=int, default=256)
    ap.add_argument("--served-pool-gib", type=int, default=32)
    ap.add_argument("--served-rates", type=float, nargs="*", default=[10, 30, 100, 300, 1000, 2000, 5000, 10000])
    ap.add_argument("--served-seconds", type=float, default=1.0)
    ap.add_argument("--batched-reps", type=int, default=5)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing candidates first")
    ap.add_argument("--graph-replay", action="store_true",
                    help="replay captured HIP graphs of the query chains instead of launching "
                         "their kernels directly (the default, measured faster)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard the volume by z-slabs of blocks, or the served "
              "volumes by GPU; that path is not built)", file=sys.stderr)
        sys.exit(2)

    import bcmc_amd
    from bcmc_amd import synth

    w = synth.WORKLOADS[args.workload]
    dims, rate = w["dims"], w["rate"]
    n_blocks = int(np.prod([(d + 3) // 4 for d in dims]))
    t0 = time.time()
    stream, sha = synth.make_stream(w)
    log(f"bench: generated {args.workload} {dims} rate {rate}: {stream.nbytes} B sha256 {sha[:16]} "
        f"in {time.time() - t0:.1f}s")

    dev = bcmc_amd.Device(0)
    dev.set_graph_replay(args.graph_replay)
    info = dev.info()
    tuning = dev.tuning() if args.no_autotune else dev.autotune()
    log(f"bench: launch shapes (workgroups per CU) {tuning}")
    cache0 = n_blocks // 4
    isos = np.linspace(0.2, 0.8, 8).astype(np.float32)

    # load-time range pass (full-volume decode) timed by itself
    sbuf = dev.buffer_from(np.concatenate([stream, np.zeros(16, np.uint8)]))
    bcmc_amd.lib().bcmc_ctx_enable_timing(dev.h, 1)
    range_ms = []
    for _ in range(5):
        bcmc_amd.block_ranges(dev, sbuf, 64 * rate, dims)
        range_ms.append(bcmc_amd.lib().bcmc_ctx_last_kernel_ms(dev.h))
    bcmc_amd.lib().bcmc_ctx_enable_timing(dev.h, 0)
    range_t = float(np.median(range_ms[1:]))
    del sbuf

    # ingest: the GPU fixed-rate encoder on the 256^3 16-Gaussian field (seed 0)
    # at rate 8, timed by itself (zeroing of the stream included) and checked
    # byte for byte against the host encoder
    ew = synth.WORKLOADS["gauss256_r8"]
    efield = synth.field(ew["kind"], ew["seed"], ew["dims"], ew["count"], ew["params"])
    fbuf = dev.buffer_from(efield)
    ebits = 64 * ew["rate"]
    ebytes = bcmc_amd.stream_bytes(ew["dims"], ebits)
    bcmc_amd.lib().bcmc_ctx_enable_timing(dev.h, 1)
    enc_ms = []
    for _ in range(5):
        ebuf = bcmc_amd.encode(dev, fbuf, ew["dims"], ebits)
        enc_ms.append(bcmc_amd.lib().bcmc_ctx_last_kernel_ms(dev.h))
    bcmc_amd.lib().bcmc_ctx_enable_timing(dev.h, 0)
    enc_t = float(np.median(enc_ms[1:]))
    enc_equal = bool(np.array_equal(ebuf.download(np.uint8, ebytes), synth.encode(efield, ebits)))
    encode_out = {"workload": "gauss256_r8: 256^3 sum of 16 Gaussians (seed 0), rate 8",
                  "ms": round(enc_t, 4), "blocks": int(np.prod([d // 4 for d in ew["dims"]])),
                  "GBps": round((efield.nbytes + ebytes) / (enc_t * 1e-3) / 1e9, 1),
                  "equal_to_host_encoder": enc_equal}
    log(f"bench: gpu encode {encode_out['ms']} ms ({encode_out['GBps']} GB/s), "
        f"equal to host encoder: {enc_equal}")
    del fbuf, ebuf, efield

    mc = bcmc_amd.CompressedMarchingCubes(dev)
    mc.setCompressedVolume(stream, rate, dims, (1, 1, 1), cache0)

    def step(records, timed=False):
        mc.lruCache.reset()
        for iso in isos:
            t = time.perf_counter()
            nv = mc.computeSurface(float(iso), timed=timed)
            dt = time.perf_counter() - t
            if records is not None:
                rec = dict(mc.lastPerf)
                rec["wall_ms"] = dt * 1e3
                rec["iso"] = float(iso)
                records.append(rec)
        return nv

    for i in range(args.warmup):
        step(None)
        log(f"bench: warmup step {i + 1}/{args.warmup} done")
    dev.sync()
    records = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(records)
    dev.sync()
    elapsed = time.perf_counter() - t_start
    log(f"bench: {args.steps} steps x {len(isos)} queries in {elapsed:.3f}s")
    queries = args.steps * len(isos)
    qps = queries / elapsed
    lat = np.array([r["wall_ms"] for r in records])
    # per-stage device times: the same steps again, timed with events (direct
    # launches with events between the stages)
    records = []
    for i in range(args.steps):
        step(records, timed=True)
    dev.sync()

    # per-stage device time and algorithmic bytes over the timed queries
    stages = ["computeActiveBlocks", "cacheUpdate", "decompression", "computeBlockHasVertices",
              "compactBlocksWithVertices", "computeVertices"]
    st_ms = {k: float(sum(r[k] for r in records)) for k in stages}
    st_bytes = {k: 0 for k in stages}
    for r in records:
        for k, b in stage_bytes(r, rate, n_blocks).items():
            st_bytes[k] += b
    dom = max((k for k in stages if k != "cacheUpdate"), key=lambda k: st_ms[k])
    gbps = {k: (st_bytes[k] / (st_ms[k] * 1e-3) / 1e9 if st_ms[k] > 0 else 0.0)
            for k in stages if k != "cacheUpdate"}
    decode_bytes_range = n_blocks * (8 * rate + 8)
    traffic, traffic_src = profiled_traffic(args.workload, dom)

    out = {
        "metric": "iso-queries/s (plain computeSurface, 1 GPU)",
        "value": round(qps, 3),
        "unit": "iso-queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {dims[0]}x{dims[1]}x{dims[2]} Fourier-mode field, "
                               f"zfp rate {rate}, 8-isovalue sweep 0.2..0.8, LRU 25% initial",
                   "queries_per_step": len(isos), "blocks": n_blocks,
                   "stream_sha256": sha, "device": info["name"], "cus": info["num_cus"]},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "mean": round(float(lat.mean()), 3)},
        "stage_ms_per_query": {k: round(v / queries, 4) for k, v in st_ms.items()},
        "stage_GBps": {k: round(v, 1) for k, v in gbps.items()},
        "zfp_decode": {
            "cache_decode_GBps": round(gbps["decompression"], 1),
            "cache_decode_frac_of_peak": round(gbps["decompression"] / HBM_PEAK_GBPS, 4),
            "blocks_decoded_per_query": round(sum(r["numBlocksDecompressed"] for r in records) / queries, 1),
            "range_pass_ms": round(range_t, 4),
            "range_pass_GBps": round(decode_bytes_range / (range_t * 1e-3) / 1e9, 1),
            "range_pass_decoded_equiv_GBps": round(256 * n_blocks / (range_t * 1e-3) / 1e9, 1),
            # per isovalue of the sweep: decoded blocks and device time of the
            # cache-decode launch (mean over the timed steps), where the
            # launch's fixed latency shows in the small miss lists
            "per_query": [{"iso": round(iso, 4),
                           "blocks": int(np.mean([r["numBlocksDecompressed"] for r in records
                                                  if r["iso"] == iso])),
                           "us": round(1e3 * float(np.mean([r["decompression"] for r in records
                                                            if r["iso"] == iso])), 1)}
                          for iso in sorted({r["iso"] for r in records})],
        },
        "zfp_encode": encode_out,
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": round(gbps[dom], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps[dom] / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": round(st_bytes[dom] / queries)},
        "counts_per_query": {k: round(float(np.mean([r[k] for r in records])), 1) for k in
                             ("numActiveBlocks", "numBlocksDecompressed", "numBlocksWithVertices",
                              "numVertices")},
        "cpu_baseline": None,
        "tuning": tuning,
        "launch": "graph replay" if args.graph_replay else "direct",
    }

    if not args.no_batched:
        out["batched"] = batched_bench(bcmc_amd, synth, dev, args)
    if not args.no_served:
        out["served"] = served_bench(bcmc_amd, synth, dev, args)

    if args.calibrate_copy_mib > 0:
        nb = args.calibrate_copy_mib << 20
        src, dst = dev.create_buffer(nb), dev.create_buffer(nb)
        src.fill_u32(1, nb // 4)
        bcmc_amd.lib().bcmc_ctx_enable_timing(dev.h, 1)
        dst.copy_from(src, nb)
        cms = bcmc_amd.lib().bcmc_ctx_last_kernel_ms(dev.h)
        bcmc_amd.lib().bcmc_ctx_enable_timing(dev.h, 0)
        out["copy_calibration"] = {"bytes": nb, "ms": round(cms, 4),
                                   "GBps": round(2 * nb / (cms * 1e-3) / 1e9, 1)}
        del src, dst

    # the 256^3 16-Gaussian field at rate 8 with every block cached, one
    # isovalue 0.3 after an LRU reset (the CPU reference configuration), on the GPU
    gw = synth.WORKLOADS["gauss256_r8"]
    gstream, _ = synth.make_stream(gw)
    g_nb = int(np.prod([(d + 3) // 4 for d in gw["dims"]]))
    gmc = bcmc_amd.CompressedMarchingCubes(dev)
    gmc.setCompressedVolume(gstream, gw["rate"], gw["dims"], (1, 1, 1), g_nb)
    g_t = []
    for _ in range(11):
        gmc.lruCache.reset()
        dev.sync()
        t = time.perf_counter()
        g_nv = gmc.computeSurface(0.3)
        g_t.append(time.perf_counter() - t)
    del gmc
    out["gauss256_r8"] = {"workload": "256^3 sum of 16 Gaussians (seed 0), zfp rate 8, every block "
                                      "cached, isovalue 0.3 after an LRU reset",
                          "gpu_qps": round(1.0 / float(np.median(g_t[1:])), 2), "vertices": int(g_nv)}

    if args.cpu_sweeps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        host = host_cpu()
        threads = args.cpu_threads or host["threads"]
        t = time.time()
        ref = oracle.MC(stream, 64 * rate, dims, initial_cache=cache0, threads=threads)
        t_load = time.time() - t
        tq = time.time()
        nq = 0
        for _ in range(args.cpu_sweeps):
            ref.lru_reset()  # one step = reset + sweep, as on the GPU
            for iso in isos:
                ref.compute_surface(float(iso))
                nq += 1
        t_q = time.time() - tq
        # single-thread sample: one sweep
        ref1 = oracle.MC(stream, 64 * rate, dims, initial_cache=cache0, threads=1)
        t1 = time.time()
        for iso in isos:
            ref1.compute_surface(float(iso))
        t_1 = time.time() - t1
        # the CPU reference configuration: cold queries after a reset, T threads and 1
        gref = {}
        for nt in (threads, 1):
            r = oracle.MC(gstream, 64 * gw["rate"], gw["dims"], initial_cache=g_nb, threads=nt)
            reps, tg = 0, 0.0
            while reps < 3 or (tg < 3.0 and reps < 50):
                r.lru_reset()
                t = time.time()
                nv_c = r.compute_surface(0.3)
                tg += time.time() - t
                reps += 1
            assert nv_c == g_nv, (nv_c, g_nv)
            gref[nt] = (reps / tg, reps, tg)
            del r
        out["gauss256_r8"]["cpu_qps"] = round(gref[threads][0], 3)
        out["gauss256_r8"]["cpu_qps_1_thread"] = round(gref[1][0], 3)
        out["gauss256_r8"]["gpu_vs_cpu"] = round(out["gauss256_r8"]["gpu_qps"] / gref[threads][0], 1)
        out["cpu_baseline"] = {
            "value": round(nq / t_q, 4), "unit": "iso-queries/s", "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_sweeps} sweeps x 8 isovalues ({nq} queries, {t_q:.1f}s) of the same "
                      f"workload on the same stream; C oracle with OpenMP over blocks, {threads} "
                      f"threads; load-time range pass ({t_load:.2f}s) excluded",
            "host": host,
            "single_thread": {"value": round(len(isos) / t_1, 4), "cores": 1,
                              "sample": f"1 sweep x 8 isovalues, {t_1:.1f}s"},
            "gauss256_r8": {"value": round(gref[threads][0], 3), "cores": threads,
                            "single_thread": round(gref[1][0], 3), "unit": "iso-queries/s",
                            "sample": f"{gref[threads][1]} cold queries at isovalue 0.3 after an LRU "
                                      f"reset ({gref[threads][2]:.1f}s); 1 thread: "
                                      f"{gref[1][1]} queries ({gref[1][2]:.1f}s)"}}
        log(f"bench: cpu oracle {nq} queries in {t_q:.2f}s with {threads} threads; "
            f"1 thread: 8 queries in {t_1:.2f}s; gauss256 r8 {gref[threads][0]:.2f} q/s "
            f"({threads} threads), {gref[1][0]:.2f} q/s (1 thread)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
