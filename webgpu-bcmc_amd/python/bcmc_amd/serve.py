"""The served form from the command line: the served set of compressed
volumes (8 blob volumes 128^3 .. 1024x1024x1080, seeds 7-14, zfp rates
2/4/8/2/4/8/2/2) and one decoded-block pool resident on the device, a request
trace replayed through Engine.serve (continuous batching), throughput and
latency printed as one JSON line per replay.

    PYTHONPATH=webgpu-bcmc_amd/python python3 -m bcmc_amd.serve \\
        [--requests N] [--rate R] [--max-batch B] [--pool-gib G] [--reps K] [--no-coalesce]

Requests draw a volume uniformly and an isovalue U[0.1, 0.9], half of them
repeating one of the last 16 isovalues of the same volume; --rate R > 0 spaces
them by Poisson arrivals at R per second, --rate 0 has them all arrive at
t = 0 (saturation).  Requests of one step with the same volume and isovalue
share one query unless --no-coalesce.  A first replay of requests of their own
draw warms the engine (buffer growth, graph capture, first decodes)."""
import argparse
import json
import sys
import time

import numpy as np

import bcmc_amd
from bcmc_amd import synth


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--requests", type=int, default=2048)
    ap.add_argument("--rate", type=float, default=0.0, help="requests per second (0: all at t = 0)")
    ap.add_argument("--max-batch", type=int, default=256)
    ap.add_argument("--pool-gib", type=float, default=32.0, help="decoded-block pool (at most 2x all blocks)")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0, help="pool + output vertices")
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--seed", type=int, default=20240607)
    ap.add_argument("--no-coalesce", action="store_true", help="every request a query of its own")
    args = ap.parse_args(argv)
    t0 = time.time()
    vols = synth.served_volumes()
    print(f"serve: volumes generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = bcmc_amd.Device(0)
    decoded = sum(int(np.prod([(d + 3) // 4 for d in v[2]])) for v in vols) * 256
    eng = bcmc_amd.Engine(dev, int(min(args.pool_gib * (1 << 30), 2 * decoded)))
    for s, rate, dims in vols:
        eng.add_volume(s, rate, dims)
    budget = int(args.hbm_budget_gb * 1e9) - eng.stats()["pool_bytes"]
    wid, wiso = bcmc_amd.request_trace(np.random.default_rng(args.seed + 1), 4 * args.max_batch, len(vols))
    eng.serve(np.zeros(len(wid)), wid, wiso, max_batch=args.max_batch, vertex_budget_bytes=budget)
    dev.sync()
    rng = np.random.default_rng(args.seed)
    for rep in range(args.reps):
        ids, isos = bcmc_amd.request_trace(rng, args.requests, len(vols))
        arr = (np.cumsum(rng.exponential(1.0 / args.rate, args.requests)) if args.rate > 0
               else np.zeros(args.requests))
        out = eng.serve(arr, ids, isos, max_batch=args.max_batch, vertex_budget_bytes=budget,
                        coalesce=not args.no_coalesce)
        lat = out["finish_s"] - out["arrival_s"]
        span = float(out["finish_s"].max() - arr[0])
        print(json.dumps({"rep": rep, "requests": args.requests, "rate_rps": args.rate,
                          "coalesce": not args.no_coalesce,
                          "qps": round(args.requests / span, 2),
                          "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                          "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                          "steps": int(out["step"].max()) + 1,
                          "mean_batch": round(float(out["batch_size"].mean()), 2),
                          "vertices": int(out["num_vertices"].sum())}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
