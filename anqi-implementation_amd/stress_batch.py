#!/usr/bin/env python3
"""Batched uneven-query stress run (configuration 5) on one GPU.

Base: n x d float32 rows from a centres-mixture (default 10,000,000 x 96,
4,000 centres uniform in [0, 128), per-dimension sigma 15, clipped to
[0, 255], not rounded).  Index: the chosen build (bf16 MFMA sweeps), G48
lifted graph, rank-8 + u8 LRQ verifier with threshold tables for k = 10, 50,
100.  Requests: queries from the same centres (separate stream), served in
batches of 512..16,384 queries; every query draws k from {10, 50, 100} and a
beam width L from 64..512.  Each batch is one replay of the captured graph
(anqi_batcher), with work areas and grid picked by anqi_search_autotune on a
sample batch.  Prints one JSON line: QPS over all batches and p50/p99 batch
latency (host wall clock from handing the batch over to holding its answers).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(*a):
    print("[stress]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--d", type=int, default=96)
    ap.add_argument("--centres", type=int, default=4000)
    ap.add_argument("--sigma", type=float, default=15.0)
    ap.add_argument("--requests", type=int, default=2_000_000)
    ap.add_argument("--min-batch", type=int, default=512)
    ap.add_argument("--max-batch", type=int, default=16384)
    ap.add_argument("--lmin", type=int, default=64)
    ap.add_argument("--lmax", type=int, default=512)
    ap.add_argument("--seed", type=int, default=5)
    a = ap.parse_args()
    os.environ.setdefault("ANQI_PROGRESS", "1")
    import torch
    import pipeline
    import pyanqi
    import workload
    if not torch.cuda.is_available():
        raise SystemExit("stress_batch.py needs a HIP device")
    ctx = pyanqi.Context(0)
    t = time.perf_counter()
    base = workload.mixture(a.n, a.d, a.centres, a.sigma, seed=0, rounded=False)
    log(f"base {a.n}x{a.d} in {time.perf_counter() - t:.1f}s")
    ix = pipeline.build(ctx, base, precision=1, log=log)
    log("build", json.dumps({k: (round(v, 3) if isinstance(v, float) else v) for k, v in ix.times.items()}))
    ks = (10, 50, 100)
    tables = torch.stack([pipeline.thresholds(ctx, ix, k) for k in ks]).contiguous()
    rng = np.random.default_rng(a.seed)
    sizes = []
    left = a.requests
    while left > 0:
        s = int(min(left, rng.integers(a.min_batch, a.max_batch + 1)))
        sizes.append(s)
        left -= s
    t = time.perf_counter()
    queries = workload.mixture(a.requests, a.d, a.centres, a.sigma, seed=1, rounded=False)
    L = rng.integers(a.lmin, a.lmax + 1, a.requests).astype(np.uint32)
    kidx = rng.integers(0, len(ks), a.requests).astype(np.uint32)
    log(f"{a.requests} requests in {len(sizes)} batches (sizes {min(sizes)}..{max(sizes)}), "
        f"generated in {time.perf_counter() - t:.1f}s")
    ns = min(4096, a.requests)
    tune, ms = ctx.search_autotune(ix.handle, tables, ctx.to_dev(queries[:ns]), ctx.to_dev(L[:ns].astype(np.int32)),
                                   ctx.to_dev(kidx[:ns].astype(np.int32)), a.lmax)
    log(f"autotune: waves/CU {tune.waves_per_cu}, visited slots {tune.visited_slots}, "
        f"FIFO {tune.result_cap} ({ms:.3f} ms for {ns} queries)")
    bat = ctx.batcher(ix.handle, tables, max_batch=a.max_batch, Lmax=a.lmax, tuning=tune)
    bat.run(queries[:a.max_batch], L[:a.max_batch], kidx[:a.max_batch])        # warm-up replay
    lat = []
    answers = 0
    pos = 0
    t0 = time.perf_counter()
    for i, s in enumerate(sizes):
        tb = time.perf_counter()
        off, ids = bat.run(queries[pos:pos + s], L[pos:pos + s], kidx[pos:pos + s])
        lat.append(time.perf_counter() - tb)
        answers += len(ids)
        pos += s
        if i % 25 == 0:
            log(f"batch {i}/{len(sizes)} size {s}: {lat[-1] * 1e3:.2f} ms")
    secs = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    res = {
        "metric": "batched RkNN queries/s (mixed k in {10,50,100}, L in [%d,%d])" % (a.lmin, a.lmax),
        "value": a.requests / secs, "unit": "queries/s",
        "p50_batch_ms": float(np.percentile(lat_ms, 50)), "p99_batch_ms": float(np.percentile(lat_ms, 99)),
        "batches": len(sizes), "requests": a.requests, "answers": int(answers),
        "tuning": {"waves_per_cu": tune.waves_per_cu, "visited_slots": tune.visited_slots,
                   "result_cap": tune.result_cap},
        "build_s": ix.times["build_s"],
        "config": {"workload": f"centres-mixture {a.n}x{a.d} ({a.centres} centres U[0,128), sigma {a.sigma}, "
                               f"clip 0..255, not rounded)", "batch_sizes": [a.min_batch, a.max_batch]},
    }
    bat.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
