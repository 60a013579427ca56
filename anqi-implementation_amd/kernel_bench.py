#!/usr/bin/env python3
"""Per-kernel timing of the build and verifier kernels on the 1,000,000 x 128
mixture, for the profiles: the chosen AKNN build (bf16 MFMA leaf sweep,
batched gather), the bf16 MFMA exact top-K on a sub-base, the LRQ
threshold decode (the verifier's unpack-and-compare stage) and the
persistent range search at a fixed L (--what search).

Prints one JSON line: per stage, its time and its algorithmic work (flop for
the MFMA sweeps, bytes for the decode) and the achieved rate.  Meant to run
under rocprofv3 (kernel trace or one PMC pass per run)."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(*a):
    print("[kernel-bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--what", default="aknn,exact,lrq")
    ap.add_argument("--exact-n", type=int, default=262144, help="rows of the exact top-K sub-base")
    ap.add_argument("--lrq-reps", type=int, default=20)
    ap.add_argument("--nq", type=int, default=10000)
    ap.add_argument("--L", type=int, default=96)
    ap.add_argument("--search-reps", type=int, default=5)
    a = ap.parse_args()
    import numpy as np
    import torch
    import lrq_fit
    import pyanqi
    import workload
    if not torch.cuda.is_available():
        raise SystemExit("kernel_bench.py needs a HIP device")
    what = set(a.what.split(","))
    ctx = pyanqi.Context(0)
    base = workload.million_base(a.n)
    x = ctx.to_dev(base)
    n, d = base.shape
    out = {"n": n, "d": d}
    ev = lambda: torch.cuda.Event(enable_timing=True)
    if what & {"aknn", "lrq"}:
        ids, keys, st = ctx.aknn_build(x, K=100, leaf=500, trees=64, ef=600, precision=1)
        torch.cuda.synchronize()
    if "aknn" in what:
        flop = 2.0 * d * 2 * st.leaf_pairs          # ordered pairs, one MFMA dot each
        out["aknn"] = {"forest_s": st.forest_s, "leaf_s": st.leaf_s, "gather_s": st.gather_s,
                       "leaf_pairs": int(st.leaf_pairs), "leaf_survivors": int(st.leaf_survivors),
                       "leaf_mfma_flop": flop, "leaf_tflops": flop / st.leaf_s / 1e12}
        log("aknn", json.dumps(out["aknn"]))
    if "exact" in what:
        m = min(a.exact_n, n)
        xs = x[:m].contiguous()
        e0, e1 = ev(), ev()
        e0.record()
        ctx.exact_knn(xs, 100, precision=1)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        flop = 2.0 * d * m * m
        out["exact"] = {"rows": m, "ms": ms, "flop": flop, "tflops": flop / (ms * 1e-3) / 1e12}
        log("exact", json.dumps(out["exact"]))
    if "lrq" in what:
        cut = torch.clamp(keys, min=0.0).contiguous()
        mu, G, cb, _ = lrq_fit.fit(cut, log=lambda *z: None)
        mu, G, cb = ctx.to_dev(mu), ctx.to_dev(G), ctx.to_dev(cb)
        coeff, codes, codes_km = ctx.lrq_encode(cut, mu, G, cb)
        thr = ctx.empty((n,), torch.float32)
        ctx.lrq_materialize(coeff, codes_km, mu, G, cb, 10, 1.0, out=thr)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        for r in range(a.lrq_reps):
            ctx.lrq_materialize(coeff, codes_km, mu, G, cb, 10, 1.0, out=thr)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / a.lrq_reps
        rankm = G.shape[0]
        byts = n * (4 * rankm + 1 + 4)                 # coefficients, one code byte, threshold out
        out["lrq_materialize"] = {"ms": ms, "bytes": byts, "GBps": byts / (ms * 1e-3) / 1e9}
        log("lrq", json.dumps(out["lrq_materialize"]))
    if "search" in what:
        import pipeline
        ix = pipeline.build(ctx, base, precision=1, log=lambda *z: None)
        thr = pipeline.thresholds(ctx, ix, 10)
        q = ctx.to_dev(workload.million_query(a.nq))
        tune, tms = ctx.search_autotune(ix.handle, thr, q, None, None, a.L)
        ctx.set_search_tuning(tune)
        bufs = ctx.search_buffers(a.nq, 4096)
        ctx.rknn_search(ix.handle, q, a.L, thr, mode=1, bufs=bufs)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        for _ in range(a.search_reps):
            ctx.rknn_search(ix.handle, q, a.L, thr, mode=1, bufs=bufs)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / a.search_reps
        out["search"] = {"L": a.L, "nq": a.nq, "ms_per_batch": ms, "qps": a.nq / (ms * 1e-3),
                         "tuning": [tune.waves_per_cu, tune.visited_slots, tune.result_cap],
                         "dist_per_query": float(bufs["nd"].double().mean() + bufs["cnt"].double().mean())}
        log("search", json.dumps(out["search"]))
        ctx.tree_free(ix.tree)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
