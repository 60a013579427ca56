"""Configuration 5 at its size: a 10,000,000 x 96 base (4,000 centres uniform
in [0, 128), sigma 15, clipped to [0, 255], not rounded; 3.84 GB) indexed by
the chosen build (G48 graph, 1.92 GB of links), rank-8 + u8 LRQ threshold
tables for k = 10, 50, 100, and query batches of 512 to 16,384 requests with
k drawn from {10, 50, 100} and beam widths from 64 to 512, each served by
one replay of the captured batcher graph with autotuned work areas.

Correctness: in batches of 512, 4,096 and 16,384, sampled queries' accepted
ids (in candidate order) equal the oracle's chain on the same lifted graph,
search tree and threshold tables -- tree entry (or_tree_entry), range search
(or_range_search, include/index.h:1535-1652) and post-filter (or_recheck_key
against the query's table, src/rknn_query.cpp:2762-2766).  Every query that
leaves the captured graph for a redo tier is counted and reported.

Throughput: 2,000,000 requests in batches of random size, QPS and p50/p99
batch latency (host clock from handing a batch over to holding its answers)
are printed."""
import ctypes
import time

import numpy as np
import pytest

import workload
from anqi_testlib import oracle, or_tree, ptr

pytestmark = pytest.mark.gpu

N, D, CENTRES, SIGMA = 10_000_000, 96, 4000, 15.0
KS = (10, 50, 100)


def say(*a):
    print("[gpu-config5]", *a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import pipeline
    import pyanqi
    ctx = pyanqi.Context(0)
    t = time.perf_counter()
    base = workload.mixture(N, D, CENTRES, SIGMA, seed=0, rounded=False)
    say(f"base {N} x {D} generated in {time.perf_counter() - t:.1f}s")
    ix = pipeline.build(ctx, base, precision=1, log=say)
    tables = pipeline.threshold_tables(ctx, ix, KS)
    t = time.perf_counter()
    q = workload.mixture(2_000_000, D, CENTRES, SIGMA, seed=1, rounded=False)
    say(f"2,000,000 requests generated in {time.perf_counter() - t:.1f}s")
    rng = np.random.default_rng(5)
    L = rng.integers(64, 513, len(q)).astype(np.uint32)
    kidx = rng.integers(0, len(KS), len(q)).astype(np.uint32)
    tune, ms = ctx.search_autotune(ix.handle, tables, ctx.to_dev(q[:4096]), ctx.to_dev(L[:4096].astype(np.int32)),
                                   ctx.to_dev(kidx[:4096].astype(np.int32)), 512)
    say(f"autotune: waves/CU {tune.waves_per_cu}, visited slots {tune.visited_slots}, FIFO {tune.result_cap} "
        f"({ms:.3f} ms for 4096 queries)")
    bat = ctx.batcher(ix.handle, tables, max_batch=16384, Lmax=512, tuning=tune)
    yield dict(ctx=ctx, ix=ix, tables=tables, base=base, q=q, L=L, kidx=kidx, bat=bat, torch=torch)
    bat.close()
    ctx.tree_free(ix.tree)
    ctx.close()


def test_sampled_answers_match_oracle(env):
    ctx, ix, bat = env["ctx"], env["ix"], env["bat"]
    O = oracle()
    t = time.perf_counter()
    lifted = ix.lifted.cpu().numpy()
    deg = ix.deg.cpu().numpy().view(np.uint32)
    links = ix.links.cpu().numpy().view(np.uint32)
    thr = env["tables"].cpu().numpy()
    ot = or_tree(ctx.tree_host(ix.tree))
    say(f"index copied to the host in {time.perf_counter() - t:.1f}s")
    n, dl = lifted.shape
    out = np.zeros(n, np.uint32)
    base, q, L, kidx = env["base"], env["q"], env["L"], env["kidx"]
    rng = np.random.default_rng(9)
    pos = 0
    checked = 0
    for size in (512, 4096, 16384):
        sl = slice(pos, pos + size)
        off, ids = bat.run(q[sl], L[sl], kidx[sl])
        retry, glob = bat.redo_counts()
        say(f"batch {size}: {len(ids)} answers, redone in the retry tier {retry}, in the global tier {glob}")
        for i in rng.choice(size, size=min(size, 100), replace=False):
            g = pos + int(i)
            qb = np.zeros(dl, np.float32)
            O.or_lift_query(ptr(q[g]), D, ix.scale, ptr(qb))
            e = O.or_tree_entry(ctypes.byref(ot), ptr(qb))
            nd = ctypes.c_uint64()
            c = O.or_range_search(ptr(lifted), n, dl, ptr(deg), ptr(links), 48, e, ptr(qb), ix.mstar2 + 1.0,
                                  int(L[g]), ptr(out), n, ctypes.byref(nd))
            tk = thr[kidx[g]]
            acc = [int(j) for j in out[:c] if np.float32(O.or_recheck_key(ptr(base[j]), ptr(q[g]), D)) <= tk[j]]
            got = ids[off[i]:off[i + 1]].tolist()
            assert got == acc, (size, int(i), int(L[g]), int(kidx[g]))
            checked += 1
        pos += size
    say(f"{checked} sampled queries equal the oracle's answers")


def test_two_million_requests_throughput(env):
    bat, q, L, kidx = env["bat"], env["q"], env["L"], env["kidx"]
    rng = np.random.default_rng(7)
    sizes = []
    left = len(q)
    while left > 0:
        s = int(min(left, rng.integers(512, 16385)))
        sizes.append(s)
        left -= s
    bat.run(q[:16384], L[:16384], kidx[:16384])          # warm-up replay
    lat, answers, retry, glob = [], 0, 0, 0
    pos = 0
    t0 = time.perf_counter()
    for s in sizes:
        tb = time.perf_counter()
        off, ids = bat.run(q[pos:pos + s], L[pos:pos + s], kidx[pos:pos + s])
        lat.append(time.perf_counter() - tb)
        r, g = bat.redo_counts()
        retry += r
        glob += g
        answers += len(ids)
        pos += s
    secs = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    say(f"{len(q)} requests in {len(sizes)} batches: {len(q) / secs:,.0f} queries/s, "
        f"p50 {np.percentile(lat_ms, 50):.2f} ms, p99 {np.percentile(lat_ms, 99):.2f} ms per batch, "
        f"{answers} answers, redone: retry tier {retry}, global tier {glob}")
    assert answers > 0
