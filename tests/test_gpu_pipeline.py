"""GPU pipeline on the first 20,000 rows of the million-row mixture (200
queries): AKNN source quality, search-tree entry quality, graph parity with
the oracle on the same AKNN source, and RkNN recall/precision against the
reference's own numbers on the same inputs (tests/golden/mix20k, within 0.5
points)."""
import ctypes
import json
import os

import numpy as np
import pytest

from anqi_testlib import GOLDEN, oracle, ptr
import workload

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-pipeline]", *a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import pipeline
    import pyanqi
    ref = json.load(open(os.path.join(GOLDEN, "mix20k", "mix20k_reference.json")))
    base = workload.million_base(ref["n"])
    q = workload.million_query(ref["nq"])
    assert workload.sha256(base) == ref["base_sha256"] and workload.sha256(q) == ref["query_sha256"]
    ctx = pyanqi.Context(0)
    ix = pipeline.build(ctx, base, log=say)
    qd = ctx.to_dev(q)
    off, gids, _ = pipeline.ground_truth(ctx, ix.base, qd, 10)
    thr = pipeline.thresholds(ctx, ix, 10)
    yield dict(ctx=ctx, ix=ix, base=base, q=q, qd=qd, off=off, gids=gids, thr=thr, ref=ref)
    ctx.tree_free(ix.tree)
    ctx.close()


def test_aknn_source_recall(env):
    ctx, ix = env["ctx"], env["ix"]
    eids, ekeys = ctx.exact_knn(ix.base, 100)
    a = ix.aknn_ids.cpu().numpy()
    e = eids.cpu().numpy()
    ek = ekeys.cpu().numpy()
    ak = ix.cut.cpu().numpy()
    below = int((ak[:, 99] < np.maximum(ek[:, 99], 0) ).sum())
    say("rows whose AKNN 100-th key is below the exact 100-th key:", below)
    assert below == 0
    n = a.shape[0]
    rec = np.mean([len(np.intersect1d(a[i], e[i])) for i in range(n)]) / 100.0
    say("aknn recall@100", rec, "reference", env["ref"]["aknn_recall_at_100"])
    keys = ix.cut.cpu().numpy()
    assert (np.diff(keys, axis=1) >= 0).all()
    assert rec >= env["ref"]["aknn_recall_at_100"] - 0.005


def _bruteforce_entries(env):
    ix = env["ix"]
    lifted = ix.lifted.cpu().numpy().astype(np.float64)
    q = env["q"].astype(np.float64)
    qb = np.zeros((len(q), lifted.shape[1]), np.float32)
    qb[:, :q.shape[1]] = (q * ix.scale).astype(np.float32)
    qb[:, q.shape[1]] = 1.0
    D = (lifted ** 2).sum(1)[None, :] - 2 * qb.astype(np.float64) @ lifted.T
    return np.argmin(D, axis=1).astype(np.int32)


def test_truth_inside_lifted_ball(env):
    """Lift property (report Theorem 1): every true RkNN_10 member lies in the
    lifted ball D <= M*^2 + 1."""
    ix = env["ix"]
    lifted = ix.lifted.cpu().numpy().astype(np.float64)
    q = env["q"]
    qb = np.zeros((len(q), lifted.shape[1]), np.float64)
    qb[:, :q.shape[1]] = (q.astype(np.float64) * ix.scale).astype(np.float32)
    qb[:, q.shape[1]] = 1.0
    off, gids = env["off"], env["gids"]
    inside = tot = 0
    for i in range(len(q)):
        T = gids[off[i]:off[i + 1]]
        D = ((lifted[T] - qb[i]) ** 2).sum(1)
        inside += int((D <= ix.mstar2 + 1.0 + 1e-9).sum())
        tot += len(T)
    say(f"truth members inside the lifted ball: {inside}/{tot}")
    assert inside >= 0.999 * tot


def test_recall_matches_reference(env):
    import pipeline
    ctx, ix = env["ctx"], env["ix"]
    ent = ctx.to_dev(_bruteforce_entries(env))
    for p in env["ref"]["points_k10"]:
        L = p["L"]
        b = ctx.rknn_search(ix.handle, env["qd"], L, env["thr"], out_cap=4096, mode=1)
        m = pipeline.metrics(b, env["off"], env["gids"])
        b2 = ctx.rknn_search(ix.handle, env["qd"], L, env["thr"], out_cap=4096, mode=1, entries=ent)
        m2 = pipeline.metrics(b2, env["off"], env["gids"])
        say(f"L={L} tree-entry recall={m['recall']:.4f} cand={m['cand']:.1f} | brute-force entry recall="
            f"{m2['recall']:.4f} cand={m2['cand']:.1f} | reference recall={p['recall']:.4f} cand={p['cand']}")
    for p in env["ref"]["points_k10"]:
        b = ctx.rknn_search(ix.handle, env["qd"], p["L"], env["thr"], out_cap=4096, mode=1)
        m = pipeline.metrics(b, env["off"], env["gids"])
        assert abs(m["recall"] - p["recall"]) <= 0.005, (p["L"], m["recall"], p["recall"])
        assert abs(m["precision"] - p["precision"]) <= 0.005


def test_graph_matches_oracle_on_gpu_source(env):
    """Warm start + LocalJoin + gather/prune + commit on the GPU's own AKNN
    source and lift, against the oracle's restatement (bit-exact links)."""
    O = oracle()
    ctx, ix = env["ctx"], env["ix"]
    x = ix.lifted.cpu().numpy()
    ids = ix.aknn_ids.cpu().numpy().view(np.uint32)
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    cap = 50
    hi = np.zeros((n, cap), np.uint32)
    hv = np.zeros((n, cap), np.float32)
    O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(ids), 100, cap, ptr(hi), ptr(hv))
    O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, d, cap, ptr(hi), ptr(hv), 6, 20, None)
    roff = np.zeros(n + 1, np.uint32)
    rids = np.zeros(n * cap, np.uint32)
    O.or_reverse_csr(ptr(hi), n, cap, ptr(roff), ptr(rids))
    Orows = np.zeros((n, 48), np.uint32)
    oc = np.zeros(n, np.uint32)
    pm = 16384
    pd = np.zeros(pm, np.float32)
    pi = np.zeros(pm, np.uint32)
    out = np.zeros(48, np.uint32)
    for i in range(n):
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi), ptr(hv), cap, ptr(roff), ptr(rids), i, 100, 800,
                        ptr(pd), ptr(pi), pm)
        dg = O.or_prune(ptr(x), ptr(nrm), d, i, n, ptr(pd), ptr(pi), m, 48, 1.2, ptr(out))
        Orows[i, :dg] = out[:dg]
        oc[i] = dg
    links = np.zeros((n, 48), np.uint32)
    deg = np.zeros(n, np.uint32)
    O.or_commit(ptr(x), ptr(nrm), n, d, ptr(Orows), ptr(oc), 48, 1.2, ptr(links), ptr(deg))
    gl = ix.links.cpu().numpy().view(np.uint32)
    gd = ix.deg.cpu().numpy().view(np.uint32)
    same = sum(int(gd[i] == deg[i] and (gl[i, :gd[i]] == links[i, :deg[i]]).all()) for i in range(n))
    say(f"graph rows identical to the oracle: {same}/{n}")
    assert same == n
