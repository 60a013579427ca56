"""Unbounded flood: queries whose flood visits and collects more nodes than
the LDS work areas hold (16,384 visited slots and FIFO entries in the retry
tier) finish in the global-memory tier (a visited bitmap of n bits and an
n-entry candidate FIFO per query), as the reference's flood does with its
n-entry visit stamps (include/index.h:33-52, 1627-1638).  The index is a
30,000 x 96 single-centre base built on the GPU; enlarging M*^2 in the index handle
widens the lifted ball R2 = M*^2 + 1 so that floods reach 15k-30k nodes.
Candidate lists (in order), #dist and accept decisions are compared with the
oracle's range search on the same graph and entries, for the per-query and
persistent forms, a too-small out_cap (status 3, the first out_cap
candidates) and the captured batcher (answers spliced from the global tier)."""
import ctypes

import numpy as np
import pytest

import workload
from anqi_testlib import oracle, ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-flood]", *a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import pipeline
    import pyanqi
    ctx = pyanqi.Context(0)
    # one centre: the graph is connected across the whole base (with many
    # well-separated centres the pruned graph has no edges between them and
    # a flood stops at its own cluster)
    base = workload.mixture(30000, 96, 1, 15.0, seed=31)
    say("build 30000 x 96")
    ix = pipeline.build(ctx, base, log=say)
    thr = pipeline.thresholds(ctx, ix, 10)
    q = workload.mixture(64, 96, 1, 15.0, seed=32)
    O = oracle()
    n, dl = ix.lifted.shape
    qb = np.zeros((len(q), dl), np.float32)
    for i in range(len(q)):
        O.or_lift_query(ptr(q[i]), q.shape[1], ix.scale, ptr(qb[i]))
    lifted = ix.lifted.cpu().numpy()
    D = (lifted.astype(np.float64) ** 2).sum(1)[None, :] - 2 * qb.astype(np.float64) @ lifted.astype(np.float64).T
    D += (qb.astype(np.float64) ** 2).sum(1)[:, None]
    yield dict(ctx=ctx, ix=ix, thr=thr, q=q, qb=qb, D=D, lifted=lifted, base=base,
               deg=ix.deg.cpu().numpy().view(np.uint32), links=ix.links.cpu().numpy().view(np.uint32),
               thr_h=thr.cpu().numpy(), torch=torch)
    ctx.tree_free(ix.tree)
    ctx.close()


def wide_index(env, mstar2):
    ix = env["ix"]
    return env["ctx"].make_index(ix.base, ix.bnorm, ix.lifted, ix.deg, ix.links, ix.tree, mstar2, ix.scale)


def oracle_search(env, i, entry, R2, L):
    O = oracle()
    n, dl = env["lifted"].shape
    out = np.zeros(n, np.uint32)
    ndist = ctypes.c_uint64()
    c = O.or_range_search(ptr(env["lifted"]), n, dl, ptr(env["deg"]), ptr(env["links"]), 48, int(entry),
                          ptr(env["qb"][i]), R2, L, ptr(out), n, ctypes.byref(ndist))
    ids = out[:c].copy()
    dec = np.array([np.float32(O.or_recheck_key(ptr(env["base"][j]), ptr(env["q"][i]), env["base"].shape[1]))
                    <= env["thr_h"][j] for j in ids], bool)
    return ids, int(ndist.value), dec


def mstar2_for(env, frac):
    """M*^2 whose lifted ball holds about frac of the rows for the median query."""
    r = np.quantile(env["D"], frac, axis=1)
    return float(np.median(r)) - 1.0


@pytest.mark.parametrize("frac,mode", [(0.6, 0), (0.6, 1), (1.5, 1)])
def test_large_floods_match_oracle(env, frac, mode):
    ctx = env["ctx"]
    n = env["lifted"].shape[0]
    m2 = mstar2_for(env, frac) if frac <= 1.0 else float(env["D"].max()) * 2.0
    ix = wide_index(env, m2)
    q = ctx.to_dev(env["q"])
    ent = ctx.search_entries(ix, q)
    ent_h = ent.cpu().numpy().view(np.uint32)
    L = 96
    say(f"frac {frac} mode {mode}: M*^2 {m2:.4g}")
    b = ctx.rknn_search(ix, q, L, env["thr"], out_cap=n, mode=mode, entries=ent)
    cnt, nd, pred = (b[k].cpu().numpy() for k in ("cnt", "nd", "pred"))
    cand, acc, st = b["cand"].cpu().numpy().view(np.uint32), b["acc"].cpu().numpy(), b["status"].cpu().numpy()
    say(f"candidates per query: min {cnt.min()} median {int(np.median(cnt))} max {cnt.max()}; "
        f"above 16384: {(cnt > 16384).sum()} of {len(cnt)}")
    assert (cnt > 16384).any(), "the case must drive queries past the LDS tiers"
    assert (st == 0).all()
    for i in range(len(env["q"])):
        ids, ndist, dec = oracle_search(env, i, ent_h[i], m2 + 1.0, L)
        c = int(cnt[i])
        assert c == len(ids) and nd[i] == ndist, (i, c, len(ids), nd[i], ndist)
        assert np.array_equal(cand[i, :c], ids), i
        assert np.array_equal(acc[i, :c].astype(bool), dec), i
        assert pred[i] == dec.sum()


def test_truncated_output_keeps_first_candidates(env):
    """out_cap below the candidate count: status 3, the full count, the
    accepted count over all candidates, and the first out_cap candidates."""
    ctx = env["ctx"]
    m2 = float(env["D"].max()) * 2.0
    ix = wide_index(env, m2)
    q = ctx.to_dev(env["q"][:16])
    ent = ctx.search_entries(ix, q)
    ent_h = ent.cpu().numpy().view(np.uint32)
    cap = 4096
    b = ctx.rknn_search(ix, q, 64, env["thr"], out_cap=cap, mode=1, entries=ent)
    cnt, pred, st = (b[k].cpu().numpy() for k in ("cnt", "pred", "status"))
    cand, acc = b["cand"].cpu().numpy().view(np.uint32), b["acc"].cpu().numpy()
    for i in range(16):
        ids, _, dec = oracle_search(env, i, ent_h[i], m2 + 1.0, 64)
        assert cnt[i] == len(ids) and st[i] == (3 if len(ids) > cap else 0)
        m = min(cap, len(ids))
        assert np.array_equal(cand[i, :m], ids[:m]) and np.array_equal(acc[i, :m].astype(bool), dec[:m])
        assert pred[i] == dec.sum()


def test_batcher_answers_from_global_tier(env):
    """The captured batcher on the widened index: every query overflows the
    tuned areas, the retry tier and out_cap; its answers (accepted ids in
    candidate order) come from the global tier's FIFO rows and equal the
    oracle's.  A second, smaller batch on the same batcher follows."""
    ctx = env["ctx"]
    torch = env["torch"]
    m2 = mstar2_for(env, 0.6)
    ix = wide_index(env, m2)
    tables = torch.stack([env["thr"]]).contiguous()
    bat = ctx.batcher(ix, tables, max_batch=64, Lmax=128, out_cap=4096)
    rng = np.random.default_rng(5)
    for nq in (64, 23):
        q = env["q"][:nq]
        L = rng.integers(16, 129, nq).astype(np.uint32)
        kidx = np.zeros(nq, np.uint32)
        off, ids = bat.run(q, L, kidx)
        ent_h = ctx.search_entries(ix, ctx.to_dev(q)).cpu().numpy().view(np.uint32)
        for i in range(nq):
            oids, _, dec = oracle_search(env, i, ent_h[i], m2 + 1.0, int(L[i]))
            assert np.array_equal(ids[off[i]:off[i + 1]], oids[dec]), i
        say(f"batch {nq}: {len(ids)} answers")
    bat.close()
