"""Stage-by-stage parity at scale (200,000 mixture rows): every graph
construction stage of the GPU path against the oracle on the same inputs,
including the rows that take the large-work-area paths (hub rows with long
reverse lists), and range search against the oracle on the GPU-built graph."""
import ctypes
import os

import numpy as np
import pytest

from anqi_testlib import oracle, ptr
import workload

pytestmark = pytest.mark.gpu
N = int(os.environ.get("ANQI_SCALE_N", "200000"))


def say(*a):
    print("[gpu-scale]", *a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import pyanqi
    ctx = pyanqi.Context(0)
    base = workload.million_base(N)
    q = workload.million_query(300)
    bd = ctx.to_dev(base)
    ids, keys, st = ctx.aknn_build(bd, K=100, leaf=500, trees=64, ef=600)
    say("aknn", st.forest_s, st.leaf_s, st.gather_s, "overflow rows", st.overflow_rows, "short", st.short_rows)
    cut = ctx.torch.clamp(keys, min=0.0)
    lifted_np, m2, sc = pyanqi.lift_host(base, cut[:, 99].contiguous().cpu().numpy())
    yield dict(ctx=ctx, base=base, bd=bd, q=q, ids=ids, cut=cut, lifted=lifted_np, m2=m2, sc=sc)
    ctx.close()


def test_aknn_cutoffs_not_below_exact(env):
    ctx = env["ctx"]
    e, ek = ctx.exact_knn(env["bd"], 100)
    e, ek = e.cpu().numpy(), ek.cpu().numpy()
    a, ak = env["ids"].cpu().numpy(), env["cut"].cpu().numpy()
    below = int((ak[:, 99] < np.maximum(ek[:, 99], 0)).sum())
    rng = np.random.default_rng(0)
    rows = rng.choice(N, 3000, replace=False)
    rec = np.mean([len(np.intersect1d(a[i], e[i])) for i in rows]) / 100
    say("aknn recall@100 (3000 rows)", rec, "rows below exact cutoff", below)
    assert below == 0 and rec > 0.95


def test_graph_stages_match_oracle(env):
    import pyanqi
    O = oracle()
    ctx = env["ctx"]
    x = env["lifted"]
    n, d = x.shape
    xd = ctx.to_dev(x)
    nd = ctx.norms(xd)
    nrm = nd.cpu().numpy()
    ids = env["ids"].cpu().numpy().view(np.uint32)
    cap = 50
    hi, hv = ctx.warm_start(xd, nd, env["ids"], cap)
    hi_o = np.zeros((n, cap), np.uint32)
    hv_o = np.zeros((n, cap), np.float32)
    O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(ids), 100, cap, ptr(hi_o), ptr(hv_o))
    assert np.array_equal(hi.cpu().numpy().view(np.uint32), hi_o), "warm start"
    hi1, hv1 = hi.clone(), hv.clone()
    passes, upd = ctx.nndescent(xd, nd, hi, hv, 6, 20)
    upd_o = np.zeros(6, np.int64)
    p_o = O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, d, cap, ptr(hi_o), ptr(hv_o), 6, 20, ptr(upd_o))
    say("localjoin", passes, upd.tolist(), "oracle", p_o, upd_o[:p_o].tolist())
    assert passes == p_o and upd.tolist() == upd_o[:p_o].tolist()
    assert np.array_equal(hi.cpu().numpy().view(np.uint32), hi_o), "localjoin"
    p1, upd1 = ctx.nndescent(xd, nd, hi1, hv1, 6, 20, form=1)      # MFMA form
    assert p1 == p_o and upd1.tolist() == upd_o[:p_o].tolist()
    assert np.array_equal(hi1.cpu().numpy(), hi.cpu().numpy()) and np.array_equal(
        hv1.cpu().numpy().view(np.uint32), hv.cpu().numpy().view(np.uint32)), "localjoin (MFMA form)"
    roff, rids = ctx.reverse_csr(hi, n)
    ro = roff.cpu().numpy().view(np.uint32)
    rev_len = np.diff(ro)
    say("reverse list length max", int(rev_len.max()), "mean", float(rev_len.mean()))
    Og, oc = ctx.gather_prune(xd, nd, hi, hv, roff, rids, 100, 800, 48, 1.2)
    Og1, oc1 = ctx.gather_prune(xd, nd, hi, hv, roff, rids, 100, 800, 48, 1.2, form=1)   # batched form
    Og, oc = Og.cpu().numpy().view(np.uint32), oc.cpu().numpy().view(np.uint32)
    assert np.array_equal(oc1.cpu().numpy().view(np.uint32), oc) and np.array_equal(
        Og1.cpu().numpy().view(np.uint32), Og), "gather + prune (batched form)"
    roff_o = np.zeros(n + 1, np.uint32)
    rids_o = np.zeros(n * cap, np.uint32)
    O.or_reverse_csr(ptr(hi_o), n, cap, ptr(roff_o), ptr(rids_o))
    rng = np.random.default_rng(1)
    hubs = np.argsort(-rev_len)[:200]
    rows = np.unique(np.concatenate([rng.choice(n, 1500, replace=False), hubs]))
    pm = 65536
    pd = np.zeros(pm, np.float32)
    pi = np.zeros(pm, np.uint32)
    out = np.zeros(48, np.uint32)
    bad = 0
    for i in rows:
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi_o), ptr(hv_o), cap, ptr(roff_o), ptr(rids_o), int(i), 100,
                        800, ptr(pd), ptr(pi), pm)
        dg = O.or_prune(ptr(x), ptr(nrm), d, int(i), n, ptr(pd), ptr(pi), m, 48, 1.2, ptr(out))
        bad += int(oc[i] != dg or Og[i, :dg].tolist() != out[:dg].tolist())
    say(f"gather+prune rows differing from the oracle: {bad}/{len(rows)} (incl. {len(hubs)} hub rows)")
    assert bad == 0
    links, deg = ctx.commit(xd, nd, ctx.to_dev(Og.view(np.int32)), ctx.to_dev(oc.view(np.int32)), 48, 1.2)
    l_o = np.zeros((n, 48), np.uint32)
    d_o = np.zeros(n, np.uint32)
    O.or_commit(ptr(x), ptr(nrm), n, d, ptr(Og), ptr(oc), 48, 1.2, ptr(l_o), ptr(d_o))
    gl, gd = links.cpu().numpy().view(np.uint32), deg.cpu().numpy().view(np.uint32)
    indeg = np.bincount(Og[np.arange(48)[None, :] < oc[:, None]], minlength=n)
    diff = [i for i in range(n) if gd[i] != d_o[i] or gl[i, :gd[i]].tolist() != l_o[i, :d_o[i]].tolist()]
    say(f"commit rows differing: {len(diff)}/{n}; max in-degree {int(indeg.max())}; "
        f"in-degree of differing rows {sorted(indeg[diff].tolist())[-5:] if diff else []}")
    assert not diff
    env["graph"] = (xd, nd, links, deg, l_o, d_o)


def test_search_matches_oracle_on_built_graph(env):
    if "graph" not in env:
        pytest.skip("graph stage failed")
    O = oracle()
    ctx = env["ctx"]
    xd, nd, links, deg, l_o, d_o = env["graph"]
    x = env["lifted"]
    n, dl = x.shape
    t = ctx.rp_tree(xd, leaf=100)
    bd = env["bd"]
    ix = ctx.make_index(bd, ctx.norms(bd), xd, deg, links, t, env["m2"], env["sc"])
    q = env["q"]
    thr = ctx.torch.full((n,), 3.0e38, dtype=ctx.torch.float32, device="cuda:0")
    qb = np.zeros((len(q), dl), np.float32)
    for i in range(len(q)):
        O.or_lift_query(ptr(q[i]), q.shape[1], env["sc"], ptr(qb[i]))
    D = (x.astype(np.float64) ** 2).sum(1)[None, :] - 2 * qb.astype(np.float64) @ x.astype(np.float64).T
    ent = np.argmin(D, axis=1).astype(np.uint32)
    R2 = env["m2"] + 1.0
    for L in (16, 64):
        b = ctx.rknn_search(ix, ctx.to_dev(q), L, thr, out_cap=8192, mode=1, entries=ctx.to_dev(ent.view(np.int32)))
        cnt, cand, ndd = b["cnt"].cpu().numpy(), b["cand"].cpu().numpy().view(np.uint32), b["nd"].cpu().numpy()
        bad = 0
        for i in range(len(q)):
            out = np.zeros(n, np.uint32)
            ndist = ctypes.c_uint64()
            c = O.or_range_search(ptr(x), n, dl, ptr(d_o), ptr(l_o), 48, int(ent[i]), ptr(qb[i]), R2, L, ptr(out), n,
                                  ctypes.byref(ndist))
            bad += int(c != cnt[i] or cand[i, :c].tolist() != out[:c].tolist() or ndd[i] != ndist.value)
        say(f"L={L}: queries differing from the oracle search: {bad}/{len(q)}; mean candidates {cnt.mean():.1f}")
        assert bad == 0
    ctx.tree_free(t)
