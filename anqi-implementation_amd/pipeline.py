"""End-to-end ANQI reverse-kNN pipeline on one GPU (used by bench.py, smoke()
and the pipeline tests).  Mirrors the reference flow:

  build_aknn_source  (tools/build_aknn_source.cpp)  -> AKNN top-100 source
  train_final_m8_u8  (tools/train_final_m8_u8.py)   -> rank-8 basis + u8 codebook
  rknn_query build   (src/rknn_query.cpp:1995-2297)  -> LRQ encode, lift, G48 graph
  rknn_query queries (src/rknn_query.cpp:2692-2846)  -> range search + postfilter,
                                                        recall / precision / F1
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

import lrq_fit
import pyanqi


def _sync(ctx):
    ctx.torch.cuda.synchronize(ctx.device)


class Index:
    pass


def build(ctx, base_np, K=100, leaf=500, trees=64, ef=600, ef_alpha=8, seed=10, precision=0,
          M=48, nn_k=50, graph_leaf=100, explore=100, iters=6, sample=20, alpha=1.2, log=print,
          metric=pyanqi.METRIC_L2):
    """Builds the whole index from host base rows; returns Index with device
    tensors and the stage timings (seconds, host clock around synchronised
    stages: the reference's end_to_end_s decomposition).  metric: L2 or the
    inner product ("ip": AKNN on 1 - <x, y>, .rec keys minus 1, the ip lift,
    per-query R2 and the -<o, q> post-filter)."""
    T = ctx.torch
    ix = Index()
    t = {}
    base = ctx.to_dev(base_np)
    _sync(ctx)
    t0 = time.perf_counter()
    ids, keys, st = ctx.aknn_build(base, K=K, leaf=leaf, trees=trees, ef=ef, ef_alpha=ef_alpha, seed=seed,
                                   precision=precision, metric=metric)
    cut, horizon = ctx.aknn_cutoffs(keys, metric)  # .rec key = max(0, d2) (ip: d - 1); lift horizon = key K
    _sync(ctx)
    t1 = time.perf_counter()
    t["aknn_source_s"] = t1 - t0
    log(f"aknn source {t1 - t0:.2f}s")
    t["aknn_detail"] = dict(forest_s=st.forest_s, leaf_s=st.leaf_s, gather_s=st.gather_s,
                            leaf_pairs=int(st.leaf_pairs), leaves=int(st.leaves), max_leaf=int(st.max_leaf),
                            short_rows=int(st.short_rows), overflow_rows=int(st.overflow_rows),
                            leaf_survivors=int(st.leaf_survivors), gather_d2=int(st.gather_d2),
                            gather_adj=int(st.gather_adj))
    mu, G, cb, fit_s = lrq_fit.fit(cut, log=lambda *a: None, codebook_dtype=T.float32)
    _sync(ctx)
    t2 = time.perf_counter()
    t["verifier_train_s"] = t2 - t1
    log(f"verifier fit {t2 - t1:.2f}s")
    mu_d, G_d, cb_d = ctx.to_dev(mu), ctx.to_dev(G), ctx.to_dev(cb)
    coeff, codes, codes_km = ctx.lrq_encode(cut, mu_d, G_d, cb_d)
    _sync(ctx)
    t3 = time.perf_counter()
    t["encode_s"] = t3 - t2
    if precision >= 1:
        # device lift (soft binary64), bit-identical to the host lift
        lifted, mstar2, scale = ctx.lift(base, horizon, metric)
    else:
        rk = horizon.cpu().numpy()
        lifted_np, mstar2, scale = pyanqi.lift_host(base_np, rk, metric)
        lifted = ctx.to_dev(lifted_np)
    _sync(ctx)
    t4 = time.perf_counter()
    t["lift_s"] = t4 - t3
    log(f"encode {t3 - t2:.2f}s lift {t4 - t3:.2f}s")
    links, deg, tree, gst = ctx.graph_build(lifted, ids, M=M, nn_k=nn_k, leaf=graph_leaf, explore=explore,
                                            ef_alpha=ef_alpha, iters=iters, sample=sample, alpha=alpha, seed=seed,
                                            form=1 if precision >= 1 else 0)
    _sync(ctx)
    t5 = time.perf_counter()
    t["graph_s"] = t5 - t4
    t["graph_detail"] = dict(tree_s=gst.tree_s, warm_s=gst.warm_s, nnd_s=gst.nnd_s, gather_s=gst.gather_s,
                             commit_s=gst.commit_s, nnd_passes=gst.nnd_passes,
                             nnd_upd=[int(gst.nnd_upd[i]) for i in range(gst.nnd_passes)],
                             overflow_rows=int(gst.overflow_rows), gather_d2=int(gst.gather_d2),
                             gather_adj=int(gst.gather_adj))
    t["build_s"] = t5 - t0
    log(f"graph {t5 - t4:.2f}s build total {t5 - t0:.2f}s")
    ix.base, ix.base_np, ix.lifted, ix.links, ix.deg, ix.tree = base, base_np, lifted, links, deg, tree
    ix.aknn_ids, ix.cut = ids, cut
    ix.mu, ix.G, ix.cb, ix.coeff, ix.codes, ix.codes_km = mu_d, G_d, cb_d, coeff, codes, codes_km
    ix.mstar2, ix.scale = mstar2, scale
    ix.bnorm = ctx.norms(base)
    ix.handle = ctx.make_index(base, ix.bnorm, lifted, deg, links, tree, mstar2, scale, metric=metric)
    ix.metric = metric
    ix.times = t
    return ix


def thresholds(ctx, ix, k, slack=1.0):
    return ctx.lrq_materialize(ix.coeff, ix.codes_km, ix.mu, ix.G, ix.cb, k, slack,
                               metric=getattr(ix, "metric", pyanqi.METRIC_L2))


def threshold_tables(ctx, ix, ks, slack=1.0):
    """One table per k in ks (len(ks) x n), decoded in one pass over the rows."""
    return ctx.lrq_materialize_multi(ix.coeff, ix.codes_km, ix.mu, ix.G, ix.cb, list(ks), slack,
                                     metric=getattr(ix, "metric", pyanqi.METRIC_L2))


def ground_truth(ctx, base, queries, k, metric=pyanqi.METRIC_L2):
    """Exact RkNN CSR (tools/dataset_gt.cpp): exact top-k thresholds, then
    membership key <= r_k (key -<x, y> for the inner product)."""
    n = base.shape[0]
    ids, keys = ctx.exact_knn(base, k, precision=1, metric=metric)   # bf16 MFMA sweep, FP32-exact output
    rk2 = keys[:, k - 1].contiguous()
    off, gids, gkeys = ctx.rknn_membership(base, queries, rk2, metric=metric)
    return off, gids.cpu().numpy().view(np.uint32), rk2


def metrics(bufs, off, gids):
    """Reference metric definitions (rknn_query.cpp:2784-2798): per query with
    a non-empty truth set, recall = |P n T|/|T|, precision = |P n T|/|P|
    (1 when P is empty); averages over those queries."""
    cnt = bufs["cnt"].cpu().numpy()
    cap = bufs["cand"].shape[1]
    cand = bufs["cand"].cpu().numpy().view(np.uint32)
    acc = bufs["acc"].cpu().numpy()
    nd = bufs["nd"].cpu().numpy()
    nq = len(cnt)
    m = np.minimum(cnt, cap)
    qi = np.repeat(np.arange(nq), m)
    mask = np.arange(cap)[None, :] < m[:, None]
    cids = cand[mask]
    cacc = acc[mask].astype(bool)
    pq, pid = qi[cacc], cids[cacc]
    tq = np.repeat(np.arange(nq), np.diff(off).astype(np.int64))
    key_p = pq.astype(np.int64) << 32 | pid.astype(np.int64)
    key_c = qi.astype(np.int64) << 32 | cids.astype(np.int64)
    key_t = tq.astype(np.int64) << 32 | gids.astype(np.int64)
    hit_p = np.isin(key_p, key_t)
    hit_c = np.isin(key_c, key_t)
    hits = np.bincount(pq[hit_p], minlength=nq)
    chits = np.bincount(qi[hit_c], minlength=nq)
    npred = np.bincount(pq, minlength=nq)
    ntruth = np.diff(off).astype(np.int64)
    ne = ntruth > 0
    rec = hits[ne] / ntruth[ne]
    prec = np.where(npred[ne] > 0, hits[ne] / np.maximum(npred[ne], 1), 1.0)
    f1 = np.where(rec + prec > 0, 2 * rec * prec / np.maximum(rec + prec, 1e-300), 0.0)
    return dict(recall=float(rec.mean()), precision=float(prec.mean()), F1=float(f1.mean()),
                cand=float(cnt[ne].mean()), pred=float(npred[ne].mean()),
                graph_dist=float(nd[ne].mean()), verify_dist=float(cnt[ne].mean()),
                dist=float((nd[ne] + cnt[ne]).mean()), cand_recall=float((chits[ne] / ntruth[ne]).mean()),
                nonempty=int(ne.sum()))


def export_reference_graph(ctx, ix, path, leaf_size=100):
    """Writes the GPU-built graph and search tree in the reference's graph
    cache format (include/index.h:1810-1863) so the reference code can load
    it."""
    import formats
    t = ix.tree
    nodes = t.nodes
    left = np.zeros(nodes, np.uint32)
    right = np.zeros(nodes, np.uint32)
    off = np.zeros(nodes, np.float32)
    hyper = np.zeros((nodes, t.d), np.float32)
    for arr, src in ((left, t.left), (right, t.right), (off, t.offset), (hyper, t.hyper)):
        ctx.check(ctx.L.anqi_d2h(ctx.ctx, pyanqi.hp(arr), ctypes.c_void_p(src), arr.nbytes))
    loff = np.ctypeslib.as_array(t.leaf_off, shape=(nodes + 1,)).copy()
    lids = np.ctypeslib.as_array(t.leaf_ids, shape=(int(loff[-1]),)).copy()
    tn = formats.tree_postorder(left, right, off, hyper, loff, lids, root=t.root)
    formats.write_graph(path, ix.base.shape[0], t.d, ix.links.shape[1], ix.deg.cpu().numpy(),
                        ix.links.cpu().numpy(), tn, leaf_size)
