"""GPU parity for the inner-product metric ("ip") on the reference's own ip
sample (tests/golden/sample_ip, 1,000 x 32, made by the reference's tools):
the lift against its _approxlift file and graph meta, the thresholds against
its materialize_k, and the range search + post-filter against its
range_search on its graph -- candidate lists in order, #dist and accept
decisions -- in every search form (per-query, persistent with whole and split
lifted rows, the no-allocation entry point, the captured batcher).  The
rknn_query drop-in on the same files prints the reference's answer lines and
writes its verifier artifact byte for byte."""
import os
import subprocess

import numpy as np
import pytest

from anqi_testlib import GOLDEN, REPO, Sample

pytestmark = pytest.mark.gpu
U32MAX = 0xFFFFFFFF


def say(*a):
    print("[gpu-ip]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a HIP device"
    import pyanqi
    c = pyanqi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def S():
    return Sample("sample_ip")


def host(t):
    return t.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_lift_matches_reference_file(ctx, S):
    import pyanqi
    rk = np.ascontiguousarray(S.rec_keys[:, 99])
    lifted, m2, sc = ctx.lift(ctx.to_dev(S.base), ctx.to_dev(rk), metric=pyanqi.METRIC_IP)
    assert np.array_equal(bits(host(lifted)), bits(S.lifted)) and m2 == S.M2 and sc == S.scale
    hl, hm2, hsc = pyanqi.lift_host(S.base, rk, metric=pyanqi.METRIC_IP)
    assert np.array_equal(bits(hl), bits(S.lifted)) and hm2 == S.M2 and hsc == S.scale


@pytest.mark.parametrize("form", [0, 1], ids=["certified", "soft"])
def test_thresholds_match_reference(ctx, S, form):
    import pyanqi
    ctx.set_lrq_form(form)
    try:
        coeff, codes, codes_km = ctx.lrq_encode(ctx.to_dev(S.rec_keys), ctx.to_dev(S.mu), ctx.to_dev(S.G),
                                                ctx.to_dev(S.cb))
        assert np.array_equal(bits(host(coeff)), bits(S.coeff)) and np.array_equal(host(codes), S.codes)
        tabs = ctx.lrq_materialize_multi(coeff, codes_km, ctx.to_dev(S.mu), ctx.to_dev(S.G), ctx.to_dev(S.cb),
                                         [10, 50, 100], metric=pyanqi.METRIC_IP)
        for t, k in enumerate((10, 50, 100)):
            assert np.array_equal(bits(host(tabs[t])), bits(S.h[f"thr_k{k}"])), k
            one = ctx.lrq_materialize(coeff, codes_km, ctx.to_dev(S.mu), ctx.to_dev(S.G), ctx.to_dev(S.cb), k,
                                      metric=pyanqi.METRIC_IP)
            assert np.array_equal(bits(host(one)), bits(S.h[f"thr_k{k}"])), k
    finally:
        ctx.set_lrq_form(0)


@pytest.mark.parametrize("precision", [0, 1])
def test_exact_knn_and_membership_match_dataset_gt(ctx, S, precision):
    """dataset_gt --metric ip (tools/dataset_gt.cpp:206-313): the exact
    top-100 table and the RkNN CSR at k = 10, 50, 100 equal the reference
    tool's, ids and key bits."""
    import pyanqi
    ids, keys = ctx.exact_knn(ctx.to_dev(S.base), 100, precision=precision, metric=pyanqi.METRIC_IP)
    assert np.array_equal(host(ids).view(np.uint32), S.gt_ids)
    assert np.array_equal(bits(host(keys)), bits(S.gt_keys))
    for k in (10, 50, 100):
        rk = ctx.to_dev(np.ascontiguousarray(S.gt_keys[:, k - 1]))
        off, cids, ckeys = ctx.rknn_membership(ctx.to_dev(S.base), ctx.to_dev(S.query), rk,
                                               metric=pyanqi.METRIC_IP)
        goff, gids, gkeys = S.csr[k]
        assert np.array_equal(off, goff) and np.array_equal(host(cids).view(np.uint32), gids), k
        assert np.array_equal(bits(host(ckeys)), bits(gkeys)), k


@pytest.mark.parametrize("case", ["gaussian-2500x128-K100", "mixture-1500x96-K50", "wide-norms-2000x128-K64"])
@pytest.mark.parametrize("precision", [0, 1])
def test_exact_knn_ip_matches_oracle(ctx, case, precision):
    """Inner-product exact top-K at widths the bf16 MFMA sweep takes (d 96 and
    128): both sweeps return the oracle's ids and key bits."""
    import pyanqi
    import workload
    from anqi_testlib import oracle, ptr
    rng = np.random.default_rng(11)
    make = {
        "gaussian-2500x128-K100": (lambda: workload.gaussian(2500, 128, seed=5), 100),
        "mixture-1500x96-K50": (lambda: workload.mixture(1500, 96, 50, 20.0, seed=7), 50),
        "wide-norms-2000x128-K64": (lambda: (rng.standard_normal((2000, 128)) *
                                             np.exp(rng.uniform(-3, 3, (2000, 1)))), 64),
    }
    mk, K = make[case]
    x = np.ascontiguousarray(mk(), np.float32)
    n, d = x.shape
    say("exact knn ip", case, "precision", precision)
    ids, keys = ctx.exact_knn(ctx.to_dev(x), K, precision=precision, metric=pyanqi.METRIC_IP)
    ri = np.zeros((n, K), np.uint32)
    rk = np.zeros((n, K), np.float32)
    oracle().or_exact_topk_metric(ptr(x), n, d, K, 1, ptr(ri), ptr(rk), 8)
    assert np.array_equal(host(ids).view(np.uint32), ri)
    assert np.array_equal(bits(host(keys)), bits(rk))


def test_dataset_gt_ip_files_equal_reference(tmp_path):
    pfx = os.path.join(GOLDEN, "sample_ip", "sample_ip")
    out = str(tmp_path / "g")
    tool = os.path.join(REPO, "anqi-implementation_amd", "bin", "dataset_gt")
    r = subprocess.run([tool, "--base", pfx + "_base.bin", "--K", "100", "--rk_k", "100", "--mode", "thresholds",
                        "--metric", "ip", "--out_prefix", out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert open(out + "_baseknn_gt.bin", "rb").read() == open(pfx + "_baseknn_gt.bin", "rb").read()
    for k in (10, 50, 100):
        o = str(tmp_path / f"q{k}")
        r = subprocess.run([tool, "--base", pfx + "_base.bin", "--query", pfx + "_query.bin", "--K", "100",
                            "--rk_k", str(k), "--mode", "ground-truth", "--metric", "ip", "--threshold-table",
                            out + "_baseknn_gt.bin", "--out_prefix", o], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert open(o + "_rknn_gt.bin", "rb").read() == open(pfx + f"_rknn_gt_k{k}.bin", "rb").read(), k


def _aknn_checks(x, ids, keys, gt_ids):
    """Row invariants of build_aknn_source.cpp:93-117, every key equal to the
    reference's "ip" distance of its pair (oracle, pinned to the reference's
    compiled function), and the recall against the exact top-K."""
    from anqi_testlib import oracle, ptr
    O = oracle()
    n, K = ids.shape
    d = x.shape[1]
    assert (ids < n).all() and (ids != np.arange(n)[:, None]).all()
    assert all(len(set(r)) == K for r in ids.tolist())
    assert (np.diff(keys, axis=1) >= 0).all()
    rows = range(0, n, max(1, n // 200))
    for i in rows:
        for j in range(K):
            ref = O.or_ipdist(ptr(x[i]), ptr(x[ids[i, j]]), d)
            assert bits(keys[i, j]) == bits(ref), (i, j)
    return np.mean([len(set(ids[i]) & set(gt_ids[i])) for i in range(n)]) / K


@pytest.mark.parametrize("precision", [0, 1])
def test_aknn_build_ip_on_reference_sample(ctx, S, precision):
    """The AKNN source for "ip" (build_aknn_source ... ip, index.h:1300-1461
    with angular splits and 1 - <x, y>): valid rows, distances bit-exact to
    the reference's function, recall against dataset_gt's exact ip top-100 not
    below the reference source's (its forest draws differ by design)."""
    import pyanqi
    ids, keys, st = ctx.aknn_build(ctx.to_dev(S.base), K=100, leaf=64, trees=8, ef=64, precision=precision,
                                   metric=pyanqi.METRIC_IP)
    ids, keys = host(ids).view(np.uint32), host(keys)
    rec = _aknn_checks(S.base, ids, keys, S.gt_ids)
    ref_rec = np.mean([len(set(S.rec_ids[i]) & set(S.gt_ids[i])) for i in range(S.n)]) / 100
    say("ip aknn recall", rec, "reference source", ref_rec, "short rows", st.short_rows)
    assert rec >= ref_rec - 0.005
    cut, hz = ctx.aknn_cutoffs(ctx.to_dev(keys), pyanqi.METRIC_IP)
    assert np.array_equal(bits(host(cut)), bits((keys - np.float32(1)).astype(np.float32)))
    assert np.array_equal(bits(host(hz)), bits(host(cut)[:, 99]))


def test_aknn_build_ip_mixture_20k(ctx):
    """A 20,000 x 96 mixture under the build's default parameters (leaf 500,
    64 trees, ef 600): valid rows, distance bits, and the recall of the GPU
    source against the GPU exact ip top-100 (itself bit-exact to dataset_gt)."""
    import pyanqi
    import workload
    x = np.ascontiguousarray(workload.mixture(20000, 96, 200, 20.0, seed=9), np.float32)
    xd = ctx.to_dev(x)
    ids, keys, st = ctx.aknn_build(xd, K=100, precision=1, metric=pyanqi.METRIC_IP)
    gt, _ = ctx.exact_knn(xd, 100, precision=1, metric=pyanqi.METRIC_IP)
    rec = _aknn_checks(x, host(ids).view(np.uint32), host(keys), host(gt).view(np.uint32))
    say("ip aknn 20k recall", rec, "gather", st.gather_s, "s")
    assert rec >= 0.9


def test_build_aknn_source_ip_tool(tmp_path, S):
    pfx = os.path.join(GOLDEN, "sample_ip", "sample_ip")
    out = str(tmp_path / "a")
    tool = os.path.join(REPO, "anqi-implementation_amd", "bin", "build_aknn_source")
    r = subprocess.run([tool, pfx + "_base.bin", out, "100", "64", "8", "64", "1", "10", "ip"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "metric=ip" in r.stdout
    rec = np.fromfile(out + "_baseknn_aknn.rec", dtype=np.dtype([("id", "<u4"), ("key", "<f4")])).reshape(S.n, 100)
    ids, keys = rec["id"].copy(), rec["key"].copy()
    # the .rec key is the distance minus 1 (build_aknn_source.cpp:125-126)
    assert (np.diff(keys, axis=1) >= 0).all() and (ids != np.arange(S.n)[:, None]).all()
    from anqi_testlib import oracle, ptr
    O = oracle()
    for i in range(0, S.n, 25):
        for j in range(100):
            dv = np.float32(O.or_ipdist(ptr(S.base[i]), ptr(S.base[ids[i, j]]), S.d))
            assert bits(keys[i, j]) == bits(np.float32(dv - np.float32(1))), (i, j)
    rec_mine = np.mean([len(set(ids[i]) & set(S.gt_ids[i])) for i in range(S.n)]) / 100
    ref_rec = np.mean([len(set(S.rec_ids[i]) & set(S.gt_ids[i])) for i in range(S.n)]) / 100
    assert rec_mine >= ref_rec - 0.005


def _index(ctx, S, split):
    import pyanqi
    t = S.tree
    ns = t["nodes"]
    left = np.where(t["left"] == U32MAX, U32MAX, t["left"]).astype(np.uint32)
    right = np.where(t["right"] == U32MAX, U32MAX, t["right"]).astype(np.uint32)
    entry = np.array([t["idx"][t["idx_off"][k]] if t["idx_off"][k + 1] > t["idx_off"][k] else 0
                      for k in range(ns)], np.uint32)
    tree = ctx.tree_upload(dict(nodes=ns, root=ns - 1, d=t["d"], left=left, right=right,
                                offset=t["offset"].astype(np.float32), hyper=t["hyper"].astype(np.float32),
                                entry=entry))
    base = ctx.to_dev(S.base)
    ix = ctx.make_index(base, ctx.norms(base), ctx.to_dev(S.lifted), ctx.to_dev(S.deg.view(np.int32)),
                        ctx.to_dev(S.links.view(np.int32)), tree, S.M2, S.scale, split=split,
                        metric=pyanqi.METRIC_IP)
    return ix, tree


def _check(S, L, k, cnt, nd, cand, acc):
    off, ids = S.h[f"cand_off_L{L}"], S.h[f"cand_ids_L{L}"]
    o128 = S.h["cand_off_L128"]
    for i in range(len(S.query)):
        exp = ids[off[i]:off[i + 1]]
        assert cnt[i] == len(exp), (L, i)
        assert cand[i, :cnt[i]].view(np.uint32).tolist() == exp.tolist(), (L, i)
        assert nd[i] == S.h[f"ndist_L{L}"][i], (L, i)
        if L == 128:
            assert acc[i, :cnt[i]].tolist() == S.h[f"dec_k{k}"][o128[i]:o128[i + 1]].tolist(), (k, i)


@pytest.mark.parametrize("mode,split", [(0, False), (1, False), (1, True)])
def test_search_matches_reference_harness(ctx, S, mode, split):
    ix, tree = _index(ctx, S, split)
    q = ctx.to_dev(S.query)
    for k in (10, 50, 100):
        thr = ctx.to_dev(S.h[f"thr_k{k}"])
        for L in (16, 32, 64, 128):
            say("search mode", mode, "split", split, "k", k, "L", L)
            b = ctx.rknn_search(ix, q, L, thr, out_cap=1024, mode=mode)
            assert (host(b["status"]) == 0).all()
            _check(S, L, k, host(b["cnt"]), host(b["nd"]), host(b["cand"]), host(b["acc"]))
    ctx.tree_free(tree)


def test_noalloc_and_batcher_match_reference_harness(ctx, S):
    ix, tree = _index(ctx, S, True)
    q = ctx.to_dev(S.query)
    nq = len(S.query)
    say("no-alloc search")
    thr = ctx.to_dev(S.h["thr_k100"])
    bufs = ctx.search_buffers(nq, 1024)
    b = ctx.rknn_search_noalloc(ix, q, 128, thr, bufs)
    _check(S, 128, 100, host(b["cnt"]), host(b["nd"]), host(b["cand"]), host(b["acc"]))
    say("batcher, mixed k and L")
    tables = ctx.to_dev(np.stack([S.h[f"thr_k{k}"] for k in (10, 50, 100)]))
    bat = ctx.batcher(ix, tables, max_batch=128, Lmax=128, out_cap=1024)
    rng = np.random.default_rng(4)
    L = rng.choice([16, 32, 64, 128], nq).astype(np.uint32)
    kidx = rng.integers(0, 3, nq).astype(np.uint32)
    off, ids = bat.run(S.query, L, kidx)
    o128, i128 = S.h["cand_off_L128"], S.h["cand_ids_L128"]
    for i in range(nq):
        h = S.h
        exp_c = h[f"cand_ids_L{L[i]}"][h[f"cand_off_L{L[i]}"][i]:h[f"cand_off_L{L[i]}"][i + 1]]
        # the accepted ids of the query's candidate list, in candidate order
        key = dict(zip(i128[o128[i]:o128[i + 1]].tolist(),
                       h[f"dec_k{(10, 50, 100)[kidx[i]]}"][o128[i]:o128[i + 1]].tolist()))
        exp = [c for c in exp_c.tolist() if key[c]]
        assert ids[off[i]:off[i + 1]].tolist() == exp, i
    bat.close()
    ctx.tree_free(tree)


PIN_ENV = {
    "ANQI_GRAPH_GEOMETRY": "anyk_lift", "ANQI_RADIUS_MODE": "rankm_resid_u8_lrq_floor",
    "ANQI_VERIFIER_PLACEMENT": "postfilter", "ANQI_RANKM": "8", "ANQI_NK": "100", "ANQI_GRAPH_NK": "100",
    "ANQI_RANKM_SLACK": "0", "ANQI_RECHECK": "0", "ANQI_RQ_OBJECTIVE": "radius_reconstruction",
    "ANQI_ALPHA": "1.2", "ANQI_LIFTNNK": "50", "ANQI_GRAPHNND": "6", "ANQI_BUILD_BREAKDOWN": "1",
    "ANQI_PREBUILT_TOPK_K": "100", "ANQI_BENCH_CALIBRATED_TIMING": "0",
}


def _answers(text):
    out = []
    for line in text.splitlines():
        if line.startswith("[rknn] L_descent="):
            out.append(line.split("| QPS")[0])
        elif line.startswith(("[rknn] DIST_BREAKDOWN", "[rknn] CANDIDATE_RECALL")):
            out.append(line)
    return out


@pytest.mark.parametrize("form", ["chosen", "plain"])
def test_rknn_query_ip_matches_reference(tmp_path, form):
    """The drop-in tool on the reference's ip files.  A first run (k = 10)
    builds here: its lift line equals the reference's and its verifier
    artifact is the reference's byte for byte.  The reference's own graph
    cache, lifted vectors and meta, put under the names this tool computes,
    are then loaded by runs at k = 10, 50, 100, which print the reference's
    answer lines (recall, precision, F1, candidates, distances) exactly."""
    import re
    import shutil
    pfx = os.path.join(GOLDEN, "sample_ip", "sample_ip")
    env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    env.update(PIN_ENV)
    env["ANQI_PREBUILT_TOPK_REC"] = pfx + "_baseknn_aknn.rec"
    env["ANQI_GRAPH_CACHE_DIR"] = g = str(tmp_path / "graph")
    env["ANQI_VERIFIER_CACHE_DIR"] = v = str(tmp_path / "verifier")
    env["ANQI_GPU_FORM"] = form
    os.makedirs(g)
    os.makedirs(v)
    tool = os.path.join(REPO, "anqi-implementation_amd", "bin", "rknn_query")

    def run(k):
        say("rknn_query ip", form, "k", k)
        r = subprocess.run([tool, "16,32,64,128", pfx, "48", "500", "64", "600", str(k)], env=env,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        return r.stdout

    first = run(10)
    ref10 = open(pfx + "_rknn_stdout_k10.txt").read()
    lift = re.compile(r"BUILD lift M\*\^2=(\S+) scale=(\S+)")
    assert lift.search(first).groups() == lift.search(ref10).groups()
    art = os.listdir(v)
    assert len(art) == 1
    assert open(os.path.join(v, art[0]), "rb").read() == open(pfx + "_verifier.anqivf2", "rb").read()
    names = os.listdir(g)
    for head, suf in {"fd_metric_l2_M_48_": "_graph.bin", "_approxlift_graph_": "_approxlift.bin",
                      "_anqi_meta_graph_": "_graph_meta.txt"}.items():
        mine = [f for f in names if f.startswith(head)]
        assert len(mine) == 1, (head, names)
        shutil.copy(pfx + suf, os.path.join(g, mine[0]))
    for k in (10, 50, 100):
        out = run(k)
        assert "[rknn] loaded cached index (M=48)" in out
        mine, theirs = _answers(out), _answers(open(pfx + f"_rknn_stdout_k{k}.txt").read())
        say("\n".join(mine))
        assert mine == theirs and len(mine) == 12, k
