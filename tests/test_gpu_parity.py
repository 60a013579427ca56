"""GPU parity: every kernel path of libanqi_hip.so against the oracle and the
reference fixtures, through the C ABI.  Integer, index and FP32-contract
outputs must be bit-exact."""
import ctypes

import numpy as np
import pytest

from anqi_testlib import Sample, or_tree, oracle, ptr
import workload

pytestmark = pytest.mark.gpu
U32MAX = 0xFFFFFFFF


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
    return Sample()


def host(t):
    return t.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def say(*a):
    print("[gpu-test]", *a, flush=True)


def test_norms_and_pair_d2_bits(ctx):
    O = oracle()
    rng = np.random.default_rng(0)
    for d in (8, 96, 128, 130, 200):
        say("norms d", d)
        x = (rng.standard_normal((700, d)) * 3).astype(np.float32)
        xd = ctx.to_dev(x)
        nrm = host(ctx.norms(xd))
        ref = np.zeros(700, np.float32)
        O.or_norms(ptr(x), 700, d, ptr(ref))
        assert np.array_equal(bits(nrm), bits(ref))
        a = rng.integers(0, 700, 5000).astype(np.uint32)
        b = rng.integers(0, 700, 5000).astype(np.uint32)
        out = ctx.empty((5000,), ctx.torch.float32)
        import pyanqi
        keep = [ctx.to_dev(ref), ctx.to_dev(a.view(np.int32)), ctx.to_dev(b.view(np.int32))]   # alive until used
        ctx.call("anqi_pair_d2", pyanqi.dp(xd), pyanqi.dp(keep[0]), d, pyanqi.dp(keep[1]), pyanqi.dp(keep[2]), 5000,
                 pyanqi.dp(out))
        got = host(out)
        exp = np.array([O.or_d2(ptr(x[i]), ptr(x[j]), ref[i], ref[j], d) for i, j in zip(a, b)], np.float32)
        assert np.array_equal(bits(got), bits(exp))


@pytest.mark.parametrize("precision", [0, 1])
def test_exact_knn_matches_dataset_gt(ctx, S, precision):
    say("exact knn sample, precision", precision)
    ids, keys = ctx.exact_knn(ctx.to_dev(S.base), 100, precision=precision)
    assert np.array_equal(host(ids).view(np.uint32), S.gt_ids)
    assert np.array_equal(bits(host(keys)), bits(S.gt_keys))


def _exact_cases():
    rng = np.random.default_rng(3)
    wide = (rng.standard_normal((2000, 128)) * np.exp(rng.uniform(-3, 3, (2000, 1)))).astype(np.float32)
    small = (rng.standard_normal((1200, 112)) * 1e-3 + 5.0).astype(np.float32)   # large offset, tiny spread
    return {
        "mixture-3000x128-K100": (lambda: workload.mixture(3000, 128, 50, 20.0, seed=7), 100),
        "mixture-1500x130-K50": (lambda: workload.mixture(1500, 130, 50, 20.0, seed=7), 50),
        "mixture-700x96-K10": (lambda: workload.mixture(700, 96, 50, 20.0, seed=7), 10),
        "gaussian-2500x128-K100": (lambda: workload.gaussian(2500, 128, seed=5), 100),
        "wide-norms-2000x128-K64": (lambda: wide, 64),
        "offset-1200x112-K128": (lambda: small, 128),
    }


@pytest.mark.parametrize("case", list(_exact_cases()))
@pytest.mark.parametrize("precision", [0, 1])
def test_exact_knn_matches_oracle(ctx, case, precision):
    """Both sweeps (FP32 tiles; bf16 MFMA filter + FP32 re-score) return the
    oracle's ids and key bits: integer data with many key ties, Gaussian
    data, rows whose norms span e^6, and rows far from the origin whose keys
    are small differences of large norms (the bf16 bound is widest there)."""
    make, K = _exact_cases()[case]
    x = np.ascontiguousarray(make(), np.float32)
    n, d = x.shape
    say("exact knn", case, "precision", precision)
    ids, keys = ctx.exact_knn(ctx.to_dev(x), K, precision=precision)
    ri = np.zeros((n, K), np.uint32)
    rk = np.zeros((n, K), np.float32)
    oracle().or_exact_topk(ptr(x), n, d, K, ptr(ri), ptr(rk), 8)
    assert np.array_equal(host(ids).view(np.uint32), ri)
    assert np.array_equal(bits(host(keys)), bits(rk))


@pytest.mark.parametrize("k", [10, 50, 100])
def test_rknn_membership_matches_dataset_gt(ctx, S, k):
    say("membership k", k)
    rk2 = ctx.to_dev(np.ascontiguousarray(S.gt_keys[:, k - 1]))
    off, ids, keys = ctx.rknn_membership(ctx.to_dev(S.base), ctx.to_dev(S.query), rk2)
    goff, gids, gkeys = S.csr[k]
    assert np.array_equal(off, goff)
    assert np.array_equal(host(ids).view(np.uint32), gids)
    assert np.array_equal(bits(host(keys)), bits(gkeys))


def test_lrq_encode_materialize_match_reference_artifact(ctx, S):
    say("lrq sample")
    coeff, codes, codes_km = ctx.lrq_encode(ctx.to_dev(S.rec_keys), ctx.to_dev(S.mu), ctx.to_dev(S.G), ctx.to_dev(S.cb))
    assert np.array_equal(bits(host(coeff)), bits(S.coeff))
    assert np.array_equal(host(codes), S.codes)
    assert np.array_equal(host(codes_km), S.codes.T)
    for k in (10, 50, 100):
        thr = ctx.lrq_materialize(coeff, codes_km, ctx.to_dev(S.mu), ctx.to_dev(S.G), ctx.to_dev(S.cb), k)
        assert np.array_equal(bits(host(thr)), bits(S.h[f"thr_k{k}"]))


def test_lrq_wide_range_matches_oracle(ctx):
    say("lrq wide range")
    O = oracle()
    rng = np.random.default_rng(3)
    n, nk, M = 4000, 100, 8
    cut = np.sort(np.exp(rng.uniform(-8, 12, (n, nk))), axis=1).astype(np.float32)
    mu = cut.mean(0).astype(np.float32)
    G = (rng.standard_normal((M, nk)) * 0.1).astype(np.float32)
    cb = np.sort(rng.standard_normal((nk, 256)) * 50, axis=1).astype(np.float32)
    c1 = np.zeros((n, M), np.float32)
    k1 = np.zeros((n, nk), np.uint8)
    O.or_lrq_encode(ptr(cut), n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, ptr(c1), ptr(k1))
    coeff, codes, codes_km = ctx.lrq_encode(ctx.to_dev(cut), ctx.to_dev(mu), ctx.to_dev(G), ctx.to_dev(cb))
    assert np.array_equal(bits(host(coeff)), bits(c1))
    assert np.array_equal(host(codes), k1)
    for k, slack in ((1, 1.0), (37, 1.0), (100, 1.0), (10, 1.25)):
        t1 = np.zeros(n, np.float32)
        O.or_lrq_materialize(ptr(c1), ptr(k1), n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, k, slack, ptr(t1))
        thr = ctx.lrq_materialize(coeff, codes_km, ctx.to_dev(mu), ctx.to_dev(G), ctx.to_dev(cb), k, slack)
        assert np.array_equal(bits(host(thr)), bits(t1))


def _ref_tree_dev(ctx, S):
    t = S.tree
    ns = t["nodes"]
    left = np.where(t["left"] == U32MAX, U32MAX, t["left"]).astype(np.uint32)
    right = np.where(t["right"] == U32MAX, U32MAX, t["right"]).astype(np.uint32)
    entry = np.array([t["idx"][t["idx_off"][k]] if t["idx_off"][k + 1] > t["idx_off"][k] else 0
                      for k in range(ns)], np.uint32)
    return ctx.tree_upload(dict(nodes=ns, root=ns - 1, d=t["d"], left=left, right=right,
                                offset=t["offset"].astype(np.float32), hyper=t["hyper"].astype(np.float32),
                                entry=entry))


@pytest.mark.parametrize("mode,split", [(0, False), (1, False), (1, True)])
def test_range_search_and_postfilter_on_reference_graph(ctx, S, mode, split):
    """Candidate lists (identical order), #dist and accept decisions against
    the reference's own range_search on the reference-built graph; the
    persistent form with whole lifted rows and with the split layout."""
    tree = _ref_tree_dev(ctx, S)
    base = ctx.to_dev(S.base)
    assert S.base.shape[1] % 16 == 0
    ix = ctx.make_index(base, ctx.norms(base), ctx.to_dev(S.lifted), ctx.to_dev(S.deg.view(np.int32)),
                        ctx.to_dev(S.links.view(np.int32)), tree, S.M2, S.scale, split=split)
    q = ctx.to_dev(S.query)
    for k in (10, 50, 100):
        thr = ctx.to_dev(S.h[f"thr_k{k}"])
        for L in (16, 32, 64, 128):
            say("search mode", mode, "k", k, "L", L)
            b = ctx.rknn_search(ix, q, L, thr, out_cap=512, mode=mode)
            cnt, nd, cand, acc = host(b["cnt"]), host(b["nd"]), host(b["cand"]), host(b["acc"])
            assert (host(b["status"]) == 0).all()
            off, ids = S.h[f"cand_off_L{L}"], S.h[f"cand_ids_L{L}"]
            decs = S.h[f"dec_k{k}"]
            o128 = S.h["cand_off_L128"]
            for i in range(len(S.query)):
                exp = ids[off[i]:off[i + 1]]
                assert cnt[i] == len(exp)
                assert cand[i, :cnt[i]].view(np.uint32).tolist() == exp.tolist()
                assert nd[i] == S.h[f"ndist_L{L}"][i]
                if L == 128:
                    assert acc[i, :cnt[i]].tolist() == decs[o128[i]:o128[i + 1]].tolist()
    ctx.tree_free(tree)


def _oracle_lift(base, rk):
    n, d = base.shape
    lifted = np.zeros((n, d + 2), np.float32)
    m2, sc = ctypes.c_double(), ctypes.c_double()
    oracle().or_lift(ptr(base), n, d, ptr(rk), ptr(lifted), ctypes.byref(m2), ctypes.byref(sc))
    return lifted, m2.value, sc.value


@pytest.mark.parametrize("case", ["sample", "mixture", "wide", "odd_dim"])
def test_device_lift_matches_oracle(ctx, S, case):
    """anqi_lift (per-row soft binary64 on the device, row-order sum on the
    host) gives the oracle's lifted rows, M*^2 and scale bit for bit: the
    reference's own sample (its lifted file), integer mixture rows, rows
    spanning many binades (with zero rows), and an odd dimension."""
    rng = np.random.default_rng(11)
    if case == "sample":
        base, rk = S.base, np.ascontiguousarray(S.rec_keys[:, 99])
    elif case == "mixture":
        base = workload.mixture(20000, 96, 50, 15.0, seed=3)
        rk = rng.uniform(50.0, 5000.0, 20000).astype(np.float32)
    elif case == "wide":
        base = (rng.standard_normal((5000, 64)) * np.exp(rng.uniform(-12, 12, (5000, 1)))).astype(np.float32)
        base[::97] = 0.0
        rk = (np.exp(rng.uniform(-20, 20, 5000))).astype(np.float32)
    else:
        base = rng.uniform(0, 255, (3001, 37)).astype(np.float32)
        rk = rng.uniform(0.0, 1e5, 3001).astype(np.float32)
    base = np.ascontiguousarray(base, np.float32)
    exp, m2, sc = _oracle_lift(base, rk)
    if case == "sample":
        assert np.array_equal(bits(exp), bits(S.lifted)) and m2 == S.M2 and sc == S.scale
    say("device lift", case, base.shape)
    got, gm2, gsc = ctx.lift(ctx.to_dev(base), ctx.to_dev(rk))
    assert gm2 == m2 and gsc == sc
    assert np.array_equal(bits(host(got)), bits(exp))


def test_lift_split_layout(ctx, S):
    """lmain = the lifted rows' first d coordinates; ltail = the FP32 sum of
    the two rounded squares (1 - h0)^2 and h1^2 (the lift coordinates' term
    of the whole-row distance to a query's (1, 0))."""
    n, dl = S.lifted.shape
    d = dl - 2
    lmain, ltail = ctx.lift_split(ctx.to_dev(S.lifted))
    assert np.array_equal(bits(host(lmain)), bits(np.ascontiguousarray(S.lifted[:, :d])))
    h0, h1 = S.lifted[:, d].astype(np.float32), S.lifted[:, d + 1].astype(np.float32)
    a0 = np.float32(1.0) - h0
    exp = (a0 * a0).astype(np.float32) + (h1 * h1).astype(np.float32)
    assert exp.dtype == np.float32
    assert np.array_equal(bits(host(ltail)), bits(exp))


def test_graph_steps_match_oracle_on_sample(ctx, S):
    _graph_steps(ctx, S.lifted, np.ascontiguousarray(S.rec_ids), "sample")


def test_graph_steps_match_oracle_on_mixture(ctx):
    base = workload.mixture(3000, 128, 40, 20.0, seed=5)
    ids = np.zeros((3000, 100), np.uint32)
    keys = np.zeros((3000, 100), np.float32)
    oracle().or_exact_topk(ptr(base), 3000, 128, 100, ptr(ids), ptr(keys), 8)
    # perturb the warm start so LocalJoin has work to do
    rng = np.random.default_rng(1)
    ids[:, 50:] = rng.integers(0, 3000, (3000, 50))
    import pyanqi
    lifted, _, _ = pyanqi.lift_host(base, np.ascontiguousarray(keys[:, 99]))
    _graph_steps(ctx, lifted, ids, "mixture")


def _graph_steps(ctx, x, warm, tag):
    O = oracle()
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    xd, nd_ = ctx.to_dev(x), ctx.to_dev(nrm)
    cap = 50
    say(tag, "warm start")
    hi_o = np.zeros((n, cap), np.uint32)
    hv_o = np.zeros((n, cap), np.float32)
    O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(warm), warm.shape[1], cap, ptr(hi_o), ptr(hv_o))
    hi, hv = ctx.warm_start(xd, nd_, ctx.to_dev(warm.view(np.int32)), cap)
    assert np.array_equal(host(hi).view(np.uint32), hi_o) and np.array_equal(bits(host(hv)), bits(hv_o))
    say(tag, "localjoin")
    upd_o = np.zeros(6, np.int64)
    hi1, hv1 = hi.clone(), hv.clone()
    it_o = O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, d, cap, ptr(hi_o), ptr(hv_o), 6, 20, ptr(upd_o))
    it, upd = ctx.nndescent(xd, nd_, hi, hv, 6, 20)
    assert it == it_o and upd.tolist() == upd_o[:it_o].tolist()
    assert np.array_equal(host(hi).view(np.uint32), hi_o) and np.array_equal(bits(host(hv)), bits(hv_o))
    # the MFMA form (bf16 gram of the sample, FP32 re-score near the roots)
    it1, upd1 = ctx.nndescent(xd, nd_, hi1, hv1, 6, 20, form=1)
    assert it1 == it_o and upd1.tolist() == upd_o[:it_o].tolist()
    assert np.array_equal(host(hi1).view(np.uint32), hi_o) and np.array_equal(bits(host(hv1)), bits(hv_o))
    say(tag, "reverse lists")
    roff, rids = ctx.reverse_csr(hi, n)
    roff_o = np.zeros(n + 1, np.uint32)
    rids_o = np.zeros(n * cap, np.uint32)
    O.or_reverse_csr(ptr(hi_o), n, cap, ptr(roff_o), ptr(rids_o))
    ro, ri = host(roff).view(np.uint32), host(rids).view(np.uint32)
    assert np.array_equal(ro, roff_o)
    for i in range(0, n, 7):
        assert sorted(ri[ro[i]:ro[i + 1]].tolist()) == rids_o[roff_o[i]:roff_o[i + 1]].tolist()
    say(tag, "gather + prune")
    Og, oc = ctx.gather_prune(xd, nd_, hi, hv, roff, rids, 100, 800, 48, 1.2)
    Og, oc = host(Og).view(np.uint32), host(oc).view(np.uint32)
    # the batched form (one expansion's first visits processed together)
    Og1, oc1 = ctx.gather_prune(xd, nd_, hi, hv, roff, rids, 100, 800, 48, 1.2, form=1)
    assert np.array_equal(host(oc1), oc) and np.array_equal(host(Og1).view(np.uint32), Og)
    O_o = np.zeros((n, 48), np.uint32)
    oc_o = np.zeros(n, np.uint32)
    pmax = 8 * n
    for i in range(n):
        pd = np.zeros(pmax, np.float32)
        pi = np.zeros(pmax, np.uint32)
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi_o), ptr(hv_o), cap, ptr(roff_o), ptr(rids_o), i, 100, 800,
                        ptr(pd), ptr(pi), pmax)
        out = np.zeros(48, np.uint32)
        dg = O.or_prune(ptr(x), ptr(nrm), d, i, n, ptr(pd), ptr(pi), m, 48, 1.2, ptr(out))
        O_o[i, :dg] = out[:dg]
        oc_o[i] = dg
        assert oc[i] == dg and Og[i, :dg].tolist() == out[:dg].tolist(), (tag, i)
    say(tag, "commit")
    links, deg = ctx.commit(xd, nd_, ctx.to_dev(O_o.view(np.int32)), ctx.to_dev(oc_o.view(np.int32)), 48, 1.2)
    l_o = np.zeros((n, 48), np.uint32)
    d_o = np.zeros(n, np.uint32)
    O.or_commit(ptr(x), ptr(nrm), n, d, ptr(O_o), ptr(oc_o), 48, 1.2, ptr(l_o), ptr(d_o))
    links, deg = host(links).view(np.uint32), host(deg).view(np.uint32)
    assert np.array_equal(deg, d_o)
    for i in range(n):
        assert links[i, :deg[i]].tolist() == l_o[i, :d_o[i]].tolist(), (tag, i)


@pytest.mark.parametrize("form", [0, 1], ids=["per-chunk", "batched"])
def test_gather_topk_matches_oracle(ctx, S, form):
    O = oracle()
    x = S.base
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    cap = 100
    hi = np.full((n, cap), U32MAX, np.uint32)
    hv = np.full((n, cap), np.finfo(np.float32).max, np.float32)
    rng = np.random.default_rng(2)
    for i in range(n):
        for j in rng.choice(n, 60, replace=False):
            if j != i:
                O.or_heap_push(ptr(hi[i]), ptr(hv[i]), cap, int(j), O.or_d2(ptr(x[i]), ptr(x[j]), nrm[i], nrm[j], d))
    roff = np.zeros(n + 1, np.uint32)
    rids = np.zeros(n * cap, np.uint32)
    O.or_reverse_csr(ptr(hi), n, cap, ptr(roff), ptr(rids))
    say("gather topk")
    ids, keys, cnt = ctx.gather_topk(ctx.to_dev(x), ctx.to_dev(nrm), ctx.to_dev(hi.view(np.int32)), ctx.to_dev(hv),
                                     ctx.to_dev(roff.view(np.int32)), ctx.to_dev(rids.view(np.int32)), 64, 512, 100,
                                     form=form)
    ids, keys, cnt = host(ids).view(np.uint32), host(keys), host(cnt)
    for i in range(n):
        pd = np.zeros(8 * n, np.float32)
        pi = np.zeros(8 * n, np.uint32)
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi), ptr(hv), cap, ptr(roff), ptr(rids), i, 64, 512,
                        ptr(pd), ptr(pi), 8 * n)
        ri = np.zeros(100, np.uint32)
        rk = np.zeros(100, np.float32)
        c = O.or_pool_topk(i, ptr(pd), ptr(pi), m, 100, ptr(ri), ptr(rk))
        assert cnt[i] == c and ids[i, :c].tolist() == ri[:c].tolist()
        assert np.array_equal(bits(keys[i, :c]), bits(rk[:c]))


def test_gather_topk_16bit_visited_set_matches_oracle(ctx):
    """The chosen AKNN gather with the build's parameters (queue 600, pool 4800,
    K 100) keeps its visited set in 16-bit entries (ids < 2^20): on 20,000
    mixture rows the set fills to ~60 % per row, so keys sit displaced from
    their home buckets; every checked row equals the oracle's gather."""
    O = oracle()
    x = workload.mixture(20000, 96, 60, 15.0, seed=9)
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    # seed lists: 40 random other rows per row with their squared distances
    # (both sides read the same arrays; slots past 40 are empty)
    cap = 100
    rng = np.random.default_rng(4)
    nb = rng.integers(0, n, (n, 40)).astype(np.uint32)
    rows = np.arange(n, dtype=np.uint32)[:, None]
    nb = np.where(nb == rows, (nb + 1) % n, nb).astype(np.uint32)
    hi = np.full((n, cap), U32MAX, np.uint32)
    hv = np.full((n, cap), np.finfo(np.float32).max, np.float32)
    hi[:, :40] = nb
    hv[:, :40] = ((x[:, None, :] - x[nb]) ** 2).sum(-1).astype(np.float32)
    roff = np.zeros(n + 1, np.uint32)
    rids = np.zeros(n * cap, np.uint32)
    O.or_reverse_csr(ptr(hi), n, cap, ptr(roff), ptr(rids))
    say("gather topk, 16-bit visited set")
    ids, keys, cnt = ctx.gather_topk(ctx.to_dev(x), ctx.to_dev(nrm), ctx.to_dev(hi.view(np.int32)), ctx.to_dev(hv),
                                     ctx.to_dev(roff.view(np.int32)), ctx.to_dev(rids.view(np.int32)), 600, 4800, 100,
                                     form=1)
    ids, keys, cnt = host(ids).view(np.uint32), host(keys), host(cnt)
    pd = np.zeros(16384, np.float32)
    pi = np.zeros(16384, np.uint32)
    for i in range(0, n, 97):
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi), ptr(hv), cap, ptr(roff), ptr(rids), i, 600, 4800,
                        ptr(pd), ptr(pi), 16384)
        ri = np.zeros(100, np.uint32)
        rk = np.zeros(100, np.float32)
        c = O.or_pool_topk(i, ptr(pd), ptr(pi), m, 100, ptr(ri), ptr(rk))
        assert cnt[i] == c and ids[i, :c].tolist() == ri[:c].tolist(), i
        assert np.array_equal(bits(keys[i, :c]), bits(rk[:c]))


def test_gather_topk_24bit_visited_set_matches_oracle(ctx):
    """With more than 2^20 rows the chosen AKNN gather keeps its visited set in
    16-bit entries of 24-bit ids plus a 32-bit overflow table (keys whose seven
    probe buckets are full): 1.1M rows, the build's queue/pool sizes, sampled
    rows against the oracle's gather."""
    O = oracle()
    n, d = 1_100_000, 32
    x = workload.mixture(n, d, 400, 15.0, seed=13)
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    cap = 100
    rng = np.random.default_rng(6)
    nb = rng.integers(0, n, (n, 40)).astype(np.uint32)
    rows = np.arange(n, dtype=np.uint32)[:, None]
    nb = np.where(nb == rows, (nb + 1) % n, nb).astype(np.uint32)
    hi = np.full((n, cap), U32MAX, np.uint32)
    hv = np.full((n, cap), np.finfo(np.float32).max, np.float32)
    hi[:, :40] = nb
    for c0 in range(0, n, 50000):
        c1 = min(n, c0 + 50000)
        hv[c0:c1, :40] = ((x[c0:c1, None, :] - x[nb[c0:c1]]) ** 2).sum(-1).astype(np.float32)
    roff = np.zeros(n + 1, np.uint32)
    rids = np.zeros(n * cap, np.uint32)
    O.or_reverse_csr(ptr(hi), n, cap, ptr(roff), ptr(rids))
    say("gather topk, 24-bit ids")
    ids, keys, cnt = ctx.gather_topk(ctx.to_dev(x), ctx.to_dev(nrm), ctx.to_dev(hi.view(np.int32)), ctx.to_dev(hv),
                                     ctx.to_dev(roff.view(np.int32)), ctx.to_dev(rids.view(np.int32)), 600, 4800, 100,
                                     form=1)
    ids, keys, cnt = host(ids).view(np.uint32), host(keys), host(cnt)
    pd = np.zeros(16384, np.float32)
    pi = np.zeros(16384, np.uint32)
    for i in list(range(0, n, 9973)) + [n - 1]:
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi), ptr(hv), cap, ptr(roff), ptr(rids), i, 600, 4800,
                        ptr(pd), ptr(pi), 16384)
        ri = np.zeros(100, np.uint32)
        rk = np.zeros(100, np.float32)
        c = O.or_pool_topk(i, ptr(pd), ptr(pi), m, 100, ptr(ri), ptr(rk))
        assert cnt[i] == c and ids[i, :c].tolist() == ri[:c].tolist(), i
        assert np.array_equal(bits(keys[i, :c]), bits(rk[:c]))


def test_search_tree_matches_reference_draws(ctx, S):
    """The device-built search tree equals the oracle's restatement of the
    reference's build (one MT19937 stream, depth-first) node for node: the
    sample tree as the reference wrote it, and every tree dataset (ties,
    degenerate deals, depth cap, stream and node-table regrowth)."""
    from anqi_testlib import oracle_rp_tree, tree_diff, tree_datasets
    say("search tree: sample")
    t = ctx.rp_tree(ctx.to_dev(S.lifted), leaf=100)
    h = ctx.tree_host(t)
    ctx.tree_free(t)
    assert tree_diff(h, dict(S.tree)) == ""
    for name, x, leaf, depth in tree_datasets():
        say("search tree:", name)
        t = ctx.rp_tree(ctx.to_dev(x), leaf=leaf, max_depth=depth)
        h = ctx.tree_host(t)
        ctx.tree_free(t)
        o = oracle_rp_tree(x, leaf, 10, depth)
        assert tree_diff(h, o) == "", name


def test_aknn_build_sample_recall(ctx, S):
    say("aknn build sample")
    ids, keys, st = ctx.aknn_build(ctx.to_dev(S.base), K=100, leaf=64, trees=8, ef=64)
    ids, keys = host(ids).view(np.uint32), host(keys)
    n = S.n
    for i in range(n):
        row = ids[i]
        assert (row < n).all() and (row != i).all() and len(set(row.tolist())) == 100
        assert (np.diff(keys[i]) >= 0).all()
    rec = np.mean([len(set(ids[i]) & set(S.gt_ids[i])) / 100 for i in range(n)])
    rec_ref = np.mean([len(set(S.rec_ids[i]) & set(S.gt_ids[i])) / 100 for i in range(n)])
    say("aknn recall", rec, "reference", rec_ref)
    assert rec >= rec_ref - 0.005


@pytest.mark.parametrize("precision", [1, 2], ids=["bf16", "fp8"])
@pytest.mark.parametrize("case", ["sample-256x128", "mixture-5000x128", "gaussian-3000x96", "wide-norms-2000x128"])
def test_aknn_mfma_leaf_sweep_equals_fp32(ctx, S, case, precision):
    """The MFMA leaf sweeps (precision 1: bf16, 2: FP8-e4m3 trial) leave
    every row's list exactly as the FP32 sweep does: identical AKNN ids and
    key bits after the whole build (leaf all-pairs over every tree, then the
    gather).  Reports the fraction of ordered leaf pairs the filter let
    through to the exact re-score."""
    rng = np.random.default_rng(11)
    x = {"sample-256x128": lambda: S.base,
         "mixture-5000x128": lambda: workload.mixture(5000, 128, 40, 20.0, seed=3),
         "gaussian-3000x96": lambda: workload.gaussian(3000, 96, seed=4),
         "wide-norms-2000x128": lambda: (rng.standard_normal((2000, 128)) *
                                         np.exp(rng.uniform(-3, 3, (2000, 1)))).astype(np.float32)}[case]()
    x = np.ascontiguousarray(x, np.float32)
    say("aknn mfma vs fp32:", case)
    xd = ctx.to_dev(x)
    K = 100 if x.shape[0] > 300 else 50
    a_ids, a_keys, _ = ctx.aknn_build(xd, K=K, leaf=200, trees=8, ef=64, precision=0)
    b_ids, b_keys, st = ctx.aknn_build(xd, K=K, leaf=200, trees=8, ef=64, precision=precision)
    say(f"  precision {precision}: {st.leaf_survivors} of {2 * st.leaf_pairs} ordered leaf pairs re-scored "
        f"({st.leaf_survivors / max(1, 2 * st.leaf_pairs):.3f}), leaf sweep {st.leaf_s * 1e3:.1f} ms")
    assert 0 < st.leaf_survivors <= 2 * st.leaf_pairs
    assert np.array_equal(host(a_ids), host(b_ids))
    assert np.array_equal(bits(host(a_keys)), bits(host(b_keys)))
