"""Pins the oracle's restatement of each hot-path function against the
reference: golden fixtures produced by the reference tools (tests/golden/sample)
and, where the harness is built, direct calls into the reference's own
compiled methods on identical inputs."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

from anqi_testlib import Sample, or_tree, oracle, ptr, ref, ref_available, GOLDEN
import workload

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference harness not built here")
U32MAX = np.uint32(0xFFFFFFFF)


@pytest.fixture(scope="module")
def S():
    return Sample()


def test_fixture_inputs_are_the_named_workload(S):
    assert workload.sha256(S.base) == workload.sha256(workload.sample_base())
    assert workload.sha256(S.query) == workload.sha256(workload.sample_query())


def test_exact_topk_matches_dataset_gt(S):
    O = oracle()
    ids = np.zeros((S.n, 100), np.uint32)
    keys = np.zeros((S.n, 100), np.float32)
    assert O.or_exact_topk(ptr(S.base), S.n, S.d, 100, ptr(ids), ptr(keys), 4) == 0
    assert np.array_equal(ids, S.gt_ids)
    assert np.array_equal(keys.view(np.uint32), S.gt_keys.view(np.uint32))


def test_exact_topk_brute_force_tiny():
    O = oracle()
    rng = np.random.default_rng(1)
    x = rng.integers(0, 4, (40, 8)).astype(np.float32)   # many ties
    ids = np.zeros((40, 10), np.uint32)
    keys = np.zeros((40, 10), np.float32)
    O.or_exact_topk(ptr(x), 40, 8, 10, ptr(ids), ptr(keys), 1)
    for i in range(40):
        dd = ((x - x[i]) ** 2).sum(1)
        order = sorted((dd[j], j) for j in range(40) if j != i)[:10]
        assert [j for _, j in order] == ids[i].tolist()


@pytest.mark.parametrize("k", [10, 50, 100])
def test_rknn_membership_matches_dataset_gt(S, k):
    O = oracle()
    rk2 = np.ascontiguousarray(S.gt_keys[:, k - 1])
    nq = len(S.query)
    off = np.zeros(nq + 1, np.uint32)
    tot = O.or_rknn_membership(ptr(S.base), S.n, ptr(S.query), nq, S.d, ptr(rk2), ptr(off), None, None, 2)
    ids = np.zeros(tot, np.uint32)
    keys = np.zeros(tot, np.float32)
    O.or_rknn_membership(ptr(S.base), S.n, ptr(S.query), nq, S.d, ptr(rk2), ptr(off), ptr(ids), ptr(keys), 2)
    goff, gids, gkeys = S.csr[k]
    assert np.array_equal(off, goff) and np.array_equal(ids, gids)
    assert np.array_equal(keys.view(np.uint32), gkeys.view(np.uint32))


def test_lift_matches_reference_file(S):
    O = oracle()
    lifted = np.zeros((S.n, S.d + 2), np.float32)
    m2 = ctypes.c_double()
    sc = ctypes.c_double()
    rk = np.ascontiguousarray(S.rec_keys[:, 99])
    O.or_lift(ptr(S.base), S.n, S.d, ptr(rk), ptr(lifted), ctypes.byref(m2), ctypes.byref(sc))
    assert m2.value == S.M2 and sc.value == S.scale
    assert np.array_equal(lifted.view(np.uint32), S.lifted.view(np.uint32))


def test_query_lift_matches_harness(S):
    O = oracle()
    for i in range(len(S.query)):
        qb = np.zeros(S.d + 2, np.float32)
        O.or_lift_query(ptr(S.query[i]), S.d, S.scale, ptr(qb))
        assert np.array_equal(qb.view(np.uint32), S.h["qb"][i].view(np.uint32))


def test_tree_entry_and_range_search_match_reference_graph(S):
    O = oracle()
    t = or_tree(S.tree)
    R2 = S.M2 + 1.0
    for i in range(len(S.query)):
        q = np.ascontiguousarray(S.h["qb"][i])
        e = O.or_tree_entry(ctypes.byref(t), ptr(q))
        assert e == S.h["entries"][i]
        for L in (16, 32, 64, 128):
            out = np.zeros(S.n, np.uint32)
            nd = ctypes.c_uint64()
            c = O.or_range_search(ptr(S.lifted), S.n, S.d + 2, ptr(S.deg), ptr(S.links), 48, e, ptr(q), R2, L,
                                  ptr(out), S.n, ctypes.byref(nd))
            off, ids = S.h[f"cand_off_L{L}"], S.h[f"cand_ids_L{L}"]
            assert out[:c].tolist() == ids[off[i]:off[i + 1]].tolist()
            assert nd.value == S.h[f"ndist_L{L}"][i]


@pytest.mark.parametrize("k", [10, 50, 100])
def test_lrq_encode_materialize_postfilter(S, k):
    O = oracle()
    n = S.n
    cut = np.ascontiguousarray(S.rec_keys)
    coeff = np.zeros((n, 8), np.float32)
    codes = np.zeros((n, 100), np.uint8)
    O.or_lrq_encode(ptr(cut), n, 100, 8, ptr(S.mu), ptr(S.G), ptr(S.cb), 256, ptr(coeff), ptr(codes))
    assert np.array_equal(coeff.view(np.uint32), S.coeff.view(np.uint32))
    assert np.array_equal(codes, S.codes)
    thr = np.zeros(n, np.float32)
    O.or_lrq_materialize(ptr(coeff), ptr(codes), n, 100, 8, ptr(S.mu), ptr(S.G), ptr(S.cb), 256, k, 1.0, ptr(thr))
    assert np.array_equal(thr.view(np.uint32), S.h[f"thr_k{k}"].view(np.uint32))
    off, ids = S.h["cand_off_L128"], S.h["cand_ids_L128"]
    dec = []
    for qi in range(len(S.query)):
        for j in range(off[qi], off[qi + 1]):
            key = O.or_recheck_key(ptr(S.base[ids[j]]), ptr(S.query[qi]), S.d)
            dec.append(np.float32(key) <= thr[ids[j]])
    assert np.array_equal(np.array(dec, np.uint8), S.h[f"dec_k{k}"])


@needs_ref
def test_lrq_encode_matches_harness_wide_range():
    """Random cutoff tables spanning many binades (fma vs mul+add matters)."""
    O, R = oracle(), ref()
    rng = np.random.default_rng(3)
    n, nk, M = 500, 100, 8
    cut = np.sort(np.exp(rng.uniform(-8, 12, (n, nk))), axis=1).astype(np.float32)
    mu = cut.mean(0).astype(np.float32)
    G = rng.standard_normal((M, nk)).astype(np.float32) * np.float32(0.1)
    cb = np.sort(rng.standard_normal((nk, 256)) * 50, axis=1).astype(np.float32)
    c1 = np.zeros((n, M), np.float32); k1 = np.zeros((n, nk), np.uint8)
    c2 = np.zeros((n, M), np.float32); k2 = np.zeros((n, nk), np.uint8)
    O.or_lrq_encode(ptr(cut), n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, ptr(c1), ptr(k1))
    R.ref_lrq_build(n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, ptr(cut), ptr(c2), ptr(k2))
    assert np.array_equal(c1.view(np.uint32), c2.view(np.uint32))
    assert np.array_equal(k1, k2)
    for k in (1, 10, 57, 100):
        t1 = np.zeros(n, np.float32); t2 = np.zeros(n, np.float32)
        O.or_lrq_materialize(ptr(c1), ptr(k1), n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, k, 1.0, ptr(t1))
        R.ref_lrq_materialize(n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, ptr(c2), ptr(k2), k, 1.0, ptr(t2))
        assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))


@needs_ref
def test_heap_push_matches_harness():
    O, R = oracle(), ref()
    rng = np.random.default_rng(4)
    for cap in (1, 2, 5, 50, 100):
        m = 400
        nodes = rng.integers(0, 300, m).astype(np.uint32)
        vals = rng.integers(0, 60, m).astype(np.float32)   # ties
        ri = np.zeros(cap, np.uint32); rv = np.zeros(cap, np.float32)
        R.ref_heap_stream(cap, ptr(nodes), ptr(vals), m, ptr(ri), ptr(rv))
        oi = np.full(cap, U32MAX, np.uint32); ov = np.full(cap, np.finfo(np.float32).max, np.float32)
        for t in range(m):
            O.or_heap_push(ptr(oi), ptr(ov), cap, int(nodes[t]), float(vals[t]))
        assert np.array_equal(oi, ri) and np.array_equal(ov, rv)


class _Idx:
    """Reference nndindex over a temporary fbin of the given rows."""

    def __init__(self, x, M=48, nn_k=50, explore=100, alpha=1.2):
        self.tmp = tempfile.TemporaryDirectory()
        p = os.path.join(self.tmp.name, "x.bin")
        workload.write_fbin(p, x)
        self.h = ref().ref_index_open(p.encode(), x.shape[0], x.shape[1], M, nn_k, explore, 8.0, alpha,
                                      (self.tmp.name + "/").encode())

    def close(self):
        ref().ref_index_close(self.h)
        self.tmp.cleanup()


@needs_ref
def test_graph_build_steps_match_harness(S):
    """warm start, sequential LocalJoin, gather, prune and commit on the
    sample's lifted rows, against the reference methods (1 thread)."""
    O, R = oracle(), ref()
    x = S.lifted
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    ix = _Idx(x)
    try:
        for i in range(0, n, 17):
            for j in range(0, n, 13):
                a = O.or_d2(ptr(x[i]), ptr(x[j]), nrm[i], nrm[j], d)
                b = R.ref_index_d2(ix.h, i, j)
                assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)
        cap = 50
        ids = np.ascontiguousarray(S.rec_ids)
        hi_o = np.zeros((n, cap), np.uint32); hv_o = np.zeros((n, cap), np.float32)
        hi_r = np.zeros((n, cap), np.uint32); hv_r = np.zeros((n, cap), np.float32)
        O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(ids), 100, cap, ptr(hi_o), ptr(hv_o))
        R.ref_warm_start(ix.h, ptr(ids), 100, cap, ptr(hi_r), ptr(hv_r))
        assert np.array_equal(hi_o, hi_r) and np.array_equal(hv_o, hv_r)
        O.or_nndescent_seq(ptr(x), ptr(nrm), n, d, cap, ptr(hi_o), ptr(hv_o), 6, 20, None)
        R.ref_nndescent(ix.h, ptr(hi_r), ptr(hv_r), cap, 6)
        assert np.array_equal(hi_o, hi_r) and np.array_equal(hv_o, hv_r)
        roff = np.zeros(n + 1, np.uint32); rids = np.zeros(n * cap, np.uint32)
        O.or_reverse_csr(ptr(hi_o), n, cap, ptr(roff), ptr(rids))
        O_rows = np.zeros((n, 48), np.uint32); ocnt = np.zeros(n, np.uint32)
        for i in range(n):
            pd = np.zeros(4 * n, np.float32); pi = np.zeros(4 * n, np.uint32)
            m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi_o), ptr(hv_o), cap, ptr(roff), ptr(rids), i, 100, 800,
                            ptr(pd), ptr(pi), 4 * n)
            rd = np.zeros(4 * n, np.float32); ri = np.zeros(4 * n, np.uint32)
            m2 = R.ref_gather(ix.h, ptr(hi_o), ptr(hv_o), cap, i, 800, ptr(rd), ptr(ri), 4 * n)
            assert m == m2 and np.array_equal(pi[:m], ri[:m]) and np.array_equal(pd[:m], rd[:m])
            out = np.zeros(48, np.uint32); out2 = np.zeros(48, np.uint32)
            dg = O.or_prune(ptr(x), ptr(nrm), d, i, n, ptr(pd.copy()), ptr(pi.copy()), m, 48, 1.2, ptr(out))
            dg2 = R.ref_prune(ix.h, i, ptr(rd), ptr(ri), m, ptr(out2))
            assert dg == dg2 and np.array_equal(out[:dg], out2[:dg2])
            O_rows[i, :dg] = out[:dg]; ocnt[i] = dg
        l1 = np.zeros((n, 48), np.uint32); d1 = np.zeros(n, np.uint32)
        l2 = np.zeros((n, 48), np.uint32); d2 = np.zeros(n, np.uint32)
        O.or_commit(ptr(x), ptr(nrm), n, d, ptr(O_rows), ptr(ocnt), 48, 1.2, ptr(l1), ptr(d1))
        R.ref_commit(ix.h, ptr(O_rows), ptr(ocnt), 48, ptr(l2), ptr(d2))
        assert np.array_equal(d1, d2)
        for i in range(n):
            assert np.array_equal(l1[i, :d1[i]], l2[i, :d2[i]])
    finally:
        ix.close()


@needs_ref
def test_aknn_gather_top100_matches_harness(S):
    """build_approx_knn tail on the base rows: queue 64, cap 512 pools."""
    O, R = oracle(), ref()
    x = S.base
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    ix = _Idx(x, M=32, nn_k=100, explore=64)
    try:
        cap = 100
        hi = np.full((n, cap), U32MAX, np.uint32); hv = np.full((n, cap), np.finfo(np.float32).max, np.float32)
        rng = np.random.default_rng(2)
        for i in range(n):
            for j in rng.choice(n, 60, replace=False):
                if j != i:
                    O.or_heap_push(ptr(hi[i]), ptr(hv[i]), cap, int(j),
                                   O.or_d2(ptr(x[i]), ptr(x[j]), nrm[i], nrm[j], d))
        roff = np.zeros(n + 1, np.uint32); rids = np.zeros(n * cap, np.uint32)
        O.or_reverse_csr(ptr(hi), n, cap, ptr(roff), ptr(rids))
        for i in range(0, n, 5):
            pd = np.zeros(8 * n, np.float32); pi = np.zeros(8 * n, np.uint32)
            m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hi), ptr(hv), cap, ptr(roff), ptr(rids), i, 64, 512,
                            ptr(pd), ptr(pi), 8 * n)
            rd = np.zeros(8 * n, np.float32); ri = np.zeros(8 * n, np.uint32)
            m2 = R.ref_gather(ix.h, ptr(hi), ptr(hv), cap, i, 512, ptr(rd), ptr(ri), 8 * n)
            assert m == m2 and np.array_equal(pi[:m], ri[:m])
            ids = np.zeros(100, np.uint32); keys = np.zeros(100, np.float32)
            c = O.or_pool_topk(i, ptr(pd), ptr(pi), m, 100, ptr(ids), ptr(keys))
            pool = sorted(zip(rd[:m].tolist(), ri[:m].tolist()))
            ref_ids = [j for _, j in pool if j != i][:100]
            assert ids[:c].tolist() == ref_ids
    finally:
        ix.close()


def test_snapshot_localjoin_invariants(S):
    """The GPU schedule (snapshot + sorted application) only improves rows and
    reaches recall close to the sequential schedule."""
    O = oracle()
    x = S.lifted
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    cap = 50
    ids = np.ascontiguousarray(S.rec_ids)
    hi_a = np.zeros((n, cap), np.uint32); hv_a = np.zeros((n, cap), np.float32)
    O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(ids), 100, cap, ptr(hi_a), ptr(hv_a))
    hi_b, hv_b = hi_a.copy(), hv_a.copy()
    before = np.sort(hv_a, 1)
    O.or_nndescent_seq(ptr(x), ptr(nrm), n, d, cap, ptr(hi_a), ptr(hv_a), 6, 20, None)
    O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, d, cap, ptr(hi_b), ptr(hv_b), 6, 20, None)
    after = np.sort(hv_b, 1)
    assert (after <= before).all()
    # exact lifted top-50 for recall
    ex = np.zeros((n, cap), np.uint32); exk = np.zeros((n, cap), np.float32)
    O.or_exact_topk(ptr(x), n, d, cap, ptr(ex), ptr(exk), 2)
    ra = np.mean([len(set(hi_a[i]) & set(ex[i])) / cap for i in range(n)])
    rb = np.mean([len(set(hi_b[i]) & set(ex[i])) / cap for i in range(n)])
    assert rb >= ra - 0.01, (ra, rb)


@needs_ref
def test_sequential_localjoin_with_updates_matches_harness(S):
    """Random warm-start ids so every pass makes updates."""
    O, R = oracle(), ref()
    x = S.lifted
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    ix = _Idx(x)
    try:
        rng = np.random.default_rng(11)
        ids = np.stack([rng.choice(n, 30, replace=False) for _ in range(n)]).astype(np.uint32)
        cap = 20
        hi_o = np.zeros((n, cap), np.uint32); hv_o = np.zeros((n, cap), np.float32)
        O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(ids), 30, cap, ptr(hi_o), ptr(hv_o))
        hi_r, hv_r = hi_o.copy(), hv_o.copy()
        upd = np.zeros(6, np.int64)
        it = O.or_nndescent_seq(ptr(x), ptr(nrm), n, d, cap, ptr(hi_o), ptr(hv_o), 6, 20, ptr(upd))
        assert upd[0] > n
        R.ref_nndescent(ix.h, ptr(hi_r), ptr(hv_r), cap, 6)
        assert np.array_equal(hi_o, hi_r) and np.array_equal(hv_o, hv_r)
    finally:
        ix.close()


def test_rp_tree_matches_reference_file_and_harness():
    """or_rp_tree_build: the sample's search tree as the reference wrote it
    (graph cache), and the reference's build_rptree on every tree dataset."""
    from anqi_testlib import oracle_rp_tree, ref_rp_tree, tree_diff, tree_datasets
    S = Sample()
    a = oracle_rp_tree(S.lifted)
    g = dict(S.tree)
    assert tree_diff(a, g) == ""
    if not ref_available():
        pytest.skip("reference harness not built")
    for name, x, leaf, depth in tree_datasets():
        o = oracle_rp_tree(x, leaf, 10, depth)
        r = ref_rp_tree(x, leaf, 10, depth)
        assert tree_diff(o, r) == "", name


def test_mt19937_known_answer():
    """The 10000th output of std::mt19937 default-seeded (5489) is 4123659995
    (C++11 [rand.predef]); the first output of seed 10 matches the harness
    tree's first split only if the engine is right, covered above."""
    from anqi_testlib import oracle
    out = np.zeros(10000, np.uint32)
    oracle().or_mt_stream(5489, ptr(out), 10000)
    assert int(out[-1]) == 4123659995
