"""Pins the oracle's inner-product ("ip") metric functions against the
reference: its compiled "ip" distance and query expressions (harness), and the
files its own tools wrote for the 1,000 x 32 ip sample (tests/golden/sample_ip,
made by tests/golden/make_golden.py): dataset_gt's exact top-100 and RkNN CSR,
rknn_query's lifted vectors, graph meta and verifier artifact, and the harness
dumps on the reference's graph (per-query R2 and candidate lists, thresholds,
post-filter keys and decisions)."""
import ctypes

import numpy as np
import pytest

from anqi_testlib import Sample, or_tree, oracle, ptr, read_csr, ref, ref_available

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference harness not built here")


@pytest.fixture(scope="module")
def S():
    return Sample("sample_ip")


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@needs_ref
@pytest.mark.parametrize("d", [1, 2, 7, 8, 9, 15, 16, 17, 23, 24, 25, 31, 32, 64, 96, 98, 128, 130, 200, 300])
def test_ipdist_and_query_lift_bits_match_reference(d):
    """include/distance.h:170-175 as GCC compiles it (vectorised products added
    in order, fused tail) and rknn_query.cpp:2719-2730 (lifted query, R2)."""
    O, R = oracle(), ref()
    rng = np.random.default_rng(100 + d)
    for t in range(200):
        sc = [1e-3, 1.0, 37.0, 1e4][t % 4]
        x = (rng.standard_normal(d) * sc).astype(np.float32)
        y = (rng.standard_normal(d) * sc).astype(np.float32)
        assert bits(O.or_ipdist(ptr(x), ptr(y), d)) == bits(R.ref_ipdist(ptr(x), ptr(y), d)), (d, t)
        assert bits(O.or_recheck_key_ip(ptr(x), ptr(y), d)) == bits(R.ref_recheck_key_ip(ptr(x), ptr(y), d))
        for ip in (0, 1):
            qa, qb = np.zeros(d + 2, np.float32), np.zeros(d + 2, np.float32)
            ra = O.or_query_lift(ptr(x), d, 0.0173 * (t + 1), 1.5 + t, ip, ptr(qa))
            rb = R.ref_query_lift(ptr(x), d, 0.0173 * (t + 1), 1.5 + t, ip, ptr(qb))
            assert ra == rb and (bits(qa) == bits(qb)).all(), (d, t, ip)


def test_ipdist_fused_tail_matters():
    """The planted variant with every step fused (or every step unfused)
    differs from the reference's evaluation on these inputs, so the
    restatement's split between the two is what the bits check."""
    O = oracle()
    rng = np.random.default_rng(7)
    fused_diff = plain_diff = 0
    for t in range(300):
        d = 27  # 16 vector products, then 8, then a fused tail of 3
        x, y = rng.standard_normal(d).astype(np.float32), rng.standard_normal(d).astype(np.float32)
        ref_v = bits(O.or_ipdist(ptr(x), ptr(y), d))
        p = np.float32(0)
        for i in range(d):
            p = np.float32(np.float64(x[i]) * np.float64(y[i]) + np.float64(p))   # fma: exact product, one rounding
        fused_diff += bits(np.float32(1) - p) != ref_v
        p = np.float32(0)
        for i in range(d):
            p = np.float32(p + np.float32(x[i] * y[i]))
        plain_diff += bits(np.float32(1) - p) != ref_v
    assert fused_diff > 0 and plain_diff > 0


def test_exact_topk_and_membership_match_dataset_gt(S):
    O = oracle()
    K = 100
    ids = np.zeros((S.n, K), np.uint32)
    keys = np.zeros((S.n, K), np.float32)
    assert O.or_exact_topk_metric(ptr(S.base), S.n, S.d, K, 1, ptr(ids), ptr(keys), 0) == 0
    assert (ids == S.gt_ids).all() and (bits(keys) == bits(S.gt_keys)).all()
    nq = len(S.query)
    for k in (10, 50, 100):
        rk = np.ascontiguousarray(S.gt_keys[:, k - 1])
        off = np.zeros(nq + 1, np.uint32)
        tot = O.or_rknn_membership_metric(ptr(S.base), S.n, ptr(S.query), nq, S.d, 1, ptr(rk), ptr(off), None,
                                          None, 0)
        cids = np.zeros(max(tot, 1), np.uint32)
        ckeys = np.zeros(max(tot, 1), np.float32)
        O.or_rknn_membership_metric(ptr(S.base), S.n, ptr(S.query), nq, S.d, 1, ptr(rk), ptr(off), ptr(cids),
                                    ptr(ckeys), 0)
        goff, gids, gkeys = S.csr[k]
        assert (off == goff).all() and (cids[:tot] == gids).all() and (bits(ckeys[:tot]) == bits(gkeys)).all(), k


def test_lift_matches_reference_files(S):
    """rknn_query.cpp:2203-2236 for ip: the reference's _approxlift file and
    graph meta (M*^2, scale), horizon = the .rec's 100th key."""
    O = oracle()
    lifted = np.zeros((S.n, S.d + 2), np.float32)
    m2, sc = ctypes.c_double(), ctypes.c_double()
    rk = np.ascontiguousarray(S.rec_keys[:, 99])
    O.or_lift_metric(ptr(S.base), S.n, S.d, ptr(rk), 1, ptr(lifted), ctypes.byref(m2), ctypes.byref(sc))
    assert (bits(lifted) == bits(S.lifted)).all()
    assert m2.value == S.M2 and sc.value == S.scale
    # the L2 lift of the same inputs differs (the planted alternative)
    O.or_lift_metric(ptr(S.base), S.n, S.d, ptr(rk), 0, ptr(lifted), ctypes.byref(m2), ctypes.byref(sc))
    assert not (bits(lifted) == bits(S.lifted)).all()


def test_search_thresholds_and_decisions_match_harness(S):
    O = oracle()
    h = S.h
    ot = or_tree(S.tree)
    nq = len(S.query)
    qb = np.zeros(S.d + 2, np.float32)
    r2 = np.zeros(nq)
    for i in range(nq):
        r2[i] = O.or_query_lift(ptr(S.query[i]), S.d, S.scale, S.M2, 1, ptr(qb))
        assert (bits(qb) == bits(h["qb"][i])).all()
        assert O.or_tree_entry(ctypes.byref(ot), ptr(qb)) == h["entries"][i]
    assert (r2 == h["R2"]).all()
    for L in (16, 32, 64, 128):
        off, ids, nd = h[f"cand_off_L{L}"], h[f"cand_ids_L{L}"], h[f"ndist_L{L}"]
        for i in range(nq):
            out = np.zeros(S.n, np.uint32)
            ndist = ctypes.c_uint64()
            c = O.or_range_search(ptr(S.lifted), S.n, S.d + 2, ptr(S.deg), ptr(S.links), 48, int(h["entries"][i]),
                                  ptr(h["qb"][i]), float(r2[i]), L, ptr(out), S.n, ctypes.byref(ndist))
            assert out[:c].tolist() == ids[off[i]:off[i + 1]].tolist() and ndist.value == nd[i], (L, i)
    off, ids = h["cand_off_L128"], h["cand_ids_L128"]
    keys = np.array([O.or_recheck_key_ip(ptr(S.base[ids[j]]), ptr(S.query[i]), S.d)
                     for i in range(nq) for j in range(off[i], off[i + 1])], np.float32)
    assert (bits(keys) == bits(h["postfilter_keys_L128"])).all()
    codes = np.ascontiguousarray(S.codes.reshape(S.n, 100))
    for k in (10, 50, 100):
        thr = np.zeros(S.n, np.float32)
        O.or_lrq_materialize_metric(ptr(S.coeff), ptr(codes), S.n, 100, 8, ptr(S.mu), ptr(S.G), ptr(S.cb), 256, k,
                                    1.0, 1, ptr(thr))
        assert (bits(thr) == bits(h[f"thr_k{k}"])).all(), k
        assert ((keys <= thr[ids]).astype(np.uint8) == h[f"dec_k{k}"]).all(), k
    # ip thresholds keep their sign: some are negative, which the L2 clamp would lose
    assert (h["thr_k10"] < 0).any()
