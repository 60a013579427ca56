"""Pins the oracle's edge policies (or_prune_policy, include/index.h:283-829)
against the reference's own prune_pool_to_ids in the harness: each policy is
configured from the environment as the reference configures it (residual
table, meta scale and lsg weights as files), and every row's gather pool of
the sample's lifted graph build is pruned by both."""
import os
import tempfile

import numpy as np
import pytest

from anqi_testlib import Sample, edge_policy_config, or_edge_policy, oracle, ptr, ref, ref_available
from test_oracle_pipeline import _Idx
from test_oracle_vamana import env

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference harness not built here")
M, CAP = 48, 50

POLICIES = {
    "residual_cover": {"ANQI_EDGE_POLICY": "residual_cover"},
    "residual_cover_ks": {"ANQI_EDGE_POLICY": "residual_cover", "ANQI_EDGE_KS": "3,10,999",
                          "ANQI_EDGE_RESID_BUDGET": "7", "ANQI_EDGE_RESID_LAMBDA": "2.5"},
    "residual_cover_off": {"ANQI_EDGE_POLICY": "residual_cover", "ANQI_EDGE_RESID_BUDGET": "0"},
    "ball_rng": {"ANQI_EDGE_POLICY": "ball_rng"},
    "ball_hybrid": {"ANQI_EDGE_POLICY": "ball_hybrid"},
    "ball_hybrid_any": {"ANQI_EDGE_POLICY": "ball_hybrid", "ANQI_EDGE_BALL_OCCLUSION": "any",
                        "ANQI_EDGE_RESID_BUDGET": "44"},
    "bubble_multik": {"ANQI_EDGE_POLICY": "bubble_multik", "ANQI_EDGE_KS": "5,20", "ANQI_EDGE_RESID_BUDGET": "8"},
    "bubble_rng_nk": {"ANQI_EDGE_POLICY": "bubble_rng", "ANQI_EDGE_NK": "30", "ANQI_EDGE_KS": "1,30"},
    "lsg_rng": {"ANQI_EDGE_POLICY": "lsg_rng"},
    "lsg_rng_half": {"ANQI_EDGE_POLICY": "lsg_rng", "ANQI_LSG_ALPHA": "0.5"},
    "lsg_bubble_hybrid": {"ANQI_EDGE_POLICY": "lsg_bubble_hybrid", "ANQI_EDGE_RESID_BUDGET": "12"},
    "bubble_auto": {"ANQI_EDGE_POLICY": "bubble_auto"},
    "bubble_auto_opts": {"ANQI_EDGE_POLICY": "bubble_auto", "ANQI_EDGE_AUTO_CANDIDATES": "20",
                         "ANQI_EDGE_AUTO_MAX_SWAPS": "5", "ANQI_EDGE_RESID_BUDGET": "0", "ANQI_EDGE_BALL_OCCLUSION": "any"},
    "bubble_auto_full": {"ANQI_EDGE_POLICY": "bubble_auto_full", "ANQI_EDGE_KS": "10,100"},
}


def pools(S):
    """The sample's lifted rows and every row's gather pool (warm start from the
    reference's .rec, no LocalJoin), as the build's first prune sees them."""
    O = oracle()
    x = np.ascontiguousarray(S.lifted)
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    hidx = np.zeros((n, CAP), np.uint32)
    hval = np.zeros((n, CAP), np.float32)
    O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(np.ascontiguousarray(S.rec_ids)), 100, CAP, ptr(hidx), ptr(hval))
    roff = np.zeros(n + 1, np.uint32)
    rids = np.zeros(n * CAP, np.uint32)
    O.or_reverse_csr(ptr(hidx), n, CAP, ptr(roff), ptr(rids))
    out = []
    for i in range(n):
        pd = np.zeros(4 * n, np.float32)
        pi = np.zeros(4 * n, np.uint32)
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hidx), ptr(hval), CAP, ptr(roff), ptr(rids), i, 100, 800,
                        ptr(pd), ptr(pi), 4 * n)
        out.append((pd[:m].copy(), pi[:m].copy()))
    return x, nrm, out


def write_inputs(folder, n, nk=100, seed=3, scale=1.0):
    """A residual table on the scale of the pools' distances (signed, so that
    the cover's positive-risk filter matters), a meta file and lsg weights."""
    rng = np.random.default_rng(seed)
    resid = (rng.standard_normal((n, nk)) * 2.0e4).astype(np.float32)
    resid.tofile(folder + "_anqi_rankm_residual.bin")
    with open(folder + "_anqi_meta.txt", "w") as f:
        f.write(f"1.5 {scale!r}\n")
    mu = rng.uniform(0.5, 2.0, n).astype(np.float32)
    mu.tofile(folder + "mu.bin")
    return resid, mu


@needs_ref
@pytest.mark.parametrize("name", list(POLICIES))
def test_policy_prune_matches_reference(name):
    S = Sample()
    x, nrm, P = pools(S)
    n, d = x.shape
    O, R = oracle(), ref()
    with tempfile.TemporaryDirectory() as td:
        folder = td + "/"
        kv = dict(POLICIES[name])
        nk = int(kv.get("ANQI_EDGE_NK", "100"))
        write_inputs(folder, n, nk)
        kv["ANQI_LSG_MU_FILE"] = folder + "mu.bin"
        f = edge_policy_config(kv, n, M, folder)
        E = or_edge_policy(f) if f else None
        ix = _Idx(x)
        # the index's folder is its own temporary directory: point the files there
        for suffix in ("_anqi_rankm_residual.bin", "_anqi_meta.txt"):
            os.link(folder + suffix, ix.tmp.name + "/" + suffix)
        try:
            with env(**kv):
                differ = 0
                for i, (pd, pi) in enumerate(P):
                    a = np.zeros(M, np.uint32)
                    if E is None:
                        da = O.or_prune(ptr(x), ptr(nrm), d, i, n, ptr(pd.copy()), ptr(pi.copy()), len(pd), M, 1.2,
                                        ptr(a))
                    else:
                        da = O.or_prune_policy(ptr(x), ptr(nrm), d, i, n, ptr(pd.copy()), ptr(pi.copy()), len(pd), M,
                                               1.2, ctypes_ref(E), ptr(a))
                    b = np.zeros(M, np.uint32)
                    db = R.ref_prune(ix.h, i, ptr(pd), ptr(pi), len(pd), ptr(b))
                    differ += not (da == db and np.array_equal(a[:da], b[:db]))
                assert differ == 0, (name, differ)
        finally:
            ix.close()


def ctypes_ref(E):
    import ctypes
    return ctypes.byref(E)


def test_policies_change_the_prune():
    """Each configured policy selects different edges than the plain prune on
    some rows (so the equalities above are not vacuous)."""
    S = Sample()
    x, nrm, P = pools(S)
    n, d = x.shape
    O = oracle()
    with tempfile.TemporaryDirectory() as td:
        folder = td + "/"
        write_inputs(folder, n)
        for name, kv in POLICIES.items():
            kv = dict(kv, ANQI_LSG_MU_FILE=folder + "mu.bin")
            if "ANQI_EDGE_NK" in kv:
                write_inputs(folder, n, int(kv["ANQI_EDGE_NK"]))
            f = edge_policy_config(kv, n, M, folder)
            if name == "residual_cover_off":
                assert f is None
                continue
            if name == "ball_hybrid_any":
                # with "any" the ball phase takes only candidates that no
                # selected edge occludes at any rank: on this sample, edges
                # the plain prune takes anyway (its pin above still holds)
                continue
            E = or_edge_policy(f)
            changed = 0
            for i, (pd, pi) in enumerate(P):
                a = np.zeros(M, np.uint32)
                b = np.zeros(M, np.uint32)
                da = O.or_prune_policy(ptr(x), ptr(nrm), d, i, n, ptr(pd.copy()), ptr(pi.copy()), len(pd), M, 1.2,
                                       ctypes_ref(E), ptr(a))
                db = O.or_prune(ptr(x), ptr(nrm), d, i, n, ptr(pd.copy()), ptr(pi.copy()), len(pd), M, 1.2, ptr(b))
                changed += not (da == db and np.array_equal(a[:da], b[:db]))
            assert changed > 0, name
            write_inputs(folder, n)


GRAPH_CASES = [
    ("bubble_auto", None), ("bubble_auto_full", "global"),
    ("residual_cover", None), ("ball_hybrid", None), ("bubble_rng_nk", None), ("lsg_bubble_hybrid", None),
    ("lsg_rng", "batch"), ("ball_rng", "global"), ("residual_cover_ks", "batch"),
]


@needs_ref
@pytest.mark.parametrize("name,vamana", GRAPH_CASES, ids=[f"{a}-{b}" for a, b in GRAPH_CASES])
def test_policy_graph_matches_reference_build_index(name, vamana):
    """The whole graph build under a policy (every prune: the gathered pools,
    commit_edges and the Vamana passes) equals the reference's single-threaded
    build_index with ANQI_REUSE_ORIG_KNN, link for link."""
    from test_oracle_vamana import oracle_graph, reference_graph, same_graph
    S = Sample()
    x, ids = np.ascontiguousarray(S.lifted), np.ascontiguousarray(S.rec_ids)
    n = x.shape[0]
    with tempfile.TemporaryDirectory() as td:
        folder = td + "/"
        kv = dict(POLICIES[name])
        write_inputs(folder, n, int(kv.get("ANQI_EDGE_NK", "100")))
        kv["ANQI_LSG_MU_FILE"] = folder + "mu.bin"
        E = or_edge_policy(edge_policy_config(kv, n, M, folder))
        vk = {} if vamana is None else dict(vamana=vamana, L=64)
        rk = {} if vamana is None else dict(ANQI_VAMANA="1", ANQI_VAMANA_STYLE=vamana, ANQI_VAMANA_L="64")
        want = oracle_graph(x, ids, policy=E, **vk)[:2]
        got = reference_graph(x, ids, files=(folder + "_anqi_rankm_residual.bin", folder + "_anqi_meta.txt"),
                              **kv, **rk)[:2]
        assert same_graph(want, got) == "", (name, vamana)
        plain = oracle_graph(x, ids, **vk)[:2]
        assert same_graph(want, plain) != "", (name, vamana)
