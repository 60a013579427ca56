"""Pins the oracle's Vamana refinement (include/index.h:831-1065, 1262-1282),
its entry choice (:190-222) and the reference's graph build without the
lifted LocalJoin (ANQI_REUSE_ORIG_KNN, :1202-1238) against the reference's own
build_index run on one thread in the harness: with the LocalJoin skipped, every
step of that build is deterministic, so the oracle's graph must equal it link
for link.  Also std::mt19937_64's known answer and libstdc++'s std::shuffle."""
import contextlib
import ctypes
import os

import numpy as np
import pytest

import workload
from anqi_testlib import Sample, or_tree, oracle, oracle_rp_tree, ptr, ref, ref_available
from test_oracle_pipeline import _Idx

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference harness not built here")
M, CAP, EXPLORE, POOL, ALPHA = 48, 50, 100, 800, 1.2


@contextlib.contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    try:
        for k, v in kv.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = str(v)
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_mt19937_64_known_answer():
    """The 10000th output of std::mt19937_64 default-seeded (5489) is
    9981545732273789042 (C++11 [rand.predef])."""
    O = oracle()
    g = ctypes.create_string_buffer(312 * 8 + 8)
    O.or_mt64_seed(g, 5489)
    for _ in range(9999):
        O.or_mt64_next(g)
    assert O.or_mt64_next(g) == 9981545732273789042


@needs_ref
@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 8, 100, 1001, 4096, 65537])
def test_shuffle_matches_libstdcxx(n):
    O, R = oracle(), ref()
    for seed in (10 + 0x9e3779b97f4a7c15, (10 + 2 * 0x9e3779b97f4a7c15) % (1 << 64), 5489):
        a = np.arange(n, dtype=np.uint32)
        g = ctypes.create_string_buffer(312 * 8 + 8)
        O.or_mt64_seed(g, seed)
        O.or_shuffle_u32(ptr(a), n, g)
        b = np.zeros(n, np.uint32)
        R.ref_shuffle(n, seed, ptr(b))
        assert np.array_equal(a, b), (n, seed)
        # a second shuffle continues the same engine (the passes keep the order)
        O.or_shuffle_u32(ptr(a), n, g)
        assert sorted(a.tolist()) == list(range(n))


def test_shuffle_pairs_differ_from_one_draw_per_element():
    """The planted alternative (Fisher-Yates with one draw per element) gives a
    different order: the restatement's paired draws are what the check sees."""
    O = oracle()
    n = 1000
    a = np.arange(n, dtype=np.uint32)
    g = ctypes.create_string_buffer(312 * 8 + 8)
    O.or_mt64_seed(g, 77)
    O.or_shuffle_u32(ptr(a), n, g)
    O.or_mt64_seed(g, 77)
    b = list(range(n))
    for i in range(1, n):
        r = O.or_mt64_next(g)
        j = (r * (i + 1)) >> 64
        b[i], b[j] = b[j], b[i]
    assert a.tolist() != b


def _rows(n=1200, d=130, seed=5):
    return workload.mixture(n, d, 40, 12.0, seed, rounded=False)


@needs_ref
def test_vamana_entry_choice_matches_reference():
    O, R = oracle(), ref()
    x = _rows(900)
    ix = _Idx(x)
    try:
        for sample in (None, 1, 7, 450, 900, 5000):
            with env(ANQI_VAMANA_ENTRY="medoid", ANQI_VAMANA_ENTRY_SAMPLE=sample):
                want = R.ref_vamana_entry(ix.h)
            s = min(900, 100000) if sample is None else sample
            assert O.or_vamana_medoid(ptr(x), 900, 130, s) == want, sample
        with env(ANQI_VAMANA_ENTRY="12345"):
            assert R.ref_vamana_entry(ix.h) == 12345 % 900
        with env(ANQI_VAMANA_ENTRY=None):
            assert R.ref_vamana_entry(ix.h) == 0
    finally:
        ix.close()
    # brute force: the sampled row nearest to the sample mean
    m = x[::3][:300].astype(np.float64).mean(axis=0)
    dd = ((x[::3][:300].astype(np.float64) - m) ** 2).sum(axis=1)
    assert O.or_vamana_medoid(ptr(x), 900, 130, 300) == 3 * int(np.argmin(dd))


def oracle_graph(x, ids, vamana=None, L=64, passes=2, frac=0.02, max_batch=1000000, base=2.0, early=1.0,
                 early_one=1, entry=0, seed=10, policy=None):
    """The reference's build_index with ANQI_REUSE_ORIG_KNN set, restated:
    search tree, warm start, gather + prune, commit, then the Vamana passes;
    every prune under the edge policy (an or_edge_policy) when given."""
    O = oracle()
    O.or_set_edge_policy(ctypes.byref(policy) if policy is not None else None)
    try:
        return _oracle_graph(O, x, ids, vamana, L, passes, frac, max_batch, base, early, early_one, entry, seed,
                             policy)
    finally:
        O.or_set_edge_policy(None)


def _oracle_graph(O, x, ids, vamana, L, passes, frac, max_batch, base, early, early_one, entry, seed, policy):
    x = np.ascontiguousarray(x, np.float32)
    ids = np.ascontiguousarray(ids, np.uint32)
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    hidx = np.zeros((n, CAP), np.uint32)
    hval = np.zeros((n, CAP), np.float32)
    O.or_warm_start(ptr(x), ptr(nrm), n, d, ptr(ids), ids.shape[1], CAP, ptr(hidx), ptr(hval))
    roff = np.zeros(n + 1, np.uint32)
    rids = np.zeros(n * CAP, np.uint32)
    O.or_reverse_csr(ptr(hidx), n, CAP, ptr(roff), ptr(rids))
    Orows = np.zeros((n, M), np.uint32)
    ocnt = np.zeros(n, np.uint32)
    pd = np.zeros(8 * n, np.float32)
    pi = np.zeros(8 * n, np.uint32)
    for i in range(n):
        m = O.or_gather(ptr(x), ptr(nrm), n, d, ptr(hidx), ptr(hval), CAP, ptr(roff), ptr(rids), i, EXPLORE, POOL,
                        ptr(pd), ptr(pi), 8 * n)
        out = np.zeros(M, np.uint32)
        if policy is None:
            ocnt[i] = O.or_prune(ptr(x), ptr(nrm), d, i, n, ptr(pd), ptr(pi), m, M, ALPHA, ptr(out))
        else:
            ocnt[i] = O.or_prune_policy(ptr(x), ptr(nrm), d, i, n, ptr(pd), ptr(pi), m, M, ALPHA,
                                        ctypes.byref(policy), ptr(out))
        Orows[i] = out
    links = np.zeros((n, M), np.uint32)
    deg = np.zeros(n, np.uint32)
    O.or_commit(ptr(x), ptr(nrm), n, d, ptr(Orows), ptr(ocnt), M, ALPHA, ptr(links), ptr(deg))
    if vamana == "global":
        ot = or_tree(oracle_rp_tree(x))
        entries = np.array([O.or_tree_entry(ctypes.byref(ot), ptr(x[i])) for i in range(n)], np.uint32)
        for _ in range(passes):
            O.or_vamana_global_pass(ptr(x), ptr(nrm), n, d, ptr(entries), ptr(hidx), CAP, L, M, ALPHA, ptr(links),
                                    ptr(deg))
    elif vamana == "batch":
        O.or_vamana_batch(ptr(x), ptr(nrm), n, d, ptr(hidx), CAP, L, passes, frac, max_batch, base, ALPHA, early,
                          early_one, entry, seed, M, ptr(links), ptr(deg))
    return links, deg, hidx, hval


def reference_graph(x, ids, files=(), **kv):
    """files: paths linked into the index's folder (the edge policies'
    residual table and meta file)."""
    R = ref()
    n = x.shape[0]
    ix = _Idx(x)
    for f in files:
        os.link(f, os.path.join(ix.tmp.name, os.path.basename(f)))
    try:
        deg = np.zeros(n, np.uint32)
        links = np.zeros((n, M), np.uint32)
        entry = ctypes.c_uint32()
        with env(ANQI_REUSE_ORIG_KNN="1", **kv):
            R.ref_build_graph(ix.h, ptr(np.ascontiguousarray(ids, np.uint32)), ids.shape[1], ptr(deg), ptr(links),
                              ctypes.byref(entry))
        return links, deg, entry.value
    finally:
        ix.close()


def same_graph(a, b):
    (la, da), (lb, db) = a, b
    if not np.array_equal(da, db):
        return f"degrees differ in {int((da != db).sum())} rows"
    bad = [i for i in range(len(da)) if not np.array_equal(la[i, :da[i]], lb[i, :db[i]])]
    return f"links differ in {len(bad)} rows" if bad else ""


def _inputs(name):
    if name == "sample":
        S = Sample()
        return S.lifted, np.ascontiguousarray(S.rec_ids)
    x = _rows()
    n = x.shape[0]
    ids = np.zeros((n, 100), np.uint32)
    keys = np.zeros((n, 100), np.float32)
    assert oracle().or_exact_topk(ptr(x), n, x.shape[1], 100, ptr(ids), ptr(keys), 0) == 0
    return x, ids


CASES = [
    ("none", {}, {}),
    ("global", dict(ANQI_VAMANA="1", ANQI_VAMANA_L="64"), dict(vamana="global", L=64)),
    ("global3", dict(ANQI_VAMANA="1", ANQI_VAMANA_STYLE="global", ANQI_VAMANA_L="40", ANQI_VAMANA_PASS="1"),
     dict(vamana="global", L=40, passes=1)),
    ("batch", dict(ANQI_VAMANA="1", ANQI_VAMANA_STYLE="batch", ANQI_VAMANA_L="64"), dict(vamana="batch", L=64)),
    ("batch_opts", dict(ANQI_VAMANA="1", ANQI_VAMANA_STYLE="batch", ANQI_VAMANA_L="48", ANQI_VAMANA_PASS="3",
                        ANQI_VAMANA_BATCH_FRAC="0.05", ANQI_VAMANA_BATCH_BASE="3", ANQI_VAMANA_EARLY_ALPHA="1.1",
                        ANQI_VAMANA_ENTRY="medoid"),
     dict(vamana="batch", L=48, passes=3, frac=0.05, base=3.0, early=1.1, entry="medoid")),
]


@needs_ref
@pytest.mark.parametrize("data", ["sample", "mixture"])
@pytest.mark.parametrize("name,refenv,kw", CASES, ids=[c[0] for c in CASES])
def test_graph_build_matches_reference_build_index(data, name, refenv, kw):
    x, ids = _inputs(data)
    kw = dict(kw)
    if kw.get("entry") == "medoid":
        kw["entry"] = oracle().or_vamana_medoid(ptr(np.ascontiguousarray(x)), x.shape[0], x.shape[1],
                                                min(x.shape[0], 100000))
    links, deg, _, _ = oracle_graph(x, ids, **kw)
    rl, rd, rentry = reference_graph(x, ids, **refenv)
    assert same_graph((links, deg), (rl, rd)) == "", (data, name)
    if kw.get("vamana") == "batch":
        assert rentry == kw.get("entry", 0)


@needs_ref
def test_vamana_changes_the_graph():
    """Each style changes the links it refines (so the equalities above are not
    vacuous), and the collect pool holds every first visit but the row."""
    x, ids = _inputs("mixture")
    base = oracle_graph(x, ids)
    for style in ("global", "batch"):
        g = oracle_graph(x, ids, vamana=style, L=64)
        assert same_graph(base[:2], g[:2]) != "", style
    x, ids = _inputs("sample")
    base = oracle_graph(x, ids)
    O = oracle()
    links, deg, _, _ = base
    n, d = x.shape
    nrm = np.zeros(n, np.float32)
    O.or_norms(ptr(x), n, d, ptr(nrm))
    pd = np.zeros(n, np.float32)
    pi = np.zeros(n, np.uint32)
    m = O.or_vamana_collect(ptr(x), ptr(nrm), n, d, ptr(deg), ptr(links), M, 5, 0, 1000, ptr(pd), ptr(pi), n)
    # a beam larger than the graph expands every reachable node: the pool is the
    # reachable set minus the row, each once
    seen, todo = {0}, [0]
    while todo:
        c = todo.pop()
        for nb in links[c, :deg[c]]:
            if int(nb) not in seen:
                seen.add(int(nb))
                todo.append(int(nb))
    assert sorted(pi[:m].tolist()) == sorted(seen - {5})
