"""Index layouts of the persistent search.  Sentinel adjacency rows
(anqi_pack_links: no degree read) and byte rows (anqi_byte_rows: d bytes per
row, the base value float(b) and the lifted coordinate
float(double(b) * scale) rebuilt from them) must give exactly the
candidates, accept flags and distance counts of the float rows with links and
degrees, on integer-valued bases of d = 128 (8 bytes per lane) and d = 96
(6); sampled queries of the byte form equal the oracle's chain
(or_range_search, include/index.h:1535-1652, and the post-filter of
src/rknn_query.cpp:2762-2766).  Byte rows are refused for any base that is
not all floats 0..255 (-0, 0.5, 256) and when the lifted rows differ from the
rebuilt coordinates by one ulp; sentinel rows hold UINT32_MAX past the
degree whatever the links hold there."""
import ctypes

import numpy as np
import pytest

import workload
from anqi_testlib import oracle, or_tree, ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-layouts]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    import pyanqi
    c = pyanqi.Context(0)
    yield c
    c.close()


def run(ctx, handle, qd, thr, L):
    b = ctx.rknn_search(handle, qd, L, thr, out_cap=4096, mode=1)
    return {k: b[k].cpu().numpy().copy() for k in ("cnt", "pred", "nd", "cand", "acc", "status")}


@pytest.mark.parametrize("d", [128, 96])
def test_layouts_give_the_same_answers(ctx, d):
    import pipeline
    base = workload.mixture(20000, d, 60, 18.0, seed=31)
    ix = pipeline.build(ctx, base, log=say)
    assert ix.handle.byte_rows, "an integer-valued base must get byte rows"
    thr = pipeline.thresholds(ctx, ix, 10)
    q = workload.mixture(3000, d, 60, 18.0, seed=32)
    qd = ctx.to_dev(q)
    outs = {}
    for pk in (False, True):
        for br in (False, True):
            h = ctx.make_index(ix.base, ix.bnorm, ix.lifted, ix.deg, ix.links, ix.tree, ix.mstar2, ix.scale,
                               packed=pk, byte_rows=br)
            assert bool(h.byte_rows) == br
            outs[(pk, br)] = run(ctx, h, qd, thr, 96)
    ref = outs[(False, False)]
    assert (ref["status"] == 0).all()
    cnt = ref["cnt"]
    say(f"d={d}: mean candidates {cnt.mean():.1f}, mean graph distances {ref['nd'].mean():.1f}")
    for key, o in outs.items():
        for f in ("cnt", "pred", "nd", "status"):
            assert np.array_equal(o[f], ref[f]), (key, f)
        for i in range(len(cnt)):
            c = int(cnt[i])
            assert np.array_equal(o["cand"][i, :c], ref["cand"][i, :c]), (key, i)
            assert np.array_equal(o["acc"][i, :c], ref["acc"][i, :c]), (key, i)
    if d != 128:
        return
    # the byte form against the oracle on sampled queries
    O = oracle()
    got = outs[(True, True)]
    lifted = ix.lifted.cpu().numpy()
    deg = ix.deg.cpu().numpy().view(np.uint32)
    links = ix.links.cpu().numpy().view(np.uint32)
    t = thr.cpu().numpy()
    ot = or_tree(ctx.tree_host(ix.tree))
    n, dl = lifted.shape
    out = np.zeros(n, np.uint32)
    for i in np.random.default_rng(5).choice(len(q), 40, replace=False):
        qb = np.zeros(dl, np.float32)
        O.or_lift_query(ptr(q[i]), d, ix.scale, ptr(qb))
        e = O.or_tree_entry(ctypes.byref(ot), ptr(qb))
        nd = ctypes.c_uint64()
        c = O.or_range_search(ptr(lifted), n, dl, ptr(deg), ptr(links), 48, e, ptr(qb), ix.mstar2 + 1.0, 96,
                              ptr(out), n, ctypes.byref(nd))
        assert int(got["cnt"][i]) == c and int(got["nd"][i]) == int(nd.value), i
        assert got["cand"][i, :c].view(np.uint32).tolist() == out[:c].tolist(), i
        acc = [int(np.float32(O.or_recheck_key(ptr(base[j]), ptr(q[i]), d)) <= t[j]) for j in out[:c]]
        assert got["acc"][i, :c].tolist() == acc, i


def test_byte_rows_only_when_exact(ctx):
    n, d = 2000, 128
    base = workload.mixture(n, d, 20, 18.0, seed=41)
    rk = ctx.to_dev(np.full(n, 50.0, np.float32))
    lifted, _, sc = ctx.lift(ctx.to_dev(base), rk)
    lmain, _ = ctx.lift_split(lifted)
    b = ctx.byte_rows(ctx.to_dev(base), lmain, sc)
    assert b is not None
    # lane order: lane l's bytes l*MC .. l*MC + MC - 1 are elements 16 t + l
    br = b.cpu().numpy().reshape(n, 16, d // 16)
    assert np.array_equal(br.transpose(0, 2, 1).reshape(n, d), base.astype(np.uint8))
    for v in (-0.0, 0.5, 256.0):
        bad = base.copy()
        bad[123, 45] = v
        assert ctx.byte_rows(ctx.to_dev(bad), lmain, sc) is None, v
    lm = lmain.cpu().numpy().copy()
    lm[7, 3] = np.nextafter(lm[7, 3], np.float32(np.inf), dtype=np.float32)
    assert ctx.byte_rows(ctx.to_dev(base), ctx.to_dev(lm), sc) is None


def test_sentinel_rows_pad_past_the_degree(ctx):
    n, M = 500, 48
    rng = np.random.default_rng(7)
    links = rng.integers(0, n, (n, M)).astype(np.int32)
    deg = rng.integers(0, M + 1, n).astype(np.int32)
    p = ctx.pack_links(ctx.to_dev(links), ctx.to_dev(deg)).cpu().numpy().view(np.uint32)
    for i in range(n):
        assert p[i, :deg[i]].tolist() == links[i, :deg[i]].astype(np.uint32).tolist()
        assert (p[i, deg[i]:] == 0xFFFFFFFF).all()
