"""anqi_rknn_search_noalloc: the persistent search with its work areas in the
library's per-device scratch array instead of hipMalloc.  Called through the
C ABI, it returns exactly what anqi_rknn_search (mode 1) returns, on two
calls in a row with the second smaller than the first: on an ordinary index,
with a tuning that makes queries overflow (redo tier), and on an index whose
lifted ball is widened so floods outgrow the LDS tiers (global tier rows in
the scratch array).  Sizes beyond its limits are refused before anything
runs."""
import numpy as np
import pytest

import workload

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import pipeline
    import pyanqi
    ctx = pyanqi.Context(0)
    base = workload.mixture(30000, 96, 1, 15.0, seed=41)
    ix = pipeline.build(ctx, base, log=say)
    thr = pipeline.thresholds(ctx, ix, 10)
    q = workload.mixture(6000, 96, 1, 15.0, seed=42)
    yield dict(ctx=ctx, ix=ix, thr=thr, q=q, torch=torch, pyanqi=pyanqi)
    ctx.tree_free(ix.tree)
    ctx.close()


def compare(a, b):
    a = {k: v.cpu().numpy() for k, v in a.items()}
    b = {k: v.cpu().numpy() for k, v in b.items()}
    for f in ("cnt", "nd", "nexp", "pred", "status"):
        assert np.array_equal(a[f], b[f]), f
    cap = a["cand"].shape[1]
    for i in range(a["cnt"].shape[0]):
        c = min(int(a["cnt"][i]), cap)
        assert np.array_equal(a["cand"][i, :c], b["cand"][i, :c]), i
        assert np.array_equal(a["acc"][i, :c], b["acc"][i, :c]), i
    return a


def run_pair(env, ix, sizes, L, out_cap):
    ctx = env["ctx"]
    for nq in sizes:
        q = ctx.to_dev(env["q"][:nq])
        a = ctx.rknn_search(ix, q, L, env["thr"], out_cap=out_cap, mode=1)
        b = ctx.search_buffers(nq, out_cap)
        for v in b.values():
            v.fill_(-1)
        ctx.rknn_search_noalloc(ix, q, L, env["thr"], b)
        r = compare(a, b)
        say(f"nq {nq} L {L}: mean candidates {r['cnt'].mean():.1f}, max {r['cnt'].max()}, "
            f"status counts {np.bincount(r['status'], minlength=4).tolist()}")


def test_noalloc_equals_search(env):
    run_pair(env, env["ix"].handle, (6000, 1500), 96, 4096)


def test_noalloc_redo_tier_equals_search(env):
    ctx, pyanqi = env["ctx"], env["pyanqi"]
    ctx.set_search_tuning(pyanqi.SearchTuning(4, 256, 64))
    try:
        run_pair(env, env["ix"].handle, (3000, 700), 128, 4096)
    finally:
        ctx.set_search_tuning(None)


def test_noalloc_global_tier_equals_search(env):
    ctx, ix = env["ctx"], env["ix"]
    wide = ctx.make_index(ix.base, ix.bnorm, ix.lifted, ix.deg, ix.links, ix.tree, ix.mstar2 * 64.0, ix.scale)
    run_pair(env, wide, (48, 17), 64, 30000)


def test_noalloc_refuses_oversize_batch(env):
    ctx, pyanqi, torch = env["ctx"], env["pyanqi"], env["torch"]
    nq = 65537
    q = torch.zeros((nq, 96), dtype=torch.float32, device="cuda:0")
    b = ctx.search_buffers(nq, 1)
    with pytest.raises(pyanqi.AnqiError) as e:
        ctx.rknn_search_noalloc(env["ix"].handle, q, 64, env["thr"], b)
    assert e.value.code == 2 and "limit" in str(e.value)
