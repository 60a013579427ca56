"""Vamana refinement and the build without the lifted LocalJoin on the GPU
(anqi_graph_build with iters 0 and vamana 1/2, include/index.h:932-1063,
1262-1282): the graph equals, link for link, the oracle's, which equals the
reference's own build_index (tests/test_oracle_vamana.py); both gather forms;
the sample, a 1,200-row and a 20,000-row mixture; the medoid entry."""
import numpy as np
import pytest

from anqi_testlib import oracle, ptr
from test_oracle_vamana import _inputs, _rows, oracle_graph, same_graph

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-vamana]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a HIP device"
    import pyanqi
    c = pyanqi.Context(0)
    yield c
    c.close()


def gpu_graph(ctx, x, ids, form=1, **kw):
    xd = ctx.to_dev(np.ascontiguousarray(x, np.float32))
    idd = ctx.to_dev(np.ascontiguousarray(ids, np.uint32).view(np.int32))
    links, deg, t, st = ctx.graph_build(xd, idd, iters=0, form=form, **kw)
    ctx.tree_free(t)
    dg = deg.cpu().numpy().view(np.uint32)
    lk = links.cpu().numpy().view(np.uint32)
    # slots past the degree hold UINT32_MAX
    assert all((lk[i, dg[i]:] == 0xFFFFFFFF).all() for i in range(0, len(dg), 97))
    return lk, dg, st


CASES = [
    ("none", {}, {}),
    ("global", dict(vamana="global", vamana_L=64), dict(vamana="global", L=64)),
    ("global_L40_P1", dict(vamana="global", vamana_L=40, vamana_passes=1), dict(vamana="global", L=40, passes=1)),
    ("global_L700", dict(vamana="global", vamana_L=700, vamana_passes=1), dict(vamana="global", L=700, passes=1)),
    ("batch", dict(vamana="batch", vamana_L=64), dict(vamana="batch", L=64)),
    ("batch_opts", dict(vamana="batch", vamana_L=48, vamana_passes=3, vamana_batch_frac=0.05, vamana_batch_base=3.0,
                        vamana_early_alpha=1.1, vamana_entry="medoid"),
     dict(vamana="batch", L=48, passes=3, frac=0.05, base=3.0, early=1.1, entry="medoid")),
]


@pytest.mark.parametrize("data", ["sample", "mixture"])
@pytest.mark.parametrize("name,gkw,okw", CASES, ids=[c[0] for c in CASES])
def test_graph_equals_oracle(ctx, data, name, gkw, okw):
    x, ids = _inputs(data)
    gkw, okw = dict(gkw), dict(okw)
    if okw.get("entry") == "medoid":
        e = oracle().or_vamana_medoid(ptr(np.ascontiguousarray(x)), x.shape[0], x.shape[1], min(x.shape[0], 100000))
        assert ctx.vamana_medoid(ctx.to_dev(np.ascontiguousarray(x))) == e
        okw["entry"] = gkw["vamana_entry"] = e
    say(data, name)
    want = oracle_graph(x, ids, **okw)[:2]
    for form in (0, 1):
        lk, dg, st = gpu_graph(ctx, x, ids, form=form, **gkw)
        assert same_graph((lk, dg), want) == "", (data, name, form)
        if okw.get("vamana"):
            assert st.vamana_s > 0


def test_medoid_entry_matches_oracle(ctx):
    x = _rows(3000)
    xd = ctx.to_dev(x)
    for sample in (1, 7, 1000, 3000, 100000):
        assert ctx.vamana_medoid(xd, sample) == oracle().or_vamana_medoid(ptr(x), 3000, 130, sample), sample


@pytest.mark.parametrize("style", ["global", "batch"])
def test_graph_equals_oracle_20k(ctx, style):
    """20,000 rows: the batch style runs 1, 2, ..., 256-row batches, then
    400-row ones (2 % of the rows), twice."""
    x = _rows(20000, seed=9)
    n = x.shape[0]
    ids = np.zeros((n, 100), np.uint32)
    keys = np.zeros((n, 100), np.float32)
    assert oracle().or_exact_topk(ptr(x), n, x.shape[1], 100, ptr(ids), ptr(keys), 0) == 0
    say("20k", style)
    want = oracle_graph(x, ids, vamana=style, L=64)[:2]
    lk, dg, st = gpu_graph(ctx, x, ids, vamana=style, vamana_L=64)
    say("vamana seconds", st.vamana_s)
    assert same_graph((lk, dg), want) == ""
