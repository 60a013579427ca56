"""Edge policies of the graph prunes on the GPU (anqi_graph_build with an
anqi_edge_policy, include/index.h:283-829): the graph equals the oracle's,
which equals the reference's own build_index (tests/test_oracle_edge.py), on
the sample and a 1,200-row mixture, with and without the Vamana passes."""
import tempfile

import numpy as np
import pytest

from anqi_testlib import Sample, edge_policy_config, or_edge_policy, oracle, ptr
from test_oracle_edge import M, POLICIES, write_inputs
from test_oracle_vamana import _rows, oracle_graph, same_graph

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-edge]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a HIP device"
    import pyanqi
    c = pyanqi.Context(0)
    yield c
    c.close()


def device_policy(ctx, f):
    import pyanqi
    keep = []
    E = pyanqi.EdgePolicy()
    E.resid_cover, E.ball, E.ball_pure, E.occlude_all, E.lsg = (f["resid_cover"], f["ball"], f["ball_pure"],
                                                               f["occlude_all"], f["lsg"])
    E.lsg_alpha, E.lambda_, E.scale2 = f["lsg_alpha"], f["lam"], f["scale2"]
    if f["resid"] is not None:
        r = ctx.to_dev(f["resid"])
        keep.append(r)
        E.resid = r.data_ptr()
    if f["mu"] is not None:
        m = ctx.to_dev(f["mu"])
        keep.append(m)
        E.mu = m.data_ptr()
    E.nk = f["nk"]
    E.nks = len(f["ks"])
    for t, k in enumerate(f["ks"]):
        E.ks[t] = k
    E.budget = f["budget"]
    E.bubble_auto, E.bubble_full = f["bubble_auto"], f["bubble_full"]
    E.auto_candidates, E.auto_max_swaps = f["auto_candidates"], f["auto_max_swaps"]
    return E, keep


def inputs(data):
    if data == "sample":
        S = Sample()
        return np.ascontiguousarray(S.lifted), np.ascontiguousarray(S.rec_ids)
    x = _rows()
    n = x.shape[0]
    ids = np.zeros((n, 100), np.uint32)
    keys = np.zeros((n, 100), np.float32)
    assert oracle().or_exact_topk(ptr(x), n, x.shape[1], 100, ptr(ids), ptr(keys), 0) == 0
    return x, ids


CASES = [("residual_cover", None), ("residual_cover_ks", "batch"), ("ball_rng", None), ("ball_hybrid", "global"),
         ("ball_hybrid_any", None), ("bubble_multik", None), ("bubble_rng_nk", "batch"), ("lsg_rng", None),
         ("lsg_rng_half", None), ("lsg_bubble_hybrid", "global"), ("bubble_auto", None),
         ("bubble_auto_opts", "batch"), ("bubble_auto_full", None)]


@pytest.mark.parametrize("data", ["sample", "mixture"])
@pytest.mark.parametrize("name,vamana", CASES, ids=[f"{a}-{b}" for a, b in CASES])
def test_policy_graph_equals_oracle(ctx, data, name, vamana):
    if name == "bubble_auto_full" and data != "sample":
        pytest.skip("the full bubble optimizer's oracle is cubic in the pool: sample only")
    x, ids = inputs(data)
    n = x.shape[0]
    with tempfile.TemporaryDirectory() as td:
        folder = td + "/"
        kv = dict(POLICIES[name])
        write_inputs(folder, n, int(kv.get("ANQI_EDGE_NK", "100")))
        kv["ANQI_LSG_MU_FILE"] = folder + "mu.bin"
        f = edge_policy_config(kv, n, M, folder)
        say(data, name, vamana)
        vk = {} if vamana is None else dict(vamana=vamana, L=64)
        want = oracle_graph(x, ids, policy=or_edge_policy(f), **vk)[:2]
        E, keep = device_policy(ctx, f)
        xd = ctx.to_dev(x)
        idd = ctx.to_dev(ids.view(np.int32))
        gk = {} if vamana is None else dict(vamana=vamana, vamana_L=64)
        links, deg, t, st = ctx.graph_build(xd, idd, iters=0, form=1, edge=E, **gk)
        ctx.tree_free(t)
        got = (links.cpu().numpy().view(np.uint32), deg.cpu().numpy().view(np.uint32))
        diff = same_graph(got, want)
        if name == "lsg_rng_half" and diff:
            # a weight power other than 1 is the device's pow: it may round
            # the last bit differently from the host library's
            say("lsg alpha 0.5:", diff)
            assert int(diff.split()[-2]) <= max(1, n // 100), diff
        else:
            if diff:
                for i in range(n):
                    a = got[0][i, :got[1][i]].tolist()
                    b = want[0][i, :want[1][i]].tolist()
                    if a != b:
                        say("row", i, "gpu", a, "oracle", b)
            assert diff == "", (data, name, vamana)
