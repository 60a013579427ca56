"""Captured query batches (configuration 5 on a small index): batches of
mixed k (threshold table per query) and uneven L, each served by one replay
of the instantiated graph, return exactly the accepted ids of the ordinary
per-call search; batches of different sizes replay back to back with eager
work in between; the autotuned work areas give the same answers."""
import numpy as np
import pytest

import workload

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-batch]", *a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import pipeline
    import pyanqi
    ctx = pyanqi.Context(0)
    base = workload.mixture(30000, 96, 120, 15.0, seed=21)
    ix = pipeline.build(ctx, base, log=say)
    tables = pipeline.threshold_tables(ctx, ix, (10, 50, 100))
    q = workload.mixture(20000, 96, 120, 15.0, seed=22)
    yield dict(ctx=ctx, ix=ix, tables=tables, q=q, torch=torch)
    ctx.tree_free(ix.tree)
    ctx.close()


def expected(env, q, L, kidx):
    ctx, torch = env["ctx"], env["torch"]
    nq = q.shape[0]
    Lt = ctx.to_dev(L.astype(np.int32))
    kt = ctx.to_dev(kidx.astype(np.int32))
    b = ctx.rknn_search(env["ix"].handle, ctx.to_dev(q), int(L.max()), env["tables"], out_cap=4096, mode=1,
                        Lq=Lt, Lmax=int(L.max()), kidx=kt)
    cnt, cand, acc = b["cnt"].cpu().numpy(), b["cand"].cpu().numpy().view(np.uint32), b["acc"].cpu().numpy()
    off = [0]
    ids = []
    for i in range(nq):
        c = int(cnt[i])
        ids.extend(cand[i, :c][acc[i, :c] != 0].tolist())
        off.append(len(ids))
    return np.array(off, np.uint32), np.array(ids, np.uint32)


def test_batches_match_per_call_search(env):
    ctx = env["ctx"]
    rng = np.random.default_rng(3)
    bat = ctx.batcher(env["ix"].handle, env["tables"], max_batch=4096, Lmax=512)
    pos = 0
    for size in (4096, 512, 1777, 3000, 600):
        q = env["q"][pos:pos + size]
        pos += size
        L = rng.integers(64, 513, size).astype(np.uint32)
        kidx = rng.integers(0, 3, size).astype(np.uint32)
        off, ids = bat.run(q, L, kidx)
        eoff, eids = expected(env, q, L, kidx)
        say(f"batch {size}: {len(ids)} answers")
        assert np.array_equal(off, eoff) and np.array_equal(ids, eids)
        env["torch"].cuda.synchronize()
        _ = (env["tables"] * 2).sum().item()        # eager work between replays
    bat.close()


def test_autotuned_batcher_same_answers(env):
    ctx = env["ctx"]
    rng = np.random.default_rng(4)
    n = 2048
    q = env["q"][:n]
    L = rng.integers(64, 513, n).astype(np.uint32)
    kidx = rng.integers(0, 3, n).astype(np.uint32)
    tune, ms = ctx.search_autotune(env["ix"].handle, env["tables"], ctx.to_dev(q), ctx.to_dev(L.astype(np.int32)),
                                   ctx.to_dev(kidx.astype(np.int32)), 512)
    say(f"autotune: waves/CU {tune.waves_per_cu} slots {tune.visited_slots} fifo {tune.result_cap} ({ms:.3f} ms)")
    bat = ctx.batcher(env["ix"].handle, env["tables"], max_batch=n, Lmax=512, tuning=tune)
    off, ids = bat.run(q, L, kidx)
    eoff, eids = expected(env, q, L, kidx)
    assert np.array_equal(off, eoff) and np.array_equal(ids, eids)
    bat.close()


@pytest.mark.parametrize("L", [1, 2, 63, 64, 65, 100, 128, 129, 256, 257, 512, 513, 1024])
def test_register_beam_equals_lds_beam(env, L):
    """The persistent form keeps the beam in registers (64 positions per
    chunk); the per-query form keeps it sorted in LDS.  Same candidate lists
    in the same order, accept decisions, #dist and #expansions at every beam
    width, including widths one past a chunk boundary and the LDS fallback
    above 512."""
    ctx = env["ctx"]
    q = ctx.to_dev(env["q"][:1500])
    thr = env["tables"][0].contiguous()
    a = ctx.rknn_search(env["ix"].handle, q, L, thr, out_cap=4096, mode=0)
    a = {k: v.cpu().numpy() for k, v in a.items()}
    b = ctx.rknn_search(env["ix"].handle, q, L, thr, out_cap=4096, mode=1)
    b = {k: v.cpu().numpy() for k, v in b.items()}
    say(f"L={L}: mean cand {a['cnt'].mean():.1f}, dist {a['nd'].mean():.1f}, expansions {a['nexp'].mean():.1f}")
    for f in ("cnt", "nd", "nexp", "pred", "status"):
        assert np.array_equal(a[f], b[f]), f
    for i in range(a["cnt"].shape[0]):
        c = int(a["cnt"][i])
        assert np.array_equal(a["cand"][i, :c], b["cand"][i, :c])
        assert np.array_equal(a["acc"][i, :c], b["acc"][i, :c])


@pytest.mark.parametrize("nq", [1, 5, 1500, 4097])
def test_work_order_covers_every_query(env, nq):
    """The persistent search takes queries in the order of their lifted
    distance to the entry (a counting sort on the device).  Every query is
    searched exactly once and its results do not depend on the order: the
    ordered run (output buffers pre-filled with sentinels) equals the run
    with caller-given entries, which keeps the batch order."""
    ctx, torch = env["ctx"], env["torch"]
    ix = env["ix"].handle
    q = ctx.to_dev(env["q"][:nq])
    thr = env["tables"][1].contiguous()
    ent = ctx.search_entries(ix, q)
    b = ctx.search_buffers(nq, 4096)
    for k, v in b.items():
        v.fill_(-1)
    ctx.rknn_search(ix, q, 96, thr, mode=1, bufs=b)
    a = ctx.rknn_search(ix, q, 96, thr, mode=1, entries=ent)
    a = {k: v.cpu().numpy() for k, v in a.items()}
    b = {k: v.cpu().numpy() for k, v in b.items()}
    assert (b["status"] == 0).all()
    for f in ("cnt", "nd", "nexp", "pred", "status"):
        assert np.array_equal(a[f], b[f]), f
    for i in range(nq):
        c = int(a["cnt"][i])
        assert np.array_equal(a["cand"][i, :c], b["cand"][i, :c])
        assert np.array_equal(a["acc"][i, :c], b["acc"][i, :c])


def test_overflowing_queries_are_redone(env):
    """A visited set far too small for the beam makes most queries overflow in
    the persistent launch; the library counts them on the device, reads the
    statuses back and reruns those queries with large work areas.  The answers
    equal the per-query form's."""
    import pyanqi
    ctx = env["ctx"]
    q = ctx.to_dev(env["q"][:2000])
    thr = env["tables"][0].contiguous()
    a = ctx.rknn_search(env["ix"].handle, q, 64, thr, out_cap=4096, mode=0)
    a = {k: v.cpu().numpy() for k, v in a.items()}
    ctx.set_search_tuning(pyanqi.SearchTuning(4, 128, 64))
    try:
        b = ctx.rknn_search(env["ix"].handle, q, 64, thr, out_cap=4096, mode=1)
        b = {k: v.cpu().numpy() for k, v in b.items()}
    finally:
        ctx.set_search_tuning(None)
    for f in ("cnt", "nd", "pred", "status"):
        assert np.array_equal(a[f], b[f]), f
    for i in range(a["cnt"].shape[0]):
        c = int(a["cnt"][i])
        assert np.array_equal(a["cand"][i, :c], b["cand"][i, :c])
        assert np.array_equal(a["acc"][i, :c], b["acc"][i, :c])


def test_batcher_redoes_overflowing_queries(env):
    """A batcher whose tuned visited set is far too small for its beams: most
    queries overflow inside the captured graph, are counted on the device and
    rerun with the retry tier's areas; the answers equal the per-call
    search's."""
    import pyanqi
    ctx = env["ctx"]
    rng = np.random.default_rng(8)
    nq = 3000
    q = env["q"][:nq]
    L = rng.integers(64, 257, nq).astype(np.uint32)
    kidx = rng.integers(0, 3, nq).astype(np.uint32)
    bat = ctx.batcher(env["ix"].handle, env["tables"], max_batch=nq, Lmax=256,
                      tuning=pyanqi.SearchTuning(4, 512, 64))
    off, ids = bat.run(q, L, kidx)
    eoff, eids = expected(env, q, L, kidx)
    assert np.array_equal(off, eoff) and np.array_equal(ids, eids)
    bat.close()


def test_batcher_rejects_bad_table_index(env):
    """A query whose table index is not below the batcher's table count is
    refused at the boundary (usage error) before anything is launched; the
    batcher keeps working afterwards."""
    import pyanqi
    ctx = env["ctx"]
    q = env["q"][:8]
    L = np.full(8, 64, np.uint32)
    bat = ctx.batcher(env["ix"].handle, env["tables"], max_batch=8, Lmax=64)
    bad = np.array([0, 1, 2, 3, 0, 0, 0, 0], np.uint32)
    with pytest.raises(pyanqi.AnqiError) as e:
        bat.run(q, L, bad)
    assert e.value.code == 2 and "table index" in str(e.value)
    off, ids = bat.run(q, L, np.zeros(8, np.uint32))
    eoff, eids = expected(env, q, L, np.zeros(8, np.uint32))
    assert np.array_equal(off, eoff) and np.array_equal(ids, eids)
    bat.close()


def test_chosen_build_equals_plain(env):
    """The chosen build (bf16 MFMA leaf sweep, batched gathers) yields the
    plain build's index bit for bit: AKNN lists, cutoffs, LRQ codes, lifted
    rows, graph links and degrees."""
    import pipeline
    ctx, a = env["ctx"], env["ix"]
    b = pipeline.build(ctx, a.base_np, precision=1, log=say)
    for f in ("aknn_ids", "cut", "codes", "lifted", "links", "deg"):
        x, y = getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy()
        assert x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), f
    say("plain vs chosen build s:", round(a.times["build_s"], 3), round(b.times["build_s"], 3))
    ctx.tree_free(b.tree)
