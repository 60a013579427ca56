"""Brute-force pin of the oracle's snapshot LocalJoin (or_nndescent_snapshot),
the schedule the GPU LocalJoin reproduces bit for bit.

The reference pass (include/index.h:1468-1487) takes, for every row u, the
first nn_sample non-sentinel slots of u's heap in storage order and pushes
d2(v, w) both ways for every pair of them with the heap rule of
HeapList::check_push (include/heap.h:83-157: reject a value above the top,
reject an id already present, else evict the top).  The reference's threads
interleave these pushes in lock order; the snapshot schedule fixes that
order: samples are taken from the heaps as they stand when the pass starts,
and each row receives its proposals in ascending (distance, id) order.

The model below is written without heaps: rows are dicts id -> distance,
and with tie-free distances "evict the top" is "drop the largest".  Two
regimes pin the two halves of the rule:
  * sample >= capacity: a row's sample is all of its entries, so a pass is a
    function of the row sets alone -- several passes, checked row set by row
    set, value bits and per-pass update counts;
  * sample < capacity: the sample is the first slots in storage order of the
    given starting heaps -- one pass from known arrays.
Planted mistakes (samples refreshed during the pass = the sequential
schedule; proposals in one direction only; the sample taken from the end of
the row) change the result, so the comparison can fail."""
import numpy as np
import pytest

from anqi_testlib import oracle, ptr

U32MAX = 0xFFFFFFFF


def d2(x, nrm, v, w):
    import ctypes
    O = oracle()
    d = x.shape[1]
    return float(np.float32(O.or_d2(ptr(x[v]), ptr(x[w]), ctypes.c_float(nrm[v]), ctypes.c_float(nrm[w]), d)))


def model_pass(rows, samples, x, nrm, cap, both=True):
    """One snapshot pass on dict rows given each u's sample list."""
    props = {}
    for u, s in enumerate(samples):
        for a in range(len(s)):
            for b in range(a + 1, len(s)):
                v, w = s[a], s[b]
                dd = d2(x, nrm, v, w)
                props.setdefault(v, []).append((dd, w))
                if both:
                    props.setdefault(w, []).append((dd, v))
    upd = 0
    for r, lst in props.items():
        row = rows[r]
        for dd, i in sorted(lst):
            top = max(row.values()) if len(row) == cap else np.inf
            if dd > top or i in row:
                continue
            if len(row) == cap:
                del row[max(row, key=row.get)]
            row[i] = dd
            upd += 1
    return upd


def start_heaps(n, cap, fill, rng, x, nrm):
    """Heap arrays with `fill` random neighbours per row (sentinel-padded),
    laid out as valid max-heaps (descending order is a max-heap)."""
    hi = np.full((n, cap), U32MAX, np.uint32)
    hv = np.full((n, cap), np.finfo(np.float32).max, np.float32)
    for i in range(n):
        nb = rng.choice([j for j in range(n) if j != i], size=fill, replace=False)
        ds = sorted(((d2(x, nrm, i, int(j)), int(j)) for j in nb), reverse=True)
        # the sentinels hold FLT_MAX, so they sit in front (heap top)
        for s, (dd, j) in enumerate(ds):
            hi[i, cap - fill + s] = j
            hv[i, cap - fill + s] = dd
    return hi, hv


def data(n, d, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d)).astype(np.float32)
    nrm = np.zeros(n, np.float32)
    oracle().or_norms(ptr(x), n, d, ptr(nrm))
    return rng, x, nrm


def rows_of(hi, hv):
    return [{int(i): float(v) for i, v in zip(hi[r], hv[r]) if i != U32MAX} for r in range(hi.shape[0])]


@pytest.mark.parametrize("n,cap,fill,seed", [(24, 6, 6, 1), (40, 8, 5, 2), (33, 10, 10, 3)])
def test_snapshot_passes_equal_set_model(n, cap, fill, seed):
    O = oracle()
    rng, x, nrm = data(n, 7, seed)
    hi, hv = start_heaps(n, cap, fill, rng, x, nrm)
    rows = rows_of(hi, hv)
    iters = 4
    upd_o = np.zeros(iters, np.int64)
    it = O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, 7, cap, ptr(hi), ptr(hv), iters, 20, ptr(upd_o))
    assert it == iters            # n / 200 == 0: no early exit
    for p in range(iters):
        samples = [list(r.keys()) for r in rows]        # sample >= cap: every entry
        assert model_pass(rows, samples, x, nrm, cap) == upd_o[p], p
    got = rows_of(hi, hv)
    for r in range(n):
        assert got[r].keys() == rows[r].keys(), r
        assert all(np.float32(got[r][i]) == np.float32(rows[r][i]) for i in got[r]), r


@pytest.mark.parametrize("n,cap,sample,seed", [(30, 12, 5, 4), (40, 16, 3, 5)])
def test_sample_is_first_slots_in_storage_order(n, cap, sample, seed):
    O = oracle()
    rng, x, nrm = data(n, 5, seed)
    hi, hv = start_heaps(n, cap, cap, rng, x, nrm)
    samples = [[int(v) for v in hi[u] if v != U32MAX][:sample] for u in range(n)]
    rows = rows_of(hi, hv)
    upd_o = np.zeros(1, np.int64)
    O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, 5, cap, ptr(hi), ptr(hv), 1, sample, ptr(upd_o))
    assert model_pass(rows, samples, x, nrm, cap) == upd_o[0]
    got = rows_of(hi, hv)
    assert all(got[r].keys() == rows[r].keys() for r in range(n))


def test_planted_schedule_mistakes_change_the_result():
    O = oracle()
    n, cap = 40, 8
    rng, x, nrm = data(n, 7, 2)
    hi0, hv0 = start_heaps(n, cap, 5, rng, x, nrm)
    hi, hv = hi0.copy(), hv0.copy()
    O.or_nndescent_snapshot(ptr(x), ptr(nrm), n, 7, cap, ptr(hi), ptr(hv), 3, 20, None)
    ref = rows_of(hi, hv)
    # (1) samples refreshed as rows change: the sequential schedule
    hs, vs = hi0.copy(), hv0.copy()
    O.or_nndescent_seq(ptr(x), ptr(nrm), n, 7, cap, ptr(hs), ptr(vs), 3, 20, None)
    assert any(a.keys() != b.keys() for a, b in zip(rows_of(hs, vs), ref))
    # (2) proposals to one side of each pair only
    rows = rows_of(hi0, hv0)
    for _ in range(3):
        model_pass(rows, [list(r.keys()) for r in rows], x, nrm, cap, both=False)
    assert any(a.keys() != b.keys() for a, b in zip(rows, ref))
    # (3) a sample from the end of the storage order (sample < cap)
    rng1, x1, n1 = data(30, 5, 4)
    hi1, hv1 = start_heaps(30, 12, 12, rng1, x1, n1)
    tail = [[int(v) for v in hi1[u] if v != U32MAX][-5:] for u in range(30)]
    rows = rows_of(hi1, hv1)
    model_pass(rows, tail, x1, n1, 12)
    O.or_nndescent_snapshot(ptr(x1), ptr(n1), 30, 5, 12, ptr(hi1), ptr(hv1), 1, 5, None)
    assert any(a.keys() != b.keys() for a, b in zip(rows, rows_of(hi1, hv1)))
