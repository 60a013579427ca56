"""The verifier's certified float-float form (the default) against the oracle
and against the all-soft-binary64 form, on inputs built to defeat the
certification: a zero basis row (every coefficient of that rank sums to an
exact zero, which is never certified), residuals that equal a codeword
exactly (the float interval straddles a codebook boundary), thresholds that
cancel to zero (mu_k = -codeword), values spread over e^-8..e^12, and a
1,000,000-row realistic case.  Coefficients, codes and thresholds must match
bit for bit; the fallback counters must show that the soft path ran on the
adversarial inputs and almost never on the realistic one.  Reference:
src/rknn_query.cpp:679-753 (build, materialize_k), :482-513 (floor code)."""
import numpy as np
import pytest

from anqi_testlib import oracle, ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-lrq]", *a, flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    import pyanqi
    c = pyanqi.Context(0)
    yield c
    c.close()


def adversarial(n, seed, zero_basis):
    """The CPU test's adversarial inputs (tests/test_lrq_cert_host.py)."""
    from test_lrq_cert_host import adversarial as make
    return make(n, seed, zero_basis)


def run_both(ctx, cut, mu, G, cb, ks):
    O = oracle()
    n, nk = cut.shape
    M = G.shape[0]
    c1 = np.zeros((n, M), np.float32)
    k1 = np.zeros((n, nk), np.uint8)
    O.or_lrq_encode(ptr(cut), n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, ptr(c1), ptr(k1))
    t1 = np.zeros((len(ks), n), np.float32)
    for t, k in enumerate(ks):
        O.or_lrq_materialize(ptr(c1), ptr(k1), n, nk, M, ptr(mu), ptr(G), ptr(cb), 256, k, 1.0, ptr(t1[t]))
    out = {}
    for form in (0, 1):
        ctx.set_lrq_form(form)
        try:
            d = [ctx.to_dev(a) for a in (cut, mu, G, cb)]
            coeff, codes, codes_km = ctx.lrq_encode(*d)
            fe = ctx.lrq_fallbacks()[0]
            thr = ctx.lrq_materialize_multi(coeff, codes_km, d[1], d[2], d[3], ks)
            fd = ctx.lrq_fallbacks()[1]
        finally:
            ctx.set_lrq_form(0)
        assert np.array_equal(bits(coeff.cpu().numpy()), bits(c1)), form
        assert np.array_equal(codes.cpu().numpy(), k1), form
        assert np.array_equal(codes_km.cpu().numpy(), k1.T), form
        assert np.array_equal(bits(thr.cpu().numpy()), bits(t1)), form
        out[form] = (fe, fd)
    say(f"n {n}: certified form recomputed {out[0][0]} encode and {out[0][1]} decode elements in soft binary64")
    return out[0]


def test_adversarial_inputs_fall_back_and_match(ctx):
    cut, mu, G, cb, _ = adversarial(20000, 11, zero_basis=False)
    fe, fd = run_both(ctx, cut, mu, G, cb, [1, 4, 10, 50, 100])
    assert fe >= 20000          # at least every row's zero coefficient


def test_boundary_rows_with_zero_basis(ctx):
    """All basis rows zero: every coefficient falls back, and so does every
    residual that lands exactly on a codeword."""
    cut, mu, G, cb, hits = adversarial(3000, 12, zero_basis=True)
    fe, fd = run_both(ctx, cut, mu, G, cb, [1, 4, 64])
    assert hits > 1000 and fe >= 3000 * 8 + hits
    assert fd >= 1              # mu_4 + cb_4[0] == 0 for the rows coded 0 at rank 4


def test_wide_range_and_realistic_rows(ctx):
    rng = np.random.default_rng(3)
    n, nk, M = 200000, 100, 8
    cut = np.sort(np.exp(rng.uniform(-8, 12, (n, nk))), axis=1).astype(np.float32)
    mu = cut.mean(0).astype(np.float32)
    G = (rng.standard_normal((M, nk)) * 0.1).astype(np.float32)
    cb = np.sort(rng.standard_normal((nk, 256)) * 50, axis=1).astype(np.float32)
    run_both(ctx, cut, mu, G, cb, [1, 10, 37, 100])


def test_million_realistic_rows(ctx):
    """Radii-like rows (ascending, 1e4..1e5 scale) with a fitted basis:
    certification covers nearly everything."""
    import torch
    import lrq_fit
    rng = np.random.default_rng(5)
    n, nk = 1_000_000, 100
    base = rng.uniform(2e4, 6e4, (n, 1))
    cut = (base + np.cumsum(rng.uniform(0, 400, (n, nk)), axis=1)).astype(np.float32)
    mu, G, cb, _ = lrq_fit.fit(torch.from_numpy(cut[:200000]).to("cuda:0"), log=lambda *a: None)
    fe, fd = run_both(ctx, cut, mu, G, cb, [10, 50, 100])
    assert fe <= n * (nk + 8) * 1e-4 and fd <= 3 * n * 1e-4
