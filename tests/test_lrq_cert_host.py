"""The verifier's certified float-float arithmetic (csrc/lrq_cert.hpp, the
functions the encode/decode kernels run per row) compiled for the host and
compared bit for bit with the oracle's binary64 restatement of
CompactLrqVerifier::build / materialize_k (src/rknn_query.cpp:679-753,
floor codes :482-513): on the reference's sample verifier, on values spread
over e^-8..e^12, on inputs built to defeat the certificates (zero basis rows,
residuals exactly on codewords, thresholds cancelling to zero) and on
radius-like rows.  The codeword search is checked against numpy's
searchsorted(side="right").  Planted mistake: the float-float roundings taken
without their certificates differ from binary64, so the comparison can
fail."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from anqi_testlib import REPO, Sample, oracle, ptr

P = ctypes.c_void_p


@pytest.fixture(scope="module")
def H(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lrq") / "liblrqcert.so")
    src = os.path.join(REPO, "tests", "lrq_cert_host.cpp")
    inc = os.path.join(REPO, "anqi-implementation_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-I", inc, src, "-o", out],
                   check=True)
    L = ctypes.CDLL(out)
    L.host_upper_bound256.restype = ctypes.c_uint32
    L.host_upper_bound256.argtypes = [P, ctypes.c_float]
    L.host_lrq_encode.restype = ctypes.c_uint64
    L.host_lrq_encode.argtypes = [P, ctypes.c_uint64, ctypes.c_uint32, P, P, P, P, P]
    L.host_lrq_decode.restype = ctypes.c_uint64
    L.host_lrq_decode.argtypes = [P, P, ctypes.c_uint64, ctypes.c_uint32, P, P, P, ctypes.c_uint32, P]
    L.host_lrq_uncertified_coeffs.restype = None
    L.host_lrq_uncertified_coeffs.argtypes = [P, ctypes.c_uint64, ctypes.c_uint32, P, P, P]
    return L


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def compare(H, cut, mu, G, cb, ks):
    O = oracle()
    n, nk = cut.shape
    c1 = np.zeros((n, 8), np.float32)
    k1 = np.zeros((n, nk), np.uint8)
    O.or_lrq_encode(ptr(cut), n, nk, 8, ptr(mu), ptr(G), ptr(cb), 256, ptr(c1), ptr(k1))
    c2 = np.zeros((n, 8), np.float32)
    k2 = np.zeros((n, nk), np.uint8)
    fe = H.host_lrq_encode(ptr(cut), n, nk, ptr(mu), ptr(G), ptr(cb), ptr(c2), ptr(k2))
    assert np.array_equal(bits(c1), bits(c2))
    assert np.array_equal(k1, k2)
    fd = 0
    for k in ks:
        t1 = np.zeros(n, np.float32)
        t2 = np.zeros(n, np.float32)
        O.or_lrq_materialize(ptr(c1), ptr(k1), n, nk, 8, ptr(mu), ptr(G), ptr(cb), 256, k, 1.0, ptr(t1))
        fd += H.host_lrq_decode(ptr(c1), ptr(k1), n, nk, ptr(mu), ptr(G), ptr(cb), k, ptr(t2))
        assert np.array_equal(bits(t1), bits(t2)), k
    return fe, fd


def test_upper_bound_matches_searchsorted(H):
    rng = np.random.default_rng(1)
    for trial in range(50):
        cb = np.sort(rng.standard_normal(256) * 50).astype(np.float32)
        if trial % 5 == 0:
            cb[100:140] = cb[100]                      # repeated codewords (cumulative maxima)
        v = np.concatenate([cb, cb - 1e-3, cb + 1e-3, [cb[0] - 1e6, cb[-1] + 1e6, -np.inf, np.inf],
                            rng.uniform(-300, 300, 200)]).astype(np.float32)
        for x in v:
            assert H.host_upper_bound256(ptr(cb), ctypes.c_float(x)) == np.searchsorted(cb, x, side="right")


def test_reference_sample_verifier(H):
    S = Sample()
    fe, fd = compare(H, np.maximum(S.rec_keys, 0.0).astype(np.float32), S.mu, S.G, S.cb, [1, 10, 50, 100])
    assert np.array_equal(bits(np.zeros(1)), bits(np.zeros(1)))


def test_wide_range(H):
    rng = np.random.default_rng(3)
    n, nk = 20000, 100
    cut = np.sort(np.exp(rng.uniform(-8, 12, (n, nk))), axis=1).astype(np.float32)
    mu = cut.mean(0).astype(np.float32)
    G = (rng.standard_normal((8, nk)) * 0.1).astype(np.float32)
    cb = np.sort(rng.standard_normal((nk, 256)) * 50, axis=1).astype(np.float32)
    fe, fd = compare(H, cut, mu, G, cb, [1, 10, 37, 100])
    assert fe <= n * 108 * 1e-3


def adversarial(n, seed, zero_basis):
    rng = np.random.default_rng(seed)
    nk = 100
    mu = rng.uniform(100, 2000, nk).astype(np.float32)
    G = (rng.standard_normal((8, nk)) * 0.1).astype(np.float32)
    G[7] = 0.0
    if zero_basis:
        G[:] = 0.0
    cb = np.sort(rng.standard_normal((nk, 256)) * 50, axis=1).astype(np.float32)
    mu[3] = -cb[3, 0]
    # values whose difference from mu is a codeword exactly
    j = rng.integers(0, 256, (n, nk))
    cut = (mu[None, :].astype(np.float64) + cb[np.arange(nk)[None, :], j]).astype(np.float32)
    exact = (cut.astype(np.float64) - mu[None, :]) == cb[np.arange(nk)[None, :], j]
    cut[n // 2:] = (mu[None, :] + rng.uniform(-300, 300, (n - n // 2, nk))).astype(np.float32)
    exact[n // 2:] = False
    return cut, mu, G, cb, int(exact.sum())


def test_adversarial_inputs(H):
    cut, mu, G, cb, _ = adversarial(8000, 11, zero_basis=False)
    fe, fd = compare(H, cut, mu, G, cb, [1, 4, 10, 100])
    assert fe >= 8000                                   # the zero-basis rank never certifies


def test_codeword_boundaries_with_zero_basis(H):
    cut, mu, G, cb, hits = adversarial(3000, 12, zero_basis=True)
    fe, fd = compare(H, cut, mu, G, cb, [1, 4, 64])
    assert hits > 1000
    assert fe >= 3000 * 8 + hits                        # every coefficient, every residual on a codeword
    assert fd >= 1                                      # mu_4 + cb_4[0] cancels to zero


def test_radius_like_rows_rarely_fall_back(H):
    rng = np.random.default_rng(5)
    n, nk = 50000, 100
    cut = (rng.uniform(2e4, 6e4, (n, 1)) + np.cumsum(rng.uniform(0, 400, (n, nk)), axis=1)).astype(np.float32)
    mu = cut.mean(0).astype(np.float32)
    c = cut - mu
    w, v = np.linalg.eigh(c.T.astype(np.float64) @ c / n)
    G = v[:, ::-1][:, :8].T.astype(np.float32).copy()
    resid = c - (c @ G.T) @ G
    cb = np.sort(np.quantile(resid, (np.arange(256) + 0.5) / 256, axis=0).T, axis=1).astype(np.float32).copy()
    fe, fd = compare(H, cut, mu, G, cb, [10, 50, 100])
    assert fe <= n * 108 * 1e-4 and fd <= 3 * n * 1e-4


def test_planted_uncertified_rounding_is_caught(H):
    """Rows whose coefficient sum is 1 + 2^-24 + 2^-50 (binary64 keeps the
    2^-50 and rounds up to 1 + 2^-23; the float-float sum drops it and sits
    on the midpoint, rounding to 1.0): taken without its certificate the
    estimate is wrong, with it the row falls back and matches."""
    nk = 100
    rows = []
    for s_, tail in ((1.0, 2.0 ** -50), (-1.0, -(2.0 ** -50)), (1.5, 2.0 ** -49), (3.0, 2.0 ** -47)):
        r = np.zeros(nk, np.float32)
        r[0], r[1], r[2] = s_, np.float32(s_ * 2.0 ** -24 if abs(s_) == 1 else 2.0 ** -23 * (s_ / 2) * 0.5), tail
        rows.append(r)
    cut = np.array(rows, np.float32)
    mu = np.zeros(nk, np.float32)
    G = np.zeros((8, nk), np.float32)
    G[0, :3] = 1.0
    cb = np.sort(np.random.default_rng(2).standard_normal((nk, 256)) * 50, axis=1).astype(np.float32)
    n = len(cut)
    c1 = np.zeros((n, 8), np.float32)
    k1 = np.zeros((n, nk), np.uint8)
    oracle().or_lrq_encode(ptr(cut), n, nk, 8, ptr(mu), ptr(G), ptr(cb), 256, ptr(c1), ptr(k1))
    assert c1[0, 0] == np.float32(1.0 + 2.0 ** -23)
    c3 = np.zeros((n, 8), np.float32)
    H.host_lrq_uncertified_coeffs(ptr(cut), n, nk, ptr(mu), ptr(G), ptr(c3))
    assert c3[0, 0] == np.float32(1.0) and not np.array_equal(bits(c1[:, 0]), bits(c3[:, 0]))
    fe, _ = compare(H, cut, mu, G, cb, [1])
    assert fe >= 1
