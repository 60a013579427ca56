"""Pins the oracle's FP32 distance contract against the reference's own
compiled functions (include/distance.h via oracle/ref_harness.cpp) and
against closed forms on integer data."""
import numpy as np
import pytest

from anqi_testlib import oracle, ptr, ref, ref_available

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference harness not built here")


def _vec(rng, d, scale):
    return (rng.standard_normal(d) * scale).astype(np.float32)


@needs_ref
@pytest.mark.parametrize("d", [1, 2, 7, 15, 16, 17, 31, 64, 96, 98, 128, 130, 200, 300])
def test_ip_l2_norm_bits_match_reference(d):
    O, R = oracle(), ref()
    rng = np.random.default_rng(d)
    for t in range(300):
        scale = [1e-3, 1.0, 37.0, 1e4][t % 4]
        x, y = _vec(rng, d, scale), _vec(rng, d, scale)
        for of, rf in (("or_ip", "ref_ip"), ("or_ip", "ref_ip_free"), ("or_l2", "ref_l2")):
            a = getattr(O, of)(ptr(x), ptr(y), d)
            b = getattr(R, rf)(ptr(x), ptr(y), d)
            assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32), (of, d, t)
        a = O.or_norm(ptr(x), d)
        b = R.ref_norm(ptr(x), d)
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)
        a = O.or_recheck_key(ptr(x), ptr(y), d)
        b = R.ref_recheck_key(ptr(x), ptr(y), d)
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)


def test_integer_data_closed_form():
    """On integer data below 2^24 every partial sum is exact: the contract
    must equal the exact integer dot product and squared distance."""
    O = oracle()
    rng = np.random.default_rng(5)
    for d in (8, 128, 130):
        for _ in range(200):
            x = rng.integers(0, 256, d).astype(np.float32)
            y = rng.integers(0, 256, d).astype(np.float32)
            assert O.or_ip(ptr(x), ptr(y), d) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))
            assert O.or_l2(ptr(x), ptr(y), d) == float(((x.astype(np.int64) - y.astype(np.int64)) ** 2).sum())


def test_lane_order_is_not_sequential():
    """A planted mistake check: a sequential float32 sum differs from the
    16-lane contract on some inputs, so the test above can fail."""
    O = oracle()
    rng = np.random.default_rng(9)
    diff = 0
    for _ in range(200):
        x, y = _vec(rng, 128, 1.0), _vec(rng, 128, 1.0)
        seq = np.float32(0)
        for a, b in zip(x, y):
            seq = np.float32(seq + np.float32(a * b))
        diff += np.float32(O.or_ip(ptr(x), ptr(y), 128)) != seq
    assert diff > 0
