"""Precision variant (configuration 3) on the sample base (256 x 128, N(0,1)):
exact top-50 and top-100 from the bf16 MFMA sweep with FP32 re-scoring are
bit-identical to the FP32 path (and to dataset_gt), and the distance error of
every returned pair against an FP64 evaluation of the same pair is reported
and held to the FP32 rounding bound of the reference's formula."""
import json
import os

import numpy as np
import pytest

from anqi_testlib import Sample

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-precision]", *a, flush=True)


@pytest.mark.parametrize("K", [50, 100])
def test_bf16_top_k_equals_fp32_and_fp64_error_per_result(K):
    import torch
    import pyanqi
    S = Sample()
    ctx = pyanqi.Context(0)
    x = ctx.to_dev(S.base)
    i1, k1 = ctx.exact_knn(x, K, precision=1)
    i0, k0 = ctx.exact_knn(x, K, precision=0)
    i1, k1, i0, k0 = (t.cpu().numpy() for t in (i1, k1, i0, k0))
    assert np.array_equal(i1, i0) and np.array_equal(k1.view(np.uint32), k0.view(np.uint32))
    assert np.array_equal(i1.view(np.uint32), S.gt_ids[:, :K])
    b = S.base.astype(np.float64)
    ids = i1.view(np.uint32).astype(np.int64)
    d64 = ((b[:, None, :] - b[ids]) ** 2).sum(axis=2)            # FP64 squared distance per result
    err = np.abs(k1.astype(np.float64) - d64)
    # FP32 bound of key = |x|^2 + |y|^2 - 2<x,y> with d-term sums: ~ (d+3) u (|x| + |y|)^2
    nx = np.sqrt((b ** 2).sum(1))
    bound = (S.d + 3) * 2.0 ** -24 * (nx[:, None] + nx[ids]) ** 2
    say(f"K={K}: |key - d_fp64| max {err.max():.3e} mean {err.mean():.3e}; relative to d: max "
        f"{(err / d64).max():.3e} mean {(err / d64).mean():.3e}; worst / bound {(err / bound).max():.3f}")
    assert (err <= bound).all()
    ctx.close()
