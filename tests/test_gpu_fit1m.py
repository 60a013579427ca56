"""Differentiable verifier fit at the headline scale: the cutoff matrix is
the chosen build's AKNN source of the 1,000,000 x 128 mixture (K = 100,
AKNN 100/500/64/600), and training uses the reference trainer's
120,000-row sample (tools/train_final_m8_u8.py:190-265: default_rng(10),
sorted).  The HIP backward is checked against the FP64 torch reference
gradient and FP64 central finite differences (eps = 1e-4) on 1,000 of those
rows; the relaxed fit starts from the Lloyd fit and its exact floor decode
is the arbiter: RMSE over all 1,000,000 rows, and coefficients, codes and
decoded thresholds (hence every accept decision) equal the reference's on
the fitted files."""
import json
import os

import numpy as np
import pytest

from relax_checks import decisions_match_reference, grad_check
import workload

pytestmark = pytest.mark.gpu

OUT = os.environ.get("ANQI_FIT1M_JSON")


def say(*a):
    print("[gpu-fit1m]", *a, flush=True)


def test_relaxed_fit_on_million_cutoffs():
    import torch
    import lrq_fit
    import pyanqi
    from lrq_relax import fit_relaxed, floor_rmse
    ctx = pyanqi.Context(0)
    base = workload.million_base(1_000_000)
    _, keys, st = ctx.aknn_build(ctx.to_dev(base), K=100, leaf=500, trees=64, ef=600, precision=1)
    cut = torch.clamp(keys, min=0.0).contiguous()
    say(f"AKNN source {st.total_s:.2f}s")
    n = cut.shape[0]
    mu, G0, cb0, fit_s = lrq_fit.fit(cut, log=say)
    mu, G0, cb0 = ctx.to_dev(mu), ctx.to_dev(G0), ctx.to_dev(cb0)
    ids = np.sort(np.random.default_rng(lrq_fit.SEED).choice(n, size=lrq_fit.TRAIN_ROWS, replace=False))
    train = cut[torch.from_numpy(ids).to(cut.device)].contiguous()
    res = {"rows": n, "train_rows": int(train.shape[0]), "lloyd_fit_s": fit_s, "grad": {}}
    for mode in (0, 1):
        rl, rg, rc, worst = grad_check(ctx, train, mu, G0, cb0, mode, say)
        res["grad"][mode] = dict(loss_rel=rl, basis_rel=rg, codebook_rel=rc, fd_worst=worst)
        assert rl <= 1e-5 and rg <= 2e-4 and rc <= 2e-4 and worst <= 1e-3
    G, cb, hist, secs = fit_relaxed(ctx, train, mu, G0, cb0, iters=200, all_rows=cut, log=say)
    t0, t1 = floor_rmse(ctx, train, mu, G0, cb0), floor_rmse(ctx, train, mu, G, cb)
    r0, r1 = floor_rmse(ctx, cut, mu, G0, cb0), floor_rmse(ctx, cut, mu, G, cb)
    say(f"relaxed fit {secs:.2f}s; floor-decoded RMSE, training rows: Lloyd {t0:.6g}, relaxed {t1:.6g}; "
        f"all {n} rows: Lloyd {r0:.6g}, relaxed {r1:.6g}")
    # the best checkpoint by training-row RMSE is kept (the start included);
    # all rows: the sample is 12% of them, drawn at random
    assert np.isfinite(t1) and t1 <= t0 * (1 + 1e-6)
    assert np.isfinite(r1) and r1 <= r0 * 1.01
    ok = decisions_match_reference(ctx, cut, mu, G, cb)
    res.update(relaxed_fit_s=secs, relaxed_loss_first=hist[0], relaxed_loss_min=min(hist),
               floor_rmse_train_lloyd=t0, floor_rmse_train_relaxed=t1, floor_rmse_lloyd=r0,
               floor_rmse_relaxed=r1, decisions_equal_reference=ok)
    say(json.dumps(res))
    if OUT:
        with open(OUT, "w") as f:
            json.dump(res, f, indent=1)
    ctx.close()
    if ok is None:
        pytest.skip("reference harness not built")
    assert ok
