"""The verifier trainer (anqi-implementation_amd/lrq_fit.py) against the
reference trainer's own output files (tools/train_final_m8_u8.py:81-265, run
by tests/golden/make_golden.py): on the 256-row sample's AKNN cutoffs and on
the synthetic 3,000 x 100 cutoff matrix with a 2,000-row training sample
(seeded row sampling, quantile interpolation, three alternating rounds).
The mean and the codebook must be bit-identical; each basis row must equal
the reference's row or its negation bit for bit (an eigenvector's sign is
arbitrary, and a flipped row flips its coefficients too, so the decoded
radii are the same).  Runs on the CPU (torch FP64, one thread)."""
import json
import os

import numpy as np
import pytest

from anqi_testlib import GOLDEN, read_basis, read_codebook, read_rec, trainer_cutoffs


def fit_cpu(cut, **kw):
    """The fit on the CPU with one torch thread: its FP64 column sums and
    products then add in numpy's order (row after row), which is what makes
    the low bits equal; more threads split those sums and change them."""
    import torch
    import lrq_fit
    prev = torch.get_num_threads()
    torch.set_num_threads(1)
    try:
        return lrq_fit.fit(torch.from_numpy(np.ascontiguousarray(cut, np.float32)), log=lambda *a: None, **kw)
    finally:
        torch.set_num_threads(prev)


def check(mu, G, cb, prefix):
    rmu, rG = read_basis(prefix + "_radbasis_M8.bin")
    rcb = read_codebook(prefix + "_learned_rq_codebook_M8_u8.bin")
    assert mu.view(np.uint32).tolist() == rmu.view(np.uint32).tolist()
    assert np.array_equal(cb.view(np.uint32), rcb.view(np.uint32))
    flipped = 0
    for m in range(rG.shape[0]):
        same = np.array_equal(G[m].view(np.uint32), rG[m].view(np.uint32))
        neg = np.array_equal((-G[m]).view(np.uint32), rG[m].view(np.uint32))
        assert same or neg, m
        flipped += int(neg and not same)
    return flipped


def test_sample_fit_equals_reference_trainer():
    prefix = os.path.join(GOLDEN, "sample", "sample")
    _, keys = read_rec(prefix + "_baseknn_aknn.rec", 256)
    mu, G, cb, _ = fit_cpu(np.maximum(keys, 0.0))
    check(mu, G, cb, prefix)


def test_sampled_fit_equals_reference_trainer():
    prefix = os.path.join(GOLDEN, "trainer3k", "trainer3k")
    meta = json.load(open(prefix + ".json"))
    cut = trainer_cutoffs(meta["n"])
    import workload
    assert workload.sha256(cut) == meta["cutoffs_sha256"]
    mu, G, cb, _ = fit_cpu(cut, train_sample=meta["train_sample"])
    check(mu, G, cb, prefix)


def test_planted_trainer_mistake_is_caught():
    """A plausible slip -- a training sample drawn with a different seed --
    changes the files, so the comparison above can fail."""
    prefix = os.path.join(GOLDEN, "trainer3k", "trainer3k")
    meta = json.load(open(prefix + ".json"))
    mu, G, cb, _ = fit_cpu(trainer_cutoffs(meta["n"]), train_sample=meta["train_sample"], seed=11)
    with pytest.raises(AssertionError):
        check(mu, G, cb, prefix)
