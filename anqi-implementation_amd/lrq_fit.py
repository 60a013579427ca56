"""Rank-M basis + u8 residual codebook fit for the LRQ verifier, on the GPU.

Restates tools/train_final_m8_u8.py (constants :16-25) with torch tensors:
  fit_basis            :81-89   mean, covariance, eigh -> top-M basis (FP64)
  reconstruct          :92-94
  train_codebook       :97-122  quantile init (i + 0.5)/256, 8 Lloyd steps,
                                cumulative-max monotone centres
  decode_nearest       :146-156
  full_residual_range  :159-173 and the endpoint extension :223-226
  main                 :190-265 (first 120,000 rows seed the basis; the
                                training sample is default_rng(10).choice,
                                sorted; 3 alternating rounds)
The basis runs in FP64; the codebooks in FP64 or FP32 (fit's codebook_dtype).
Outputs are the radbasis and codebook arrays (float32) in the reference file
layouts.
"""
from __future__ import annotations

import struct
import time

import numpy as np

RANK_M = 8
K_MAX = 100
BITS = 8
CODE_SIZE = 1 << BITS
TRAIN_ROWS = 120_000
RECURSIVE_ITERS = 3
LLOYD_ITERS = 8
SEED = 10
MAGIC = b"ANQIRQ1\0"


def _torch():
    import torch
    return torch


def fit_basis(rows):
    torch = _torch()
    mean = rows.mean(dim=0)
    c = rows - mean
    cov = c.T @ c / max(1, rows.shape[0])
    w, v = torch.linalg.eigh(cov)
    order = torch.argsort(w, descending=True)
    basis = v[:, order[:RANK_M]].T.contiguous()
    explained = float(w[order[:RANK_M]].sum() / torch.clamp(w.sum(), min=1e-30))
    return mean, basis, explained


def reconstruct(rows, mean, basis):
    return mean + ((rows - mean) @ basis.T) @ basis


def _quantile_linear(sorted_cols, q):
    """numpy's default ('linear') quantile of each row of sorted_cols."""
    torch = _torch()
    n = sorted_cols.shape[1]
    pos = q * (n - 1)
    lo = torch.floor(pos).long()
    hi = torch.clamp(lo + 1, max=n - 1)
    t = pos - lo.to(pos.dtype)
    a = sorted_cols[:, lo]
    b = sorted_cols[:, hi]
    diff = b - a
    out = a + diff * t
    alt = b - diff * (1 - t)
    return torch.where(t >= 0.5, alt, out)


def train_codebooks(residual, iterations=LLOYD_ITERS):
    """One 256-level 1-D Lloyd codebook per column (columns in parallel)."""
    torch = _torch()
    vals = torch.sort(residual.T.contiguous(), dim=1).values          # (K, N)
    k, n = vals.shape
    dt = vals.dtype
    q = (torch.arange(CODE_SIZE, dtype=dt, device=vals.device) + 0.5) / CODE_SIZE
    centres = _quantile_linear(vals, q)
    prefix = torch.cat([torch.zeros(k, 1, dtype=dt, device=vals.device), torch.cumsum(vals, dim=1)], 1)
    active = torch.ones(k, dtype=torch.bool, device=vals.device)
    for _ in range(iterations):
        bounds = (centres[:, :-1] + centres[:, 1:]) * 0.5
        cuts = torch.searchsorted(vals, bounds.contiguous(), right=True)
        starts = torch.cat([torch.zeros(k, 1, dtype=cuts.dtype, device=vals.device), cuts], 1)
        ends = torch.cat([cuts, torch.full((k, 1), n, dtype=cuts.dtype, device=vals.device)], 1)
        counts = ends - starts
        sums = torch.gather(prefix, 1, ends) - torch.gather(prefix, 1, starts)
        upd = torch.where(counts > 0, sums / torch.clamp(counts, min=1).to(dt), centres)
        upd = torch.cummax(upd, dim=1).values
        done = (upd - centres).abs().amax(dim=1) < 1e-12
        centres = torch.where(active[:, None], upd, centres)
        active = active & ~done
        if not bool(active.any()):
            break
    same = vals[:, 0] == vals[:, -1]
    if bool(same.any()):
        centres[same] = vals[same, :1].expand(-1, CODE_SIZE)
    return centres                                                    # (K, 256)


def decode_nearest(residual, codebooks):
    torch = _torch()
    out = torch.empty_like(residual)
    v = residual.T.contiguous()                                      # (K, N)
    ins = torch.searchsorted(codebooks.contiguous(), v, right=False)
    hi = torch.clamp(ins, 0, CODE_SIZE - 1)
    lo = torch.clamp(ins - 1, 0, CODE_SIZE - 1)
    ch = torch.gather(codebooks, 1, hi)
    cl = torch.gather(codebooks, 1, lo)
    take_hi = (ch - v).abs() < (v - cl).abs()
    out = torch.where(take_hi, ch, cl).T.contiguous()
    return out


def fit(thresholds, train_sample=TRAIN_ROWS, seed=SEED, recursive_iters=RECURSIVE_ITERS,
        lloyd_iters=LLOYD_ITERS, chunk=65536, log=print, codebook_dtype=None):
    """thresholds: (n, 100) float32 tensor (the AKNN cutoff keys).
    Returns (mean[100] f32, basis[8,100] f32, codebook[100,256] f32) as numpy.
    The basis (mean, covariance, eigenvectors) and the rank-M projection run
    in FP64, as the reference's numpy does.  codebook_dtype sets the
    arithmetic of the residual codebooks (quantile start, Lloyd rounds,
    nearest decode): FP64 (the default, the reference's arithmetic, bit-equal
    to its files on one CPU thread) or FP32 (what the GPU pipeline uses, so
    that FP64 on the device stays in the least-squares/eigen step; the exact
    floor decode of the encode judges whichever codebook results)."""
    torch = _torch()
    t0 = time.perf_counter()
    n = thresholds.shape[0]
    if n < CODE_SIZE:
        raise RuntimeError(f"at least {CODE_SIZE} rows are required for u8 training, got {n}")
    sample_n = min(train_sample, n)
    seed_rows = thresholds[:sample_n].to(torch.float64)
    mean, basis, explained = fit_basis(seed_rows)
    rng = np.random.default_rng(seed)
    ids = np.sort(rng.choice(n, size=sample_n, replace=False))
    rows = thresholds[torch.from_numpy(ids).to(thresholds.device)].to(torch.float64)
    cdt = codebook_dtype or torch.float64
    codebooks = None
    for it in range(recursive_iters):
        recon = reconstruct(rows, mean, basis)
        resid = rows - recon
        codebooks = train_codebooks(resid.to(cdt), lloyd_iters)
        dec = decode_nearest(resid.to(cdt), codebooks).to(torch.float64)
        rmse = float(torch.sqrt(torch.mean((recon + dec - rows) ** 2)))
        log(f"[verifier-train] round={it + 1}/{recursive_iters} sample_rmse={rmse:.6g} explained={explained:.6f}")
        if it + 1 < recursive_iters:
            mean, basis, explained = fit_basis(rows - dec)
    codebooks = codebooks.to(torch.float64)
    lo = torch.full((K_MAX,), float("inf"), dtype=torch.float64, device=thresholds.device)
    hi = torch.full((K_MAX,), float("-inf"), dtype=torch.float64, device=thresholds.device)
    for s in range(0, n, chunk):
        blk = thresholds[s:s + chunk].to(torch.float64)
        r = blk - reconstruct(blk, mean, basis)
        lo = torch.minimum(lo, r.amin(0))
        hi = torch.maximum(hi, r.amax(0))
    codebooks[:, 0] = torch.minimum(codebooks[:, 0], lo)
    codebooks[:, -1] = torch.maximum(codebooks[:, -1], hi)
    codebooks = torch.cummax(codebooks, dim=1).values
    mean32 = mean.float().cpu().numpy()
    basis32 = basis.float().cpu().numpy()
    cb32 = codebooks.float().cpu().numpy()
    if not (np.isfinite(mean32).all() and np.isfinite(basis32).all() and np.isfinite(cb32).all()):
        raise RuntimeError("trained verifier contains non-finite values")
    if np.any(np.diff(cb32, axis=1) < 0):
        raise RuntimeError("trained codebook is not sorted")
    return mean32, basis32, cb32, time.perf_counter() - t0


def write_basis(path, mean, basis):
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", K_MAX, RANK_M))
        f.write(np.asarray(mean, "<f4").tobytes())
        f.write(np.asarray(basis, "<f4").tobytes())


def write_codebook(path, cb):
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<iiii", K_MAX, RANK_M, BITS, CODE_SIZE))
        f.write(np.asarray(cb, "<f4").tobytes())
