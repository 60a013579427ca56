"""Differentiable LRQ verifier fit: the radius-reconstruction objective of
tools/train_final_m8_u8.py (:81-156) with the residual codeword choice relaxed
to a softmax, optimised by gradient descent on the rank-M basis and the
residual codebook.

RelaxedLoss is a torch.autograd.Function whose forward and backward are the
hand-written HIP kernels of libanqi_hip.so (anqi_lrq_relax,
csrc/k_lrq_relax.hip).  relaxed_loss_reference is the same objective in
plain torch, used by the tests (FP64, CPU) as the reference gradient and for
finite differences.  fit_relaxed trains from a starting basis/codebook and
finishes the codebook as the reference trainer does (endpoint extension to
the full residual range, :223-226; monotone rows), so the exact floor decode
(anqi_lrq_encode / materialize) stays the arbiter of the result.
"""
from __future__ import annotations

import math
import time

import numpy as np


def _torch():
    import torch
    return torch


def relaxed_loss_reference(cut, mu, G, cb, tau, mode=0):
    """mean_ik (r - mu - c G - q(e))^2 with c = (r - mu) G^T, e = y - c G and
    mode 0: q = sum_j softmax_j(-(e - cb_j)^2 / tau) cb_j   (nearest codeword)
    mode 1: q = cb_0 + sum_{j>=1} sigmoid((e - cb_j)/tau) (cb_j - cb_{j-1})
            (the floor decode, exact as tau -> 0 for sorted codewords).
    Any dtype/device (torch)."""
    torch = _torch()
    y = cut - mu
    c = y @ G.T
    e = y - c @ G
    if mode == 1:
        D = cb[:, 1:] - cb[:, :-1]                                # (nk, ncode-1)
        sg = torch.sigmoid((e.unsqueeze(2) - cb[:, 1:].unsqueeze(0)) / tau)
        q = cb[:, 0].unsqueeze(0) + (sg * D.unsqueeze(0)).sum(dim=2)
    else:
        s = -(e.unsqueeze(2) - cb.unsqueeze(0)) ** 2 / tau       # (n, nk, ncode)
        w = torch.softmax(s, dim=2)
        q = (w * cb.unsqueeze(0)).sum(dim=2)
    return ((e - q) ** 2).mean()


class RelaxedLoss:
    """torch.autograd.Function over anqi_lrq_relax (device FP32)."""

    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, G, cb, cut, mu, tau, mode, actx):
                    loss, gG, gcb = actx.lrq_relax(cut, mu, G.detach().contiguous(), cb.detach().contiguous(), tau,
                                                   mode)
                    ctx.save_for_backward(gG, gcb)
                    return torch.tensor(loss, dtype=torch.float32, device=G.device)

                @staticmethod
                def backward(ctx, gout):
                    gG, gcb = ctx.saved_tensors
                    return gG * gout, gcb * gout, None, None, None, None, None

            cls._fn = _F
        return cls._fn

    @classmethod
    def apply(cls, G, cb, cut, mu, tau, actx, mode=0):
        return cls.fn().apply(G, cb, cut, mu, tau, mode, actx)


def default_tau(cb):
    """Squared median codeword spacing (one spacing ~ one unit of the
    softmax's length scale)."""
    torch = _torch()
    sp = torch.diff(cb, dim=1)
    med = float(torch.median(sp[sp > 0])) if bool((sp > 0).any()) else 1.0
    return med * med


def _finish_codebook(cb, G, mu, rows):
    """Sorted rows, ends extended to the full residual range of rows
    (train_final_m8_u8.py:223-226), monotone (cummax)."""
    torch = _torch()
    cb = torch.sort(cb, dim=1).values
    y = rows - mu
    r = y - (y @ G.T) @ G
    cb[:, 0] = torch.minimum(cb[:, 0], r.amin(0))
    cb[:, -1] = torch.maximum(cb[:, -1], r.amax(0))
    return torch.cummax(cb, dim=1).values.contiguous()


def fit_relaxed(actx, cut, mu, G0, cb0, iters=200, lr_G=1e-3, lr_cb=None, tau0=None, tau1=None,
                all_rows=None, check_every=20, mode=1, log=print):
    """Adam on (G, cb) of the relaxed loss (in RMS-normalised radius units;
    tau0/tau1 in those units), tau annealed geometrically from tau0 to tau1.  cut: (n, nk) device float32 training rows; all_rows: the
    full cutoff matrix for the endpoint extension (defaults to cut).
    Every check_every steps the candidate (G, finished codebook) is decoded
    exactly (floor codes, anqi_lrq_encode/materialize); the best checkpoint by
    floor-decoded RMSE, the starting point included, is returned.
    Returns (G, cb) float32 device tensors, the relaxed-loss history, seconds."""
    torch = _torch()
    t0 = time.perf_counter()
    # work in RMS-normalised radius units (G is scale free)
    sc = float(torch.sqrt((cut.float() ** 2).mean()))
    cut_n = (cut / sc).contiguous()
    mu_n = (mu / sc).contiguous()
    G = G0.detach().clone().float().requires_grad_(True)
    cb = (cb0.detach().clone().float() / sc).requires_grad_(True)
    base_tau = default_tau(cb.detach())
    if mode == 1:   # logistic width: from a few codeword spacings down to a tenth of one
        tau0 = tau0 if tau0 is not None else 4.0 * math.sqrt(base_tau)
        tau1 = tau1 if tau1 is not None else 0.1 * math.sqrt(base_tau)
    else:
        tau0 = tau0 if tau0 is not None else max(1e-4, 16.0 * base_tau)
        tau1 = tau1 if tau1 is not None else 16.0 * base_tau
    opt = torch.optim.Adam([{"params": [G], "lr": lr_G}, {"params": [cb], "lr": 1.0}])
    rows = cut if all_rows is None else all_rows
    # the exact floor decode rules on every checkpoint; the start is one of them
    best_G, best_cb = G0.detach().clone().float(), cb0.detach().clone().float()
    best = floor_rmse(actx, cut, mu, best_G, best_cb)
    log(f"[relaxed-fit] start: floor-decoded RMSE {best:.6g}")
    hist = []
    for it in range(iters):
        tau = tau0 * (tau1 / tau0) ** (it / max(1, iters - 1))
        # codeword steps follow the softmax length scale
        width = tau if mode == 1 else math.sqrt(tau)
        opt.param_groups[1]["lr"] = lr_cb if lr_cb is not None else 0.05 * width
        opt.zero_grad()
        loss = RelaxedLoss.apply(G, cb, cut_n, mu_n, tau, actx, mode)
        loss.backward()
        opt.step()
        if mode == 1:
            with torch.no_grad():     # the floor relaxation assumes sorted codewords
                cb.copy_(torch.sort(cb, dim=1).values)
        hist.append(float(loss.detach()))
        if (it + 1) % check_every == 0 or it + 1 == iters:
            Gc = G.detach().clone().contiguous()
            cbc = _finish_codebook(cb.detach() * sc, Gc, mu, rows)
            r = floor_rmse(actx, cut, mu, Gc, cbc)
            log(f"[relaxed-fit] iter={it} tau={tau:.4g} relaxed loss={hist[-1]:.6g} floor-decoded RMSE={r:.6g}")
            if r < best:
                best, best_G, best_cb = r, Gc, cbc
    return best_G.contiguous(), best_cb.contiguous(), hist, time.perf_counter() - t0


def floor_rmse(actx, cut, mu, G, cb):
    """RMSE of the exact floor-decoded radii (anqi_lrq_encode + decode of
    every rank) against the cutoffs."""
    torch = _torch()
    coeff, codes, codes_km = actx.lrq_encode(cut, mu, G, cb)
    nk = cut.shape[1]
    err = 0.0
    for k in range(1, nk + 1):
        thr = actx.lrq_materialize(coeff, codes_km, mu, G, cb, k, 1.0)
        d = thr.cpu().double() - cut[:, k - 1].clamp(min=0).cpu().double()
        err += float((d ** 2).sum())
    return math.sqrt(err / (cut.shape[0] * nk))
