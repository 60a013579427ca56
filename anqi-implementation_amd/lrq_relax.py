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
(anqi_lrq_encode / materialize) stays the judge of the result.
"""
from __future__ import annotations

import math
import time

import numpy as np


def _torch():
    import torch
    return torch


def relaxed_loss_reference(cut, mu, G, cb, tau):
    """mean_ik (r - mu - c G - sum_j w_j cb_j)^2 with c = (r - mu) G^T and
    w = softmax_j(-(e - cb_j)^2 / tau); any dtype/device (torch)."""
    torch = _torch()
    y = cut - mu
    c = y @ G.T
    e = y - c @ G
    s = -(e.unsqueeze(2) - cb.unsqueeze(0)) ** 2 / tau          # (n, nk, ncode)
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
                def forward(ctx, G, cb, cut, mu, tau, actx):
                    loss, gG, gcb = actx.lrq_relax(cut, mu, G.detach().contiguous(), cb.detach().contiguous(), tau)
                    ctx.save_for_backward(gG, gcb)
                    return torch.tensor(loss, dtype=torch.float32, device=G.device)

                @staticmethod
                def backward(ctx, gout):
                    gG, gcb = ctx.saved_tensors
                    return gG * gout, gcb * gout, None, None, None, None

            cls._fn = _F
        return cls._fn

    @classmethod
    def apply(cls, G, cb, cut, mu, tau, actx):
        return cls.fn().apply(G, cb, cut, mu, tau, actx)


def default_tau(cb):
    """Squared median codeword spacing (one spacing ~ one unit of the
    softmax's length scale)."""
    torch = _torch()
    sp = torch.diff(cb, dim=1)
    med = float(torch.median(sp[sp > 0])) if bool((sp > 0).any()) else 1.0
    return med * med


def fit_relaxed(actx, cut, mu, G0, cb0, iters=200, lr_G=1e-3, lr_cb=None, tau0=None, tau1=None,
                all_rows=None, log=print):
    """Adam on (G, cb) of the relaxed loss (in RMS-normalised radius units;
    tau0/tau1 in those units), tau annealed geometrically from tau0 to tau1.  cut: (n, nk) device float32 training rows; all_rows: the
    full cutoff matrix for the endpoint extension (defaults to cut).
    Returns (G, cb) float32 device tensors and a history of losses."""
    torch = _torch()
    t0 = time.perf_counter()
    # work in RMS-normalised radius units (G is scale free)
    sc = float(torch.sqrt((cut.float() ** 2).mean()))
    cut_n = (cut / sc).contiguous()
    mu_n = (mu / sc).contiguous()
    G = G0.detach().clone().float().requires_grad_(True)
    cb = (cb0.detach().clone().float() / sc).requires_grad_(True)
    base_tau = default_tau(cb.detach())
    tau0 = tau0 if tau0 is not None else 4.0 * base_tau
    tau1 = tau1 if tau1 is not None else 0.25 * base_tau
    lr_cb = lr_cb if lr_cb is not None else 0.05 * math.sqrt(base_tau)
    opt = torch.optim.Adam([{"params": [G], "lr": lr_G}, {"params": [cb], "lr": lr_cb}])
    hist = []
    for it in range(iters):
        tau = tau0 * (tau1 / tau0) ** (it / max(1, iters - 1))
        opt.zero_grad()
        loss = RelaxedLoss.apply(G, cb, cut_n, mu_n, tau, actx)
        loss.backward()
        opt.step()
        hist.append(float(loss))
        if it % 50 == 0 or it + 1 == iters:
            log(f"[relaxed-fit] iter={it} tau={tau:.4g} loss={hist[-1]:.6g}")
    G = G.detach()
    cb = torch.sort(cb.detach(), dim=1).values * sc
    # endpoint extension to the full residual range (train_final_m8_u8.py:223-226)
    rows = cut if all_rows is None else all_rows
    y = rows - mu
    r = y - (y @ G.T) @ G
    cb[:, 0] = torch.minimum(cb[:, 0], r.amin(0))
    cb[:, -1] = torch.maximum(cb[:, -1], r.amax(0))
    cb = torch.cummax(cb, dim=1).values
    return G.contiguous(), cb.contiguous(), hist, time.perf_counter() - t0


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
