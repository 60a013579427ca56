"""Checks shared by the differentiable-verifier-fit GPU tests: the HIP
backward against the FP64 torch reference gradient and FP64 central finite
differences, and the fitted verifier's exact floor decode against the
reference's LRQ build and threshold decode (oracle/_ref harness)."""
import numpy as np

from anqi_testlib import ptr, ref, ref_available


def grad_check(ctx, cut, mu, G, cb, mode, say, rows=1000, eps=1e-4, coords=12):
    """Relaxed loss on the first `rows` rows, in RMS-normalised radius units,
    at a point perturbed away from the Lloyd/PCA optimum (where the gradient
    would be a small difference of large terms).  tau = 1e-4 (softmax length
    scale 0.01) or a logistic width of 1e-2 keeps eps-sized steps inside the
    smooth range.  Returns (loss rel err, basis grad rel err, codebook grad
    rel err, worst FD error / max|grad|)."""
    import torch
    from lrq_relax import default_tau, relaxed_loss_reference
    sc = float(torch.sqrt((cut.float() ** 2).mean()))
    sub = (cut[:rows] / sc).contiguous()
    mu, cb = (mu / sc).contiguous(), (cb / sc).contiguous()
    tau = 1e-4 if mode == 0 else 1e-2
    gen = torch.Generator(device="cpu").manual_seed(5)
    G = (G + 0.02 * torch.randn(G.shape, generator=gen).to(G.device)).contiguous()
    cb = torch.sort(cb + 0.5 * default_tau(cb) ** 0.5 * torch.randn(cb.shape, generator=gen).to(cb.device),
                    dim=1).values.contiguous()
    loss, gG, gcb = ctx.lrq_relax(sub, mu, G, cb, tau, mode=mode)
    S64, mu64 = sub.double().cpu(), mu.double().cpu()
    G64 = G.double().cpu().requires_grad_(True)
    cb64 = cb.double().cpu().requires_grad_(True)
    L = relaxed_loss_reference(S64, mu64, G64, cb64, tau, mode=mode)
    L.backward()
    rl = abs(loss - L.item()) / L.item()
    rg = float((gG.double().cpu() - G64.grad).norm() / G64.grad.norm())
    rc = float((gcb.double().cpu() - cb64.grad).norm() / cb64.grad.norm())
    say(f"mode {mode}: loss hip {loss:.9g} fp64 {L.item():.9g}; grad rel err basis {rg:.2e} codebook {rc:.2e}")
    rng = np.random.default_rng(1)
    worst = 0.0
    with torch.no_grad():
        for which, P, grad in (("G", G64, gG), ("cb", cb64, gcb)):
            scale = float(grad.abs().max())
            for _ in range(coords):
                idx = tuple(int(rng.integers(0, s)) for s in P.shape)
                Pp, Pm = P.detach().clone(), P.detach().clone()
                Pp[idx] += eps
                Pm[idx] -= eps
                args_p = (Pp, cb64.detach()) if which == "G" else (G64.detach(), Pp)
                args_m = (Pm, cb64.detach()) if which == "G" else (G64.detach(), Pm)
                fd = (relaxed_loss_reference(S64, mu64, *args_p, tau, mode=mode) -
                      relaxed_loss_reference(S64, mu64, *args_m, tau, mode=mode)).item() / (2 * eps)
                worst = max(worst, abs(fd - float(grad[idx])) / scale)
    say(f"finite differences (eps {eps:g}, {2 * coords} coordinates): worst error / max|grad| = {worst:.2e}")
    return rl, rg, rc, worst


def decisions_match_reference(ctx, cut, mu, G, cb, ks=(10, 50, 100)):
    """LRQ coefficients, codes and the decoded thresholds of every row
    (hence every accept decision key <= t_k) equal the reference's on the
    fitted (mu, G, cb).  Returns None when the harness is not built."""
    if not ref_available():
        return None
    R = ref()
    n, nk = cut.shape
    c_np = cut.cpu().numpy()
    mu_np, G_np, cb_np = (t.cpu().numpy().astype(np.float32) for t in (mu, G, cb))
    coeff, codes, codes_km = ctx.lrq_encode(cut, mu, G, cb)
    rc = np.zeros((n, G_np.shape[0]), np.float32)
    rk = np.zeros((n, nk), np.uint8)
    R.ref_lrq_build(n, nk, G_np.shape[0], ptr(mu_np), ptr(G_np), ptr(cb_np), cb_np.shape[1], ptr(c_np), ptr(rc),
                    ptr(rk))
    ok = np.array_equal(coeff.cpu().numpy().view(np.uint32), rc.view(np.uint32))
    ok = ok and np.array_equal(codes.cpu().numpy(), rk)
    for k in ks:
        t_gpu = ctx.lrq_materialize(coeff, codes_km, mu, G, cb, k, 1.0).cpu().numpy()
        t_ref = np.zeros(n, np.float32)
        R.ref_lrq_materialize(n, nk, G_np.shape[0], ptr(mu_np), ptr(G_np), ptr(cb_np), cb_np.shape[1], ptr(rc),
                              ptr(rk), k, 1.0, ptr(t_ref))
        ok = ok and np.array_equal(t_gpu.view(np.uint32), t_ref.view(np.uint32))
    return ok
