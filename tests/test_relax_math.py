"""The backward formulas of the relaxed LRQ objective (the ones the HIP
kernels csrc/k_lrq_relax.hip implement) restated in numpy FP64 and checked
against torch autograd of the forward on the CPU, and against central finite
differences.  Pins the derivation independently of the GPU."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def manual_grads_floor(cut, mu, G, cb, beta):
    y = cut - mu
    c = y @ G.T
    e = y - c @ G
    D = cb[:, 1:] - cb[:, :-1]
    sg = 1 / (1 + np.exp(-(e[:, :, None] - cb[None, :, 1:]) / beta))      # s_j, j >= 1
    q = cb[None, :, 0] + (sg * D[None]).sum(axis=2)
    z = e - q
    g = 2 * z / z.size
    dq = (sg * (1 - sg) * D[None]).sum(axis=2) / beta
    ge = g * (1 - dq)
    J = cb.shape[1]
    dqd = np.zeros((z.shape[0], z.shape[1], J))
    dqd[:, :, 0] = 1 - sg[:, :, 0]
    dqd[:, :, 1:] = sg - sg * (1 - sg) * D[None] / beta
    dqd[:, :, 1:J - 1] -= sg[:, :, 1:]
    gcb = (-g[:, :, None] * dqd).sum(0)
    gp = -ge
    gc = gp @ G.T
    gG = c.T @ gp + gc.T @ y
    return (z ** 2).mean(), gG, gcb


def manual_grads(cut, mu, G, cb, tau):
    y = cut - mu
    c = y @ G.T
    e = y - c @ G
    s = -(e[:, :, None] - cb[None]) ** 2 / tau
    s -= s.max(axis=2, keepdims=True)
    w = np.exp(s)
    w /= w.sum(axis=2, keepdims=True)
    q = (w * cb[None]).sum(axis=2)
    var = (w * cb[None] ** 2).sum(axis=2) - q ** 2
    z = e - q
    scale = 1.0 / z.size
    g = 2 * z * scale
    ge = g * (1 - 2 / tau * var)
    gcb = (-g[:, :, None] * w * (1 + 2 / tau * (cb[None] - q[:, :, None]) * (e[:, :, None] - cb[None]))).sum(0)
    gp = -ge
    gc = gp @ G.T
    gG = c.T @ gp + gc.T @ y
    return (z ** 2).mean(), gG, gcb


def test_manual_backward_equals_autograd_and_finite_differences():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "anqi-implementation_amd"))
    from lrq_relax import relaxed_loss_reference
    rng = np.random.default_rng(0)
    n, nk, m, nc = 40, 100, 8, 256
    cut = np.cumsum(rng.gamma(2.0, 50.0, (n, nk)), axis=1)      # increasing radii per row, like cutoffs
    cut /= np.sqrt((cut ** 2).mean())                             # the fit works in RMS-normalised units
    mu = cut.mean(0)
    u, _, vt = np.linalg.svd(cut - mu, full_matrices=False)
    G = vt[:m].copy()
    resid = (cut - mu) - ((cut - mu) @ G.T) @ G
    cb = np.sort(np.quantile(resid, (np.arange(nc) + 0.5) / nc, axis=0).T, axis=1) + rng.normal(0, 1e-3, (nk, nc))
    tau = float(np.median(np.diff(cb, axis=1))) ** 2 * 4
    L, gG, gcb = manual_grads(cut, mu, G, cb, tau)
    Gt = torch.tensor(G, requires_grad=True)
    cbt = torch.tensor(cb, requires_grad=True)
    Lt = relaxed_loss_reference(torch.tensor(cut), torch.tensor(mu), Gt, cbt, tau)
    Lt.backward()
    assert abs(L - Lt.item()) <= 1e-12 * abs(L)
    assert np.abs(gG - Gt.grad.numpy()).max() <= 1e-9 * np.abs(gG).max()
    assert np.abs(gcb - cbt.grad.numpy()).max() <= 1e-9 * np.abs(gcb).max()
    eps = 1e-4
    for (a, b) in [(0, 0), (3, 17), (7, 99), (5, 50)]:
        P = G.copy(); P[a, b] += eps
        M_ = G.copy(); M_[a, b] -= eps
        fd = (manual_grads(cut, mu, P, cb, tau)[0] - manual_grads(cut, mu, M_, cb, tau)[0]) / (2 * eps)
        assert abs(fd - gG[a, b]) <= 1e-5 * max(1.0, abs(gG[a, b])), (a, b, fd, gG[a, b])
    for (a, b) in [(0, 0), (10, 128), (99, 255), (50, 3)]:
        P = cb.copy(); P[a, b] += eps
        M_ = cb.copy(); M_[a, b] -= eps
        fd = (manual_grads(cut, mu, G, P, tau)[0] - manual_grads(cut, mu, G, M_, tau)[0]) / (2 * eps)
        assert abs(fd - gcb[a, b]) <= 1e-5 * max(1.0, abs(gcb[a, b])), (a, b, fd, gcb[a, b])


def test_floor_relaxation_backward_equals_autograd_and_finite_differences():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "anqi-implementation_amd"))
    from lrq_relax import relaxed_loss_reference
    rng = np.random.default_rng(1)
    n, nk, m, nc = 40, 100, 8, 256
    cut = np.cumsum(rng.gamma(2.0, 50.0, (n, nk)), axis=1)
    cut /= np.sqrt((cut ** 2).mean())
    mu = cut.mean(0)
    _, _, vt = np.linalg.svd(cut - mu, full_matrices=False)
    G = vt[:m].copy()
    resid = (cut - mu) - ((cut - mu) @ G.T) @ G
    cb = np.sort(np.quantile(resid, (np.arange(nc) + 0.5) / nc, axis=0).T + rng.normal(0, 1e-4, (nk, nc)), axis=1)
    beta = 4 * float(np.median(np.diff(cb, axis=1)))
    L, gG, gcb = manual_grads_floor(cut, mu, G, cb, beta)
    Gt = torch.tensor(G, requires_grad=True)
    cbt = torch.tensor(cb, requires_grad=True)
    Lt = relaxed_loss_reference(torch.tensor(cut), torch.tensor(mu), Gt, cbt, beta, mode=1)
    Lt.backward()
    assert abs(L - Lt.item()) <= 1e-12 * abs(L)
    assert np.abs(gG - Gt.grad.numpy()).max() <= 1e-9 * np.abs(gG).max()
    assert np.abs(gcb - cbt.grad.numpy()).max() <= 1e-9 * np.abs(gcb).max()
    # the floor decode is the tau -> 0 limit
    y = cut - mu
    e = y - (y @ G.T) @ G
    idx = np.array([[max(0, np.searchsorted(cb[k], e[i, k], side="right") - 1) for k in range(nk)]
                    for i in range(n)])
    floor = cb[np.arange(nk)[None, :], idx]
    q_small = relaxed_loss_reference(torch.tensor(cut), torch.tensor(mu), torch.tensor(G), torch.tensor(cb),
                                     1e-12, mode=1).item()
    assert abs(q_small - ((e - floor) ** 2).mean()) <= 1e-9 * max(1.0, q_small)
    eps = 1e-6
    for (a, b) in [(0, 0), (3, 17)]:
        P = G.copy(); P[a, b] += eps
        M_ = G.copy(); M_[a, b] -= eps
        fd = (manual_grads_floor(cut, mu, P, cb, beta)[0] - manual_grads_floor(cut, mu, M_, cb, beta)[0]) / (2 * eps)
        assert abs(fd - gG[a, b]) <= 1e-4 * max(1e-3, abs(gG[a, b])), (a, b, fd, gG[a, b])
    for (a, b) in [(0, 0), (10, 128), (99, 255)]:
        P = cb.copy(); P[a, b] += eps
        M_ = cb.copy(); M_[a, b] -= eps
        fd = (manual_grads_floor(cut, mu, G, P, beta)[0] - manual_grads_floor(cut, mu, G, M_, beta)[0]) / (2 * eps)
        assert abs(fd - gcb[a, b]) <= 1e-4 * max(1e-3, abs(gcb[a, b])), (a, b, fd, gcb[a, b])


def test_relaxations_reduce_to_the_exact_decodes_as_the_width_vanishes():
    """As the width goes to 0 the relaxed objective becomes the exact decode
    it relaxes: mode 1 the floor decode (largest codeword <= e, the first
    codeword below the range: src/rknn_query.cpp:638-720's floor policy),
    mode 0 the nearest codeword.  Residuals are placed away from codeword
    midpoints/edges so the limit is reached at a finite width."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "anqi-implementation_amd"))
    from lrq_relax import relaxed_loss_reference
    rng = np.random.default_rng(2)
    n, nk, m, nc = 30, 20, 4, 16
    cb = np.sort(rng.uniform(-10, 10, (nk, nc)), axis=1)
    cb[:, 1:] = np.maximum(cb[:, 1:], cb[:, :-1] + 0.5)        # spacing >= 0.5
    G = np.zeros((m, nk))          # no basis part: e = cut - mu exactly (the basis is pinned above)
    # residuals e in a known place: 70% of the way from codeword j to j+1
    # (floor picks j, nearest picks j+1; >= 0.15 from either)
    j = rng.integers(0, nc - 1, (n, nk))
    cbb = np.broadcast_to(cb, (n, nk, nc))
    lo = np.take_along_axis(cbb, j[:, :, None], 2)[:, :, 0]
    hi = np.take_along_axis(cbb, j[:, :, None] + 1, 2)[:, :, 0]
    e = lo + 0.7 * (hi - lo)
    e[:, :3] = cb[None, :3, 0] - 1.0                           # some below the first codeword
    mu = rng.uniform(50, 60, nk)
    cut = mu + e
    T = lambda a: torch.tensor(a, dtype=torch.float64)
    floor_q = np.empty_like(e)
    near_q = np.empty_like(e)
    for k in range(nk):
        f = np.searchsorted(cb[k], e[:, k], side="right") - 1
        floor_q[:, k] = cb[k, np.maximum(f, 0)]
        near_q[:, k] = cb[k, np.abs(e[:, k][:, None] - cb[k][None]).argmin(1)]
    exact_floor = ((e - floor_q) ** 2).mean()
    exact_near = ((e - near_q) ** 2).mean()
    l1 = float(relaxed_loss_reference(T(cut), T(mu), T(G), T(cb), 1e-4, mode=1))
    l0 = float(relaxed_loss_reference(T(cut), T(mu), T(G), T(cb), 1e-6, mode=0))
    assert abs(l1 - exact_floor) <= 1e-9 * max(1.0, exact_floor)
    assert abs(l0 - exact_near) <= 1e-9 * max(1.0, exact_near)
    # and they differ from each other here (the relaxations are not interchangeable)
    assert abs(exact_floor - exact_near) > 1e-3
