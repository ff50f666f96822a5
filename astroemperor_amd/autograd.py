"""torch.autograd.Function over the adjoint kernel, and a batched MAP search.

``LogLikelihood.apply(theta, model)`` returns lnL[B] for a float64 CUDA
tensor theta[B, ndim]; its backward uses dlnL/dtheta from the hand-written
adjoint kernel (emp_lnl_grad_batch), computed in the same launch as lnL.

``map_search`` runs a batched L-BFGS (Nocedal 1980 two-loop recursion, one
independent problem per start with its own curvature history and Armijo
backtracking) and returns the posterior optimum reached from every start; it
seeds PT walkers close to the posterior modes.
"""
from __future__ import annotations

import numpy as np
import torch


class LogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, model):
        lnl, grad = model.lnl_grad_batch(theta)
        ctx.save_for_backward(grad)
        return lnl

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g[:, None] * grad, None


def log_prior_torch(model, theta):
    """Differentiable restatement of the prior terms used on the device
    (Uniform/Jeffreys: constant inside the box, Normal: Gaussian, Beta: log
    density) plus the derived e = S^2 + C^2 support; -inf outside."""
    lim = torch.as_tensor(model.limits_free(), dtype=theta.dtype, device=theta.device)
    inside = ((theta >= lim[:, 0]) & (theta <= lim[:, 1])).all(dim=1)
    lp = torch.zeros(theta.shape[0], dtype=theta.dtype, device=theta.device)
    free = [p for p in model.params() if p.fixed is None]
    for j, p in enumerate(free):
        if p.prior == "Normal":
            mu, s, logZ = p.prargs
            lp = lp - 0.5 * ((theta[:, j] - mu) / s) ** 2 - float(np.log(s * np.sqrt(2 * np.pi))) - logZ
        elif p.prior == "Beta":
            from scipy.special import betaln
            a, b = p.prargs[0], p.prargs[1]
            x = theta[:, j].clamp(1e-300, 1 - 1e-16)
            lp = lp + (a - 1) * torch.log(x) + (b - 1) * torch.log1p(-x) - float(betaln(a, b))
        elif p.prior in ("Uniform", "Jeffreys"):
            lp = lp + float(p.prargs if p.prior == "Uniform" else np.log(1 / (p.limits[1] - p.limits[0])))
    if model.parameterisation in (1, 2, 4, 7):
        full_idx = {int(c): j for j, c in enumerate(model.C_)}
        for k in range(model.kplan__):
            s_i, c_i = full_idx.get(5 * k + 3), full_idx.get(5 * k + 4)
            if s_i is not None and c_i is not None:
                inside = inside & (theta[:, s_i] ** 2 + theta[:, c_i] ** 2 <= 1)
    return torch.where(inside, lp, torch.full_like(lp, -np.inf))


def _prior_grad(model, theta):
    """d ln prior / d theta inside the support (Normal and Beta terms)."""
    g = torch.zeros_like(theta)
    free = [p for p in model.params() if p.fixed is None]
    for j, p in enumerate(free):
        if p.prior == "Normal":
            mu, s, _ = p.prargs
            g[:, j] = -(theta[:, j] - mu) / s ** 2
        elif p.prior == "Beta":
            a, b = p.prargs[0], p.prargs[1]
            x = theta[:, j].clamp(1e-12, 1 - 1e-12)
            g[:, j] = (a - 1) / x - (b - 1) / (1 - x)
    return g


def map_search(model, starts, iters=100, history=10, device="cuda"):
    """Batched L-BFGS on f = -(lnL + ln prior), one independent problem per row.

    Every row keeps its own curvature pairs and its own backtracking (Armijo)
    step; iterates are projected into the prior box, and a row whose step
    leaves the support (e.g. S^2 + C^2 > 1) backtracks like any other failed
    step.  Returns (theta [B, ndim], log posterior [B]) as numpy arrays."""
    x = torch.tensor(np.asarray(starts, dtype=np.float64), device=device)
    B, D = x.shape
    lim = torch.as_tensor(model.limits_free(), dtype=torch.float64, device=device)
    lo, hi = lim[:, 0], lim[:, 1]

    def proj(z):
        return torch.maximum(torch.minimum(z, hi), lo)

    def fg(z):
        lnl, gl = model.lnl_grad_batch(z)
        lp = log_prior_torch(model, z)
        f = -(lnl + lp)
        g = -(gl + _prior_grad(model, z))
        bad = ~torch.isfinite(f)
        f = torch.where(bad, torch.full_like(f, np.inf), f)
        g = torch.where(bad[:, None], torch.zeros_like(g), g)
        return f, g

    x = proj(x)
    f, g = fg(x)
    S = torch.zeros(history, B, D, dtype=torch.float64, device=device)
    Y = torch.zeros_like(S)
    rho = torch.zeros(history, B, dtype=torch.float64, device=device)
    n_hist = 0
    for it in range(iters):
        # two-loop recursion, per row
        q = g.clone()
        alpha = torch.zeros(history, B, dtype=torch.float64, device=device)
        k = min(n_hist, history)
        newest = history - 1                      # pairs live in slots [history - k, history)
        for i in range(newest, history - k - 1, -1):
            alpha[i] = rho[i] * (S[i] * q).sum(1)
            q = q - alpha[i][:, None] * Y[i]
        if k:
            yy = (Y[newest] * Y[newest]).sum(1)
            sy_n = (S[newest] * Y[newest]).sum(1)
            gamma = torch.where((yy > 0) & (sy_n > 0), sy_n / torch.where(yy > 0, yy, torch.ones_like(yy)),
                                torch.full_like(yy, 1e-6))
            q = q * gamma[:, None]
        else:
            q = q / torch.clamp(g.norm(dim=1, keepdim=True), min=1.0) * 1e-3 * (hi - lo)
        for i in range(history - k, history):
            beta = rho[i] * (Y[i] * q).sum(1)
            q = q + S[i] * (alpha[i] - beta)[:, None]
        d = -q
        slope = (g * d).sum(1)
        d = torch.where((slope < 0)[:, None], d, -g * 1e-3 * (hi - lo))
        slope = (g * d).sum(1)
        step = torch.ones(B, dtype=torch.float64, device=device)
        active = torch.isfinite(f)
        xn, fn, gn = x.clone(), f.clone(), g.clone()
        for _ in range(30):
            if not active.any():
                break
            xt = proj(x + step[:, None] * d)
            ft, gt = fg(xt)
            ok = active & (ft <= f + 1e-4 * step * slope)
            xn[ok], fn[ok], gn[ok] = xt[ok], ft[ok], gt[ok]
            active = active & ~ok
            step = torch.where(active, step * 0.5, step)
        s_vec, y_vec = xn - x, gn - g
        sy = (s_vec * y_vec).sum(1)
        good = sy > 1e-12 * s_vec.norm(dim=1) * y_vec.norm(dim=1)
        S = torch.roll(S, -1, 0)
        Y = torch.roll(Y, -1, 0)
        rho = torch.roll(rho, -1, 0)
        S[-1] = torch.where(good[:, None], s_vec, torch.zeros_like(s_vec))
        Y[-1] = torch.where(good[:, None], y_vec, torch.zeros_like(y_vec))
        rho[-1] = torch.where(good, 1.0 / torch.where(good, sy, torch.ones_like(sy)), torch.zeros_like(sy))
        n_hist += 1
        moved = (xn - x).abs().max()
        x, f, g = xn, fn, gn
        if moved == 0:
            break
    return x.cpu().numpy(), (-f).cpu().numpy()


def _prior_hess_diag(model, theta):
    """d^2 ln prior / d theta^2 inside the support (Normal and Beta terms)."""
    h = np.zeros_like(theta)
    free = [p for p in model.params() if p.fixed is None]
    for j, p in enumerate(free):
        if p.prior == "Normal":
            h[j] = -1.0 / p.prargs[1] ** 2
        elif p.prior == "Beta":
            a, b = p.prargs[0], p.prargs[1]
            x = min(max(theta[j], 1e-12), 1 - 1e-12)
            h[j] = -(a - 1) / x ** 2 - (b - 1) / (1 - x) ** 2
    return h


def _hessian(model, theta, h):
    """Hessian of lnL + ln prior at theta: central differences of the
    adjoint gradient with per-parameter steps h (all 2 D points in one
    batch), symmetrised, plus the prior's curvature."""
    D = len(theta)
    pts = np.repeat(theta[None], 2 * D, axis=0)
    pts[np.arange(D), np.arange(D)] += h
    pts[D + np.arange(D), np.arange(D)] -= h
    _, g = model.lnl_grad_batch(pts)
    H = (g[:D] - g[D:]) / (2 * h[:, None])
    return 0.5 * (H + H.T) + np.diag(_prior_hess_diag(model, theta))


def laplace_scales(model, theta, h_rel=1e-8, h_sigma=1e-2):
    """Conditional posterior scales at a mode, sigma_j = (-H_jj)^-1/2, H the
    Hessian of lnL + ln prior.  The posterior can be far narrower than the
    parameters' size (a period known to 1e-5 d), so H is taken twice: with
    steps h_rel max(|theta|, 1), then with steps h_sigma sigma_j.  Only the
    diagonal is used: the full inverse is ill-conditioned where parameters
    are nearly degenerate (period and phase over a long baseline), while
    independent draws at these scales lower lnL by D/2 on average for a
    Gaussian posterior.  Returns None when a curvature is not negative."""
    theta = np.asarray(theta, dtype=np.float64)
    d = -np.diag(_hessian(model, theta, h_rel * np.maximum(np.abs(theta), 1.0)))
    if not np.all(d > 0):
        return None
    d = -np.diag(_hessian(model, theta, h_sigma / np.sqrt(d)))
    if not np.all(d > 0):
        return None
    return 1.0 / np.sqrt(d)


def map_init(model, ntemps, nwalkers, rng, betas=None, starts=256, iters=100, spread=1.0, max_repeats=100):
    """Initial positions [T, W, D] around the posterior mode, as an
    alternative to EMPEROR's set_init/test_init (emp/emp.py:617-682): a
    batched L-BFGS search (map_search) from `starts` prior-valid draws, then
    at each temperature W draws of independent normals with the mode's
    conditional posterior scales (laplace_scales; 1e-6 of the prior widths
    when a curvature is not negative) times spread / sqrt(beta_t), beta_t
    floored at 1e-3; draws outside the prior support are drawn again, up to
    `max_repeats` rounds, then replaced by set_init draws.  Returns
    (positions, mode, ln posterior at the mode)."""
    from .sampler import default_ladder
    from .simulation import set_init, test_init
    D = model.ndim__
    x0 = test_init(model, 1, starts, rng)[0]
    x, lp = map_search(model, x0, iters=iters)
    k = int(np.nanargmax(np.where(np.isfinite(lp), lp, -np.inf)))
    mode = x[k]
    sig = laplace_scales(model, mode)
    if sig is None:
        lim = model.limits_free()
        sig = 1e-6 * (lim[:, 1] - lim[:, 0])
    betas = default_ladder(ntemps, D) if betas is None else np.asarray(betas, dtype=np.float64)
    pos = np.empty((ntemps, nwalkers, D))
    for t in range(ntemps):
        sc = spread / np.sqrt(max(float(betas[t]), 1e-3))
        todo = np.arange(nwalkers)
        for _ in range(max_repeats):
            pos[t, todo] = mode + sc * sig * rng.normal(size=(len(todo), D))
            lpr = model.lnprior_batch(pos[t, todo])
            todo = todo[~np.isfinite(lpr)]
            if not len(todo):
                break
        if len(todo):
            fill = set_init(model, 1, len(todo), rng)[0]
            pos[t, todo] = fill
    return pos, mode, float(lp[k])
