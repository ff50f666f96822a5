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
