"""FP64 torch restatement of the reference lnL (CPU, autograd) for gradient checks.

Follows the literal template formula (support/models/kep0*.model):
f = 2 atan(sqrt((1+e)/(1-e)) tan(E/2)), mu += A (cos(f + w) + e cos w),
err2 = yerr^2 + J^2, lnL = -1/2 sum(r^2/err2 + ln err2) - N/2 ln 2 pi.
E comes from the oracle's solver as a constant starting point and is then
refined by two Newton steps in torch, so autograd sees dE/dM and dE/de of the
converged root (the implicit-function derivative), not of the iterations
that found it.
"""
import numpy as np
import torch

from oracle import emp_oracle as orc


def _kepler_torch(M, e):
    Md = M.detach()
    E0 = torch.as_tensor(orc.kepler_solve(Md.numpy(), np.broadcast_to(e.detach().numpy(), M.shape)))
    E0 = E0 + (Md - torch.remainder(Md, 2 * np.pi))     # same branch as M (the solver returns [0, 2 pi))
    E = E0 + (M - (E0 - e * torch.sin(E0))) / (1 - e * torch.cos(E0))
    E = E - (E - e * torch.sin(E) - M) / (1 - e * torch.cos(E))
    return E


def lnl_torch(model, theta):
    """theta: 1-D float64 tensor of free parameters (requires_grad allowed)."""
    d = model.data
    X = torch.as_tensor(d.X)
    Y = torch.as_tensor(d.Y)
    s2 = torch.as_tensor(d.YERR) ** 2
    flag = torch.as_tensor(d.FLAG - 1)
    full = [None] * model.ndim_full
    for j, c in enumerate(model.C_):
        full[int(c)] = theta[j]
    for a in model.A_:
        full[int(a)] = torch.tensor(float(model.mod_fixed[int(a)]), dtype=torch.float64)
    mu = torch.zeros_like(X)
    p = model.parameterisation
    for k in range(model.kplan__):
        a0, a1, a2, a3, a4 = full[5 * k:5 * k + 5]
        if p in (2, 6, 7):
            per = torch.exp(a0)
        else:
            per = a0
        if p == 2:
            A = a1 ** 2 + a2 ** 2
            phase = torch.atan2(a1, a2) % (2 * np.pi)
        else:
            A = a1
            phase = a2
        if p in (1, 2, 4, 7):
            e = a3 ** 2 + a4 ** 2
            w = torch.atan2(a3, a4)
        else:
            e, w = a3, a4
        M = 2 * np.pi / per * (X - a2) if p in (3, 4) else 2 * np.pi / per * X + phase
        E = _kepler_torch(M, e)
        f = torch.atan(torch.sqrt((1 + e) / (1 - e)) * torch.tan(E / 2)) * 2
        mu = mu + A * (torch.cos(f + w) + e * torch.cos(w))
    off = model.block("Offset").slice.start
    gam = torch.stack(full[off:off + model.nins__])
    mu = mu + gam[flag]
    err2 = s2
    jb = model.block("Jitter")
    if jb is not None:
        J = torch.stack(full[jb.slice.start:jb.slice.start + model.nins__])
        err2 = s2 + J[flag] ** 2
    acc = model.block("Acceleration")
    if acc is not None:
        coefs = full[acc.slice.start:acc.slice.start + len(acc)]
        x = X - X[0]
        v = torch.zeros_like(X)
        for c in coefs:
            v = v * x + c
        mu = mu + v * x
    r = Y - mu
    ma = model.block("MOAV")
    if ma is not None and (model.semantics == "intended" or ma.is_global):
        # MA(q) (moav01.model:7-15): eps_i = r_i - sum_c c_c exp(-|t_i - t_{i-1-c}|/tau_c) eps_{i-1-c}
        # along the time-ordered series (global) or each instrument's series.
        q = ma.moav
        s0 = ma.slice.start
        eps = [None] * len(X)
        groups = [np.arange(len(X))] if ma.is_global else \
            [np.where(d.FLAG == i + 1)[0] for i in range(model.nins__)]
        for gi, idx in enumerate(groups):
            th = full[s0 + 2 * q * gi:s0 + 2 * q * (gi + 1)]
            for pos, i in enumerate(idx):
                e_i = r[i]
                for c in range(q):
                    if pos > c:
                        p = idx[pos - 1 - c]
                        e_i = e_i - th[2 * c] * torch.exp(-torch.abs(X[i] - X[p]) / th[2 * c + 1]) * eps[p]
                eps[i] = e_i
        r = torch.stack(eps)
    # StellarActivity, Sinusoid and MagneticCycle follow MOAV in block order:
    # added after the MA.
    sa = model.block("StellarActivity")
    if sa is not None and len(sa):
        zeta = torch.stack(full[sa.slice.start:sa.slice.start + len(sa)])
        r = r - torch.as_tensor(d.SA) @ zeta
    sn = model.block("Sinusoid")
    if sn is not None:
        per, A, ph = full[sn.slice.start:sn.slice.start + 3]
        r = r - A * torch.cos(2 * np.pi / per * X + ph)
    mc = model.block("MagneticCycle")
    if mc is not None:
        per, A1, A2, ph1, ph2 = full[mc.slice.start:mc.slice.start + 5]
        r = r - (A1 * torch.cos(np.pi / per * X + ph1) + A2 * torch.cos(2 * np.pi / per * X + ph2))
    return -0.5 * torch.sum(r ** 2 / err2 + torch.log(err2)) - 0.5 * np.log(2 * np.pi) * len(X)
