"""Evidence from a device run against a closed form.

The offsets-only model (three instruments, jitter fixed, flat priors on
[-200, 200]) has log Z = sum over instruments of a truncated Gaussian
integral.  A 24-temperature geometric ladder from 1 to 1e-7 plus beta = 0,
3000 sweeps: stepping stones land within 0.1 nats, the PCHIP integral within
its every-other-temperature error, the trapezoid (biased on a coarse ladder)
within its own; the device's per-temperature moments equal numpy's on the
recorded history.
"""
import numpy as np
import pytest
from scipy.special import ndtr

from tests.test_gpu_pt import _gaussian_model

pytestmark = pytest.mark.gpu


def _exact_log_z(m):
    d = m.data
    e = d.YERR ** 2 + np.array([1.0, 2.0, 0.5])[d.FLAG - 1] ** 2
    lz = -0.5 * np.sum(np.log(e)) - 0.5 * len(e) * np.log(2 * np.pi)
    for k in (1, 2, 3):
        sel = d.FLAG == k
        A, B, C = np.sum(1 / e[sel]), np.sum(d.Y[sel] / e[sel]), np.sum(d.Y[sel] ** 2 / e[sel])
        mu = B / A
        lz += (np.log(1 / 400) - 0.5 * (C - B * B / A) + 0.5 * np.log(2 * np.pi / A)
               + np.log(ndtr((200 - mu) * np.sqrt(A)) - ndtr((-200 - mu) * np.sqrt(A))))
    return lz


def test_evidence_gaussian_closed_form():
    from astroemperor_amd import PTSampler
    rng = np.random.default_rng(3)
    m, mean, var = _gaussian_model(rng)
    exact = _exact_log_z(m)
    T, W, S, disc = 24, 64, 3000, 500
    betas = np.concatenate([np.geomspace(1.0, 1e-7, T - 1), [0.0]])
    p0 = np.stack([mean + rng.normal(size=(W, 3)) * np.sqrt(var / max(b, 1e-7)) * 0.5 for b in betas])
    s = PTSampler(W, 3, m, ntemps=T, betas=betas, adapt_mode=-1, seed=7)
    s.run_mcmc(np.clip(p0, -199, 199), nsweeps=S)
    ll = s.get_log_like(discard=disc)                       # [T, S', W]
    mdev, lse, v = s.lnl_moments(disc)
    mean_np = ll.reshape(T, -1).mean(1)
    lse_np = np.array([0.0] + [np.log(np.mean(np.exp((betas[t - 1] - betas[t]) * ll[t].ravel())))
                               for t in range(1, T)])
    assert np.allclose(mdev, mean_np, rtol=1e-12, atol=0)
    assert np.allclose(lse, lse_np, rtol=0, atol=1e-12)
    ss, ss_err = s.get_evidence_ss(disc)
    tp, tp_err = s.get_evidence_ti(disc, pchip=True)
    tc, tc_err = s.get_evidence_ti(disc, pchip=False)
    print(f"exact {exact:.3f}: SS {ss:.3f} +- {ss_err:.3f}, TI pchip {tp:.3f} +- {tp_err:.3f}, "
          f"TI trapezoid {tc:.3f} +- {tc_err:.3f}", flush=True)
    assert abs(ss - exact) < 0.1
    assert abs(tp - exact) < max(tp_err, 0.3)
    assert abs(tc - exact) < tc_err
    with pytest.raises(NotImplementedError):
        s.get_evidence_hybrid(disc)
