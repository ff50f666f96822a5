"""Pins the torch FP64 gradient reference to the oracle (CPU)."""
import numpy as np
import pytest
import torch

from tests.conftest import make_sim, oracle_for
from tests.torch_reference import lnl_torch


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 6, 7])
def test_torch_reference_equals_oracle(p):
    rng = np.random.default_rng(p)
    sim = make_sim("GJ876", 2, parameterisation=p, acceleration=1)
    m = sim.model
    o = oracle_for(m)
    lim = m.limits_free()
    for _ in range(3):
        th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(0.2, 0.8, size=len(lim))
        if p in (1, 2, 4, 7):
            for k in range(2):
                th[5 * k + 3:5 * k + 5] *= 0.5
        ref = o.my_likelihood(th)
        got = lnl_torch(m, torch.tensor(th)).item()
        assert abs(got - ref) <= 1e-10 * abs(ref)


def test_torch_reference_gradient_vs_fd():
    """autograd of the torch reference against oracle central differences."""
    sim = make_sim("51Peg", 1, parameterisation=1)
    m = sim.model
    o = oracle_for(m)
    th = np.array([4.2307, 55.0, 1.7, 0.1, 0.06, 3.0, 3.0])
    t = torch.tensor(th, requires_grad=True)
    lnl_torch(m, t).backward()
    g = t.grad.numpy()
    for j in range(len(th)):
        h = 1e-6 * max(abs(th[j]), 1.0)
        a, b = th.copy(), th.copy()
        a[j] += h
        b[j] -= h
        fd = (o.my_likelihood(a) - o.my_likelihood(b)) / (2 * h)
        assert abs(fd - g[j]) <= 1e-5 * max(abs(g[j]), 1.0), (j, fd, g[j])


@pytest.mark.parametrize("glob", [False, True])
def test_torch_reference_ma1_equals_oracle(glob):
    """The MA(1) recurrence of the torch reference against the oracle's
    (per-instrument series with the intended semantics, or the global one)."""
    rng = np.random.default_rng(11 + glob)
    sim = make_sim("GJ876", 1, parameterisation=1, moav={"order": 1, "global": glob}, semantics="intended")
    m = sim.model
    o = oracle_for(m)
    lim = m.limits_free()
    for _ in range(3):
        th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(0.2, 0.8, size=len(lim))
        th[3:5] *= 0.5
        ref = o.my_likelihood(th)
        got = lnl_torch(m, torch.tensor(th)).item()
        assert abs(got - ref) <= 1e-10 * abs(ref), (got, ref)


@pytest.mark.parametrize("moav", [None, {"order": 1, "global": True}, {"order": 1, "global": False}])
def test_torch_reference_cosine_terms_equal_oracle(moav):
    """Sinusoid and magnetic-cycle terms (after the MA) in the torch reference."""
    rng = np.random.default_rng(21)
    sim = make_sim("GJ876", 1, parameterisation=1, moav=moav, semantics="intended", sinusoid=1,
                   magnetic_cycle=1)
    m = sim.model
    o = oracle_for(m)
    lim = m.limits_free()
    for _ in range(3):
        th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(0.2, 0.8, size=len(lim))
        th[3:5] *= 0.5
        ref = o.my_likelihood(th)
        got = lnl_torch(m, torch.tensor(th)).item()
        assert abs(got - ref) <= 1e-10 * abs(ref), (got, ref)


@pytest.mark.parametrize("q,glob", [(2, True), (3, False)])
def test_torch_reference_maq_equals_oracle(q, glob):
    """MA(q) of the torch reference against the oracle's literal loop."""
    rng = np.random.default_rng(30 + q)
    sim = make_sim("GJ876", 1, parameterisation=1, moav={"order": q, "global": glob}, semantics="intended")
    m = sim.model
    o = oracle_for(m)
    lim = m.limits_free()
    for _ in range(2):
        th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(0.2, 0.8, size=len(lim))
        th[3:5] *= 0.5
        ref = o.my_likelihood(th)
        got = lnl_torch(m, torch.tensor(th)).item()
        assert abs(got - ref) <= 1e-10 * abs(ref), (got, ref)
