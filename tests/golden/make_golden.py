"""Writes the golden vectors under tests/golden/ with 50-digit mpmath.

Run:  python tests/golden/make_golden.py

Nothing here imports the reference, the oracle or the product: every value
comes from the mathematics of the reference templates evaluated in
arbitrary precision, so that the oracle and the GPU kernels are both checked
against an independent third party.

* kepler_mpmath.npz -- roots of E - e sin E = M on a grid of (e, M),
  solved by bisection in 60-digit arithmetic (E reduced to [0, 2 pi)).
* lnl_mpmath.json   -- per-epoch model and lnL for small data sets under
  every supported Keplerian parameterisation, evaluated with the literal
  template formula f = 2 atan(sqrt((1+e)/(1-e)) tan(E/2)),
  mu += A (cos(f + w) + e cos w)  (support/models/kep0*.model),
  err2 = yerr^2 + J^2 (jitter00.model:3),
  lnL = -1/2 sum(r^2/err2 + ln err2) - N/2 ln 2 pi (likelihoods/00.like:1-5).
* kat.json          -- the reference's recorded known answers (notebook
  outputs), with their file:line sources.
"""
import json
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 60


def kepler_exact(M, e):
    M = mp.mpf(M)
    e = mp.mpf(e)
    two_pi = 2 * mp.pi
    m = M - two_pi * mp.floor(M / two_pi)
    upper = m > mp.pi
    if upper:
        m = two_pi - m
    if e == 0 or m == 0:
        E = m
    else:
        lo, hi = m, min(m + e, mp.pi)
        for _ in range(230):          # bisection: 2^-230 << 1e-60
            mid = (lo + hi) / 2
            if mid - e * mp.sin(mid) - m > 0:
                hi = mid
            else:
                lo = mid
        E = (lo + hi) / 2
    if upper:
        E = two_pi - E
    return E


def make_kepler():
    es = [0.0, 1e-8, 1e-4, 0.01, 0.1, 0.3, 0.5, 0.7, 0.9, 0.95, 0.99, 0.995]
    Ms = list(np.linspace(-4 * np.pi, 4 * np.pi, 61)) + [0.0, 1e-9, 1e-5, 1e-3, 0.05,
                                                        np.pi - 1e-7, np.pi, -np.pi / 3,
                                                        1000.123456, -2500.5, 12566.0]
    M_all, e_all, E_all = [], [], []
    for e in es:
        for M in Ms:
            M_all.append(M)
            e_all.append(e)
            E_all.append(float(kepler_exact(M, e)))
    np.savez(os.path.join(HERE, "kepler_mpmath.npz"), M=np.array(M_all), e=np.array(e_all),
             E=np.array(E_all))
    print("kepler_mpmath.npz:", len(M_all), "roots")


def read_vels(path):
    d = np.loadtxt(path)
    return d[:, 0], d[:, 1], d[:, 2]


def planet_mu(theta5, X, p):
    """Literal template formula in mpmath for parameterisation p."""
    a = [mp.mpf(v) for v in theta5]
    def omega_sc(S, C, thresh):
        ecc = S ** 2 + C ** 2
        if ecc < thresh:
            return ecc, mp.mpf(0)
        w = mp.acos(C / mp.sqrt(ecc))
        if S < 0:
            w = 2 * mp.pi - w
        return ecc, w
    tp = None
    if p == 0:
        per, A, phase, ecc, w = a
    elif p == 1:
        per, A, phase = a[:3]
        ecc, w = omega_sc(a[3], a[4], mp.mpf("1e-6"))
    elif p == 2:
        per = mp.exp(a[0])
        A = a[1] ** 2 + a[2] ** 2
        ecc, w = omega_sc(a[3], a[4], mp.mpf("1e-5"))
        phase = mp.acos(a[2] / mp.sqrt(A))
        if a[1] < 0:
            phase = 2 * mp.pi - phase
    elif p == 3:
        per, A, tp, ecc, w = a
    elif p == 4:
        per, A, tp = a[:3]
        ecc, w = omega_sc(a[3], a[4], mp.mpf("1e-5"))
    elif p == 6:
        per = mp.exp(a[0])
        A, phase, ecc, w = a[1:]
    elif p == 7:
        per = mp.exp(a[0])
        A, phase = a[1:3]
        ecc, w = omega_sc(a[3], a[4], mp.mpf("1e-6"))
    out = []
    freq = 2 * mp.pi / per
    for x in X:
        x = mp.mpf(float(x))
        M = freq * (x - tp) if tp is not None else freq * x + phase
        E = kepler_exact(M, ecc)
        f = mp.atan(mp.sqrt((1 + ecc) / (1 - ecc)) * mp.tan(E / 2)) * 2
        out.append(A * (mp.cos(f + w) + ecc * mp.cos(w)))
    return out


def lnl_case(X, Y, YERR, FLAG, nins, planets, p, offsets, jitters):
    N = len(X)
    mu = [mp.mpf(0)] * N
    for th in planets:
        contrib = planet_mu(th, X, p)
        mu = [m + c for m, c in zip(mu, contrib)]
    chi = mp.mpf(0)
    for i in range(N):
        k = int(FLAG[i]) - 1
        mu[i] += mp.mpf(offsets[k])
        err2 = mp.mpf(float(YERR[i])) ** 2 + mp.mpf(jitters[k]) ** 2
        r = mp.mpf(float(Y[i])) - mu[i]
        chi += r ** 2 / err2 + mp.log(err2)
    lnl = -chi / 2 - mp.mpf(N) / 2 * mp.log(2 * mp.pi)
    return float(lnl), [float(m) for m in mu]


def make_lnl():
    rng = np.random.default_rng(20261202)
    cases = []
    # Data set A: first 24 epochs of 51 Peg (one instrument), mean-subtracted
    # over the full file as the loader does, times relative to the first epoch.
    bjd, rv, erv = read_vels(os.path.join(HERE, "datafiles/51Peg/RV/51peg.vels"))
    rv = rv - rv.mean()
    order = np.argsort(bjd, kind="stable")
    bjd, rv, erv = bjd[order][:24], rv[order][:24], erv[order][:24]
    XA, YA, EA, FA = bjd - bjd.min(), rv, erv, np.ones(24, dtype=int)
    # Data set B: three GJ876 instruments, 8 epochs each, merged in time order.
    parts = []
    for k, name in enumerate(["GJ876_1_LICK.vels", "GJ876_5_APF.vels", "GJ876_6_PFS.vels"]):
        b, r, e = read_vels(os.path.join(HERE, "datafiles/GJ876/RV", name))
        r = r - r.mean()
        parts.append((b[:8], r[:8], e[:8], np.full(8, k + 1)))
    b = np.concatenate([q[0] for q in parts]); r = np.concatenate([q[1] for q in parts])
    e = np.concatenate([q[2] for q in parts]); fl = np.concatenate([q[3] for q in parts])
    o = np.argsort(b, kind="stable")
    XB, YB, EB, FB = b[o] - b.min(), r[o], e[o], fl[o]

    def draw(p, span):
        P = float(rng.uniform(2.0, min(span, 400.0)))
        K = float(rng.uniform(5.0, 60.0))
        ph = float(rng.uniform(0, 2 * np.pi))
        ecc = float(rng.choice([0.0, 5e-7, 0.05, 0.3, 0.6, 0.85]))
        w = float(rng.uniform(0, 2 * np.pi))
        S, C = np.sqrt(ecc) * np.sin(w), np.sqrt(ecc) * np.cos(w)
        tp = float(rng.uniform(-50, 50))
        if p == 0: return [P, K, ph, ecc, w]
        if p == 1: return [P, K, ph, float(S), float(C)]
        if p == 2:
            sk = np.sqrt(K)
            return [float(np.log(P)), float(sk * np.sin(ph)), float(sk * np.cos(ph)), float(S), float(C)]
        if p == 3: return [P, K, tp, ecc, w]
        if p == 4: return [P, K, tp, float(S), float(C)]
        if p == 6: return [float(np.log(P)), K, ph, ecc, w]
        if p == 7: return [float(np.log(P)), K, ph, float(S), float(C)]

    for dname, (X, Y, E, F, nins) in {"51peg24": (XA, YA, EA, FA, 1),
                                      "gj876x3": (XB, YB, EB, FB, 3)}.items():
        span = float(X.max() - X.min())
        for p in [0, 1, 2, 3, 4, 6, 7]:
            for k in [1, 2]:
                planets = [draw(p, span) for _ in range(k)]
                offsets = [float(v) for v in rng.uniform(-10, 10, nins)]
                jitters = [float(v) for v in rng.uniform(0.5, 6, nins)]
                lnl, mu = lnl_case(X, Y, E, F, nins, planets, p, offsets, jitters)
                cases.append(dict(data=dname, parameterisation=p, planets=planets,
                                  offsets=offsets, jitters=jitters, lnl=lnl, mu=mu))
    datasets = {"51peg24": dict(X=XA.tolist(), Y=YA.tolist(), YERR=EA.tolist(), FLAG=FA.tolist(), nins=1),
                "gj876x3": dict(X=XB.tolist(), Y=YB.tolist(), YERR=EB.tolist(), FLAG=FB.tolist(), nins=3)}
    with open(os.path.join(HERE, "lnl_mpmath.json"), "w") as f:
        json.dump(dict(datasets=datasets, cases=cases), f, indent=0)
    print("lnl_mpmath.json:", len(cases), "cases")


def make_kat():
    kat = {
        "note": "lnL at the max-posterior sample recorded in the reference's notebook outputs",
        "51Peg_k0": {"values": [-1308.795, -1308.794, -1308.794],
                     "sources": ["tests/00_mini_test.ipynb:267", "tests/01_51peg_basic.ipynb:323",
                                 "tests/02_51peg_advanced.ipynb:401"]},
        "51Peg_k1_param1": {"values": [-869.480, -869.493],
                            "sources": ["tests/00_mini_test.ipynb:560", "tests/01_51peg_basic.ipynb:616"]},
        "51Peg_k1_param0": {"values": [-869.552], "sources": ["tests/02_51peg_advanced.ipynb:681"]},
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    make_kepler()
    make_lnl()
    make_kat()
