#!/usr/bin/env python3
"""PT sweeps/s and lnL evals/s of the device engine, FP32-mixed vs FP64.

Workload (default): 3 Keplerians (ln P, K, M0, e, w), 4 instruments with
80/150/400/900 epochs over 6000 d (1530 epochs), per-instrument offset,
jitter and MA(1) (fixed 5 d timescale), 10 temperatures x 256 walkers,
synthetic data from astroemperor_amd.synth.c2.  One step = one PT sweep
(one red/blue stretch step at every temperature + one swap round + ladder
adaptation + chain record), inputs resident in device memory.

    python bench.py [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  `value` is the FP32-mixed path's PT sweeps/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "PT sweeps/s + lnL evals/s, FP32-mixed vs FP64 on 1 GPU at |ΔlnL|<1e-6"

# Algorithmic work per unit (fixed convention, FP64-equivalent lane-ops; an
# FMA counts 1 op, an elementary function 20 FP64 ops, FP32 arithmetic 1/2
# (packed rate) and an FP32 elementary function 4 FP64-op slots):
#   per planet-epoch : 24 arithmetic + 3 FP64 phase ops + 5 elementary
#                      (two sin/cos pairs, one reciprocal)
#   per epoch        : 6 arithmetic + 2 elementary (reciprocal, log)
#   per MA epoch     : 3 arithmetic + 1 elementary (exp)
# FP64 path: everything at FP64 rates.  Mixed path: the 24 arithmetic and 4
# of the 5 elementary ops of the Kepler term at FP32 rates, the rest FP64.
OPS_KEP = {"fp64": 24 + 3 + 20 * 5, "mixed": 24 * 0.5 + 3 + 4 * 4 + 20 * 1}
OPS_EPOCH = 6 + 20 * 2
OPS_MA = 3 + 20 * 1
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector, FMA = 2 flops
# HBM bytes per launch of k_pt_half<mixed, W waves per walker> on the default
# workload: rocprofv3 2.00 x FETCH_SIZE + 1.00 x WRITE_SIZE in KiB per
# dispatch, the factors from the same counters on a 256 MiB copy (the summary
# is profiles/k_pt_half_k_sys_sweep_hbm_traffic.md).
TRAFFIC_KIB = {1: (507.3, 94.7), 2: (511.6, 94.7), 4: (527.1, 95.6), 8: (591.3, 107.5)}


def traffic_bytes(wpt):
    f, w = TRAFFIC_KIB[wpt]
    return int((2.0 * f + 1.0 * w) * 1024)


def flops_per_theta(model, precision):
    d = model.data
    kep = model.kplan__ * d.ndat
    ma = d.ndat if (model.block("MOAV") is not None and
                    (model.semantics == "intended" or model.block("MOAV").is_global)) else 0
    ops = kep * OPS_KEP[precision] + d.ndat * OPS_EPOCH + ma * OPS_MA
    return 2.0 * ops


def run_gpu(work, precision, steps, warmup, seed, half_launches):
    half_launches += half_launches & 1          # whole sweeps: two likelihood launches each
    from astroemperor_amd import PTSampler
    from astroemperor_amd.simulation import test_init
    m = work["model"]
    T, W, _, nsteps = work["setup"]
    p0 = test_init(m, T, W, np.random.default_rng(seed))
    s = PTSampler(W, m.ndim__, m, ntemps=T, seed=seed, precision=precision, adapt_tau=1000, adapt_nu=1)
    cap = warmup + steps + (half_launches + 1) // 2 + 4
    s.run_mcmc(p0, nsweeps=max(warmup, 1), capacity=cap)
    wpt, _ = s.autotune(10)
    n0 = s.n_lnl_evaluated
    t0 = time.perf_counter()
    s.run_mcmc(None, nsweeps=steps)
    t1 = time.perf_counter()
    n1 = s.n_lnl_evaluated
    ms_sweep_events, _ = s.last_timing()
    wall = t1 - t0
    ms_half = s.time_halfstep(half_launches)
    n2 = s.n_lnl_evaluated
    return dict(sampler=s, seconds=wall, sweeps_per_s=steps / wall, ms_per_sweep=1e3 * wall / steps,
                ms_per_sweep_events=ms_sweep_events,
                proposals_per_s=steps * T * W * nsteps / wall,
                lnl_evaluated_per_s=(n1 - n0) / wall,
                accept_in_support=(n1 - n0) / (steps * T * W),
                ms_halfstep=ms_half, evals_per_halfstep=(n2 - n1) / half_launches,
                T=T, W=W, waves_per_walker=wpt)


def dlnl_report(work, samplers):
    """|lnL_mixed - lnL_fp64| on positions the samplers visited."""
    m = work["model"]
    thetas = []
    for s in samplers:
        ch = s.get_chain()                              # [T, S, W, D]
        S = ch.shape[1]
        pick = np.linspace(0, S - 1, min(S, 16)).astype(int)
        thetas.append(ch[:, pick].reshape(-1, ch.shape[-1]))
    th = np.concatenate(thetas)
    a = m.lnl_batch(th, "fp64")
    b = m.lnl_batch(th, "mixed")
    fin = np.isfinite(a) & np.isfinite(b)
    d = np.abs(a[fin] - b[fin])
    return dict(n=int(fin.sum()), max=float(d.max()), p99=float(np.percentile(d, 99)),
                mismatched_finiteness=int((np.isfinite(a) != np.isfinite(b)).sum()))


def run_systems(n_systems, steps, warmup, seed, precisions=("mixed", "fp64"), trial_sweeps=2):
    """Batched uneven systems (synth.c3): every system's own 8 x 64 PT
    ensemble advanced by one persistent kernel launch per sweep, replayed
    from a hipGraph, launch shape from the autotuner."""
    from astroemperor_amd import SystemsPT, synth
    t0 = time.perf_counter()
    work = synth.c3(n_systems)
    models = work["models"]
    T, W, _, nsteps = work["setup"]
    rng = np.random.default_rng(seed)
    p0 = [synth.initial_positions(m, T, W, rng) for m in models]
    build_s = time.perf_counter() - t0
    N = np.array([m.data.ndat for m in models])
    K = np.array([m.kplan__ for m in models])
    flops_sys = np.array([flops_per_theta(m, "mixed") for m in models])
    out = {"workload": f"{len(models)} independent systems, epochs lognormal(median 200) clipped to "
                       "[50, 2000], 1-4 Keplerians, 1-3 instruments, each 8 temps x 64 walkers",
           "n_systems": len(models), "epochs_total": int(N.sum()), "epochs_max": int(N.max()),
           "planet_epochs_total": int((N * K).sum()), "ntemps": T, "nwalkers": W, "host_build_s": round(build_s, 1)}
    states = {}
    for prec in precisions:
        t1 = time.perf_counter()
        s = SystemsPT(models, T, W, p0, seed=seed, precision=prec, adapt_tau=1000, adapt_nu=1)
        create_s = time.perf_counter() - t1
        s.run(1)
        tune = s.autotune(trial_sweeps=trial_sweeps)
        s.run(max(warmup, 1))
        n0 = s.n_lnl_evaluated
        ta = time.perf_counter()
        s.run(steps, use_graph=True)
        tb = time.perf_counter()
        n1 = s.n_lnl_evaluated
        wall = tb - ta
        ev_ms = s.last_ms_per_sweep()
        props = steps * T * W * nsteps * len(models)
        frac_eval = (n1 - n0) / props
        # algorithmic FP64-equivalent flops of one sweep: every evaluated lnL
        # at its system's cost (evaluation fraction taken as uniform)
        flops = float(flops_sys.sum()) * T * W * nsteps * frac_eval
        achieved = flops / (ev_ms * 1e-3) / 1e12
        out[prec] = {"sweeps_per_s": round(steps / wall, 3), "ms_per_sweep": round(1e3 * wall / steps, 4),
                     "ms_per_sweep_events": round(ev_ms, 4),
                     "lnl_evals_per_s": round(props / wall, 1), "lnl_evaluated_per_s": round((n1 - n0) / wall, 1),
                     "tuning": tune, "create_s": round(create_s, 2),
                     "roofline": {"bound": "valu", "kernel": f"k_sys_sweep<{prec}>", "achieved": round(achieved, 3),
                                  "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                  "frac": round(achieved / PEAK_FP64_TFLOPS, 4)}}
        print(f"[bench] systems {prec}: {steps / wall:.2f} sweeps/s ({props / wall / 1e6:.1f} M lnL/s), "
              f"tuning {tune}", file=sys.stderr, flush=True)
        states[prec] = s
    if "mixed" in states and "fp64" in states:
        out["mixed_over_fp64"] = round(out["mixed"]["sweeps_per_s"] / out["fp64"]["sweeps_per_s"], 3)
        # |dlnL| on the states the mixed run visited, a spread of systems
        s = states["mixed"]
        dmax, n = 0.0, 0
        for i in np.linspace(0, len(models) - 1, 24).astype(int):
            x, _, _ = s.state(int(i))
            th = x.reshape(-1, x.shape[-1])
            a = models[i].lnl_batch(th, "fp64")
            b = models[i].lnl_batch(th, "mixed")
            fin = np.isfinite(a) & np.isfinite(b)
            if fin.any():
                dmax = max(dmax, float(np.abs(a[fin] - b[fin]).max()))
                n += int(fin.sum())
            models[i]._device = None
        out["dlnl_mixed_vs_fp64"] = {"n": n, "max": dmax}
    for s in states.values():
        s.close()
    return out


def _state_dlnl(model, coords):
    """|lnL_mixed - lnL_fp64| over a set of positions [..., D]."""
    th = coords.reshape(-1, coords.shape[-1])
    a = model.lnl_batch(th, "fp64")
    b = model.lnl_batch(th, "mixed")
    fin = np.isfinite(a) & np.isfinite(b)
    d = np.abs(a[fin] - b[fin])
    return {"n": int(fin.sum()), "max": float(d.max()) if d.size else None,
            "p99": float(np.percentile(d, 99)) if d.size else None}


def run_c1(seed, cpu_cores, with_cpu):
    """1 Keplerian, 1 instrument, 256 epochs; 5 temps x 100 walkers x 500
    sweeps, the full run on the GPU (FP64 and mixed) and on the host cores."""
    from astroemperor_amd import PTSampler, synth
    from astroemperor_amd.simulation import test_init
    work = synth.c1()
    m = work["model"]
    T, W, S, _ = work["setup"]
    p0 = test_init(m, T, W, np.random.default_rng(seed))
    out = {"workload": f"1 Keplerian, 1 instrument, {m.data.ndat} epochs, {T} temps x {W} walkers x {S} sweeps"}
    last = None
    for prec in ("fp64", "mixed"):
        s = PTSampler(W, m.ndim__, m, ntemps=T, seed=seed, precision=prec)
        s.run_mcmc(p0, nsweeps=1, capacity=S + 1)
        t0 = time.perf_counter()
        s.run_mcmc(None, nsweeps=S - 1)
        wall = time.perf_counter() - t0
        out[prec] = {"sweeps_per_s": round((S - 1) / wall, 1), "lnl_evals_per_s": round((S - 1) * T * W / wall, 1),
                     "seconds": round(wall, 3)}
        last = s.get_last_sample().coords
        s.close()
    out["dlnl_mixed_vs_fp64"] = _state_dlnl(m, last)
    if with_cpu:
        from oracle import cpu_reference
        from astroemperor_amd.sampler import default_ladder
        d = m.data
        r = cpu_reference.run(m.spec(), d.X, d.Y, d.YERR, d.FLAG, p0, default_ladder(T, m.ndim__), S,
                              ncores=cpu_cores, seed=seed)
        out["cpu"] = {"sweeps_per_s": round(r["sweeps_per_s"], 2), "lnl_evals_per_s": round(r["lnl_evals_per_s"], 1),
                      "cores": r["cores"], "seconds": round(r["seconds"], 2), "kind": "port",
                      "sample": f"the full run ({S} sweeps)"}
        out["fp64_over_cpu"] = round(out["fp64"]["sweeps_per_s"] / r["sweeps_per_s"], 1)
    return out


def run_c4(seed, batch=65536, reps=5, map_starts=1024):
    """Adjoint gradients: 5 Keplerians, N = 2000, 2 instruments, a batch of
    parameter vectors drawn from the prior; throughput of lnL + gradient,
    a central finite-difference check and the batched L-BFGS MAP search."""
    import torch
    from astroemperor_amd import synth
    from astroemperor_amd.autograd import map_search
    work = synth.c4()
    m = work["model"]
    rng = np.random.default_rng(seed)
    th = synth.initial_positions(m, 1, batch, rng)[0]
    tt = torch.tensor(th, device="cuda")
    m.lnl_grad_batch(tt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        m.lnl_grad_batch(tt)
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        m.lnl_batch(tt, "fp64")
    torch.cuda.synchronize()
    tl = (time.perf_counter() - t0) / reps
    # FD check (FP64, h = 1e-6 max(|theta|, 1)) on 8 vectors.  The central
    # difference carries a rounding error ~ 2^-52 |lnL| / h, large for draws
    # far from the data; entries are judged where that error is below 1e-7 |fd|.
    # Richardson's extrapolation of the h and h/2 differences removes their
    # O(h^2) truncation error, which h = 1e-6 |theta| leaves near 1e-5 on the
    # periods (the phase derivative grows as t / P^2).
    sub = th[:8]
    f0, g = m.lnl_grad_batch(sub)
    errs, errs_r, noisy = [], [], 0

    def cdiff(i, j, h):
        xp, xm = sub[i].copy(), sub[i].copy()
        xp[j] += h
        xm[j] -= h
        return (m.lnl_batch(xp[None])[0] - m.lnl_batch(xm[None])[0]) / (2 * h)

    for i in range(len(sub)):
        for j in range(sub.shape[1]):
            h = 1e-6 * max(abs(sub[i, j]), 1.0)
            fd = cdiff(i, j, h)
            if 2.0 ** -52 * abs(f0[i]) / h > 1e-7 * abs(fd):
                noisy += 1
                continue
            fr = (4.0 * cdiff(i, j, 0.5 * h) - fd) / 3.0
            errs.append(abs(g[i, j] - fd) / abs(fd))
            errs_r.append(abs(g[i, j] - fr) / abs(fr))
    t0 = time.perf_counter()
    _, fbest = map_search(m, th[:map_starts], iters=50)
    tmap = time.perf_counter() - t0
    return {"workload": f"5 Keplerians, 2 instruments, {m.data.ndat} epochs, batch {batch} from the prior",
            "lnl_grad_evals_per_s": round(batch / tg, 1), "lnl_evals_per_s": round(batch / tl, 1),
            "grad_over_lnl_time": round(tg / tl, 2),
            "fd_check": {"entries": len(errs), "skipped_fd_rounding": noisy, "max_rel": float(np.max(errs)),
                         "median_rel": float(np.median(errs)), "richardson_max_rel": float(np.max(errs_r)),
                         "richardson_median_rel": float(np.median(errs_r))},
            "map_search": {"starts": map_starts, "iters": 50, "seconds": round(tmap, 2),
                           "best_lnpost": float(np.max(fbest))}}


def run_c5(seed, steps, warmup):
    """Large-scale mixed: 7 Keplerians, N = 10000 over 5 instruments with
    MA(1), 16 temps x 1024 walkers, gradient sweeps mixed 1:4 with stretch
    sweeps; a shortened timed run.  Each sweep kind is also timed alone."""
    from astroemperor_amd import GradientMove, PTSampler, StretchMove, synth
    from astroemperor_amd.simulation import test_init
    work = synth.c5()
    m = work["model"]
    T, W, _, _ = work["setup"]
    p0 = test_init(m, T, W, np.random.default_rng(seed))
    mix = [(StretchMove(), 4.0), (GradientMove(), 1.0)]
    out = {"workload": f"7 Keplerians, 5 instruments, {m.data.ndat} epochs, MA(1), {T} temps x {W} walkers, "
                       f"{steps} timed sweeps, gradient (Langevin) sweeps mixed 1:4 with stretch sweeps",
           "gradient_step": GradientMove().scale(m.ndim__)}
    last = None
    for prec in ("mixed", "fp64"):
        extra = 2 * 6 + 3          # the per-move timings and the timed half-steps
        s = PTSampler(W, m.ndim__, m, ntemps=T, seed=seed, precision=prec, record=False, moves=mix)
        s.run_mcmc(p0, nsweeps=max(warmup, 1), capacity=warmup + steps + extra + 1)
        wpt, _ = s.autotune(3)
        n0 = s.n_lnl_evaluated
        t0 = time.perf_counter()
        s.run_mcmc(None, nsweeps=steps)
        wall = time.perf_counter() - t0
        n1 = s.n_lnl_evaluated
        h = s.time_halfstep(6)
        n2 = s.n_lnl_evaluated
        flops = flops_per_theta(m, prec) * (n2 - n1) / 6
        ach = flops / (h * 1e-3) / 1e12
        kinds = {}
        for name, mv in (("stretch", [(StretchMove(), 1.0)]), ("gradient", [(GradientMove(), 1.0)])):
            s.set_moves(mv)
            s.run_mcmc(None, nsweeps=1)            # captures this kind's graph
            a1 = s._get(6, np.int64).sum()
            s.run_mcmc(None, nsweeps=5)
            kinds[name] = {"ms_per_sweep": round(s.last_timing()[0], 3),
                           "acceptance": round(float(s._get(6, np.int64).sum() - a1) / (5 * T * W), 4)}
        s.set_moves(mix)
        out[prec] = {"sweeps_per_s": round(steps / wall, 2), "lnl_evals_per_s": round(steps * T * W / wall, 1),
                     "lnl_evaluated_per_s": round((n1 - n0) / wall, 1), "waves_per_walker": wpt,
                     "by_move": kinds,
                     "roofline": {"bound": "valu", "kernel": f"k_pt_half<{prec}>", "achieved": round(ach, 3),
                                  "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4)}}
        if prec == "mixed":
            last = s.get_last_sample().coords
        s.close()
    out["mixed_over_fp64"] = round(out["mixed"]["sweeps_per_s"] / out["fp64"]["sweeps_per_s"], 3)
    out["dlnl_mixed_vs_fp64"] = _state_dlnl(m, last)
    return out


def run_lnl(n, reps, seed):
    """Batched lnL alone (emp_lnl_batch, device pointers) on n prior draws of
    the default workload's model, reps times per precision: the subject of
    the instruction-mix counter runs."""
    import torch
    from astroemperor_amd import synth
    m = synth.c2()["model"]
    th = torch.tensor(synth.initial_positions(m, 1, n, np.random.default_rng(seed))[0], device="cuda")
    out = {}
    for prec in ("mixed", "fp64"):
        m.lnl_batch(th, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            m.lnl_batch(th, prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[prec] = {"batch": n, "ms_per_batch": round(1e3 * dt, 4), "lnl_evals_per_s": round(n / dt, 1)}
    return out


def run_copy(mib, reps):
    """A plain device copy of known size (torch's copy kernel), the unit the
    HBM counters (FETCH_SIZE, WRITE_SIZE) are calibrated against."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"copy_bytes_each_way": n * 8, "reps": reps}


def run_cpu(work, sweeps, cores, seed):
    from oracle import cpu_reference
    from astroemperor_amd.simulation import test_init
    from astroemperor_amd.sampler import default_ladder
    m = work["model"]
    T, W, _, _ = work["setup"]
    p0 = test_init(m, T, W, np.random.default_rng(seed))
    betas = default_ladder(T, m.ndim__)
    d = m.data
    r = cpu_reference.run(m.spec(), d.X, d.Y, d.YERR, d.FLAG, p0, betas, sweeps, ncores=cores, seed=seed)
    r.pop("final")
    return r


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed PT sweeps")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-sweeps", type=int, default=20)
    ap.add_argument("--cpu-cores", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--half-launches", type=int, default=40,
                    help="likelihood-kernel launches timed with events (an even number: whole sweeps)")
    ap.add_argument("--systems", type=int, default=4096, help="batched-systems leg (0 = skip)")
    ap.add_argument("--systems-steps", type=int, default=30)
    ap.add_argument("--systems-warmup", type=int, default=5)
    ap.add_argument("--no-configs", action="store_true", help="skip the other workloads (c1, c4, c5)")
    ap.add_argument("--c5-steps", type=int, default=20)
    ap.add_argument("--only", choices=["all", "c2", "systems", "configs", "c5", "lnl", "copy"], default="all",
                    help="run one leg only (profiling); 'lnl': batched lnL alone on 20480 prior draws; "
                         "'copy': the HBM-counter calibration copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by independent systems or by groups of temperatures)",
              file=sys.stderr)
        sys.exit(2)

    from astroemperor_amd import synth, device_info
    info = device_info()
    if args.only == "copy":
        print(json.dumps({"calibration": run_copy(256, 10), "device": info}), flush=True)
        return
    if args.only == "lnl":
        print(json.dumps({"lnl_batch": run_lnl(20480, 8, args.seed), "device": info}), flush=True)
        return
    if args.only == "systems":
        print(json.dumps({"batched_systems": run_systems(args.systems, args.systems_steps, args.systems_warmup,
                                                         args.seed), "device": info}), flush=True)
        return
    if args.only == "c5":
        print(json.dumps({"c5": run_c5(args.seed, args.c5_steps, 3), "device": info}), flush=True)
        return
    cores = args.cpu_cores or min(os.cpu_count() or 1, 16)
    if args.only == "configs":
        print(json.dumps({"c1": run_c1(args.seed, cores, not args.no_cpu), "c4": run_c4(args.seed),
                          "c5": run_c5(args.seed, args.c5_steps, 3), "device": info}), flush=True)
        return

    work = synth.c2()
    m = work["model"]
    res = {}
    for prec in ("fp64", "mixed"):
        res[prec] = run_gpu(work, prec, args.steps, args.warmup, args.seed, args.half_launches)
        print(f"[bench] {prec}: {res[prec]['sweeps_per_s']:.1f} sweeps/s, "
              f"half-step kernel {res[prec]['ms_halfstep']:.4f} ms", file=sys.stderr, flush=True)
    dl = dlnl_report(work, [res["fp64"]["sampler"], res["mixed"]["sampler"]])
    for r in res.values():
        r["sampler"].close()
    full = args.only == "all"
    systems = run_systems(args.systems, args.systems_steps, args.systems_warmup, args.seed) \
        if full and args.systems else None
    configs = None
    if full and not args.no_configs:
        configs = {"c1": run_c1(args.seed, cores, not args.no_cpu)}
        print(f"[bench] c1: {configs['c1']}", file=sys.stderr, flush=True)
        configs["c4"] = run_c4(args.seed)
        print(f"[bench] c4: {configs['c4']}", file=sys.stderr, flush=True)
        configs["c5"] = run_c5(args.seed, args.c5_steps, 3)
        print(f"[bench] c5: {configs['c5']}", file=sys.stderr, flush=True)

    mixed, fp64 = res["mixed"], res["fp64"]
    flops = flops_per_theta(m, "mixed") * mixed["evals_per_halfstep"]
    achieved = flops / (mixed["ms_halfstep"] * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_pt_half<mixed>", "achieved": round(achieved, 3),
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": traffic_bytes(mixed["waves_per_walker"]),
            "convention": "FP64-equivalent flops per lnL eval = 2 x (planet-epochs x ops_kep + epochs x 46"
                          " + MA epochs x 23); ops_kep = 127 (fp64) / 51 (mixed)",
            "flops_per_launch": flops, "launch_ms": mixed["ms_halfstep"],
            "fp64_path": {"achieved": round(flops_per_theta(m, "fp64") * fp64["evals_per_halfstep"]
                                            / (fp64["ms_halfstep"] * 1e-3) / 1e12, 3)}}
    roof["fp64_path"]["frac"] = round(roof["fp64_path"]["achieved"] / PEAK_FP64_TFLOPS, 4)

    cpu = None
    if not args.no_cpu:
        r = run_cpu(work, args.cpu_sweeps, cores, args.seed)
        cpu = {"value": round(r["sweeps_per_s"], 4), "unit": "PT sweeps/s", "cores": r["cores"],
               "kind": "port",
               "sample": f"{args.cpu_sweeps} sweeps of the same workload (10 temps x 256 walkers, "
                         f"{r['sweeps'] * 2560} proposals) on a {r['cores']}-process pool",
               "lnl_evals_per_s": round(r["lnl_evals_per_s"], 1), "cpu_model": r["cpu_model"],
               "seconds": round(r["seconds"], 2),
               "label": "reconstructed reference CPU path (numpy + oracle C Kepler + multiprocessing)"}

    T, W = mixed["T"], mixed["W"]
    line = {
        "metric": METRIC, "value": round(mixed["sweeps_per_s"], 2), "unit": "PT sweeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_sweep"], 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64 (mixed)", "data": "synthetic",
        "config": {"workload": "3 Keplerians, 4 instruments (80/150/400/900 epochs), MA(1), "
                               f"{T} temps x {W} walkers, 1 stretch step per sweep",
                   "epochs": int(m.data.ndat), "planets": int(m.kplan__), "instruments": int(m.nins__),
                   "ndim": int(m.ndim__), "ntemps": T, "nwalkers": W},
        "lnl_evals_per_s": round(mixed["proposals_per_s"], 1),
        "lnl_evaluated_per_s": round(mixed["lnl_evaluated_per_s"], 1),
        "waves_per_walker": mixed["waves_per_walker"],
        "fp64": {"value": round(fp64["sweeps_per_s"], 2), "lnl_evals_per_s": round(fp64["proposals_per_s"], 1),
                 "ms_per_step": round(fp64["ms_per_sweep"], 5), "halfstep_ms": round(fp64["ms_halfstep"], 5),
                 "waves_per_walker": fp64["waves_per_walker"]},
        "mixed_over_fp64": round(mixed["sweeps_per_s"] / fp64["sweeps_per_s"], 3),
        "dlnl_mixed_vs_fp64": dl,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_systems": systems,
        "configs": configs,
        "device": info,
    }
    if cpu:
        line["fp64_over_cpu"] = round(fp64["sweeps_per_s"] / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
