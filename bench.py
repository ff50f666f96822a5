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


def flops_per_theta(model, precision):
    d = model.data
    kep = model.kplan__ * d.ndat
    ma = d.ndat if (model.block("MOAV") is not None and
                    (model.semantics == "intended" or model.block("MOAV").is_global)) else 0
    ops = kep * OPS_KEP[precision] + d.ndat * OPS_EPOCH + ma * OPS_MA
    return 2.0 * ops


def run_gpu(work, precision, steps, warmup, seed, half_launches):
    from astroemperor_amd import PTSampler
    from astroemperor_amd.simulation import test_init
    m = work["model"]
    T, W, _, nsteps = work["setup"]
    p0 = test_init(m, T, W, np.random.default_rng(seed))
    s = PTSampler(W, m.ndim__, m, ntemps=T, seed=seed, precision=precision, adapt_tau=1000, adapt_nu=1)
    cap = warmup + steps + 4
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
    ap.add_argument("--half-launches", type=int, default=40)
    ap.add_argument("--systems", type=int, default=4096, help="batched-systems leg (0 = skip)")
    ap.add_argument("--systems-steps", type=int, default=30)
    ap.add_argument("--systems-warmup", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path shards by independent systems or temperature groups; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    from astroemperor_amd import synth, device_info
    work = synth.c2()
    m = work["model"]
    info = device_info()
    res = {}
    for prec in ("fp64", "mixed"):
        res[prec] = run_gpu(work, prec, args.steps, args.warmup, args.seed, args.half_launches)
        print(f"[bench] {prec}: {res[prec]['sweeps_per_s']:.1f} sweeps/s, "
              f"half-step kernel {res[prec]['ms_halfstep']:.4f} ms", file=sys.stderr, flush=True)
    dl = dlnl_report(work, [res["fp64"]["sampler"], res["mixed"]["sampler"]])
    for r in res.values():
        r["sampler"].close()
    systems = run_systems(args.systems, args.systems_steps, args.systems_warmup, args.seed) if args.systems else None

    mixed, fp64 = res["mixed"], res["fp64"]
    flops = flops_per_theta(m, "mixed") * mixed["evals_per_halfstep"]
    achieved = flops / (mixed["ms_halfstep"] * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_pt_half<mixed>", "achieved": round(achieved, 3),
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": None,
            "convention": "FP64-equivalent flops per lnL eval = 2 x (planet-epochs x ops_kep + epochs x 46"
                          " + MA epochs x 23); ops_kep = 127 (fp64) / 51 (mixed)",
            "flops_per_launch": flops, "launch_ms": mixed["ms_halfstep"],
            "fp64_path": {"achieved": round(flops_per_theta(m, "fp64") * fp64["evals_per_halfstep"]
                                            / (fp64["ms_halfstep"] * 1e-3) / 1e12, 3)}}
    roof["fp64_path"]["frac"] = round(roof["fp64_path"]["achieved"] / PEAK_FP64_TFLOPS, 4)

    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or min(os.cpu_count() or 1, 16)
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
        "device": info,
    }
    if cpu:
        line["fp64_over_cpu"] = round(fp64["sweeps_per_s"] / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
