"""Benchmark: person-days per second of the batched dengue-IBM day loop on one
MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sims 256] [--no-cpu] [--autotune-days D]
    python bench.py --workload plain     # the single 100x100-lattice simulation instead

Workload: the 256-simulation batch (communities log-uniform 400..10,000 squared
to a lattice, 200..2,000 humans per community, per-simulation transmission,
long-distance and seasonal parameters, 1,825 days).  One step = one simulated
day of every simulation of the batch: the demography kernel and the
event-application kernel over every community.  W warm-up days, then K timed
days (W + K = 1,825 by default, the whole configured run).  The first warm-up
days time every candidate launch tuning for D days each and keep the fastest
(results do not depend on the tuning).  Inputs are synthetic and resident in
HBM before timing starts.

With --workload plain: one simulation on the 100x100 lattice (1e7 humans,
Beta(2,5) initial immunity, 3,650 days), the day loop replayed as hipGraphs
of 8 days (10,000 communities: one wave each, too few to fill the GPU); the
roofline comes from a second run of the same days with per-kernel HIP
events.

The JSON line carries:
  roofline       the demography kernel (the dominant launch of the day) against
                 HBM bandwidth.  Algorithmic bytes per launch = 4 B per human
                 (the packed census word, read) + 4 B per mosquito slot (the
                 packed 16-bit word, read and written back) + 32 B per
                 (community, serotype) of counters; `achieved` divides them by
                 the kernel's mean launch time, measured with HIP events on the
                 engine's stream.  `traffic` is the kernel's measured HBM bytes
                 per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated
                 against a plain copy), read from the committed profile summary
                 of this same command;
  step_roofline  the whole day step priced at 7 B per human and per mosquito
                 slot + 32 B per (community, serotype), the per-person-day
                 figure of a census that stores an age, over the measured
                 ms_per_step;
  cpu_baseline   the CPU oracle (OpenMP, and serial beside it) on a bounded
                 sample of the same workload, on this machine's host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "person-days/sec on 1 MI355X (plain, batched); MCMC ESS/hour plain vs surrogate"
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join("profiles", "k_day_demography_batch256.md")


def profiled_traffic() -> float | None:
    """HBM bytes per launch of the demography kernel from the committed
    rocprofv3 summary (FETCH_SIZE and WRITE_SIZE passes, copy-calibrated)."""
    try:
        with open(os.path.join(ROOT, PROFILE)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def _oracle_rate(lib, specs, days: int, threads: int):
    from arbo_gpu_fit_amd.engine import Engine

    eng = Engine(lib, threads=threads)
    eng.init(specs, days)
    t0 = time.perf_counter()
    eng.step(days)
    dt = time.perf_counter() - t0
    eng.close()
    humans = sum(int(s.humans.sum()) for s in specs)
    return humans * days / dt, humans, dt


def cpu_baseline(specs, days: int, threads: int, serial_days: int, what: str = "") -> dict:
    """The CPU oracle on a bounded sample: the first simulations of the batch
    (or `what`) for the first `days` days on `threads` OpenMP threads, and
    simulation 0 alone on one thread for `serial_days` days."""
    from tests.oracle_lib import load_oracle

    lib = load_oracle()
    rate, humans, dt = _oracle_rate(lib, specs, days, threads)
    srate, shumans, sdt = _oracle_rate(lib, specs[:1], serial_days, 1)
    return {
        "value": rate,
        "unit": "person-days/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (g++ -O2, OpenMP {threads} threads): "
                  + (f"{what} ({humans} humans)" if what else
                     f"simulations 0..{len(specs) - 1} of the batch ({humans} humans)")
                  + f", days 0..{days - 1}, {dt:.1f} s",
        "serial": {
            "value": srate,
            "cores": 1,
            "sample": f"simulation 0 ({shumans} humans), days 0..{serial_days - 1}, one thread, {sdt:.1f} s",
        },
    }


def mcmc_bench(args) -> dict:
    """MCMC ESS/hour, plain batched adaptive Metropolis vs surrogate delayed
    acceptance, on the 6-parameter model (1e6 humans, 1,024 communities, 260
    weekly counts, synthetic data at the true parameters):
      1. `train_runs` simulations at prior draws -> weekly incidence;
      2. the surrogate MLP trained on them (Poisson loss, Adam, `epochs`);
      3. a surrogate-only pilot chain gives both samplers the same starting
         states near the posterior;
      4. `iters` iterations of plain adaptive Metropolis and of Langevin
         delayed acceptance, `screen_iters` of screened delayed acceptance
         (cheaper iterations: only proposals the surrogate passes are
         simulated); 64 chains, one simulation per proposal; ESS over the
         iterations after the first quarter, per hour of that sampler's own
         wall clock."""
    import numpy as np

    from arbo_gpu_fit_amd import abi, configs, fit, mcmc
    from arbo_gpu_fit_amd.engine import Engine

    lib = abi.load_hip_library()
    eng = Engine(lib, flags=abi.F_GRAPH)
    data, truth_weekly = configs.mcmc_data(eng)
    log(f"[mcmc] data: {int(truth_weekly.sum())} cases simulated, {int(data.sum())} observed")
    t0 = time.perf_counter()
    if args.load:
        z = np.load(args.load)
        thetas, weekly = z["thetas"], z["weekly"]
    else:
        thetas = fit.prior_sample(args.train_runs, seed=11)
        weekly = fit.simulate_weekly(eng, thetas, seed=12)
    t_data = time.perf_counter() - t0
    if args.save:
        np.savez_compressed(args.save, thetas=thetas, weekly=weekly)
    t0 = time.perf_counter()
    sur, hist = fit.train_surrogate(thetas, weekly, epochs=args.epochs, lib=lib)
    t_train = time.perf_counter() - t0
    log(f"[mcmc] surrogate trained in {t_train:.1f} s: loss {hist[0]:.4g} -> {hist[-1]:.4g}")
    # surrogate-only pilot: delayed acceptance whose second stage is the
    # surrogate itself, so every surrogate move is kept
    g, humans = configs.mcmc_shape()
    cfg = mcmc.McmcConfig(n_chains=args.chains, n_weeks=len(data), priors=configs.mcmc_priors(),
                          adapt_after=100 * mcmc.D, seed=99)
    pilot = mcmc.Mcmc(cfg, g, g, humans, configs.base_params(), data, lib=lib)

    def sur_loglik(th):
        lp, _ = pilot.surrogate_logpost(sur, th)
        return lp - np.array([pilot.log_prior(t) for t in th])

    pilot.run_da_target(sur_loglik, sur, 200, sub_steps=5)
    theta0 = pilot.history()[0][-1]
    pilot.close()
    runs = {}
    for name, s, mode, iters in (("plain", None, mcmc.DA_LANGEVIN, args.iters),
                                 ("surrogate", sur, mcmc.DA_LANGEVIN, args.iters),
                                 ("screened", sur, mcmc.DA_SCREEN, args.screen_iters)):
        _, runs[name] = fit.run_fit(eng, data, iters, chains=args.chains, seed=args.mcmc_seed, sur=s,
                                    sub_steps=args.sub_steps, theta0=theta0, da_mode=mode)
        log(f"[mcmc] {name}: ESS min {runs[name]['ess_min']:.1f}, {runs[name]['ess_per_hour']:.0f}/h, "
            f"R-hat max {runs[name]['rhat_max']:.3f}")
    best = max(("surrogate", "screened"), key=lambda k: runs[k]["ess_per_hour"])
    gain = runs[best]["ess_per_hour"] / runs["plain"]["ess_per_hour"]
    with_training = {k: runs[k]["ess_min"] / ((runs[k]["wall_s"] * 0.75 + t_data + t_train) / 3600)
                     for k in ("surrogate", "screened")}
    out = {
        "metric": "MCMC ESS/hour plain vs surrogate",
        "value": gain,
        "unit": f"ratio of effective samples per hour ({best} delayed acceptance / plain adaptive Metropolis)",
        "n_gpus": 1,
        "plain": runs["plain"],
        "surrogate": runs["surrogate"],
        "screened": runs["screened"],
        "ratios": {k: runs[k]["ess_per_hour"] / runs["plain"]["ess_per_hour"] for k in ("surrogate", "screened")},
        "ess_per_hour_incl_training": with_training,
        "training": {"runs": int(len(thetas)), "simulate_s": t_data, "train_s": t_train, "epochs": args.epochs,
                     "loss_first": float(hist[0]), "loss_last": float(hist[-1])},
        "truth": dict(zip(mcmc.PARAM_NAMES, configs.MCMC_TRUTH.tolist())),
        "data": "synthetic (the model at the true parameters, NB(10) observations)",
        "config": {"workload": f"MCMC fit: {args.chains} chains x 1 simulation per proposal, 1e6 humans in "
                               f"32x32 communities, 1,819 days, 260 weekly counts; {args.iters} iterations plain "
                               f"and Langevin delayed acceptance, {args.screen_iters} screened",
                   "sub_steps": args.sub_steps},
    }
    sur.close()
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed days (batch 1,800; plain 3,600)")
    ap.add_argument("--warmup", type=int, default=None, help="warm-up days (batch 25; plain 50)")
    ap.add_argument("--workload", choices=("batch", "plain"), default="batch")
    ap.add_argument("--sims", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sims", type=int, default=8)
    ap.add_argument("--cpu-days", type=int, default=365)
    ap.add_argument("--cpu-serial-days", type=int, default=365)
    ap.add_argument("--graph", action="store_true", help="replay the day loop as hipGraphs (no per-kernel timing)")
    ap.add_argument("--no-graph", action="store_true",
                    help="plain workload: launch every kernel (per-kernel HIP events) instead of graph replay")
    ap.add_argument("--autotune-days", type=int, default=2,
                    help="days per candidate launch tuning, timed within the warm-up (0: default tuning)")
    ap.add_argument("--mcmc", action="store_true", help="MCMC ESS/hour, plain vs surrogate, instead")
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--iters", type=int, default=160)
    ap.add_argument("--screen-iters", type=int, default=480, help="iterations of the screened delayed acceptance")
    ap.add_argument("--mcmc-seed", type=int, default=1)
    ap.add_argument("--train-runs", type=int, default=50000)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--sub-steps", type=int, default=10)
    ap.add_argument("--save", default="", help="save the surrogate's training set (npz)")
    ap.add_argument("--load", default="", help="load the surrogate's training set (npz)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build "
                                   "(simulations shard by simulation with no per-day traffic)"}))
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.int32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} MiB x3")
        return

    if args.mcmc:
        print(json.dumps(mcmc_bench(args)), flush=True)
        return
    plain = args.workload == "plain"
    if args.steps is None:
        args.steps = 3600 if plain else 1800
    if args.warmup is None:
        args.warmup = 50 if plain else 25
    if plain and not args.no_graph:
        args.graph = True

    from arbo_gpu_fit_amd import abi, configs
    from arbo_gpu_fit_amd.engine import Engine

    lib = abi.load_hip_library()
    t0 = time.perf_counter()
    specs, days_cfg = configs.lattice100() if plain else configs.batch256(args.sims)
    humans = sum(int(s.humans.sum()) for s in specs)
    comms = sum(s.communities for s in specs)
    sero_comms = sum(s.communities * s.params.n_serotypes for s in specs)
    log(f"[bench] generated {len(specs)} simulations, {comms} communities, {humans} humans "
        f"in {time.perf_counter() - t0:.1f} s")
    days = args.warmup + args.steps
    flags = abi.F_GRAPH if args.graph else abi.F_TIMING
    eng = Engine(lib, device=0, flags=flags)
    t0 = time.perf_counter()
    eng.init(specs, days)
    slots = sum(eng.sizes(i)[2] for i in range(len(specs)))
    log(f"[bench] init {time.perf_counter() - t0:.1f} s; {slots} mosquito slots")

    done = 0
    chunk = 100
    n_cand = len(eng.tuning_candidates())
    tune_ms = None
    if args.autotune_days and n_cand * args.autotune_days <= args.warmup:
        chosen, tune_ms = eng.autotune(args.autotune_days)
        done = n_cand * args.autotune_days
        log(f"[bench] autotuned over {n_cand} launch tunings x {args.autotune_days} days: {chosen}, "
            f"{tune_ms:.3f} ms/day")
    while done < args.warmup:
        n = min(chunk, args.warmup - done)
        eng.step(n)
        done += n
    eng.sync()
    eng.reset_timing()
    log(f"[bench] warm-up {args.warmup} days done")

    t_start = time.perf_counter()
    done = 0
    last = t_start
    while done < args.steps:
        n = min(chunk, args.steps - done)
        eng.step(n)
        done += n
        if time.perf_counter() - last > 20.0:
            eng.sync()
            last = time.perf_counter()
            log(f"[bench] {done}/{args.steps} days, {last - t_start:.1f} s")
    eng.sync()
    elapsed = time.perf_counter() - t_start

    person_days = humans * args.steps
    value = person_days / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "person-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": (f"plain single simulation: 100x100 lattice, {humans} humans (1,000 per community), "
                         f"{slots} mosquito slots, Beta(2,5) initial immunity, {days} days" if plain else
                         f"batched plain simulation: {len(specs)} simulations, {comms} communities "
                         f"(400-10,000 per simulation), {humans} humans (200-2,000 per community), "
                         f"{slots} mosquito slots, {days} days"),
            "simulations": len(specs),
            "communities": comms,
            "humans": humans,
            "mosquito_slots": slots,
            "days": days,
            "graph": bool(args.graph),
            "tuning": eng.tuning(),
            "autotuned": tune_ms is not None,
        },
    }
    step_bytes = 7 * (humans + slots) + 32 * sero_comms
    step_gbs = step_bytes / (elapsed / args.steps) / 1e9
    out["step_roofline"] = {"bytes_per_step": step_bytes, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS}
    if plain and args.graph:   # kernel means from a separate pass with per-kernel HIP events, same tuning
        tun = eng.tuning()
        eng.close()
        eng = Engine(lib, device=0, flags=abi.F_TIMING)
        eng.init(specs, days)
        eng.set_tuning(**tun)
        eng.step(args.warmup)
        eng.sync()
        eng.reset_timing()
        eng.step(args.steps)   # the same days as the timed run
        eng.sync()
    if not args.graph or plain:
        traffic = profiled_traffic() if (args.sims == 256 and not plain) else None
        ms_a, n_a = eng.kernel_time(0)
        ms_b, n_b = eng.kernel_time(1)
        bytes_a = 4 * humans + 4 * slots + 32 * sero_comms
        mean_a = ms_a / max(n_a, 1)
        achieved = bytes_a / (mean_a * 1e-3) / 1e9
        out["roofline"] = {
            "bound": "hbm",
            "kernel": "k_day_demography",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bytes_a,
            "mean_launch_ms": mean_a,
            "traffic_source": PROFILE if traffic is not None else None,
        }
        if plain and args.graph:
            out["roofline"]["pass"] = (f"a second run of the same days ({args.warmup}..{days - 1}) with per-kernel "
                                       "HIP events")
        out["kernels_ms"] = {"k_day_demography": ms_a, "k_day_events": ms_b, "launches": n_a}
    eng.close()

    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] CPU oracle baseline on {threads} threads")
        if plain:   # a bounded sample: 730 days on the threads, 60 days on one
            out["cpu_baseline"] = cpu_baseline(specs, 730, threads, 60, what="the 100x100-lattice simulation")
        else:
            out["cpu_baseline"] = cpu_baseline(specs[:args.cpu_sims], args.cpu_days, threads, args.cpu_serial_days)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
