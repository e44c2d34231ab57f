"""Benchmark: person-days per second of the dengue-IBM day loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sims 256] [--no-cpu] [--no-plain]
    python bench.py --workload plain     # only the single 100x100-lattice simulation
    python bench.py --mcmc               # MCMC ESS/hour, plain vs surrogate delayed acceptance
    python bench.py --mcmc-noise         # the fit's likelihood noise under common seeds
    python bench.py --surrogate          # surrogate training throughput

Headline workload: the 256-simulation batch (communities log-uniform
400..10,000 squared to a lattice, 200..2,000 humans per community,
per-simulation transmission, long-distance and seasonal parameters, 1,825 days
configured).  One step = one simulated day of every simulation of the batch:
the demography, reduce, long-distance and event-application kernels over
every community.  W warm-up days, then K timed days (W + K <= 1,825; by
default 25 and 1,800).  With --autotune-days D and at least 10 D warm-up days
the warm-up first times every launch tuning for D days each and keeps the
fastest (results do not depend on the tuning).  Inputs are synthetic and
resident in HBM before timing starts; `value` is person-days per second over
the timed days, whole job.

The same line carries a `plain` record: the single simulation of the
100x100 lattice (1e7 humans, Beta(2,5) initial immunity, 3,650 days), its day
loop replayed as hipGraphs of 8 days, timed in the same run (50 warm-up days,
3,600 timed), with its own roofline and CPU baseline.

roofline (per workload): the demography kernel, the dominant launch of the
day, against HBM bandwidth.
  * Algorithmic bytes of a launch are the census and per-community bytes the
    kernel's algorithm must move, counted from what it scanned
    (arbo_scan_counts): 16 B per human census vector read (4 packed words;
    a community's humans are skipped, exactly, on days before its earliest
    death or state change) + 32 B per mosquito vector (8 packed slots, read
    and written back; slots above the highest alive one are skipped) + 4 B
    per 32 mosquito slots of alive bitmap written + per community 64 B of
    descriptor, 32 B of summary read and written, 96 B of partials and 40 B
    per CSR entry (its row read, its inbox counts written).  `achieved`
    divides them by the kernel's mean launch time, measured with HIP events
    on the engine's stream.
  * `traffic` is the kernel's HBM bytes per launch from rocprofv3's
    FETCH_SIZE and WRITE_SIZE passes over this same command (calibrated
    against a plain copy), read from the committed profile; `traffic_frac`
    is that rate over the peak.  `full_scan_bytes` prices a scan of every
    member every day (4 B per human, 4 B per slot, 32 B per community and
    serotype) for comparison.
cpu_baseline: the CPU oracle built -O3 -march=native on this host, OpenMP
(and serial beside it), resumed from the GPU's census at the first timed
day and run over the timed days (at most 60 for the batch, 600 for the
plain simulation) for a sample of simulations (64 of the batch, 128 when
at most 60 days are timed): about 10-15 s of CPU work.
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
PROFILES = {"batch": os.path.join("profiles", "k_day_demography_batch256.md"),
            "plain": os.path.join("profiles", "k_day_demography_plain100.md")}


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(workload: str, warmup: int, steps: int) -> float | None:
    """HBM bytes per launch of the demography kernel from the committed
    rocprofv3 summary of this workload and day window (FETCH_SIZE and
    WRITE_SIZE passes, copy-calibrated): a line
    `traffic workload=<w> warmup=<W> steps=<K> bytes_per_launch=<B>`."""
    try:
        with open(os.path.join(ROOT, PROFILES[workload])) as f:
            for line in f:
                if not line.startswith("traffic "):
                    continue
                kv = dict(x.split("=", 1) for x in line.split()[1:])
                if kv.get("workload") == workload and int(kv["warmup"]) == warmup and int(kv["steps"]) == steps:
                    return float(kv["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        pass
    return None


def row_lengths(spec) -> list[int]:
    """CSR entries of every community of a simulation: its lattice
    neighbours (Moore or von Neumann) plus itself when it is a dispersal
    target or keeps a staying fraction."""
    p = spec.params
    own = 1 if (p.include_self or p.no_distance > 0.0) else 0
    out = []
    for y in range(spec.grid_h):
        ny = 1 + (y > 0) + (y < spec.grid_h - 1)
        for x in range(spec.grid_w):
            nx = 1 + (x > 0) + (x < spec.grid_w - 1)
            n = nx * ny - 1 if p.neighbourhood == 8 else (nx - 1) + (ny - 1)
            out.append(max(1, n + own))
    return out


def host_cpu() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(specs, states, day: int, days: int, threads: int, serial_days: int, what: str) -> dict:
    """The CPU oracle (-O3 -march=native, built here) resumed at `day` from
    the GPU's census of the simulations `specs` (states[i]) and run `days`
    days on `threads` OpenMP threads; simulation 0 alone on one thread for
    `serial_days` days."""
    from arbo_gpu_fit_amd.engine import Engine
    from tests.oracle_lib import load_native_oracle

    lib, flags = load_native_oracle()
    T = day + max(days, serial_days)

    def timed(sp, st, n, thr):
        eng = Engine(lib, threads=thr)
        eng.init(sp, T)
        eng.set_day(day)
        for i, s in enumerate(st):
            eng.write_state(i, s)
        t0 = time.perf_counter()
        eng.step(n)
        dt = time.perf_counter() - t0
        eng.close()
        humans = sum(int(s.humans.sum()) for s in sp)
        return humans * n / dt, humans, dt

    rate, humans, dt = timed(specs, states, days, threads)
    srate, shumans, sdt = timed(specs[:1], states[:1], serial_days, 1)
    model = host_cpu()
    return {
        "value": rate,
        "unit": "person-days/s",
        "cores": threads,
        "kind": "port",
        "cpu": model,
        "build": flags,
        "sample": f"CPU oracle ({flags}, OpenMP {threads} threads, {model}): {what} ({humans} humans) resumed "
                  f"from the GPU's census at day {day}, days {day}..{day + days - 1}, {dt:.1f} s",
        "serial": {
            "value": srate,
            "cores": 1,
            "sample": f"the first of them ({shumans} humans), days {day}..{day + serial_days - 1}, one thread, "
                      f"{sdt:.1f} s",
        },
    }


def sim_bench(workload: str, args, lib) -> dict:
    """Person-days/s of one workload ("batch" or "plain") with its roofline
    and CPU baseline."""
    from arbo_gpu_fit_amd import abi, configs
    from arbo_gpu_fit_amd.engine import Engine

    plain = workload == "plain"
    steps = args.plain_steps if plain else args.steps
    warmup = args.plain_warmup if plain else args.warmup
    t0 = time.perf_counter()
    specs, days_cfg = configs.lattice100() if plain else configs.batch256(args.sims)
    humans = sum(int(s.humans.sum()) for s in specs)
    comms = sum(s.communities for s in specs)
    sero_comms = sum(s.communities * s.params.n_serotypes for s in specs)
    log(f"[bench] {workload}: {len(specs)} simulations, {comms} communities, {humans} humans "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    days = warmup + steps
    if days > days_cfg:
        raise SystemExit(f"{workload}: warm-up + steps = {days} exceeds the configured {days_cfg} days")
    graph = (plain and not args.no_graph) or args.graph
    eng = Engine(lib, device=0, flags=abi.F_GRAPH if graph else abi.F_TIMING)
    t0 = time.perf_counter()
    eng.init(specs, days)
    slots = sum(eng.sizes(i)[2] for i in range(len(specs)))
    log(f"[bench] {workload}: init {time.perf_counter() - t0:.1f} s; {slots} mosquito slots")

    done = 0
    n_cand = len(eng.tuning_candidates())
    tune_ms = None
    if args.autotune_days and n_cand * args.autotune_days <= warmup:
        chosen, tune_ms = eng.autotune(args.autotune_days)
        done = n_cand * args.autotune_days
        log(f"[bench] {workload}: autotuned over {n_cand} launch tunings x {args.autotune_days} days: {chosen}")
    if warmup > done:
        eng.step(warmup - done)
    eng.sync()
    # the CPU baseline resumes from the GPU's census at the first timed day
    cpu_sims = list(range(min(len(specs), args.cpu_sims if not plain else 1)))
    states = [eng.state(i) for i in cpu_sims] if not args.no_cpu else []
    eng.reset_timing()
    log(f"[bench] {workload}: warm-up {warmup} days done")

    t_start = time.perf_counter()
    done = 0
    last = t_start
    while done < steps:
        n = min(100, steps - done)
        eng.step(n)
        done += n
        if time.perf_counter() - last > 20.0:
            eng.sync()
            last = time.perf_counter()
            log(f"[bench] {workload}: {done}/{steps} days, {last - t_start:.1f} s")
    eng.sync()
    elapsed = time.perf_counter() - t_start
    tuning = eng.tuning()
    if graph:   # kernel means and scan counts from a second pass over the same days with per-kernel events
        eng.close()
        eng = Engine(lib, device=0, flags=abi.F_TIMING)
        eng.init(specs, days)
        eng.set_tuning(**tuning)
        eng.step(warmup)
        eng.sync()
        eng.reset_timing()
        eng.step(steps)
        eng.sync()
    ms_a, n_a = eng.kernel_time(0)
    ms_b, _ = eng.kernel_time(1)
    hv, mv = eng.scan_counts()
    eng.close()

    mean_a = ms_a / max(n_a, 1)
    bitmap = 4 * ((slots + 31) // 32)
    per_comm = sum(64 + 32 + 96 + 40 * rl for s in specs for rl in row_lengths(s))
    alg = (16 * hv + 32 * mv) / max(n_a, 1) + bitmap + per_comm
    achieved = alg / (mean_a * 1e-3) / 1e9
    traffic = profiled_traffic(workload, warmup, steps)
    full = 4 * humans + 4 * slots + 32 * sero_comms
    value = humans * steps / elapsed
    out = {
        "value": value,
        "unit": "person-days/s",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "config": {
            "workload": (f"plain single simulation: 100x100 lattice, {humans} humans (1,000 per community), "
                         f"{slots} mosquito slots, Beta(2,5) initial immunity, days {warmup}..{days - 1} timed"
                         if plain else
                         f"batched plain simulation: {len(specs)} simulations, {comms} communities "
                         f"(400-10,000 per simulation), {humans} humans (200-2,000 per community), "
                         f"{slots} mosquito slots, days {warmup}..{days - 1} timed"),
            "simulations": len(specs),
            "communities": comms,
            "humans": humans,
            "mosquito_slots": slots,
            "days": days,
            "graph": graph,
            "tuning": tuning,
            "autotuned": tune_ms is not None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_day_demography",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_frac": (traffic / (mean_a * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "traffic_source": PROFILES[workload] if traffic else None,
            "bytes_per_launch": alg,
            "census_vectors_per_launch": {"human": hv / max(n_a, 1), "mosquito": mv / max(n_a, 1)},
            "full_scan_bytes": full,
            "mean_launch_ms": mean_a,
            "launches": n_a,
        },
        "kernels_ms": {"k_day_demography": ms_a, "k_day_events_and_rest": ms_b, "launches": n_a},
    }
    if graph:
        out["roofline"]["pass"] = f"a second run of the same days ({warmup}..{days - 1}) with per-kernel HIP events"
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cdays = min(steps, 600 if plain else 60)   # about 10-15 s of CPU work either way
        log(f"[bench] {workload}: CPU oracle baseline on {threads} threads")
        what = ("the 100x100-lattice simulation" if plain else
                f"simulations 0..{len(cpu_sims) - 1} of the batch")
        out["cpu_baseline"] = cpu_baseline([specs[i] for i in cpu_sims], states, warmup, cdays, threads,
                                           min(cdays, 10 if plain else 20), what)
    return out


def mcmc_bench(args) -> dict:
    """MCMC ESS/hour, plain batched adaptive Metropolis vs surrogate delayed
    acceptance, on the 7-parameter model (five epidemic parameters, the
    reporting rate and the NB dispersion; 1e6 humans, 1,024 communities, 260
    weekly counts, synthetic data at the true parameters):
      1. `train_runs` simulations at prior draws -> weekly incidence;
      2. the surrogate MLP trained on them (Poisson loss, Adam, `epochs`);
      3. a surrogate-only pilot chain gives both samplers the same starting
         states near the posterior;
      4. `iters` iterations of plain adaptive Metropolis and of Langevin
         delayed acceptance, `screen_iters` of screened delayed acceptance
         (cheaper iterations: only proposals the surrogate passes are
         simulated); 64 chains, `particles` simulations per proposal
         (their likelihoods averaged; with `blocks` > 0 the particle seeds
         are kept in the chain state and one block refreshed per iteration);
         ESS over the iterations after the first quarter, per hour of that
         sampler's own wall clock."""
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
    pilot = mcmc.Mcmc(cfg, g, g, humans, configs.base_params(), data, lib=lib)  # surrogate only

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
                                    n_particles=args.particles, n_blocks=args.blocks,
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
        "config": {"workload": f"MCMC fit: {args.chains} chains x {args.particles} simulation(s) per proposal"
                               f"{f' in {args.blocks} seed blocks' if args.blocks else ''}, 1e6 humans in "
                               f"32x32 communities, 1,819 days, 260 weekly counts; {args.iters} iterations plain "
                               f"and Langevin delayed acceptance, {args.screen_iters} screened",
                   "sub_steps": args.sub_steps},
    }
    sur.close()
    eng.close()
    return out


def mcmc_noise_bench(args) -> dict:
    """Noise of the fit's likelihood estimate, the quantity that decides
    whether pseudo-marginal chains can move: at the true parameters and at
    1, 3 and 10 fixed-proposal steps along each of the five epidemic
    parameters and along 3 random directions (3 steps), `noise_seeds`
    simulations each under the same seeds (common random numbers), binned by
    week and scored with arbo_nb_loglik.  Reports the sd over seeds of the
    log-likelihood at the truth and, per step, of ll(theta + step) and of
    ll(theta + step) - ll(theta)."""
    import numpy as np

    from arbo_gpu_fit_amd import abi, configs, mcmc
    from arbo_gpu_fit_amd.engine import Engine, SimSpec

    lib = mcmc.bind(abi.load_hip_library())
    eng = Engine(lib, flags=abi.F_GRAPH)
    data, _ = configs.mcmc_data(eng)
    W = configs.MCMC_WEEKS
    T = 7 * W - 1
    g, humans = configs.mcmc_shape()
    truth = configs.MCMC_TRUTH.copy()
    D = len(truth)
    # the fixed proposal's sd per parameter: 0.1 prior sd / sqrt(D)
    psd = []
    for kind, a, b in configs.mcmc_priors():
        if kind in (mcmc.PRIOR_UNIFORM, mcmc.PRIOR_CIRCULAR):
            psd.append((b - a) / np.sqrt(12.0))
        elif kind == mcmc.PRIOR_BETA:
            psd.append(np.sqrt(a * b / ((a + b) ** 2 * (a + b + 1))))
        else:
            psd.append(0.0)
    sd0 = 0.1 * np.array(psd) / np.sqrt(D)
    steps = []
    for i in range(5):
        for m in (1.0, 3.0, 10.0):
            d = np.zeros(D)
            d[i] = m * sd0[i]
            steps.append((f"{mcmc.PARAM_NAMES[i]} x{m:g}", d))
    rng = np.random.default_rng(1)
    for r in range(3):
        d = np.zeros(D)
        d[:5] = rng.normal(size=5) * sd0[:5] * 3
        steps.append((f"random{r} x3", d))
    thetas = [("truth", np.zeros(D))] + steps
    seeds = [(s << 32) | 77 for s in range(args.noise_seeds)]
    specs = [SimSpec(g, g, humans, mcmc.sim_params(lib, truth + d, configs.base_params()), sd)
             for _, d in thetas for sd in seeds]
    t0 = time.perf_counter()
    ll = np.zeros((len(thetas), len(seeds)))
    for b0 in range(0, len(specs), 64):
        chunk = specs[b0:b0 + 64]
        eng.init(chunk, T)
        eng.step(T)
        for k in range(len(chunk)):
            i, j = divmod(b0 + k, len(seeds))
            wk = mcmc.weekly_incidence(lib, eng.series(k)["n_incidence"], W)
            ll[i, j] = mcmc.nb_loglik(lib, wk, data, 1, truth[5], 1.0 / truth[6])
        log(f"[noise] {b0 + len(chunk)}/{len(specs)} simulations, {time.perf_counter() - t0:.1f} s")
    eng.close()
    rows = []
    for i, (name, _) in enumerate(thetas[1:], 1):
        diff = ll[i] - ll[0]
        rows.append({"step": name, "mean_diff": float(diff.mean()), "sd_diff": float(diff.std(ddof=1)),
                     "sd_ll": float(ll[i].std(ddof=1))})
    return {"metric": "log-likelihood noise of one simulation per proposal", "unit": "nats",
            "value": float(ll[0].std(ddof=1)), "seeds": len(seeds), "proposal_sd": sd0.tolist(),
            "ll_truth": ll[0].tolist(), "rows": rows,
            "config": {"workload": f"the fit's model (1e6 humans, 32x32 communities, {T} days, {W} weeks), "
                                   f"{len(thetas)} parameter vectors x {len(seeds)} common seeds"}}


def surrogate_bench(args) -> dict:
    """Training throughput of the surrogate MLP 6 -> 256^4 -> 260 (GELU) on
    the MFMA kernels: `train_rows` rows of synthetic data (inputs uniform on
    [-1, 1]^6, Poisson counts of a fixed random teacher network), minibatches
    of 512, Adam, `sur_epochs` epochs after one untimed warm-up epoch.  The
    matrix-product FLOPs per training row are those of the float32 network
    (forward, weight and input gradients); the MFMA units execute three
    bfloat16 products for each (the operands' hi/lo split)."""
    import numpy as np

    from arbo_gpu_fit_amd import abi
    from arbo_gpu_fit_amd.surrogate import Surrogate
    from tests.mlp_ref import forward64, random_layers

    widths = [6, 256, 256, 256, 256, 260]
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(args.train_rows, widths[0])).astype(np.float32)
    teacher = random_layers(widths, 7)
    mu = np.exp(np.clip(0.3 * forward64(teacher, x)[0] + 2.0, -3, 6))
    y = rng.poisson(mu).astype(np.float32)
    sur = Surrogate(widths, max_batch=4096, seed=1, lib=abi.load_hip_library())
    sur.train(x, y, epochs=1, batch=512, lr=1e-3, seed=2)   # warm-up
    t0 = time.perf_counter()
    hist = sur.train(x, y, epochs=args.sur_epochs, batch=512, lr=1e-3, seed=3)
    dt = time.perf_counter() - t0
    sur.close()
    rows = args.train_rows * args.sur_epochs
    steps = -(-args.train_rows // 512) * args.sur_epochs
    mm = sum(a * b for a, b in zip(widths[:-1], widths[1:]))
    flop_row = 2 * mm * 3 - 2 * widths[0] * widths[1]   # forward + weight gradients + input gradients (not of layer 0)
    tf = flop_row * rows / dt / 1e12
    return {
        "metric": "surrogate training throughput (6-256^4-260 MLP, bfloat16 MFMA)",
        "value": rows / dt,
        "unit": "training rows/s",
        "n_gpus": 1,
        "steps": steps,
        "ms_per_step": dt * 1e3 / steps,
        "higher_is_better": True,
        "dtype": "bf16x3 (float32 operands split into two bfloat16 halves), float32 accumulation",
        "data": "synthetic (a random teacher network's Poisson counts)",
        "config": {"workload": f"{args.train_rows} rows x {args.sur_epochs} epochs, minibatch 512, Adam"},
        "roofline": {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": tf / 2500.0,
                     "mfma_issued": 3 * tf, "mfma_issued_frac": 3 * tf / 2500.0, "flop_per_row": flop_row},
        "loss_first_last": [float(hist[0]), float(hist[-1])],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1800, help="timed days of the batch")
    ap.add_argument("--warmup", type=int, default=25, help="warm-up days of the batch")
    ap.add_argument("--plain-steps", type=int, default=3600, help="timed days of the plain workload")
    ap.add_argument("--plain-warmup", type=int, default=50, help="warm-up days of the plain workload")
    ap.add_argument("--workload", choices=("batch", "plain"), default="batch",
                    help="plain: only the 100x100-lattice simulation as the headline")
    ap.add_argument("--no-plain", action="store_true", help="leave out the plain record")
    ap.add_argument("--sims", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sims", type=int, default=0,
                    help="simulations of the batch in the CPU sample (0: 128 for windows up to 60 days, else 64)")
    ap.add_argument("--graph", action="store_true", help="batch: replay the day loop as hipGraphs")
    ap.add_argument("--no-graph", action="store_true",
                    help="plain: launch every kernel on its own (for profilers; the same kernels and days)")
    ap.add_argument("--autotune-days", type=int, default=2,
                    help="days per candidate launch tuning, timed within the warm-up (0: default tuning)")
    ap.add_argument("--mcmc", action="store_true", help="MCMC ESS/hour, plain vs surrogate, instead")
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--iters", type=int, default=160)
    ap.add_argument("--screen-iters", type=int, default=480, help="iterations of the screened delayed acceptance")
    ap.add_argument("--mcmc-seed", type=int, default=1)
    ap.add_argument("--particles", type=int, default=1, help="simulations per proposal")
    ap.add_argument("--blocks", type=int, default=0, help="pseudo-marginal seed blocks (0: fresh seeds)")
    ap.add_argument("--train-runs", type=int, default=50000)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--sub-steps", type=int, default=10)
    ap.add_argument("--save", default="", help="save the surrogate's training set (npz)")
    ap.add_argument("--load", default="", help="load the surrogate's training set (npz)")
    ap.add_argument("--surrogate", action="store_true", help="surrogate training throughput instead")
    ap.add_argument("--mcmc-noise", action="store_true",
                    help="log-likelihood noise of the fit's one-simulation estimate under common seeds instead")
    ap.add_argument("--noise-seeds", type=int, default=8, help="common seeds per parameter vector (--mcmc-noise)")
    ap.add_argument("--train-rows", type=int, default=50000)
    ap.add_argument("--sur-epochs", type=int, default=20)
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
    if args.surrogate:
        print(json.dumps(surrogate_bench(args)), flush=True)
        return
    if args.mcmc_noise:
        print(json.dumps(mcmc_noise_bench(args)), flush=True)
        return
    if args.cpu_sims == 0:
        args.cpu_sims = 128 if args.steps <= 60 else 64   # about 10-15 s of CPU work

    from arbo_gpu_fit_amd import abi

    lib = abi.load_hip_library()
    head = sim_bench(args.workload, args, lib)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "person-days/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": head["warmup"],
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": head["config"],
        "roofline": head["roofline"],
        "kernels_ms": head["kernels_ms"],
    }
    if "cpu_baseline" in head:
        out["cpu_baseline"] = head["cpu_baseline"]
    if args.workload == "batch" and not args.no_plain:
        out["plain"] = sim_bench("plain", args, lib)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
