"""Benchmark: person-days per second of the batched dengue-IBM day loop on one
MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sims 256] [--no-cpu]

Workload: the 256-simulation batch (communities log-uniform 400..10,000 squared
to a lattice, 200..2,000 humans per community, per-simulation transmission,
long-distance and seasonal parameters, 1,825 days).  One step = one simulated
day of every simulation of the batch: the demography kernel and the
event-application kernel over every community.  W warm-up days, then K timed
days (W + K = 1,825 by default, the whole configured run).  Inputs are
synthetic and resident in HBM before timing starts.

The JSON line carries:
  roofline      the demography kernel against HBM bandwidth: algorithmic bytes
                per launch (4 B per human word read, 2 B per mosquito word read
                and 2 B written back) over its mean launch time, measured with
                HIP events on the engine's stream;
  cpu_baseline  the CPU oracle (OpenMP) on a bounded sample of the same
                workload, on this machine's host cores.
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


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(specs, days: int, threads: int) -> dict:
    """The CPU oracle on a bounded sample: the first simulations of the batch
    for the first `days` days."""
    from arbo_gpu_fit_amd.engine import Engine
    from tests.oracle_lib import load_oracle

    lib = load_oracle()
    eng = Engine(lib, threads=threads)
    eng.init(specs, days)
    t0 = time.perf_counter()
    eng.step(days)
    dt = time.perf_counter() - t0
    humans = sum(int(s.humans.sum()) for s in specs)
    eng.close()
    return {
        "value": humans * days / dt,
        "unit": "person-days/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (g++ -O2, OpenMP {threads} threads): simulations 0..{len(specs) - 1} of the "
                  f"batch ({humans} humans), days 0..{days - 1}, {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1800)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--sims", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sims", type=int, default=4)
    ap.add_argument("--cpu-days", type=int, default=60)
    ap.add_argument("--graph", action="store_true", help="replay the day loop as hipGraphs (no per-kernel timing)")
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

    from arbo_gpu_fit_amd import abi, configs
    from arbo_gpu_fit_amd.engine import Engine

    lib = abi.load_hip_library()
    t0 = time.perf_counter()
    specs, days_cfg = configs.config2(args.sims)
    humans = sum(int(s.humans.sum()) for s in specs)
    comms = sum(s.communities for s in specs)
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
            "workload": f"batched plain simulation: {len(specs)} simulations, {comms} communities "
                        f"(400-10,000 per simulation), {humans} humans (200-2,000 per community), "
                        f"{slots} mosquito slots, {days} days",
            "simulations": len(specs),
            "communities": comms,
            "humans": humans,
            "mosquito_slots": slots,
            "days": days,
            "graph": bool(args.graph),
        },
    }
    if not args.graph:
        ms_a, n_a = eng.kernel_time(0)
        ms_b, n_b = eng.kernel_time(1)
        bytes_a = 4 * humans + 4 * slots
        mean_a = ms_a / max(n_a, 1)
        achieved = bytes_a / (mean_a * 1e-3) / 1e9
        out["roofline"] = {
            "bound": "hbm",
            "kernel": "k_day_demography",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_a,
            "mean_launch_ms": mean_a,
        }
        out["kernels_ms"] = {"k_day_demography": ms_a, "k_day_events": ms_b, "launches": n_a}
    eng.close()

    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] CPU oracle baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(specs[:args.cpu_sims], args.cpu_days, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
