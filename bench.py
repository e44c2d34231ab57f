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


def cpu_baseline(specs, days: int, threads: int, serial_days: int) -> dict:
    """The CPU oracle on a bounded sample: the first simulations of the batch
    for the first `days` days on `threads` OpenMP threads, and simulation 0
    alone on one thread for `serial_days` days."""
    from tests.oracle_lib import load_oracle

    lib = load_oracle()
    rate, humans, dt = _oracle_rate(lib, specs, days, threads)
    srate, shumans, sdt = _oracle_rate(lib, specs[:1], serial_days, 1)
    return {
        "value": rate,
        "unit": "person-days/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (g++ -O2, OpenMP {threads} threads): simulations 0..{len(specs) - 1} of the "
                  f"batch ({humans} humans), days 0..{days - 1}, {dt:.1f} s",
        "serial": {
            "value": srate,
            "cores": 1,
            "sample": f"simulation 0 ({shumans} humans), days 0..{serial_days - 1}, one thread, {sdt:.1f} s",
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1800)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--sims", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sims", type=int, default=8)
    ap.add_argument("--cpu-days", type=int, default=365)
    ap.add_argument("--cpu-serial-days", type=int, default=365)
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
    specs, days_cfg = configs.batch256(args.sims)
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
    step_bytes = 7 * (humans + slots) + 32 * sero_comms
    step_gbs = step_bytes / (elapsed / args.steps) / 1e9
    out["step_roofline"] = {"bytes_per_step": step_bytes, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS}
    if not args.graph:
        traffic = profiled_traffic() if args.sims == 256 else None
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
        out["kernels_ms"] = {"k_day_demography": ms_a, "k_day_events": ms_b, "launches": n_a}
    eng.close()

    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] CPU oracle baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(specs[:args.cpu_sims], args.cpu_days, threads, args.cpu_serial_days)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
