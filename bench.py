#!/usr/bin/env python3
"""Throughput of the per-hour agent step on one MI355X.

Workload (default): 1,000,000 agents on a 2000x2000 grid, 70% working, 20% public
transport, 10 initial exposed, 2,399 simulated hours, seed 42, deterministic mode
(tests/golden/configs/cfg2_1m.json). A "step" is one simulated hour of the hot path over
all agents (hour-of-day-0 hospital scan, agent step, commit, counts, interventions).

    python bench.py [--gpus 1] [--steps 2399] [--warmup 240]

Prints one JSON line. `value` is agent-hours/s = agents x hours stepped / wall seconds of
the hour loop, inputs resident in HBM (population synthesis and upload happen before the
timed region, as the reference's timer starts after Epidemiology::new,
engine/src/epidemiology_simulation.rs:212-272).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
DEFAULT_CONFIG = os.path.join(ROOT, "tests", "golden", "configs", "cfg2_1m.json")
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_hour(n, g):
    # 32 B of agent state read + written per agent-hour, plus a 4-byte occupancy cell read
    # and written per grid cell-hour: the roofline model this benchmark is quoted against.
    return 32.0 * n + 8.0 * g * g


def cpu_baseline(cfg, target_seconds):
    """The CPU oracle (single-threaded C restatement of the reference step) on the first
    hours of the same workload, bounded to roughly target_seconds."""
    from oracle import epicpu as O
    n = cfg["population"]["Auto"]["number_of_agents"]
    t0 = time.perf_counter()
    sim = O.OracleSim(cfg)
    init_s = time.perf_counter() - t0
    hours = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < target_seconds and hours < cfg["hours"] - 1:
        rows, stopped = sim.step(4)
        hours += len(rows)
        if stopped:
            break
    dt = time.perf_counter() - t0
    sim.close()
    return {"value": n * hours / dt, "unit": "agent-hours/s", "cores": 1, "kind": "port",
            "sample": "oracle/epicpu.c, hours 1..%d of this workload (%.1f s; init %.1f s excluded)"
                      % (hours, dt, init_s)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2399, help="simulated hours timed")
    ap.add_argument("--warmup", type=int, default=240, help="untimed hours on a separate instance")
    ap.add_argument("--config", default=DEFAULT_CONFIG)
    ap.add_argument("--block", type=int, default=0, help="threads per agent block (0 = default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (one region per GPU; "
              "see DESIGN.md for how the step would shard).", file=sys.stderr)
        sys.exit(2)

    import torch
    import epirust_amd as G

    with open(args.config) as f:
        cfg = json.load(f)
    n = cfg["population"]["Auto"]["number_of_agents"]
    g = cfg["geography_parameters"]["grid_size"]
    run_hours = cfg["hours"] - 1

    def make():
        s = G.Simulation(cfg)
        if args.block:
            s.set_block_size(args.block)
        return s

    if args.warmup > 0:
        w = make()
        w.step(min(args.warmup, run_hours))
        w.close()
    sims = [make() for _ in range(max(1, math.ceil(args.steps / run_hours)))]

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hours_done = 0
    remaining = args.steps
    for s in sims:
        rows, stopped = s.step(min(remaining, run_hours))
        hours_done += len(rows)
        remaining -= min(remaining, run_hours)
        if remaining <= 0:
            break
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for s in sims:
        s.close()
    value = n * hours_done / elapsed

    # kernel durations from HIP events on the simulation's stream, over a separate run
    p = make()
    prof = p.profile(min(args.steps, run_hours))
    p.close()
    scan_ms, n_scan = prof["hospital_scan"]
    step_ms, n_step = prof["step"]
    commit_ms, _ = prof["commit"]
    hour_ms = step_ms + commit_ms + (scan_ms * n_scan / max(n_step, 1))
    achieved = algorithmic_bytes_per_hour(n, g) / (hour_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_cfg2_1m.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_hour")

    result = {
        "metric": "agent-hours/s @1M agents",
        "value": value,
        "unit": "agent-hours/s",
        "n_gpus": 1,
        "steps": hours_done,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / max(hours_done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic population (deterministic Philox synthesis of the reference's Population::Auto)",
        "config": {"workload": "EpiRust standalone hour step, %d agents, %dx%d grid, %d hours, seed %s, "
                               "deterministic mode" % (n, g, g, run_hours, cfg.get("seed", 0)),
                   "agents": n, "grid_size": g, "hours": run_hours,
                   "kernel_ms": {"hospital_scan": scan_ms, "step": step_ms, "commit": commit_ms}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "hour step = k_step + k_commit (+ k_hospital_scan once a day)"},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
