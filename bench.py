#!/usr/bin/env python3
"""Throughput of the per-hour agent step on one MI355X.

Headline workload: 1,000,000 agents on a 2000x2000 grid, 70% working, 20% public
transport, 10 initial exposed, 2,399 simulated hours, seed 42, deterministic mode
(tests/golden/configs/cfg2_1m.json). A "step" is one simulated hour of the hot path over
all agents (hour-of-day-0 hospital scan, agent step, epilogue with counts and
interventions, commit).

    python bench.py [--gpus 1] [--steps 2399] [--warmup 2399] [--skip-extra]

Prints one JSON line. `value` is agent-hours/s = agents x hours stepped / wall seconds of
the hour loop with inputs resident in HBM (population synthesis and upload happen before
the timed region, as the reference's timer starts after Epidemiology::new,
engine/src/epidemiology_simulation.rs:212-272). Two further workloads are reported in the
same line: 64 mixed replicates in one batched launch (replicates/s,
tests/golden/configs/cfg4_replicates.json) and 10M agents with lockdown, hospital growth
and daily vaccination over 4,799 hours (tests/golden/configs/cfg3_10m_interventions.json).

The default warm-up is one whole run on a separate instance. In a fresh process, day-graph
launches hold the host longer than the GPU takes to run a day until many of them have
been made: after a 240-hour warm-up the first timed run measured 32.6-32.9 us of wall
time per hour against 30.0-30.3 us of device time, and after a whole-run warm-up the two
agree.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
CONFIGS = os.path.join(ROOT, "tests", "golden", "configs")
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_hour(n, g):
    # 32 B of agent state read + written per agent-hour, plus a 4-byte occupancy cell read
    # and written per grid cell-hour: the roofline model this benchmark is quoted against.
    return 32.0 * n + 8.0 * g * g


def tuned_shape_of(cfg, block=None):
    """(home, office) threads per block a handle of this config runs with."""
    from epirust_amd import autotune
    if block:
        return block, block
    default = 512 if cfg["population"]["Auto"]["number_of_agents"] >= (1 << 19) else 256
    return autotune.tuned_shape([cfg]) or (default, default)


def measured_traffic(name):
    """HBM bytes per simulated hour from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_hour")


def cpu_threads():
    # a one-GPU host's CPU share is 16 cores; use at most that many oracle threads
    return max(1, min(16, os.cpu_count() or 1))


def refcpu_rate(cfgs, target_seconds, threads=None, samples=1, window_hours=0):
    """refcpu (oracle/librefcpu.so: the reference's step structure restated in C: citizens
    held by value in two FNV hash maps keyed by cell, a `threads`-thread map over the
    current map's buckets, a single-thread merge in bucket order, swap and clear,
    allocation_map.rs:44-134) on the first hours of the given configs: exactly
    window_hours hours if given, else whole hours until about target_seconds have passed.
    Population synthesis is outside the timed region, as the reference's timer starts
    after Epidemiology::new. With several samples (fresh simulations over the same hours)
    the median rate is reported. Returns (agent-hours/s, text)."""
    from oracle import epicpu as O
    threads = threads or cpu_threads()
    rates, hours_by, spent_all = [], [], 0.0
    for _ in range(samples):
        agent_hours, spent, hours_by = 0, 0.0, []
        for cfg in cfgs:
            sim = O.RefCpuSim(cfg)
            sim.set_threads(threads)
            n = cfg["population"]["Auto"]["number_of_agents"]
            budget = target_seconds / len(cfgs)
            t0 = time.perf_counter()
            h = 0
            while (h < window_hours if window_hours else time.perf_counter() - t0 < budget) and h < cfg["hours"] - 1:
                rows, stopped = sim.step(window_hours - h if window_hours else 1)
                h += len(rows)
                if stopped:
                    break
            dt = time.perf_counter() - t0
            sim.close()
            agent_hours += n * h
            spent += dt
            hours_by.append(h)
        rates.append(agent_hours / spent)
        spent_all += spent
    rates.sort()
    return rates[len(rates) // 2], (
        "refcpu (oracle/librefcpu.so, the reference's hash-map step structure restated in C), %d threads, "
        "first %s hours of %d config(s), median of %d samples (%s agent-h/s), %.1f s in all" % (
            threads, "/".join(str(h) for h in hours_by[:4]) + ("/..." if len(hours_by) > 4 else ""), len(cfgs),
            samples, ", ".join("%.3g" % r for r in rates), spent_all))


_hip = None


def device_sync():
    """Device-wide synchronisation (hipDeviceSynchronize) on both sides of a timed region."""
    global _hip
    import ctypes
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
    rc = _hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError("hipDeviceSynchronize failed: %d" % rc)


def timed_steps(handle, hours):
    device_sync()
    t0 = time.perf_counter()
    rows, stopped = handle.step(hours)
    device_sync()
    return time.perf_counter() - t0, rows


def bench_headline(args, G):
    with open(args.config) as f:
        cfg = json.load(f)
    n = cfg["population"]["Auto"]["number_of_agents"]
    g = cfg["geography_parameters"]["grid_size"]
    run_hours = cfg["hours"] - 1

    from epirust_amd import autotune

    def make():
        s = G.Simulation(cfg)
        autotune.apply_shape(s, [cfg], block=args.block)
        return s

    if args.warmup > 0:
        w = make()
        w.step(min(args.warmup, run_hours))
        w.close()
    sims = [make() for _ in range(max(1, math.ceil(args.steps / run_hours)))]
    device_sync()
    t0 = time.perf_counter()
    hours_done, remaining = 0, args.steps
    device_ms, device_hours = 0.0, 0
    for s in sims:
        k = min(remaining, run_hours)
        rows, _ = s.step(k)
        hours_done += len(rows)
        ms, hrs = s.last_run_ms()  # HIP events around this call's hour loop on the handle's stream
        device_ms += ms
        device_hours += hrs
        remaining -= k
        if remaining <= 0:
            break
    device_sync()
    elapsed = time.perf_counter() - t0
    for s in sims:
        s.close()
    p = make()
    prof = p.profile(min(args.steps, run_hours))
    p.close()
    prof["device_us_per_hour"] = 1e3 * device_ms / max(device_hours, 1)
    return cfg, n, g, run_hours, hours_done, elapsed, prof


def bench_replicates(G, cpu_seconds):
    from epirust_amd import autotune
    with open(os.path.join(CONFIGS, "cfg4_replicates.json")) as f:
        cfgs = json.load(f)
    b = G.Batch(cfgs)
    shape = autotune.apply_shape(b, cfgs) or (256, 256)
    dt, rows = timed_steps(b, 2399)
    b.close()
    agent_hours = sum(c["population"]["Auto"]["number_of_agents"] * len(r) for c, r in zip(cfgs, rows))
    out = {"metric": "replicates/s @64 mixed", "value": len(cfgs) / dt, "unit": "replicates/s",
           "seconds": dt, "agent_hours_per_s": agent_hours / dt, "replicates": len(cfgs),
           "total_agents": sum(c["population"]["Auto"]["number_of_agents"] for c in cfgs),
           "block": shape[0], "office_block": shape[1], "hours_per_replicate": 2399}
    if cpu_seconds > 0:
        rate, sample = refcpu_rate(cfgs[:4], cpu_seconds)
        full = sum(c["population"]["Auto"]["number_of_agents"] * 2399 for c in cfgs)
        out["cpu_baseline"] = {"value": len(cfgs) / (full / rate), "unit": "replicates/s", "cores": cpu_threads(),
                               "kind": "port", "sample": sample + "; replicates run serially, projected "
                               "from agent-hours/s over the batch's %.3g agent-hours" % full}
    return out


def bench_interventions(G, cpu_seconds):
    with open(os.path.join(CONFIGS, "cfg3_10m_interventions.json")) as f:
        cfg = json.load(f)
    from epirust_amd import autotune
    s = G.Simulation(cfg)
    autotune.apply_shape(s, [cfg])
    dt, rows = timed_steps(s, cfg["hours"] - 1)
    device_ms, device_hours = s.last_run_ms()
    ev = s.events()
    s.close()
    n = cfg["population"]["Auto"]["number_of_agents"]
    g = cfg["geography_parameters"]["grid_size"]
    hour_us = 1e3 * device_ms / max(device_hours, 1)
    achieved = algorithmic_bytes_per_hour(n, g) / (hour_us * 1e-6) / 1e9
    out = {"metric": "agent-hours/s @10M agents with interventions", "value": n * len(rows) / dt,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_hour": algorithmic_bytes_per_hour(n, g),
                        "basis": "HIP events on the handle's stream around the run's hour loop: %.1f us of device "
                                 "time per simulated hour" % hour_us,
                        "traffic": measured_traffic("traffic_cfg3_10m.json")},
           "unit": "agent-hours/s", "hours": len(rows), "seconds": dt,
           "events": {"vaccination": sum(1 for e in ev if e[1] == 0), "lockdown": sum(1 for e in ev if e[1] in (1, 2)),
                      "build_new_hospital": sum(1 for e in ev if e[1] == 3)},
           "peak_infected": int(rows[:, 3].max()) if len(rows) else 0}
    if cpu_seconds > 0:
        rate, sample = refcpu_rate([cfg], cpu_seconds)
        out["cpu_baseline"] = {"value": rate, "unit": "agent-hours/s", "cores": cpu_threads(), "kind": "port",
                               "sample": sample}
    return out


def bench_hour_at_a_time(G, hours=240):
    """The integration path that hands every hour's row to the reference's listeners as it
    is made: epi_step(sim, 1, ...) called once per hour (a stream drain and a read-back per
    call) over the first `hours` hours of the 1M-agent run, against the same hours in one
    call."""
    with open(os.path.join(CONFIGS, "cfg2_1m.json")) as f:
        cfg = json.load(f)
    from epirust_amd import autotune
    n = cfg["population"]["Auto"]["number_of_agents"]
    out = {"agents": n, "hours": hours}
    for name, per_call in (("one_call", hours), ("hour_at_a_time", 1)):
        s = G.Simulation(cfg)
        autotune.apply_shape(s, [cfg])
        device_sync()
        t0 = time.perf_counter()
        done = 0
        while done < hours:
            rows, _ = s.step(per_call)
            done += len(rows)
        device_sync()
        dt = time.perf_counter() - t0
        s.close()
        out[name] = {"agent_hours_per_s": n * done / dt, "us_per_hour": 1e6 * dt / done}
    return out


def bench_init(G):
    """epi_create time of the 10M-agent config (allocation + population synthesis + grid),
    with the population synthesized by the device kernels and by the host restatement
    (multithreaded C++), after one untimed device create."""
    with open(os.path.join(CONFIGS, "cfg3_10m_interventions.json")) as f:
        cfg = json.load(f)
    out = {"agents": cfg["population"]["Auto"]["number_of_agents"]}
    runs = [("warm", G.EPI_FLAG_DETERMINISTIC), ("device_s", G.EPI_FLAG_DETERMINISTIC),
            ("host_s", G.EPI_FLAG_DETERMINISTIC | G.EPI_FLAG_HOST_INIT)]
    for name, flags in runs:
        t0 = time.perf_counter()
        s = G.Simulation(cfg, flags=flags)
        device_sync()
        dt = time.perf_counter() - t0
        s.close()
        if name != "warm":
            out[name] = dt
    return out


def bench_surrogate(n_train, n_eval, seeds, epochs):
    """The learned surrogate on a reduced, stated sweep: n_train training simulations
    (100,000 agents, 632x632, 2,400 hours, Sobol points over the ranges of
    tests/golden/configs/cfg5_surrogate.json), two networks trained with the bf16-MFMA
    kernels, then n_eval held-out points answered with the 5%-error simulator fallback and
    timed against the plain sweep of `seeds`-seed simulation ensembles, which also judges
    the error (peak infected and final size). The full sweep (20,000 / 500 / 32) is
    `python -m epirust_amd.surrogate`."""
    import tempfile
    from epirust_amd import surrogate
    with tempfile.TemporaryDirectory() as d:
        m = surrogate.run_pipeline(d, n_train=n_train, n_eval=n_eval, eval_seeds=seeds, epochs=epochs,
                                   log=lambda line: print(line, file=sys.stderr, flush=True))
    cal = m["calibrated"]
    return {"metric": "surrogate speedup @<=5% err", "value": m["speedup_vs_plain_sweep"], "unit": "x",
            "surrogate_points_per_s": m["surrogate_points_per_s"], "plain_points_per_s": m["plain_points_per_s"],
            "fraction_within_5pct": m["with_fallback_fraction_within"], "fallback_fraction": m["fallback_fraction"],
            "surrogate_alone_fraction_within_5pct": m["surrogate_fraction_within"],
            "nearest_neighbour_fraction_within_5pct": m["nn_fraction_within"],
            "calibrated_gate": {"speedup": cal["speedup_vs_plain_sweep"],
                                "fraction_within_5pct": cal["test_fraction_within"],
                                "fallback_fraction": cal["test_fallback_fraction"]},
            "config": {"train_simulations": n_train, "eval_points": n_eval, "seeds_per_point": seeds,
                       "epochs": epochs, "agents": m["agents"], "grid_size": m["grid_size"], "hours": m["hours"],
                       "note": "reduced sweep; the full one is 20,000 training simulations and 500 points"},
            "seconds": {"train_simulations": m["train_sim_seconds"], "training": m["train_seconds"],
                        "plain_sweep": m["plain_sweep_seconds"], "surrogate_sweep": m["surrogate_sweep_seconds"]}}


def profile_driver(args, G):
    import ctypes
    with open(args.config) as f:
        cfg = json.load(f)
    from epirust_amd import autotune
    s = G.Simulation(cfg)
    autotune.apply_shape(s, [cfg], block=args.block)
    prof = s.profile(args.profile_hours)
    s.close()
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 64 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"profile_hours": prof["hours"], "kernel_ms": prof, "calibration_copy_bytes": nbytes}))


def profile_gemm(shape, reps):
    """Profiling driver for the surrogate's MFMA GEMM (k_gemm): `reps` calls of
    epi_gemm_bf16 at M x N x K with random bf16 operands (host buffers; rocprofv3's kernel
    trace gives the kernel's own duration)."""
    import ctypes
    import numpy as np
    import epirust_amd as G
    m, n, k = shape
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((m, k)).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    b = (rng.standard_normal((k, n)).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    c = np.zeros((m, n), dtype=np.float32)
    L = G.lib()
    for _ in range(reps):
        rc = L.epi_gemm_bf16(a.ctypes.data, b.ctypes.data, c.ctypes.data, m, n, k, 0, 0)
        if rc != 0:
            raise RuntimeError(L.epi_mlp_last_error().decode())
    print(json.dumps({"gemm": {"M": m, "N": n, "K": k, "reps": reps, "flop_per_call": 2.0 * m * n * k}}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2399, help="simulated hours timed")
    ap.add_argument("--warmup", type=int, default=2399, help="untimed hours on a separate instance")
    ap.add_argument("--config", default=os.path.join(CONFIGS, "cfg2_1m.json"))
    ap.add_argument("--block", type=int, default=0, help="threads per agent block (0 = default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--surrogate-train", type=int, default=6144, help="training simulations of the surrogate leg")
    ap.add_argument("--surrogate-eval", type=int, default=96, help="held-out points of the surrogate leg")
    ap.add_argument("--skip-surrogate", action="store_true")
    ap.add_argument("--profile-hours", type=int, default=0,
                    help="profiling driver for rocprofv3: run this many hours as separate kernel dispatches "
                         "(no graphs, no torch), then one 64 MiB device-to-device copy as the known-size "
                         "calibration dispatch for FETCH_SIZE / WRITE_SIZE, and exit")
    ap.add_argument("--profile-gemm", default="", help="M,N,K: profiling driver for the surrogate's MFMA GEMM")
    ap.add_argument("--profile-gemm-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path; the step would shard as one "
              "region per GPU with halo and claim exchange over RCCL.", file=sys.stderr)
        sys.exit(2)

    import epirust_amd as G

    if args.profile_gemm:
        profile_gemm([int(x) for x in args.profile_gemm.split(",")], args.profile_gemm_reps)
        return
    if args.profile_hours:
        profile_driver(args, G)
        return

    cfg, n, g, run_hours, hours_done, elapsed, prof = bench_headline(args, G)
    value = n * hours_done / elapsed
    scan_ms, n_scan = prof["hospital_scan"]
    step_ms, n_step = prof["step"]
    commit_ms, n_hours = prof["commit"]
    # device time of one simulated hour over the timed region (all of an hour's launches:
    # the commit launch every hour, the step at hours of day 0 and 7..23, the hospital scan
    # at hour of day 0, and the gaps between them)
    hour_us = prof["device_us_per_hour"]
    achieved = algorithmic_bytes_per_hour(n, g) / (hour_us * 1e-6) / 1e9
    traffic = measured_traffic("traffic_cfg2_1m.json")
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
        "config": {"workload": "EpiRust standalone hour step, %d agents, %dx%d grid, seed %s, deterministic "
                               "mode: %d simulated hours timed (hours 1..%d of runs of %d hours)" % (
                                   n, g, g, cfg.get("seed", 0), hours_done, min(hours_done, run_hours), run_hours),
                   "agents": n, "grid_size": g, "hours": run_hours, "hours_timed": hours_done,
                   "threads_per_block": dict(zip(("home", "office"), tuned_shape_of(cfg, args.block))),
                   "kernel_ms": {"hospital_scan": scan_ms, "step": step_ms, "commit": commit_ms}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "one simulated hour: k_commit with its folded epilogue every hour, k_step at hours of day 0 and 7..23, k_hospital_scan once a day",
                     "basis": "HIP events on the handle's stream around the timed hour loop: %.2f us of device time per simulated hour" % hour_us,
                     "algorithmic_bytes_per_hour": algorithmic_bytes_per_hour(n, g)},
    }
    cpu_s = 0.0 if args.no_cpu_baseline else args.cpu_seconds
    if cpu_s > 0:
        # one whole simulated day per sample (hours 1..24: sleep, commute, office and evening hours)
        rate, sample = refcpu_rate([cfg], 0, samples=3, window_hours=24)
        result["cpu_baseline"] = {"value": rate, "unit": "agent-hours/s", "cores": cpu_threads(), "kind": "port",
                                  "sample": sample}
        # the reference engine's default rayon pool is 4 threads (engine-app/src/main.rs:80-82)
        rate4, sample4 = refcpu_rate([cfg], 0, threads=4, samples=3, window_hours=24)
        result["cpu_baseline_4_threads"] = {"value": rate4, "unit": "agent-hours/s", "cores": 4, "kind": "port",
                                            "sample": sample4}
    if not args.skip_extra:
        result["replicates"] = bench_replicates(G, cpu_s / 3)
        result["interventions_10m"] = bench_interventions(G, cpu_s / 3)
        result["create_10m"] = bench_init(G)
        result["hour_at_a_time"] = bench_hour_at_a_time(G)
        if not args.skip_surrogate:
            result["surrogate"] = bench_surrogate(args.surrogate_train, args.surrogate_eval, 32, 60)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
