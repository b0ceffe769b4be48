"""Summaries of rocprofv3 output for the hour-step kernels.

    python -m epirust_amd.prof_summary trace <kernel_stats.csv> [<kernel_trace.csv> --hours H] [--command "..."]
    python -m epirust_amd.prof_summary traffic <fetch.csv> <write.csv> --agents N --grid G --hours H [--command "..."]
    python -m epirust_amd.prof_summary counters <counter_collection.csv>... [--command "..."]
    python -m epirust_amd.prof_summary hourly <counter_collection.csv>... [--command "..."]
    python -m epirust_amd.prof_summary gemm <kernel_stats.csv> --shape M,N,K [--command "..."]

`trace` turns a `--kernel-trace --stats` kernel_stats.csv into a markdown table (calls,
average / min / max duration per kernel) and, given the kernel_trace.csv, a table of the
hour kernels' mean durations per hour of day. `traffic` reads two `--pmc FETCH_SIZE` and
`--pmc WRITE_SIZE` counter_collection.csv files of `bench.py --profile-hours H` (one
dispatch per kernel and hour, then a 64 MiB device-to-device copy). It scales each
kernel's KiB by the ratio the copy dispatch needs to read/write exactly 64 MiB, reports
HBM bytes per simulated hour, and prints the JSON that bench.py reads as roofline.traffic.
"""
import argparse
import collections
import csv
import json
import re
import sys

HOUR_KERNELS = ("k_hospital_scan", "k_relayout", "k_step", "k_commit", "k_quiet_hours", "k_night")
# dispatches that close several hours: k_quiet_hours runs hours of day 3..6, k_night 1..6
SPAN = {"k_commit": 1, "k_quiet_hours": 4, "k_night": 6}


def short(name):
    m = re.search(r"\b(k_\w+)(?:<[^>]*>)?\(", name)
    if m:
        return m.group(1)
    if "copyBuffer" in name:
        return "copy (calibration)"
    return name[:40]


def by_hour_of_day(trace_csv, hours):
    """Mean duration per kernel for each hour of day over the first `hours` hours of the
    trace (an hour is [k_hospital_scan] [k_relayout] [k_step] k_commit: hours are counted by
    k_commit dispatches; a k_quiet_hours dispatch stands for the four hours of day 3..6 and
    a k_night dispatch for the six of 1..6, reported at their first hour), with the wall
    time from one hour's first kernel to the next hour's (a dispatch that closes several
    hours has its wall time split evenly over them)."""
    ks = []
    for r in csv.DictReader(open(trace_csv)):
        k = short(r["Kernel_Name"])
        if k in HOUR_KERNELS:
            ks.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), k))
    ks.sort()
    dur = collections.defaultdict(lambda: collections.defaultdict(list))
    start, span = {}, {}
    hour = 1
    for s, e, k in ks:
        if hour > hours:
            break
        start.setdefault(hour, s)
        dur[hour % 24][k].append((e - s) / 1e3)
        if k in SPAN:
            n = SPAN[k]
            span[hour] = n
            for q in range(1, n):
                for kk in HOUR_KERNELS:
                    dur[(hour + q) % 24][kk]  # listed hours, no dispatches
            hour += n
    wall = collections.defaultdict(list)
    for h in sorted(start):
        n = span.get(h, 1)
        if h + n in start and h + n <= hours:
            for q in range(n):
                wall[(h + q) % 24].append((start[h + n] - start[h]) / 1e3 / n)
    out = ["| hour of day | " + " | ".join(HOUR_KERNELS) + " | kernels µs | hour-to-hour µs |",
           "|---" * (len(HOUR_KERNELS) + 3) + "|"]
    tot = 0.0
    for hd in range(24):
        means = [sum(dur[hd][k]) / len(dur[hd][k]) if dur[hd][k] else 0.0 for k in HOUR_KERNELS]
        w = sorted(wall[hd])[len(wall[hd]) // 2] if wall[hd] else 0.0
        tot += sum(means)
        out.append("| %d | " % hd + " | ".join("%.1f" % m for m in means) + " | %.1f | %.1f |" % (sum(means), w))
    out.append("")
    out.append("Mean kernel time per simulated hour: %.1f µs. Hour-to-hour times include the profiler's "
               "own per-dispatch overhead; the untraced hour is shorter." % (tot / 24))
    return "\n".join(out) + "\n"


def trace(path, command, trace_csv=None, hours=0):
    rows = list(csv.DictReader(open(path)))
    out = ["# Kernel trace summary", "", "Command: `%s`" % command, "",
           "| kernel | calls | average µs | min µs | max µs | total ms | share |", "|---|---|---|---|---|---|---|"]
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append("| %s | %s | %.2f | %.2f | %.2f | %.3f | %.1f%% |" % (
            short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
            float(r["MaxNs"]) / 1e3, float(r["TotalDurationNs"]) / 1e6, 100.0 * float(r["TotalDurationNs"]) / total))
    text = "\n".join(out) + "\n"
    if trace_csv:
        text += "\n## By hour of day (first %d simulated hours)\n\n" % hours + by_hour_of_day(trace_csv, hours)
    return text


def _counter(path, counter):
    per = collections.defaultdict(float)
    names = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        per[d] += float(r["Counter_Value"])
        names[d] = r["Kernel_Name"]
    return per, names


def traffic(fetch_csv, write_csv, agents, grid, hours, command):
    fetch, names = _counter(fetch_csv, "FETCH_SIZE")
    write, names_w = _counter(write_csv, "WRITE_SIZE")
    copies = [d for d in fetch if "copyBuffer" in names[d]]
    cal = max(copies)  # the calibration copy is the last copy dispatch
    kib = 64 * 1024.0
    f_scale = kib / fetch[cal]
    copies_w = [d for d in write if "copyBuffer" in names_w[d]]
    w_scale = kib / write[max(copies_w)]
    per_kernel = collections.defaultdict(lambda: [0.0, 0.0, 0])
    for d, v in fetch.items():
        k = short(names[d])
        if k in HOUR_KERNELS:
            per_kernel[k][0] += v * f_scale * 1024
            per_kernel[k][2] += 1
    for d, v in write.items():
        k = short(names_w[d])
        if k in HOUR_KERNELS:
            per_kernel[k][1] += v * w_scale * 1024
    hour_bytes = sum(f + w for f, w, _ in per_kernel.values()) / hours
    algo = 32.0 * agents + 8.0 * grid * grid
    lines = ["# HBM traffic per simulated hour", "", "Commands:", ""] + ["    " + c for c in command.split(" && ")] + [
        "", "Calibration: the 64 MiB device-to-device copy read %.0f KiB (FETCH_SIZE) and wrote %.0f KiB "
            "(WRITE_SIZE); kernel figures are scaled by %.3f (fetch) and %.3f (write)." % (
                fetch[cal], write[max(copies_w)], f_scale, w_scale), "",
        "| kernel | dispatches | fetch MB / dispatch | write MB / dispatch |", "|---|---|---|---|"]
    for k, (f, w, n) in sorted(per_kernel.items()):
        lines.append("| %s | %d | %.2f | %.2f |" % (k, n, f / n / 1e6, w / n / 1e6))
    lines += ["", "Over %d hours: %.1f MB of HBM traffic per simulated hour; the roofline model's algorithmic "
                  "bytes are 32·N + 8·g² = %.1f MB per hour." % (hours, hour_bytes / 1e6, algo / 1e6)]
    return "\n".join(lines) + "\n", {"hbm_bytes_per_hour": hour_bytes, "algorithmic_bytes_per_hour": algo,
                                     "hours": hours, "fetch_scale": f_scale, "write_scale": w_scale}


def counters(paths, command, clock_ghz=2.4):
    """Per-kernel means per dispatch of the counters in one or more counter_collection.csv
    files (each a separate --pmc pass), with derived fractions where their inputs are
    present: VALU busy = SQ_ACTIVE_INST_VALU x 4 cycles over SQ_BUSY_CYCLES x the SIMDs
    per SQ; waves waiting = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES; L2 hit rate =
    TCC_HIT / (TCC_HIT + TCC_MISS); atomics past the L2 = TCC_EA0_ATOMIC / TCC_ATOMIC."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    n = collections.defaultdict(set)
    for path in paths:
        for r in csv.DictReader(open(path)):
            k = short(r["Kernel_Name"])
            if k not in HOUR_KERNELS:
                continue
            per[k][r["Counter_Name"]] += float(r["Counter_Value"])
            n[(k, r["Counter_Name"])].add(int(r["Dispatch_Id"]))
    names = sorted({c for d in per.values() for c in d})
    out = ["# Hardware counters per dispatch", "", "Commands:", ""] + ["    " + c for c in command.split(" && ")] + [
        "", "| kernel | " + " | ".join(names) + " | derived |", "|---" * (len(names) + 2) + "|"]
    for k in HOUR_KERNELS:
        if k not in per:
            continue
        mean = {c: per[k][c] / max(len(n[(k, c)]), 1) for c in per[k]}
        der = []
        if "SQ_WAVES" in mean and "SQ_INSTS_VALU" in mean:
            der.append("%.0f VALU and %.0f SALU instructions per wave" % (
                mean["SQ_INSTS_VALU"] / mean["SQ_WAVES"], mean.get("SQ_INSTS_SALU", 0) / mean["SQ_WAVES"]))
        if "SQ_WAIT_INST_ANY" in mean and "SQ_WAVE_CYCLES" in mean:
            der.append("waves waiting %.0f%% of their cycles" % (100 * mean["SQ_WAIT_INST_ANY"] / mean["SQ_WAVE_CYCLES"]))
        if "SQ_ACTIVE_INST_VALU" in mean and "SQ_BUSY_CYCLES" in mean:
            # SQ_BUSY_CYCLES sums the busy cycles of the 32 shader engines (32 SIMDs each; it
            # reads ~32 x the kernel's duration in cycles); a wave64 VALU instruction
            # occupies its SIMD 4 cycles
            der.append("VALU busy %.0f%% of SIMD cycles" % (
                100 * 4 * mean["SQ_ACTIVE_INST_VALU"] / (mean["SQ_BUSY_CYCLES"] * 32)))
        if "TCC_HIT_sum" in mean and "TCC_MISS_sum" in mean:
            der.append("L2 hit %.0f%%" % (100 * mean["TCC_HIT_sum"] / (mean["TCC_HIT_sum"] + mean["TCC_MISS_sum"])))
        if mean.get("TCC_ATOMIC_sum"):
            der.append("atomics past the L2 %.0f%%" % (100 * mean.get("TCC_EA0_ATOMIC_sum", 0) / mean["TCC_ATOMIC_sum"]))
        out.append("| %s | " % k + " | ".join("%.4g" % mean.get(c, 0) for c in names) + " | " + "; ".join(der) + " |")
    return "\n".join(out) + "\n"


def hourly(paths, command):
    """Per hour of day, the mean per dispatch of each counter of k_step and k_commit, from
    counter_collection.csv files of `bench.py --profile-hours H` (every hour dispatched on
    its own; hours are counted by k_commit dispatches), with the L2 hit rate where
    TCC_HIT / TCC_MISS are present."""
    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(list)))
    names = set()
    for path in paths:
        disp = collections.OrderedDict()
        for r in csv.DictReader(open(path)):
            d = int(r["Dispatch_Id"])
            disp.setdefault(d, [short(r["Kernel_Name"]), collections.defaultdict(float)])
            disp[d][1][r["Counter_Name"]] += float(r["Counter_Value"])
        hour = 1
        for d in sorted(disp):
            k, vals = disp[d]
            if k in ("k_step", "k_commit"):
                for c, v in vals.items():
                    per[hour % 24][k][c].append(v)
                    names.add(c)
            if k == "k_commit":
                hour += 1
    names = sorted(names)
    hit = "TCC_HIT_sum" in names and "TCC_MISS_sum" in names
    out = ["# Counters per hour of day", "", "Commands:", ""] + ["    " + c for c in command.split(" && ")] + [
        "", "| hour of day | kernel | " + " | ".join(names) + (" | L2 hit |" if hit else " |"),
        "|---" * (len(names) + 2 + (1 if hit else 0)) + "|"]
    for hd in list(range(7, 24)) + [0]:
        for k in ("k_step", "k_commit"):
            if not per[hd][k]:
                continue
            mean = {c: sum(v) / len(v) for c, v in per[hd][k].items()}
            row = "| %d | %s | " % (hd, k) + " | ".join("%.4g" % mean.get(c, 0) for c in names)
            if hit:
                row += " | %.0f%%" % (100 * mean["TCC_HIT_sum"] / max(mean["TCC_HIT_sum"] + mean["TCC_MISS_sum"], 1))
            out.append(row + " |")
    return "\n".join(out) + "\n"


def gemm(stats_csv, m, n, k, command, peak_tflops=2500.0):
    """k_gemm's mean duration from a kernel_stats.csv, as TFLOP/s and a fraction of the
    bf16 dense MFMA peak."""
    for r in csv.DictReader(open(stats_csv)):
        if short(r["Name"]) == "k_gemm":
            us = float(r["AverageNs"]) / 1e3
            tf = 2.0 * m * n * k / (us * 1e-6) / 1e12
            return ("# k_gemm (bf16 MFMA, fp32 accumulation)\n\nCommand: `%s`\n\n"
                    "| M x N x K | calls | mean µs | min µs | TFLOP/s | fraction of %.0f TFLOP/s |\n|---|---|---|---|---|---|\n"
                    "| %d x %d x %d | %s | %.2f | %.2f | %.1f | %.4f |\n" % (
                        command, peak_tflops, m, n, k, r["Calls"], us, float(r["MinNs"]) / 1e3, tf, tf / peak_tflops))
    raise SystemExit("no k_gemm in %s" % stats_csv)


def main(argv):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "traffic", "counters", "hourly", "gemm"])
    ap.add_argument("--shape", default="4096,512,512", help="gemm: M,N,K")
    ap.add_argument("paths", nargs="+")
    ap.add_argument("--command", default="")
    ap.add_argument("--agents", type=int, default=1000000)
    ap.add_argument("--grid", type=int, default=2000)
    ap.add_argument("--hours", type=int, default=48)
    ap.add_argument("--out", default="")
    ap.add_argument("--json", default="")
    a = ap.parse_args(argv)
    if a.mode == "trace":
        text = trace(a.paths[0], a.command, a.paths[1] if len(a.paths) > 1 else None, a.hours)
    elif a.mode == "counters":
        text = counters(a.paths, a.command)
    elif a.mode == "hourly":
        text = hourly(a.paths, a.command)
    elif a.mode == "gemm":
        m, n, k = (int(x) for x in a.shape.split(","))
        text = gemm(a.paths[0], m, n, k, a.command)
    else:
        text, js = traffic(a.paths[0], a.paths[1], a.agents, a.grid, a.hours, a.command)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(js, f, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)


if __name__ == "__main__":
    main(sys.argv[1:])
