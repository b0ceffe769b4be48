"""Launch-shape autotuner for the hour step.

The tunable is the number of threads per agent block (64..1024). For a workload (one or
many simulation configs) each candidate is timed over whole captured days on a fresh
handle, the fastest is kept, and the choice is stored in tuning.json under a size class
key derived from the simulation count and total agents, so later handles of the same
class start with the tuned shape:

    python -m epirust_amd.autotune tests/golden/configs/cfg4_replicates.json
"""
import json
import math
import os
import sys
import time

import epirust_amd as G

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CANDIDATES = (64, 128, 256, 512, 1024)


def size_class(configs):
    total = sum(c["population"]["Auto"]["number_of_agents"] for c in configs)
    return "sims%d-agents2^%d" % (len(configs), max(0, int(round(math.log2(max(total, 1))))))


def load_table(path=TUNING_PATH):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tuned_block(configs, path=TUNING_PATH):
    """Tuned threads per block for this size class, or None."""
    entry = load_table(path).get(size_class(configs))
    return entry["block"] if entry else None


def _open(configs, seeds):
    if len(configs) == 1:
        return G.Simulation(configs[0], seed=seeds[0])
    return G.Batch(configs, seeds)


def tune(configs, seeds=None, hours=48, repeats=2, path=TUNING_PATH, candidates=CANDIDATES, verbose=True):
    """Time each candidate block size on `hours` hours (whole days) and persist the best."""
    seeds = seeds or [c.get("seed", 0) for c in configs]
    results = {}
    for block in candidates:
        best = float("inf")
        for _ in range(repeats):
            h = _open(configs, seeds)
            h.set_block_size(block)
            t0 = time.perf_counter()
            h.step(hours)
            best = min(best, time.perf_counter() - t0)
            h.close()
        results[block] = best
        if verbose:
            print("block %4d: %.3f ms per hour" % (block, 1e3 * best / hours), flush=True)
    block = min(results, key=results.get)
    table = load_table(path)
    table[size_class(configs)] = {"block": block, "ms_per_hour": {str(k): 1e3 * v / hours for k, v in results.items()},
                                  "hours": hours}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return block, results


def main(argv):
    paths = argv or []
    for p in paths:
        with open(p) as f:
            cfg = json.load(f)
        configs = cfg if isinstance(cfg, list) else [cfg]
        block, _ = tune(configs)
        print("%s -> %s: block %d" % (p, size_class(configs), block))


if __name__ == "__main__":
    main(sys.argv[1:])
