"""Launch-shape autotuner for the hour step.

Two launch sizes are tuned: the threads per agent block of the home layout (which also
sets the night launch's) and those of the office layout used at hours of day 8..15, whose
worker blocks pack whole offices into that many slots, so the second size also sets how
many offices share a block's claim filter. For a workload (one or many simulation
configs) the search is by coordinates: every home size with the office size equal to it,
then every office size beside the best home size. Each candidate is timed over whole
captured days on a fresh handle, the fastest pair is kept, and it is stored in tuning.json
under a size class key derived from the simulation count and total agents, so later
handles of the same class start with the tuned shape:

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


def tuned_shape(configs, path=TUNING_PATH):
    """(home, office) threads per block tuned for this size class, or None."""
    entry = load_table(path).get(size_class(configs))
    if not entry:
        return None
    return entry["block"], entry.get("office_block", entry["block"])


def tuned_block(configs, path=TUNING_PATH):
    """Tuned threads per block of the home layout for this size class, or None."""
    shape = tuned_shape(configs, path)
    return shape[0] if shape else None


def apply_shape(handle, configs, block=None, path=TUNING_PATH):
    """Give a handle the tuned shape of its size class; `block`, if given, sets both
    layouts instead. Returns the (home, office) pair applied, or None for the defaults."""
    if block:
        handle.set_block_size(block)
        return block, block
    shape = tuned_shape(configs, path)
    if shape:
        handle.set_block_size(shape[0])
        if shape[1] != shape[0]:
            handle.set_office_block_size(shape[1])
    return shape


def _open(configs, seeds):
    if len(configs) == 1:
        return G.Simulation(configs[0], seed=seeds[0])
    return G.Batch(configs, seeds)


def _time(configs, seeds, hours, repeats, block, office):
    best = float("inf")
    for _ in range(repeats):
        h = _open(configs, seeds)
        h.set_block_size(block)
        if office != block:
            h.set_office_block_size(office)
        t0 = time.perf_counter()
        h.step(hours)
        best = min(best, time.perf_counter() - t0)
        h.close()
    return best


def tune(configs, seeds=None, hours=48, repeats=2, path=TUNING_PATH, candidates=CANDIDATES, verbose=True):
    """Time the candidate shapes on `hours` hours (whole days) and persist the best pair.
    Returns (home, office, {(home, office): seconds})."""
    seeds = seeds or [c.get("seed", 0) for c in configs]
    results = {}

    def run(block, office):
        if (block, office) not in results:
            results[(block, office)] = _time(configs, seeds, hours, repeats, block, office)
            if verbose:
                print("home %4d office %4d: %.3f ms per hour" % (block, office, 1e3 * results[(block, office)] / hours),
                      flush=True)
        return results[(block, office)]

    block = min(candidates, key=lambda b: run(b, b))
    office = min(candidates, key=lambda o: run(block, o))
    table = load_table(path)
    table[size_class(configs)] = {
        "block": block, "office_block": office, "hours": hours,
        "ms_per_hour": {str(b): 1e3 * results[(b, b)] / hours for b in candidates},
        "office_ms_per_hour": {str(o): 1e3 * results[(block, o)] / hours for o in candidates}}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return block, office, results


def main(argv):
    paths = argv or []
    for p in paths:
        with open(p) as f:
            cfg = json.load(f)
        configs = cfg if isinstance(cfg, list) else [cfg]
        block, office, _ = tune(configs)
        print("%s -> %s: home block %d, office block %d" % (p, size_class(configs), block, office))


if __name__ == "__main__":
    main(sys.argv[1:])
