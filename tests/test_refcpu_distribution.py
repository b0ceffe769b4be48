"""The reference-structure CPU engine (oracle/librefcpu.so) and the distributional check
at the 10,000-agent config.

refcpu runs the oracle's routine on the reference's store: citizens held by value in two
FNV hash maps keyed by cell (allocation_map.rs:44-65), a threaded map over the current
map's buckets and a single-thread merge in bucket order (allocation_map.rs:82-134), so
collision winners follow hash order, and every agent-hour draws from one sequential
stream in the reference's call order. It stands in for the unmodified engine, which
cannot be built here.

Criterion (the 10,000-agent / 250x250 / 2,399-hour config, seeds 1..32 in each mode): the
deterministic-mode ensemble's means of the peak day and of the final size lie inside the
95% interval of the refcpu ensemble's mean (mean +- t(0.975, 31) * sd / sqrt(32)).
  peak day   = index of the largest `infected` among the daily rows, hour % 24 == 1
               (engine/plot/plot.py:33), a run that stopped early carried forward;
  final size = infected + hospitalized + recovered + deceased of the last row
               (engine/plot/update_total_infections.py:19).
Deterministic mode is the GPU's: the GPU parity tests show it bit-exact with the oracle.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import epicpu as O
from tests.conftest import load_config

SEEDS = range(1, 33)
T_975_31 = 2.0395  # Student t quantile, 31 degrees of freedom


def run_one(args):
    kind, seed = args
    cfg = load_config("cfg1_10k.json")
    s = (O.RefCpuSim if kind == "refcpu" else O.OracleSim)(cfg, seed=seed)
    rows, _ = s.step(cfg["hours"])
    s.close()
    return kind, seed, rows


def stats(rows):
    daily = rows[rows[:, 0] % 24 == 1]
    return float(np.argmax(daily[:, 3])), float(rows[-1, 3:7].sum())


@pytest.fixture(scope="module")
def ensembles():
    jobs = [(k, s) for s in SEEDS for k in ("det", "refcpu")]
    out = {"det": {}, "refcpu": {}}
    with cf.ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for kind, seed, rows in ex.map(run_one, jobs):
            out[kind][seed] = rows
    return out


def test_refcpu_conserves_agents_and_cells(cfg1):
    cfg = dict(cfg1)
    cfg["population"] = {"Auto": dict(cfg1["population"]["Auto"], number_of_agents=3000)}
    cfg["geography_parameters"] = dict(cfg1["geography_parameters"], grid_size=120)
    cfg["disease"] = dict(cfg1["disease"], regular_transmission_rate=0.2, high_transmission_rate=0.5)
    s = O.RefCpuSim(cfg, seed=11)
    s.set_threads(4)
    rows, _ = s.step(300)
    assert (rows[:, 1:].sum(axis=1) == 3000).all()
    a = s.agents()
    cells = set(zip(a["x"].tolist(), a["y"].tolist()))
    assert len(cells) == 3000  # one citizen per key of the map
    assert rows[-1, 2:].sum() > rows[0, 2:].sum()  # the epidemic spreads
    s.close()


def test_refcpu_starts_from_the_same_population(cfg1):
    """Population synthesis is shared: the two stores differ only from the first hour on."""
    a, b = O.OracleSim(cfg1, seed=3), O.RefCpuSim(cfg1, seed=3)
    ga, gb = a.agents(), b.agents()
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k
    a.close()
    b.close()


def test_refcpu_results_do_not_depend_on_threads(cfg1):
    cfg = dict(cfg1)
    cfg["population"] = {"Auto": dict(cfg1["population"]["Auto"], number_of_agents=6000)}
    cfg["disease"] = dict(cfg1["disease"], regular_transmission_rate=0.2, high_transmission_rate=0.5)
    out = []
    for t in (1, 5):
        s = O.RefCpuSim(cfg, seed=2)
        s.set_threads(t)
        out.append(s.step(200)[0])
        s.close()
    assert np.array_equal(out[0], out[1])


def test_deterministic_mode_inside_refcpu_interval_at_10k_agents(ensembles):
    det = np.array([stats(ensembles["det"][s]) for s in SEEDS])
    ref = np.array([stats(ensembles["refcpu"][s]) for s in SEEDS])
    half = T_975_31 * ref.std(axis=0, ddof=1) / np.sqrt(len(ref))
    lo, hi = ref.mean(axis=0) - half, ref.mean(axis=0) + half
    m = det.mean(axis=0)
    detail = {"peak day": (m[0], lo[0], hi[0]), "final size": (m[1], lo[1], hi[1])}
    assert np.all((lo <= m) & (m <= hi)), detail
