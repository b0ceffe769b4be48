"""Deterministic mode against a reference-structured run, in distribution.

Deterministic mode changes how randomness is consumed: fixed Philox counters per decision
site, an OR over per-neighbour Bernoullis, and lowest-id-wins collisions. The reference
instead consumes one thread_rng stream in call order: rand 0.8.5 gen_range with rejection
(area.rs:76-81), IteratorRandom::choose (citizen/mod.rs:424-429), a short-circuit find with
a gen_bool per infectious neighbour (default_disease_handler.rs:72-79). It also merges
collisions in hash order (allocation_map.rs:82-125). The oracle's reference-order mode
restates that structure (a uniformly random merge order stands in for the hash order), so
the two modes must give the same epicurve distribution. Checked over 48 seeds per mode:
peak infected, final size (infected + hospitalized + recovered + deceased at the last
hour), the hour of the peak and the run length, each by a two-sample z statistic
(|z| < 3.5).

The check's power is shown by planting a mistake in the oracle's input: a 25% higher
transmission rate in deterministic mode moves at least one statistic beyond |z| > 5.
"""
import copy

import numpy as np
import pytest

from oracle import epicpu as O
from tests.conftest import load_config

SEEDS = 48


def moderate(cfg1, scale=1.0):
    c = copy.deepcopy(cfg1)
    c["population"]["Auto"]["number_of_agents"] = 2000
    c["geography_parameters"]["grid_size"] = 90
    c["hours"] = 3600
    c["disease"]["regular_transmission_rate"] = 0.05 * scale
    c["disease"]["high_transmission_rate"] = 0.3 * scale
    c["starting_infections"]["exposed"] = 10
    return c


def ensemble(cfg, reference_order, seeds):
    out = []
    for sd in seeds:
        s = O.OracleSim(cfg, seed=sd)
        s.set_threads(8)
        if reference_order:
            s.set_reference_order(True)
        rows, _ = s.step(cfg["hours"])
        s.close()
        infected = rows[:, 3].astype(np.float64)
        out.append((infected.max(), rows[-1, 3:7].sum(), rows[int(np.argmax(infected)), 0], len(rows)))
    return np.array(out, dtype=np.float64)


def z_scores(a, b):
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    return (a.mean(axis=0) - b.mean(axis=0)) / se


@pytest.fixture(scope="module")
def reference_ensemble():
    return ensemble(moderate(load_config("cfg1_10k.json")), True, range(1001, 1001 + SEEDS))


def test_reference_order_mode_is_a_different_stream(cfg1):
    cfg = moderate(cfg1)
    a, b = O.OracleSim(cfg, seed=3), O.OracleSim(cfg, seed=3)
    b.set_reference_order(True)
    ra, _ = a.step(400)
    rb, _ = b.step(400)
    assert not np.array_equal(ra, rb)
    assert (rb[:, 1:].sum(axis=1) == 2000).all()  # same conservation law


def test_deterministic_mode_matches_reference_structure_in_distribution(cfg1, reference_ensemble):
    det = ensemble(moderate(cfg1), False, range(1, 1 + SEEDS))
    z = z_scores(det, reference_ensemble)
    names = ["peak infected", "final size", "peak hour", "hours run"]
    assert np.all(np.abs(z) < 3.5), dict(zip(names, np.round(z, 2)))


def test_planted_transmission_error_is_detected(cfg1, reference_ensemble):
    wrong = ensemble(moderate(cfg1, scale=1.25), False, range(1, 1 + SEEDS))
    z = z_scores(wrong, reference_ensemble)
    assert np.max(np.abs(z)) > 5.0, np.round(z, 2)
