"""Population synthesis on the device (init.hip) against the oracle's hour-0 population,
bit-exact, on the cases where the init steps have edges: a binding public-transport cap
(rank test over a scan), starting infections in all four groups with empty groups between
them (radix selection of up to four ranks), every agent infected, fewer agents than houses,
full houses, one agent, Population::Csv, lockdown essential-worker draws; and the host
restatement (EPI_FLAG_HOST_INIT) against the device at 1M and 10M agents."""
import copy
import json
import os
import time

import numpy as np
import pytest

import epirust_amd as G
from oracle import epicpu as O
from tests.gpu_helpers import FIELDS, assert_agents_equal, say

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def variant(cfg1, n=None, g=None, pt=None, wp=None, si=None, interventions=None, seed=None):
    c = copy.deepcopy(cfg1)
    if n is not None:
        c["population"]["Auto"]["number_of_agents"] = n
    if g is not None:
        c["geography_parameters"]["grid_size"] = g
    if pt is not None:
        c["population"]["Auto"]["public_transport_percentage"] = pt
    if wp is not None:
        c["population"]["Auto"]["working_percentage"] = wp
    if si is not None:
        c["starting_infections"] = dict(zip(["exposed", "infected_mild_asymptomatic", "infected_mild_symptomatic",
                                             "infected_severe"], si))
    if interventions is not None:
        c["interventions"] = interventions
    if seed is not None:
        c["seed"] = seed
    return c


CASES = {
    "transport cap binds": dict(pt=1.0, wp=1.0),
    "four groups": dict(si=(700, 500, 300, 200)),
    "empty groups between": dict(si=(0, 40, 0, 9)),
    "only severe": dict(si=(0, 0, 0, 17)),
    "everyone infected": dict(n=3000, g=120, si=(1000, 1000, 500, 500)),
    "fewer agents than houses": dict(n=37),
    "full houses": dict(n=1000, g=50),
    "one agent": dict(n=1, si=(1, 0, 0, 0)),
    "no starting infections": dict(si=(0, 0, 0, 0)),
    "essential workers": dict(interventions=[{"Lockdown": {"at_number_of_infections": 10,
                                                           "essential_workers_population": 0.35}}]),
    "other seed": dict(seed=987654321987),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_init_matches_oracle(cfg1, name):
    cfg = variant(cfg1, **CASES[name])
    say("[init] %s" % name)
    gpu = G.Simulation(cfg)
    cpu = O.OracleSim(cfg)
    a, b = gpu.agents(), cpu.agents()
    assert_agents_equal(a, b, name)
    assert gpu.counts() == cpu.counts()
    if name == "transport cap binds":  # the cap must actually bind for this case to test it
        users = int(((np.asarray(b["flags"]) & 16) != 0).sum())
        assert 0 < users < cfg["population"]["Auto"]["number_of_agents"]
    # the grid built on the device: a few hours stepped in lockstep exercise it
    g_rows, _ = gpu.step(30)
    c_rows, _ = cpu.step(30)
    assert np.array_equal(g_rows, c_rows)
    assert_agents_equal(gpu.agents(), cpu.agents(), name + " hour 30")
    gpu.close()
    cpu.close()


def test_device_init_csv_population(cfg1, tmp_path):
    from tests.test_csv_population import csv_config, write_population
    cfg = csv_config(cfg1, write_population(str(tmp_path / "pop.csv"), 9000))
    cfg["starting_infections"] = {"exposed": 5, "infected_mild_asymptomatic": 2, "infected_mild_symptomatic": 0,
                                  "infected_severe": 4}
    say("[init] csv population")
    gpu = G.Simulation(cfg)
    cpu = O.OracleSim(cfg)
    assert_agents_equal(gpu.agents(), cpu.agents(), "csv")
    gpu.close()
    cpu.close()


def test_batch_device_init_matches_single_oracle_runs(cfg1):
    cfgs = [variant(cfg1, **CASES[k]) for k in ("four groups", "fewer agents than houses", "transport cap binds")]
    say("[init] batch of %d" % len(cfgs))
    b = G.Batch(cfgs)
    for i, c in enumerate(cfgs):
        cpu = O.OracleSim(c)
        assert_agents_equal(b.agents(i), cpu.agents(), "batch member %d" % i)
        cpu.close()
    b.close()


@pytest.mark.parametrize("config", ["cfg2_1m.json", "cfg3_10m_interventions.json"])
def test_device_init_equals_host_restatement_at_scale(config):
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", config)))
    cfg["starting_infections"] = {"exposed": 5000, "infected_mild_asymptomatic": 300,
                                  "infected_mild_symptomatic": 200, "infected_severe": 100}
    say("[init] device vs host restatement, %s" % config)
    t0 = time.perf_counter()
    dev = G.Simulation(cfg)
    t1 = time.perf_counter()
    host = G.Simulation(cfg, flags=G.EPI_FLAG_DETERMINISTIC | G.EPI_FLAG_HOST_INIT)
    t2 = time.perf_counter()
    say("[init] %s: create with device init %.3f s, with host init %.3f s" % (config, t1 - t0, t2 - t1))
    a, b = dev.agents(), host.agents()
    for f in FIELDS:
        assert np.array_equal(np.asarray(a[f]), np.asarray(b[f])), f
    assert dev.counts() == host.counts()
    r1, _ = dev.step(24)
    r2, _ = host.step(24)
    assert np.array_equal(r1, r2)
    dev.close()
    host.close()
