"""Bit-exact parity of the HIP step (through the C ABI) with the CPU oracle, in
deterministic mode: every agent's state, position and counters, hour by hour."""
import copy
import json
import os

import numpy as np
import pytest

import epirust_amd as G
from oracle import epicpu as O
from tests.gpu_helpers import assert_agents_equal, assert_one_agent_per_cell, interventions_config, say

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def run_lockstep(cfg, hours, label, every=1, block=None, seed=None, office_block=None):
    say("[parity] %s: lockstep %d hours" % (label, hours))
    gpu = G.Simulation(cfg, seed=seed)
    if block:
        gpu.set_block_size(block)
    if office_block:
        gpu.set_office_block_size(office_block)
    cpu = O.OracleSim(cfg, seed=seed)
    assert_agents_equal(gpu.agents(), cpu.agents(), "%s hour 0" % label)
    h = 0
    while h < hours:
        k = min(every, hours - h)
        g_rows, g_stop = gpu.step(k)
        c_rows, c_stop = cpu.step(k)
        assert np.array_equal(g_rows, c_rows), "%s rows differ after hour %d" % (label, h)
        assert g_stop == c_stop
        h += len(c_rows)
        assert_agents_equal(gpu.agents(), cpu.agents(), "%s hour %d" % (label, h))
        if c_stop:
            break
    assert gpu.events() == cpu.events()
    gpu.close()
    cpu.close()
    return h


def test_initial_population_matches(cfg1):
    say("[parity] initial population cfg1")
    gpu = G.Simulation(cfg1)
    cpu = O.OracleSim(cfg1)
    assert_agents_equal(gpu.agents(), cpu.agents(), "init")
    assert gpu.counts() == cpu.counts()
    (rects, ver) = gpu.geometry()
    c_rects, _ = cpu.geometry()
    assert rects[:4] == c_rects[:4] and ver == 0


def test_cfg1_every_hour_bit_exact(cfg1):
    """10,000 agents on 250x250, 2,399 hours: all agents compared after every hour."""
    h = run_lockstep(cfg1, 2399, "cfg1")
    assert h > 0


def test_interventions_every_hour_bit_exact(cfg1):
    cfg = interventions_config(cfg1)
    run_lockstep(cfg, 900, "interventions")


def test_hospital_growth_and_full_hospital(cfg1):
    """Tiny hospital (beds 0) fills up, so admissions fall back to a home cell
    (allocation_map.rs:148); a BuildNewHospital then grows it to (x0,0)-(g,g)."""
    cfg = interventions_config(cfg1, n=6000, g=200)
    cfg["geography_parameters"]["hospital_beds_percentage"] = 0.0
    cfg["disease"]["percentage_severe_infected_population"] = 0.9
    cfg["interventions"] = [{"BuildNewHospital": {"spread_rate_threshold": 60}}]
    run_lockstep(cfg, 1500, "hospital", every=3)


def test_zero_transmission_closed_forms_through_gpu(cfg1):
    for asym, sev, death in ((1.0, 0.0, 0.0), (0.0, 1.0, 1.0), (0.0, 1.0, 0.0)):
        cfg = copy.deepcopy(cfg1)
        d = cfg["disease"]
        d.update(regular_transmission_rate=0.0, high_transmission_rate=0.0, percentage_asymptomatic_population=asym,
                 percentage_severe_infected_population=sev, death_rate=death)
        cfg["starting_infections"]["exposed"] = 5
        run_lockstep(cfg, 700, "zero-transmission %s" % ((asym, sev, death),), every=7)


@pytest.mark.parametrize("block", [64, 128, 512, 1024])
def test_block_shapes_do_not_change_results(cfg1, block):
    cfg = interventions_config(cfg1, n=3000, g=140)
    run_lockstep(cfg, 300, "block%d" % block, every=5, block=block)


@pytest.mark.parametrize("home,office", [(512, 64), (128, 1024)])
def test_office_block_size_apart_from_home_does_not_change_results(cfg1, home, office):
    """The office layout's blocks (hours of day 8..15) launched at another size than the
    home layout's. This config has 28 offices of ~75 workers: 64 slots split each of them
    over two blocks, 1,024 slots pack 13 whole offices into one."""
    cfg = interventions_config(cfg1, n=3000, g=140)
    run_lockstep(cfg, 300, "home%d office%d" % (home, office), every=5, block=home, office_block=office)


def test_sparse_offices_fill_blocks_to_the_filter_limit(cfg1):
    """3,000 agents on 600x600: 720 offices of ~3 workers each, so each office-layout block
    of workers holds the claim filter's maximum of 20 offices (the other blocks' lanes run
    the agents that do not work in an office)."""
    cfg = interventions_config(cfg1, n=3000, g=600)
    run_lockstep(cfg, 240, "sparse offices", every=8)


def test_whole_run_with_day_graphs_matches(cfg1):
    """One epi_step call covering the run: whole days replay the captured hipGraph."""
    say("[parity] cfg1 graph-replayed run")
    cfg = copy.deepcopy(cfg1)
    cfg["disease"]["regular_transmission_rate"] = 0.1
    cfg["disease"]["high_transmission_rate"] = 0.4
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 50, "essential_workers_population": 0.1}},
                            {"Vaccinate": {"at_hour": 480, "percent": 0.2}}]
    gpu = G.Simulation(cfg)
    cpu = O.OracleSim(cfg)
    g_rows, g_stop = gpu.step(2399)
    c_rows, c_stop = cpu.step(2399)
    assert g_stop == c_stop
    assert np.array_equal(g_rows, c_rows)
    assert_agents_equal(gpu.agents(), cpu.agents(), "after whole run")
    assert gpu.events() == cpu.events()


def test_seeds_change_the_run(cfg1):
    say("[parity] seeds")
    a = G.Simulation(cfg1, seed=1)
    b = G.Simulation(cfg1, seed=2)
    ra, _ = a.step(300)
    rb, _ = b.step(300)
    assert not np.array_equal(a.agents()["x"], b.agents()["x"])
    c = O.OracleSim(cfg1, seed=2)
    rc, _ = c.step(300)
    assert np.array_equal(rb, rc)


def test_batch_of_mixed_replicates_matches_single_runs(cfg1):
    """Replicates of different sizes and parameters in one launch: each equals its own run."""
    rng = np.random.default_rng(7)
    cfgs, seeds = [], []
    for r in range(6):
        c = copy.deepcopy(cfg1)
        n = int(rng.integers(2000, 9000))
        c["population"]["Auto"]["number_of_agents"] = n
        c["geography_parameters"]["grid_size"] = int(round((n / 0.25) ** 0.5))
        c["disease"]["regular_transmission_rate"] = float(rng.uniform(0.05, 0.2))
        c["disease"]["high_transmission_rate"] = float(rng.uniform(0.2, 0.5))
        c["disease"]["death_rate"] = float(rng.uniform(0.05, 0.3))
        c["hours"] = int(rng.integers(300, 700))
        c["starting_infections"]["exposed"] = 5
        if r % 2:
            c["interventions"] = [{"Lockdown": {"at_number_of_infections": 30, "essential_workers_population": 0.1}}]
        cfgs.append(c)
        seeds.append(1000 + r)
    say("[parity] batch of %d" % len(cfgs))
    batch = G.Batch(cfgs, seeds)
    rows, stopped = batch.step(800)
    for i, c in enumerate(cfgs):
        cpu = O.OracleSim(c, seed=seeds[i])
        c_rows, c_stop = cpu.step(800)
        assert np.array_equal(rows[i], c_rows), "replicate %d rows" % i
        assert stopped[i] == c_stop
        assert_agents_equal(batch.agents(i), cpu.agents(), "replicate %d" % i)
        assert batch.events(i) == cpu.events()


def test_night_launch_in_a_batch_with_sweeps(cfg1):
    """Day graphs run hours of day 1..6 as one launch when no vaccination falls at 1..5 and
    no run ends at 1..6: a batch of runs ending at hours of day 0, 7 and 23, with a
    lockdown (decided at hour of day 0, applied in the night) and vaccinations at hours of
    day 0 and 6. Each equals its own run."""
    cfgs, seeds = [], []
    for r, hours in enumerate((24 * 12 + 1, 24 * 15 + 8, 24 * 20)):
        c = copy.deepcopy(cfg1)
        c["hours"] = hours
        c["disease"]["regular_transmission_rate"] = 0.1 + 0.05 * r
        c["disease"]["high_transmission_rate"] = 0.4
        c["starting_infections"]["exposed"] = 5
        c["interventions"] = [{"Lockdown": {"at_number_of_infections": 30, "essential_workers_population": 0.1}},
                              {"Vaccinate": {"at_hour": 240, "percent": 0.3}},
                              {"Vaccinate": {"at_hour": 270, "percent": 0.2}}]
        cfgs.append(c)
        seeds.append(2000 + r)
    say("[parity] night launch: batch of %d" % len(cfgs))
    batch = G.Batch(cfgs, seeds)
    rows, stopped = batch.step(500)
    for i, c in enumerate(cfgs):
        cpu = O.OracleSim(c, seed=seeds[i])
        c_rows, c_stop = cpu.step(500)
        assert np.array_equal(rows[i], c_rows), "run %d rows" % i
        assert stopped[i] == c_stop
        assert_agents_equal(batch.agents(i), cpu.agents(), "run %d" % i)
        assert batch.events(i) == cpu.events()


def test_quiet_hours_launch_with_a_vaccination_at_hour_of_day_1(cfg1):
    """A vaccination at hour of day 1 keeps the night launch off (its sweep is pending at
    hour of day 2), but hours of day 3..6 still run as one epilogue-only launch in the day
    graphs. Compared with the oracle over one graph-replayed run."""
    cfg = interventions_config(cfg1)
    cfg["hours"] = 24 * 8
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 20, "essential_workers_population": 0.3}},
                            {"Vaccinate": {"at_hour": 49, "percent": 0.3}},
                            {"Vaccinate": {"at_hour": 97, "percent": 0.2}}]
    say("[parity] quiet hours")
    gpu = G.Simulation(cfg)
    cpu = O.OracleSim(cfg)
    g_rows, g_stop = gpu.step(24 * 7)
    c_rows, c_stop = cpu.step(24 * 7)
    assert np.array_equal(g_rows, c_rows) and g_stop == c_stop
    assert_agents_equal(gpu.agents(), cpu.agents(), "quiet hours")
    assert gpu.events() == cpu.events()


def test_cfg2_1m_first_two_days_bit_exact():
    """1,000,000 agents on 2000x2000: the first 48 hours compared agent by agent."""
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", "cfg2_1m.json")))
    run_lockstep(cfg, 48, "cfg2", every=12)


def test_cfg2_full_run_rows_match_oracle_fixture():
    """Full 2,399-hour run at 1M agents: the epicurve rows and a checksum of the final
    agent state equal the oracle's, stored as a fixture (tests/golden/make_golden.py)."""
    path = os.path.join(HERE, "golden", "cfg2_1m_oracle.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    gold = np.load(path)
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", "cfg2_1m.json")))
    say("[parity] cfg2 full run vs fixture")
    gpu = G.Simulation(cfg)
    rows, stopped = gpu.step(2399)
    assert np.array_equal(rows, gold["rows"])
    a = gpu.agents()
    from tests.golden.make_golden import state_digest
    assert state_digest(a) == str(gold["digest"])


def test_cfg3_10m_first_day_bit_exact():
    """10,000,000 agents on 6000x6000 with interventions configured: 30 hours agent by agent."""
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", "cfg3_10m_interventions.json")))
    run_lockstep(cfg, 30, "cfg3", every=15)


def test_cfg3_full_run_matches_oracle_fixture():
    """The whole 4,799-hour run at 10,000,000 agents with lockdown, hospital growth and 170
    daily vaccinations: every row and every intervention event equal the oracle's, and the
    digest of all agents' state equals the oracle's after the hour that follows each
    lockdown / unlock / hospital-growth event, after hours 504, 505 and 721, and after the
    last hour (tests/golden/make_golden.py, cfg3)."""
    path = os.path.join(HERE, "golden", "cfg3_10m_oracle.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    gold = np.load(path)
    from tests.golden.make_golden import state_digest
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", "cfg3_10m_interventions.json")))
    say("[parity] cfg3 full run vs fixture, digests at %s" % gold["digest_hours"].tolist())
    gpu = G.Simulation(cfg)
    rows, h = [], 0
    for target, digest in zip(gold["digest_hours"].tolist(), gold["digests"].tolist()):
        r, stopped = gpu.step(target - h)
        rows.append(r)
        h += len(r)
        assert h == target, (h, target)
        a = gpu.agents()
        assert_one_agent_per_cell(a, "cfg3 hour %d" % h)
        assert state_digest(a) == digest, "cfg3 state digest after hour %d" % h
        say("[parity] cfg3 digest after hour %d ok" % h)
    rows = np.concatenate(rows)
    assert np.array_equal(rows, gold["rows"])
    ev = np.array(gpu.events(), dtype=np.uint32).reshape(-1, 2)
    assert np.array_equal(ev, gold["events"])
    kinds = ev[:, 1].tolist()
    assert kinds.count(0) == 170 and kinds.count(1) >= 1 and kinds.count(3) == 1
    gpu.close()


def test_cfg4_batched_replicates_match_their_own_runs():
    """The 64-replicate batch in one launch: 8 of its replicates compared with their own
    oracle runs (rows and agents) after 72 hours, then every row of all 64 replicates over
    the whole 2,399 hours and each one's final state digest against its own oracle run
    (tests/golden/make_golden.py, cfg4)."""
    path = os.path.join(HERE, "golden", "cfg4_replicates_oracle.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    gold = np.load(path)
    from tests.golden.make_golden import state_digest
    cfgs = json.load(open(os.path.join(HERE, "golden", "configs", "cfg4_replicates.json")))
    say("[parity] cfg4 batch of %d, first 72 hours" % len(cfgs))
    batch = G.Batch(cfgs)
    rows, stopped = batch.step(72)
    for i in range(0, 64, 8):
        cpu = O.OracleSim(cfgs[i])
        c_rows, c_stop = cpu.step(72)
        assert np.array_equal(rows[i], c_rows), "replicate %d rows" % i
        assert_agents_equal(batch.agents(i), cpu.agents(), "replicate %d" % i)
        cpu.close()
    say("[parity] cfg4 full run vs fixture")
    rest, stopped = batch.step(2399 - 72)
    for i, c in enumerate(cfgs):
        r = np.concatenate([rows[i], rest[i]]) if len(rest[i]) else rows[i]
        k = int(gold["n_rows"][i])
        assert len(r) == k and np.array_equal(r, gold["rows"][i, :k]), "replicate %d rows" % i
        assert stopped[i]
        a = batch.agents(i)
        assert_one_agent_per_cell(a, "replicate %d" % i)
        assert state_digest(a) == str(gold["digests"][i]), "replicate %d final state" % i


def test_csv_population_bit_exact(cfg1, tmp_path):
    """Population::Csv read by the C++ host and by the oracle's Python reader."""
    from tests.test_csv_population import csv_config, write_population
    path = write_population(str(tmp_path / "pop.csv"), 8000)
    cfg = csv_config(cfg1, path)
    cfg["disease"]["regular_transmission_rate"] = 0.2
    cfg["disease"]["high_transmission_rate"] = 0.5
    run_lockstep(cfg, 400, "csv population", every=8)


def test_sweeps_at_sleep_hours_and_at_the_last_hour(cfg1):
    """Sweeps are applied by the next hour's step (and re-applied before read-back): three
    vaccinations on consecutive sleep hours (3, 4, 5 of day 2), a lockdown, and one
    vaccination at the very last hour. Compared with the oracle both through one
    graph-replayed run and hour by hour."""
    cfg = interventions_config(cfg1)
    cfg["hours"] = 600
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 20, "essential_workers_population": 0.3}},
                            {"Vaccinate": {"at_hour": 51, "percent": 0.2}},
                            {"Vaccinate": {"at_hour": 52, "percent": 0.3}},
                            {"Vaccinate": {"at_hour": 53, "percent": 0.1}},
                            {"Vaccinate": {"at_hour": 599, "percent": 0.5}}]
    say("[parity] sweeps: whole run")
    gpu = G.Simulation(cfg)
    cpu = O.OracleSim(cfg)
    g_rows, g_stop = gpu.step(599)
    c_rows, c_stop = cpu.step(599)
    assert np.array_equal(g_rows, c_rows) and g_stop == c_stop
    assert_agents_equal(gpu.agents(), cpu.agents(), "sweeps whole run")
    assert gpu.events() == cpu.events()
    gpu.close()
    cpu.close()
    h = run_lockstep(cfg, 80, "sweeps hourly", every=1)
    assert h == 80


@pytest.mark.parametrize("chunk", [20, 7])
def test_partial_days_match_day_graphs_and_plain_launches(cfg1, chunk):
    """Step calls that end inside a day enqueue their hours directly, with hours of day 1..6
    (or 3..6) as one launch when the call holds all of them; EPI_FLAG_NO_GRAPH launches
    every hour on its own. Both equal the oracle, with a lockdown decided at an hour of
    day 0 and vaccinations at hours of day 0 and 6 (so the night launch is allowed)."""
    cfg = interventions_config(cfg1)
    cfg["hours"] = 24 * 14
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 30, "essential_workers_population": 0.2}},
                            {"Vaccinate": {"at_hour": 96, "percent": 0.3}},
                            {"Vaccinate": {"at_hour": 150, "percent": 0.2}}]
    say("[parity] partial days, %d-hour calls" % chunk)
    for flags in (G.EPI_FLAG_DETERMINISTIC, G.EPI_FLAG_DETERMINISTIC | G.EPI_FLAG_NO_GRAPH):
        gpu = G.Simulation(cfg, flags=flags)
        cpu = O.OracleSim(cfg)
        h = 0
        while h < cfg["hours"] - 1:
            g_rows, g_stop = gpu.step(chunk)
            c_rows, c_stop = cpu.step(chunk)
            assert np.array_equal(g_rows, c_rows), "flags %d: rows after hour %d" % (flags, h)
            assert g_stop == c_stop
            h += len(c_rows)
            if h % (3 * chunk) == 0 or c_stop:
                assert_agents_equal(gpu.agents(), cpu.agents(), "flags %d hour %d" % (flags, h))
            if c_stop:
                break
        assert gpu.events() == cpu.events()
        gpu.close()
        cpu.close()
