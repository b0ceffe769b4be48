"""Pins the CPU oracle against the reference's own known-answer tests.

Each test names the reference test it restates (paths relative to the reference
repository). The oracle is the checker for the HIP product, so it is pinned here
first, on the CPU.
"""
import ctypes
import copy
import json
import os

import numpy as np
import pytest

from oracle import epicpu as E

HERE = os.path.dirname(os.path.abspath(__file__))
I32 = ctypes.c_int32


def rects(g):
    out = (I32 * 16)()
    E.lib().epicpu_define_geography(g, out)
    return [tuple(out[4 * i:4 * i + 4]) for i in range(4)]


def test_philox_known_answers():
    # Random123 published known-answer vectors for Philox4x32-10 (kat_vectors).
    assert E.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert E.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert E.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_define_geography_g10():
    # engine/src/geography/mod.rs:76-90 should_define_geography
    housing, transport, work, hospital = rects(10)
    assert housing == (0, 0, 3, 10)
    assert transport == (4, 0, 5, 10)
    assert work == (6, 0, 7, 10)
    assert hospital == (8, 0, 8, 10)


def test_define_geography_g100():
    # engine/src/geography/grid.rs:377-385 (start offsets used by the hospital tests)
    housing, transport, work, hospital = rects(100)
    assert hospital[:2] == (80, 0)
    assert hospital[2:] == (89, 100)
    assert housing == (0, 0, 39, 100)


def test_area_iteration_order():
    # engine/src/geography/area.rs:165-186 should_iterate_over_points_in_area
    out = (I32 * 32)()
    n = E.lib().epicpu_area_iter((I32 * 4)(0, 0, 2, 2), out, 16)
    pts = [tuple(out[2 * i:2 * i + 2]) for i in range(n)]
    assert pts == [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2)]
    n = E.lib().epicpu_area_iter((I32 * 4)(1, 1, 2, 2), out, 16)
    assert [tuple(out[2 * i:2 * i + 2]) for i in range(n)] == [(1, 1), (2, 1), (1, 2), (2, 2)]


def test_area_factory():
    # engine/src/geography/area.rs:199-225 should_create_areas
    out = (I32 * 400)()
    n = E.lib().epicpu_area_factory(10, 0, 21, 10, 3, out, 100)
    b = [tuple(out[4 * i:4 * i + 4]) for i in range(n)]
    assert n == 12
    assert b[0] == (10, 0, 12, 2)
    assert b[5] == (13, 3, 15, 5)
    assert b[-1] == (19, 6, 21, 8)

    def contains(r, x, y):
        return r[0] <= x <= r[2] and r[1] <= y <= r[3]

    for y in (0, 1, 2):
        assert contains(b[0], 12, y) and not contains(b[1], 12, y) and contains(b[1], 13, y)


def test_neighbours_in_area():
    # engine/src/geography/area.rs:228-234 should_get_neighbor_of
    out = (I32 * 16)()
    n = E.lib().epicpu_neighbors_in_area(4, 5, (I32 * 4)(0, 0, 5, 5), out)
    pts = [tuple(out[2 * i:2 * i + 2]) for i in range(n)]
    assert n == 5
    assert (4, 4) in pts and (4, 6) not in pts


def test_moore_order():
    # engine/src/geography/point.rs:85-102 should_iterate_over_neighbor_cells
    out = (I32 * 16)()
    n = E.lib().epicpu_neighbors_in_area(1, 1, (I32 * 4)(-10, -10, 10, 10), out)
    pts = [tuple(out[2 * i:2 * i + 2]) for i in range(n)]
    assert pts == [(0, 0), (1, 0), (2, 0), (0, 1), (2, 1), (0, 2), (1, 2), (2, 2)]


def test_hospital_resize_and_increase():
    # engine/src/geography/grid.rs:377-421
    r = (I32 * 4)(*rects(100)[3])
    E.lib().epicpu_increase_hospital_size(r, 120)
    assert tuple(r) == (80, 0, 120, 120)
    r = (I32 * 4)(*rects(100)[3])
    E.lib().epicpu_resize_hospital(r, 1000, 0.02, 0.01)
    assert tuple(r) == (80, 0, 89, 3)
    r = (I32 * 4)(*rects(100)[3])
    E.lib().epicpu_resize_hospital(r, 50000, 0.02, 0.01)
    assert tuple(r) == (80, 0, 89, 100)


def test_first_vacant_hospital_cell():
    # engine/src/allocation_map.rs:444-495 (goto_hospital and the full-hospital fallback)
    out = (I32 * 2)()
    occupied = (I32 * 4)(0, 1, 1, 0)
    assert E.lib().epicpu_first_vacant((I32 * 4)(2, 2, 4, 4), occupied, 2, out) == 1
    assert tuple(out) == (2, 2)
    full = [(x, y) for y in range(2, 5) for x in range(2, 5)]
    flat = (I32 * 18)(*[v for p in full for v in p])
    assert E.lib().epicpu_first_vacant((I32 * 4)(2, 2, 4, 4), flat, 9, out) == 0
    # occupied first cell -> the next one in row-major order
    flat = (I32 * 2)(2, 2)
    assert E.lib().epicpu_first_vacant((I32 * 4)(2, 2, 4, 4), flat, 1, out) == 1 and tuple(out) == (3, 2)


def _small_pox_config():
    with open(os.path.join(HERE, "golden", "reference", "diseases_small_pox.json")) as f:
        sp = json.load(f)["small_pox"]
    with open(os.path.join(HERE, "golden", "configs", "cfg1_10k.json")) as f:
        cfg = json.load(f)
    cfg["disease"] = sp
    return E.config_from_json(cfg)


def test_transmission_windows_small_pox():
    # common/src/disease/mod.rs:145-163 get_current_transmission_rate / to_be_quarantined
    c = _small_pox_config()
    L = E.lib()
    assert L.epicpu_transmission_rate(ctypes.byref(c), 12) == 0.05
    assert L.epicpu_transmission_rate(ctypes.byref(c), 22) == 0.5
    assert not L.epicpu_is_to_be_hospitalized(ctypes.byref(c), 12)
    assert L.epicpu_is_to_be_hospitalized(ctypes.byref(c), 22)
    # window edges: (start, end] for both windows, 0 outside
    assert L.epicpu_transmission_rate(ctypes.byref(c), 10) == 0.0
    assert L.epicpu_transmission_rate(ctypes.byref(c), 16) == 0.05
    assert L.epicpu_transmission_rate(ctypes.byref(c), 17) == 0.5
    assert L.epicpu_transmission_rate(ctypes.byref(c), 23) == 0.0
    # (day + immunity) wraps as u32 when negative (citizen/mod.rs:184): rate 0
    assert L.epicpu_transmission_rate(ctypes.byref(c), 0xFFFFFFFF) == 0.0


def test_lockdown_triggers():
    # engine/src/interventions/lockdown.rs:126-191
    L = E.lib()
    assert not L.epicpu_lockdown_should_apply(0, 20, 0, 1)
    assert not L.epicpu_lockdown_should_apply(0, 20, 22, 20)
    assert not L.epicpu_lockdown_should_apply(0, 20, 28, 21)
    assert L.epicpu_lockdown_should_apply(0, 20, 48, 21)
    assert not L.epicpu_lockdown_should_apply(1, 20, 48, 25)  # already locked down
    until = 48 + 7 * 24
    assert not any(L.epicpu_lockdown_should_unlock(1, until, h) for h in range(48, until))
    assert L.epicpu_lockdown_should_unlock(1, until, until + 21 * 24)
    assert L.epicpu_lockdown_should_unlock(1, 28, 532)
    assert not L.epicpu_lockdown_should_unlock(0, 28, 532)


def test_hospital_trigger():
    # engine/src/interventions/hospital.rs:96-160
    L = E.lib()
    n = L.epicpu_hospital_counts_updated(0, 0, 0)
    assert not L.epicpu_hospital_should_apply(0, 1, 10, n, 0)
    n = L.epicpu_hospital_counts_updated(n, 24, 20)
    assert L.epicpu_hospital_should_apply(0, 1, 10, n, 0)
    assert not L.epicpu_hospital_should_apply(0, 0, 10, n, 0)       # intervention absent
    assert not L.epicpu_hospital_should_apply(1, 1, 10, n, 0)       # already applied
    n = L.epicpu_hospital_counts_updated(0, 24, 5)
    assert not L.epicpu_hospital_should_apply(0, 1, 10, n, 0)       # below threshold
    # the alternating quirk: n := infected - n (saturating) every start of day
    assert L.epicpu_hospital_counts_updated(20, 48, 25) == 5
    assert L.epicpu_hospital_counts_updated(20, 48, 15) == 0
    assert L.epicpu_hospital_counts_updated(20, 47, 15) == 20


def test_bernoulli_threshold_matches_rand():
    # rand 0.8.5 Bernoulli::new: p_int = (p * 2^64) as u64; p == 1 -> ALWAYS_TRUE
    L = E.lib()
    assert L.epicpu_bernoulli_threshold(0.0) == 0
    assert L.epicpu_bernoulli_threshold(0.5) == 1 << 63
    assert L.epicpu_bernoulli_threshold(0.25) == 1 << 62
    assert L.epicpu_bernoulli_threshold(1.0) == 0xFFFFFFFFFFFFFFFF
    p = 0.025
    assert L.epicpu_bernoulli_threshold(p) == int(np.floor(np.float64(p) * np.float64(2.0 ** 64)))


def test_config_parse_auto_pop():
    # common/src/config/mod.rs:160-189 should_read_config_with_auto_population
    with open(os.path.join(HERE, "golden", "reference", "auto_pop.json")) as f:
        c = E.config_from_json(f.read())
    assert c.number_of_agents == 10000
    assert c.public_transport_percentage == 0.2 and c.working_percentage == 0.7
    assert (c.regular_transmission_start_day, c.high_transmission_start_day, c.last_day) == (5, 20, 40)
    assert c.death_rate == 0.035 and c.exposed_duration == 48
    assert c.grid_size == 250 and c.hospital_beds_percentage == 0.003 and c.hours == 10000
    assert (c.si_infected_mild_asymptomatic, c.si_infected_mild_symptomatic, c.si_infected_severe, c.si_exposed) == \
        (2, 3, 4, 5)
    # engine/src/interventions/vaccination.rs:77-98: {5000: 0.2}
    assert c.n_vaccinate == 1 and c.vaccinate_at_hour[0] == 5000 and c.vaccinate_percent[0] == 0.2


def test_starting_infections_split(cfg1):
    # engine/src/citizen/citizen_factory.rs:189-215 should_set_starting_infections (2/3/4/5)
    cfg = copy.deepcopy(cfg1)
    cfg["starting_infections"] = {"infected_mild_asymptomatic": 2, "infected_mild_symptomatic": 3,
                                  "infected_severe": 4, "exposed": 5}
    s = E.OracleSim(cfg)
    a = s.agents()
    inf = a["state"] == 2
    assert (a["state"] == 1).sum() == 5
    assert (inf & (a["severity"] == 1)).sum() == 2
    assert (inf & (a["severity"] == 2)).sum() == 3
    assert (inf & (a["severity"] == 3)).sum() == 4
    assert (a["infection_day"][inf] == 1).all()
    # utils/util.rs:45-51 counts_at_start: infected (not hospitalized) counted as I
    assert s.counts() == [0, 10000 - 14, 5, 9, 0, 0, 0]


def test_generate_population_invariants():
    # engine/src/geography/grid.rs:350-375 should_generate_population
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", "cfg1_10k.json")))
    cfg["population"]["Auto"] = {"number_of_agents": 10, "public_transport_percentage": 0.2,
                                 "working_percentage": 0.2}
    cfg["geography_parameters"]["grid_size"] = 100
    s = E.OracleSim(cfg)
    a = s.agents()
    (housing, transport, work, _, _), _ = s.geometry()
    assert len(a["x"]) == 10
    for x, y in zip(a["x"], a["y"]):
        assert housing[0] <= x <= housing[2] and housing[1] <= y <= housing[3]


def test_house_capacity_error(cfg1):
    # engine/src/geography/grid.rs:140-142 (panic text mirrored as an error)
    cfg = copy.deepcopy(cfg1)
    cfg["population"]["Auto"]["number_of_agents"] = 6250 * 4 + 1
    with pytest.raises(RuntimeError, match="agents assigned to a house, but house capacity is 4"):
        E.OracleSim(cfg)


def test_too_many_starting_infections(cfg1):
    # engine/src/citizen/citizen_factory.rs:113-115
    cfg = copy.deepcopy(cfg1)
    cfg["population"]["Auto"]["number_of_agents"] = 3
    cfg["starting_infections"]["exposed"] = 4
    with pytest.raises(RuntimeError, match="There are 4 people set to infect, but only 3 agents available"):
        E.OracleSim(cfg)


def test_transport_point_cap():
    # grid.rs:96-100 + area.rs:64-74: ceil(N(pt+.1)(wp+.1)) requested, capped by the
    # distinct x/y choices of the transport strip (sqrt in f32).
    cfg = json.load(open(os.path.join(HERE, "golden", "configs", "cfg1_10k.json")))
    s = E.OracleSim(cfg)
    assert s.geometry()[1] == [6250, 125, 2400]
    cfg["population"]["Auto"]["number_of_agents"] = 1000000
    cfg["geography_parameters"]["grid_size"] = 2000
    s = E.OracleSim(cfg)
    (_, transport, _, hosp, _), counts = s.geometry()
    assert transport == (800, 0, 1199, 2000)
    assert counts[2] == 400 * 490  # 240000 requested; x capped at the 400-wide strip
    assert hosp == (1600, 0, 1799, 25)  # beds = ceil(1e6*0.003 + 1e6*0.002) = 5000; 5000 // 199 = 25
