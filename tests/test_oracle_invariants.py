"""Closed forms and per-hour invariants that pin the oracle's step where no
reference known-answer test exists (the reference seeds from the OS, so no
reference run can be replayed bit for bit).

The closed forms follow from the reference's rules, not from the oracle's code:
  * infection, E->I and Pre->symptomatic are evaluated only at hours of day 7..22
    (engine/src/citizen/mod.rs:238-253);
  * exposed_duration 48 with a random factor in {-1,0,1}
    (default_disease_handler.rs:52-62);
  * infection_day increments at hour-of-day 0 (citizen/mod.rs:240-242);
  * asymptomatic recover at day 9, severe resolve at last_day at hour-of-day 23
    (default_disease_handler.rs:88-102, constants.rs:48-50);
  * with all transmission rates 0 a severe agent is always "to be hospitalized"
    because 0.0 >= 0.0 (common/src/disease/mod.rs:97-99).
"""
import copy

import numpy as np
import pytest

from oracle import epicpu as E


def zero_transmission(cfg, asym, severe, death, exposed=5):
    c = copy.deepcopy(cfg)
    d = c["disease"]
    d["regular_transmission_rate"] = 0.0
    d["high_transmission_rate"] = 0.0
    d["percentage_asymptomatic_population"] = asym
    d["percentage_severe_infected_population"] = severe
    d["death_rate"] = death
    c["starting_infections"]["exposed"] = exposed
    return c


def test_asymptomatic_course_closed_form(cfg1):
    s = E.OracleSim(zero_transmission(cfg1, asym=1.0, severe=0.0, death=0.0))
    rows, stopped = s.step(2399)
    by_hour = {int(r[0]): r for r in rows}
    # E through hour 54 (first evaluated hour with h >= 49 is 55 = day 2, hour-of-day 7)
    assert all(by_hour[h][2] == 5 for h in range(1, 55))
    assert by_hour[55][2] == 0 and by_hour[55][3] == 5
    # day 1 at h=72 ... day 9 at h=264; recovered at hour-of-day 23 of that day
    assert by_hour[286][3] == 5 and by_hour[286][5] == 0
    assert by_hour[287][5] == 5 and by_hour[287][3] == 0
    assert stopped and rows[-1][0] == 287  # E = I = H = 0 stops the run
    assert all(r[1] == 9995 for r in rows)  # S constant: nothing transmits


@pytest.mark.parametrize("death,col", [(1.0, 6), (0.0, 5)])
def test_severe_course_closed_form(cfg1, death, col):
    s = E.OracleSim(zero_transmission(cfg1, asym=0.0, severe=1.0, death=death))
    rows, stopped = s.step(2399)
    by_hour = {int(r[0]): r for r in rows}
    assert by_hour[54][2] == 5 and by_hour[55][3] == 5      # Pre from 55
    assert by_hour[119][3] == 5 and by_hour[119][4] == 0    # Severe from 103, not yet admitted
    # first hour-of-day 0 after becoming severe: every severe agent is flagged hospitalized,
    # even those that lose the single first-vacant hospital cell (allocation_map.rs:144-147)
    assert by_hour[120][4] == 5 and by_hour[120][3] == 0
    assert by_hour[622][4] == 5
    # day 23 (= last_day) is reached at h = 72 + 22*24 = 600; resolved at 623
    assert by_hour[623][col] == 5 and by_hour[623][4] == 0
    assert stopped and rows[-1][0] == 623
    a = s.agents()
    assert (a["flags"] & 1).sum() == 0  # hospitalized cleared on R/D (citizen/mod.rs:409-411)


def test_single_hospital_cell_collision(cfg1):
    """All same-hour admissions target one cell: only the lowest id enters it."""
    s = E.OracleSim(zero_transmission(cfg1, asym=0.0, severe=1.0, death=0.0))
    s.step(119)
    before = s.agents()
    s.step(1)
    after = s.agents()
    (_, _, _, _, hosp), _ = s.geometry()
    in_h = (after["x"] >= hosp[0]) & (after["x"] <= hosp[2]) & (after["y"] >= hosp[1]) & (after["y"] <= hosp[3])
    sev = np.nonzero(after["state"] == 2)[0]
    assert len(sev) == 5
    moved = sev[in_h[sev]]
    assert list(moved) == [sev.min()]
    assert (after["x"][moved[0]], after["y"][moved[0]]) == (hosp[0], hosp[1])
    stay = sev[~in_h[sev]]
    assert (after["x"][stay] == before["x"][stay]).all() and (after["y"][stay] == before["y"][stay]).all()


def _check_hour(prev, cur, h, n, geo):
    (housing, transport, work, hosp0, hosp1) = geo
    occ = cur["x"].astype(np.int64) * 100003 + cur["y"]
    assert len(np.unique(occ)) == n, "two agents share a cell at hour %d" % h
    moved = (prev["x"] != cur["x"]) | (prev["y"] != cur["y"])
    hd = h % 24
    if 1 <= hd <= 6:
        assert not moved.any(), "movement during sleep hours at %d" % h
    if hd == 0:
        assert (cur["state"][moved] == 2).all() and (cur["severity"][moved] == 3).all()
    # susceptible count never grows; recovered/deceased never shrink
    assert ((prev["state"] != 0) <= (cur["state"] != 0)).all()
    assert ((prev["state"] >= 3) <= (cur["state"] >= 3)).all()
    # vaccinated agents stay susceptible
    vac = (prev["flags"] & 2) != 0
    assert (cur["state"][vac & (prev["state"] == 0)] == 0).all()
    # deceased never move
    assert not (moved & (prev["state"] == 4)).any()
    if 7 <= hd <= 22:
        blocked = ((prev["state"] == 2) & ((prev["severity"] == 2) | (prev["severity"] == 3))) | \
                  ((prev["flags"] & (1 | 4)) != 0) | (prev["state"] == 4)
        override = (hd == 16) & (prev["work_status"] <= 1) & (prev["state"] == 2) & (prev["severity"] >= 2)
        assert not (moved & blocked & ~override).any(), "an agent that cannot move moved at %d" % h
    # every mover lands in its (new) current area, its home strip, or a hospital
    x, y = cur["x"][moved], cur["y"][moved]
    ca = cur["current_area"][moved]
    ok = (x >= ca[:, 0]) & (x <= ca[:, 2]) & (y >= ca[:, 1]) & (y <= ca[:, 3])
    for r in (housing, hosp0, hosp1):
        ok |= (x >= r[0]) & (x <= r[2]) & (y >= r[1]) & (y <= r[3])
    assert ok.all()


def test_invariants_every_hour_with_interventions(cfg1):
    cfg = copy.deepcopy(cfg1)
    cfg["population"]["Auto"]["number_of_agents"] = 4000
    cfg["geography_parameters"]["grid_size"] = 160
    cfg["disease"]["regular_transmission_rate"] = 0.3
    cfg["disease"]["high_transmission_rate"] = 0.6
    cfg["starting_infections"] = {"exposed": 10, "infected_mild_asymptomatic": 3, "infected_mild_symptomatic": 3,
                                  "infected_severe": 3}
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 40, "essential_workers_population": 0.2}},
                            {"BuildNewHospital": {"spread_rate_threshold": 5}},
                            {"Vaccinate": {"at_hour": 100, "percent": 0.3}}]
    s = E.OracleSim(cfg)
    rects, _ = s.geometry()
    prev = s.agents()
    n = len(prev["x"])
    for h in range(1, 700):
        rows, stopped = s.step(1)
        if len(rows) == 0:
            break
        cur = s.agents()
        assert rows[0][0] == h and rows[0][1:].sum() == n
        rects, _ = s.geometry()
        _check_hour(prev, cur, h, n, rects)
        prev = cur
        if stopped:
            break
    kinds = {k for _, k in s.events()}
    assert 0 in kinds and 1 in kinds and 3 in kinds  # vaccination, lockdown, new hospital all fired


def test_lockdown_isolates_non_essential(cfg1):
    cfg = copy.deepcopy(cfg1)
    cfg["disease"]["regular_transmission_rate"] = 0.5
    cfg["disease"]["high_transmission_rate"] = 0.9
    cfg["starting_infections"]["infected_mild_asymptomatic"] = 30
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 20, "essential_workers_population": 0.3}}]
    s = E.OracleSim(cfg)
    a0 = s.agents()
    normal_or_ess = a0["work_status"] <= 1
    frac = (a0["work_status"][normal_or_ess] == 1).mean()
    assert 0.25 < frac < 0.35  # essential draw over Normal agents only
    assert (a0["work_status"] == 2).sum() > 0  # hospital staff exist (p = 0.002 of workers)
    rows, _ = s.step(24)
    ev = s.events()
    assert ev and ev[0] == (24, 1)  # locked at the first hour-of-day 0 with I > 20
    a = s.agents()
    iso = (a["flags"] & 4) != 0
    assert (iso == (a["work_status"] != 1)).all()


def test_standalone_lockdown_unlocks_only_at_504(cfg1):
    """zero_infection_hour stays 0 in standalone, so unlock fires only at h = 504."""
    cfg = copy.deepcopy(cfg1)
    cfg["disease"]["regular_transmission_rate"] = 0.0
    cfg["disease"]["high_transmission_rate"] = 0.0
    cfg["disease"]["percentage_asymptomatic_population"] = 1.0
    cfg["starting_infections"] = {"exposed": 0, "infected_mild_asymptomatic": 30, "infected_mild_symptomatic": 0,
                                  "infected_severe": 0}
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 20, "essential_workers_population": 0.1}}]
    s = E.OracleSim(cfg)
    rows, stopped = s.step(2399)
    # asymptomatic from day 1: day 9 from h = 8*24 = 192, recovered at its hour-of-day 23 (h = 215);
    # locked at 24 and never unlocked
    assert s.events() == [(24, 1)]
    assert stopped and rows[-1][0] == 215
