"""Known-answer tests of the oracle's citizen routine, hour by hour of the day.

Each case builds a handful of agents on a 50x50 grid (oracle/epicpu.c
`epicpu_create_scenario`), runs one hour and compares the outcome with what the reference
text fixes. The expected cells and flags below are worked out by hand from
engine/src/citizen/mod.rs:227-454 and engine/src/models/constants.rs:27-44, not read from
the oracle:

  hour of day 0        increment_infection_day, then hospitalize          mod.rs:240-243, 351-365
  1..6                 non-staff: current_area := home                    mod.rs:244-248
  7 and 17             Normal/Essential: public transport -> goto transport area, else move_agent_from
  8                    Normal/Essential: goto work; NA: goto housing area; staff: hospital rules
  12                   NA: goto home                                      mod.rs:326-333
  16                   Normal/Essential: goto home; staff: work_quarantined := true
  23                   deceased(): routine-end outcome, Recovered agents go home  mod.rs:397-413
  staff, h - work_start_at == 336: work_quarantined := true, return       mod.rs:292-295
  staff, h - work_start_at == 672: goto home, work_start_at := h + 336, return   mod.rs:297-304

Where the reference draws a random point or a random neighbour, the cell is predicted from
the deterministic mode's definition (Philox4x32-10 words of site MOVE for (agent, hour):
x = x0 + (w0 * width) >> 32, y = y0 + (w1 * height) >> 32, neighbour = candidates[(w2 * n) >> 32],
candidates in Moore order, point.rs:57-61). Most cases are built so that only one outcome
is possible whatever the draw.
"""
import copy

import pytest

from oracle import epicpu as O

SEED = 1234
G = 50
# geography/mod.rs:33-48 at grid_size 50: strip widths ceil(0.4g)=20, ceil(0.2g)=10, 10, ceil(0.1g)=5
HOUSING = (0, 0, 19, 50)
TRANSPORT = (20, 0, 29, 50)
WORK = (30, 0, 39, 50)
# grid.rs:240-261: beds = ceil(n*0.003 + n*0.002) = 1 for n <= 200 agents; 1 <= (44-40)*(50-0), so
# y1 = 1 / (44 - 40) = 0
HOSPITAL = (40, 0, 44, 0)
# area.rs:95-117: 2x2 houses and 10x10 offices laid out row-major from the strip's start
HOUSE0, HOUSE1 = (0, 0, 1, 1), (2, 0, 3, 1)
OFFICE0 = (30, 0, 39, 9)

S, E, I, R, D = 0, 1, 2, 3, 4
PRE, ASYM, MILD, SEVERE = 0, 1, 2, 3
NORMAL, ESSENTIAL, STAFF, NA = 0, 1, 2, 3
HOSPITALIZED, VACCINATED, ISOLATED, QUARANTINED, PT = 1, 2, 4, 8, 16
MOORE = [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]


@pytest.fixture
def cfg(cfg1):
    c = copy.deepcopy(cfg1)
    c["geography_parameters"]["grid_size"] = G
    c["hours"] = 100000
    c["interventions"] = []
    d = c["disease"]
    d["regular_transmission_rate"] = 0.0
    d["high_transmission_rate"] = 0.0
    return c


def agent(x, y, home=HOUSE0, work=OFFICE0, area=None, state=S, severity=0, ws=NORMAL, flags=0, day=0, at=0,
          wsa=0, immunity=0):
    return {"x": x, "y": y, "home": home, "work": work, "current_area": area or home, "state": state,
            "severity": severity, "work_status": ws, "flags": flags, "infection_day": day, "at_hour": at,
            "work_start_at": wsa, "immunity": immunity}


# Keeps a run from stopping (an infected, hospitalized agent counts as H) without touching
# anyone: it cannot move, is not infectious (hospitalized) and stands on row y = g, outside the
# grid, where no one sees it as a neighbour.
KEEPER = agent(44, 50, ws=NA, state=I, severity=SEVERE, flags=HOSPITALIZED, day=1)
DEAD = dict(state=D, ws=NA)  # a blocker: deceased agents never move (can_move is false)


def run_hour(cfg, agents, hour, hospital=None):
    """Runs hour `hour` for these agents (plus the keeper, last) and returns the agents after it."""
    sim = O.OracleSim.scenario(cfg, agents + [KEEPER], hour - 1, hospital=hospital, seed=SEED)
    rows, _ = sim.step(1)
    assert len(rows) == 1 and rows[0][0] == hour
    out = sim.agents()
    sim.close()
    return [{k: (tuple(v[i]) if k == "current_area" else int(v[i])) for k, v in out.items()} for i in range(len(agents))]


def move_words(aid, hour):
    return O.philox([aid, hour, 1, 0], [SEED & 0xFFFFFFFF, SEED >> 32])


def random_point(aid, hour, r):
    w = move_words(aid, hour)
    return (r[0] + ((w[0] * (r[2] - r[0] + 1)) >> 32), r[1] + ((w[1] * (r[3] - r[1] + 1)) >> 32))


def inside(r, p):
    return r[0] <= p[0] <= r[2] and r[1] <= p[1] <= r[3]


def pos(a):
    return (a["x"], a["y"])


def at_hod(hd, day=3):
    return 24 * day + hd


def test_scenario_geometry(cfg):
    sim = O.OracleSim.scenario(cfg, [KEEPER], 10, seed=SEED)
    rects, counts = sim.geometry()
    sim.close()
    assert rects[0] == HOUSING and rects[1] == TRANSPORT and rects[2] == WORK and rects[4] == HOSPITAL
    assert counts[:2] == [10 * 25, 5]  # houses 20/2 x 51/2; offices 10/10 x 51/10


# ---------------------------------------------------------------- Normal / Essential workers

@pytest.mark.parametrize("hd", [7, 17])
def test_public_transport_user_goes_to_transport_area(cfg, hd):
    h = at_hod(hd)
    area = OFFICE0 if hd == 17 else HOUSE0
    start = (35, 5) if hd == 17 else (0, 0)
    a = run_hour(cfg, [agent(*start, area=area, flags=PT)], h)[0]
    assert a["current_area"] == TRANSPORT
    assert pos(a) == random_point(0, h, TRANSPORT)
    assert inside(TRANSPORT, pos(a))


@pytest.mark.parametrize("hd", [7, 17])
def test_non_transport_worker_steps_locally(cfg, hd):
    # house 0 with (1,0) and (0,1) taken: the only vacant Moore neighbour of (0,0) in the house is (1,1)
    h = at_hod(hd)
    ag = [agent(0, 0), agent(1, 0, **DEAD), agent(0, 1, **DEAD)]
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == (1, 1) and a["current_area"] == HOUSE0


@pytest.mark.parametrize("ws", [NORMAL, ESSENTIAL])
def test_worker_goes_to_office_at_8(cfg, ws):
    h = at_hod(8)
    a = run_hour(cfg, [agent(0, 0, ws=ws)], h)[0]
    assert a["current_area"] == OFFICE0
    assert pos(a) == random_point(0, h, OFFICE0)


def test_goto_refused_when_target_taken_area_still_set(cfg):
    h = at_hod(8)
    target = random_point(0, h, OFFICE0)
    a, blocker = run_hour(cfg, [agent(0, 0), agent(*target, **DEAD)], h)[:2]
    assert pos(a) == (0, 0)               # goto_area: target not vacant -> stay
    assert a["current_area"] == OFFICE0   # current_area is assigned even when the move fails
    assert pos(blocker) == target


def test_outside_current_area_teleports_then_steps(cfg):
    # A worker whose goto work failed stands at home with current_area = office. At hour of
    # day 9, move_agent_from picks a random point of the office (no vacancy check) and moves
    # to a vacant Moore neighbour of it inside the office, never to the point itself.
    h = at_hod(9)
    origin = random_point(0, h, OFFICE0)
    cands = [(origin[0] + dx, origin[1] + dy) for dx, dy in MOORE]
    cands = [p for p in cands if inside(OFFICE0, p)]
    a = run_hour(cfg, [agent(0, 0, area=OFFICE0)], h)[0]
    w = move_words(0, h)
    assert pos(a) == cands[(w[2] * len(cands)) >> 32]
    assert pos(a) != origin and max(abs(pos(a)[0] - origin[0]), abs(pos(a)[1] - origin[1])) == 1


def test_teleport_origin_with_no_vacant_neighbour_stays(cfg):
    # office replaced by a 1x2 area: the random origin has exactly one neighbour in the
    # area, the other cell; block that cell and the origin's only neighbour is taken
    h = at_hod(9)
    area = (30, 0, 30, 1)
    origin = random_point(0, h, area)
    other = (30, 1 - origin[1])
    a = run_hour(cfg, [agent(0, 0, area=area, work=area), agent(*other, **DEAD)], h)[0]
    assert pos(a) == (0, 0)


def test_worker_goes_home_at_16(cfg):
    h = at_hod(16)
    a = run_hour(cfg, [agent(35, 5, area=OFFICE0)], h)[0]
    assert a["current_area"] == HOUSE0
    assert pos(a) == random_point(0, h, HOUSE0)


@pytest.mark.parametrize("hd", [9, 12, 15])
def test_worker_wanders_in_office(cfg, hd):
    # corner (30,0) of office 0 with (31,0) and (30,1) taken: the only candidate is (31,1)
    h = at_hod(hd)
    ag = [agent(30, 0, area=OFFICE0), agent(31, 0, **DEAD), agent(30, 1, **DEAD)]
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == (31, 1) and a["current_area"] == OFFICE0


def test_symptomatic_worker_goes_home_from_work_at_16(cfg):
    # goto_area's override: a Normal/Essential agent standing in its office, mild or severe,
    # going home may move although it cannot move otherwise (mod.rs:367-384)
    h = at_hod(16)
    for sev in (MILD, SEVERE):
        a = run_hour(cfg, [agent(35, 5, area=OFFICE0, state=I, severity=sev, day=3)], h)[0]
        assert pos(a) == random_point(0, h, HOUSE0) and a["current_area"] == HOUSE0


def test_symptomatic_override_needs_the_office_cell(cfg):
    # same agent standing at home (goto work failed): no override, so it stays
    h = at_hod(16)
    a = run_hour(cfg, [agent(0, 0, area=OFFICE0, state=I, severity=MILD, day=3)], h)[0]
    assert pos(a) == (0, 0) and a["current_area"] == HOUSE0


def test_isolated_worker_does_not_go_to_work(cfg):
    h = at_hod(8)
    a = run_hour(cfg, [agent(0, 0, flags=ISOLATED)], h)[0]
    assert pos(a) == (0, 0) and a["current_area"] == OFFICE0


def test_symptomatic_worker_stays_in_office(cfg):
    h = at_hod(10)
    ag = [agent(30, 0, area=OFFICE0, state=I, severity=MILD, day=3)]
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == (30, 0)


# ---------------------------------------------------------------- NA (not working)

def test_na_at_8_steps_inside_home_then_takes_housing_area(cfg):
    # goto_area(housing) for an agent that is not working is move_agent_from with the
    # current area still home (mod.rs:390-394); the housing area is assigned afterwards
    h = at_hod(8)
    ag = [agent(0, 0, ws=NA, work=HOUSE0), agent(1, 0, **DEAD), agent(0, 1, **DEAD)]
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == (1, 1)
    assert a["current_area"] == HOUSING


def test_na_wanders_in_housing_area(cfg):
    # (10,10) in the housing strip with all neighbours but (11,11) taken
    h = at_hod(10)
    ag = [agent(10, 10, ws=NA, work=HOUSE0, area=HOUSING)]
    for dx, dy in MOORE:
        if (dx, dy) != (1, 1):
            ag.append(agent(10 + dx, 10 + dy, **DEAD))
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == (11, 11) and a["current_area"] == HOUSING


def test_na_at_12_takes_home_area_but_steps_in_housing(cfg):
    h = at_hod(12)
    ag = [agent(10, 10, ws=NA, work=HOUSE0, area=HOUSING)]
    for dx, dy in MOORE:
        if (dx, dy) != (-1, 0):
            ag.append(agent(10 + dx, 10 + dy, **DEAD))
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == (9, 10)
    assert a["current_area"] == HOUSE0


def test_na_at_13_returns_into_home(cfg):
    # outside its current area (home): teleport origin in the house, then a vacant
    # neighbour inside the house; with two house cells taken the step is forced
    h = at_hod(13)
    origin = random_point(0, h, HOUSE0)
    free = [p for p in [(0, 0), (1, 0), (0, 1), (1, 1)] if p != origin]
    blockers = free[:2]
    expect = free[2]
    ag = [agent(10, 10, ws=NA, work=HOUSE0, area=HOUSE0)] + [agent(*b, **DEAD) for b in blockers]
    a = run_hour(cfg, ag, h)[0]
    assert pos(a) == expect and a["current_area"] == HOUSE0


def test_symptomatic_na_has_no_override(cfg):
    h = at_hod(12)
    a = run_hour(cfg, [agent(10, 10, ws=NA, work=HOUSE0, area=HOUSING, state=I, severity=MILD, day=3)], h)[0]
    assert pos(a) == (10, 10) and a["current_area"] == HOUSE0


# ---------------------------------------------------------------- hospital staff

def staff(x, y, area, wsa, flags=0, **kw):
    return agent(x, y, ws=STAFF, area=area, wsa=wsa, flags=flags, **kw)


def test_staff_first_shift_goes_to_hospital(cfg):
    # created with work_start_at = ROUTINE_WORK_TIME = 8 (mod.rs:442-447); at h = 8 the staff
    # member is not in the hospital area and work_start_at <= h
    h = 8
    a = run_hour(cfg, [staff(0, 0, HOUSE0, 8, flags=QUARANTINED)], h)[0]
    assert pos(a) == random_point(0, h, HOSPITAL) and a["current_area"] == HOSPITAL
    assert a["work_start_at"] == 8 and not a["flags"] & QUARANTINED


def test_staff_next_days_at_8_stay_and_clear_quarantine(cfg):
    h = at_hod(8, day=5)
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, 8, flags=QUARANTINED)], h)[0]
    assert pos(a) == (40, 0) and a["work_start_at"] == 8 and not a["flags"] & QUARANTINED


def test_staff_at_16_quarantines_without_moving(cfg):
    h = at_hod(16)
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, 8)], h)[0]
    assert pos(a) == (40, 0) and a["flags"] & QUARANTINED


@pytest.mark.parametrize("quarantined", [False, True])
def test_staff_moves_in_hospital_unless_quarantined(cfg, quarantined):
    # (40,0) in the one-row hospital: its only neighbour in the area is (41,0)
    h = at_hod(11)
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, 8, flags=QUARANTINED if quarantined else 0)], h)[0]
    assert pos(a) == ((40, 0) if quarantined else (41, 0))


def test_staff_sleep_hours_keep_the_hospital_area(cfg):
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, 8), agent(35, 5, area=OFFICE0)], at_hod(3))
    assert a[0]["current_area"] == HOSPITAL
    assert a[1]["current_area"] == HOUSE0 and pos(a[1]) == (35, 5)


RIPE_E = dict(state=E, at=0)  # exposed at hour 0: h - 0 >= 48 + 1 at any later evaluated hour


def test_staff_quarantine_after_14_days(cfg):
    # h - work_start_at = 24 * 14: work_quarantined := true and return, before the disease update
    wsa = 8
    h = wsa + 24 * 14
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, wsa, **RIPE_E)], h)[0]
    assert a["flags"] & QUARANTINED
    assert pos(a) == (40, 0) and a["state"] == E  # no disease update in this branch


def test_staff_13_days_is_an_ordinary_shift_start(cfg):
    wsa = 8
    h = wsa + 24 * 13
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, wsa, flags=QUARANTINED, **RIPE_E)], h)[0]
    assert not a["flags"] & QUARANTINED and a["state"] == I


def test_staff_goes_home_after_28_days(cfg):
    wsa = 8
    h = wsa + 24 * 28
    a = run_hour(cfg, [staff(40, 0, HOSPITAL, wsa, **RIPE_E)], h)[0]
    assert pos(a) == random_point(0, h, HOUSE0) and a["current_area"] == HOUSE0
    assert a["work_start_at"] == h + 24 * 14
    assert a["state"] == E


def test_staff_waits_for_work_start_at(cfg):
    # back home with work_start_at in the future: the shift start is skipped, quarantine cleared
    h = at_hod(8, day=30)
    a = run_hour(cfg, [staff(0, 0, HOUSE0, h + 24 * 3, flags=QUARANTINED)], h)[0]
    assert pos(a) == (0, 0) and a["current_area"] == HOUSE0 and not a["flags"] & QUARANTINED


def test_staff_returns_when_work_start_at_arrives(cfg):
    h = at_hod(8, day=30)
    a = run_hour(cfg, [staff(0, 0, HOUSE0, h)], h)[0]
    assert a["current_area"] == HOSPITAL and pos(a) == random_point(0, h, HOSPITAL)
    assert a["work_start_at"] == h


# ---------------------------------------------------------------- sleep, hour 0, hour 23

@pytest.mark.parametrize("hd", [1, 6])
def test_sleep_hours_reset_area_without_moving(cfg, hd):
    ag = [agent(35, 5, area=OFFICE0), agent(10, 10, ws=NA, area=HOUSING, **RIPE_E)]
    a = run_hour(cfg, ag, at_hod(hd))
    assert [pos(x) for x in a] == [(35, 5), (10, 10)]
    assert a[0]["current_area"] == HOUSE0 and a[1]["current_area"] == HOUSE0
    assert a[1]["state"] == E  # no disease update while asleep


def hosp_cfg(cfg):
    # is_to_be_hospitalized is rate(day + immunity) >= high_transmission_rate
    # (common/src/disease/mod.rs:97-99): with both rates 0 every day would qualify
    c = copy.deepcopy(cfg)
    c["disease"]["regular_transmission_rate"] = 0.025
    c["disease"]["high_transmission_rate"] = 0.25
    return c


def test_hour_0_increments_day_and_admits_to_first_vacant_cell(cfg):
    # day 20 + 1 = 21 is in the high window (20, 23]: to be hospitalized. Two such agents both
    # target the first vacant hospital cell in row-major order, (41,0) because (40,0) is
    # taken; the lower id wins it and the other stays, still flagged hospitalized.
    h = at_hod(0, day=4)
    sev = dict(state=I, severity=SEVERE, day=20)
    ag = [agent(0, 0, **sev), agent(1, 0, **sev), agent(40, 0, **DEAD), agent(2, 0, home=HOUSE1, state=I, severity=MILD, day=5)]
    a = run_hour(hosp_cfg(cfg), ag, h)
    assert pos(a[0]) == (41, 0) and a[0]["flags"] & HOSPITALIZED and a[0]["infection_day"] == 21
    assert pos(a[1]) == (1, 0) and a[1]["flags"] & HOSPITALIZED
    assert a[3]["infection_day"] == 6 and pos(a[3]) == (2, 0) and not a[3]["flags"] & HOSPITALIZED


def test_hour_0_immunity_shifts_the_hospitalization_day(cfg):
    # day 19 + 1 = 20 is not in (20, 23], but with immunity +1 the effective day is 21
    h = at_hod(0, day=4)
    sev = dict(state=I, severity=SEVERE, day=19)
    a = run_hour(hosp_cfg(cfg), [agent(0, 0, **sev), agent(1, 0, immunity=1, **sev)], h)
    assert not a[0]["flags"] & HOSPITALIZED and pos(a[0]) == (0, 0)
    assert a[1]["flags"] & HOSPITALIZED and pos(a[1]) == (40, 0)


def test_hour_0_full_hospital_falls_back_to_home(cfg):
    h = at_hod(0, day=4)
    full = [agent(x, 0, **DEAD) for x in range(40, 45)]
    a = run_hour(hosp_cfg(cfg), [agent(35, 5, area=OFFICE0, state=I, severity=SEVERE, day=20)] + full, h)[0]
    assert not a["flags"] & HOSPITALIZED
    assert pos(a) == random_point(0, h, HOUSE0)


def test_hour_23_outcomes(cfg):
    # recovery days are the constants 9 and 12 (constants.rs:48-50); the config's own
    # asymptomatic_last_day / mild_infected_last_day are read but never used
    c = copy.deepcopy(cfg)
    c["disease"]["death_rate"] = 1.0
    c["disease"]["asymptomatic_last_day"] = 5
    c["disease"]["mild_infected_last_day"] = 6
    h = at_hod(23)
    ag = [agent(35, 5, area=OFFICE0, state=I, severity=ASYM, day=9),
          agent(36, 5, home=HOUSE1, area=OFFICE0, state=I, severity=MILD, day=12),
          agent(41, 0, state=I, severity=SEVERE, day=23, flags=HOSPITALIZED, home=(4, 0, 5, 1)),
          agent(37, 5, home=(6, 0, 7, 1), area=OFFICE0, state=I, severity=ASYM, day=8),
          agent(38, 5, home=(8, 0, 9, 1), area=OFFICE0, state=I, severity=MILD, day=9),
          agent(34, 5, home=(10, 0, 11, 1), area=OFFICE0, state=I, severity=ASYM, day=5),
          agent(33, 5, home=(12, 0, 13, 1), area=OFFICE0, state=I, severity=MILD, day=6)]
    a = run_hour(c, ag, h)
    assert a[0]["state"] == R and pos(a[0]) == random_point(0, h, HOUSE0)   # ASYMPTOMATIC_LAST_DAY = 9
    assert a[1]["state"] == R and pos(a[1]) == random_point(1, h, HOUSE1)   # MILD_INFECTED_LAST_DAY = 12
    assert a[2]["state"] == D and pos(a[2]) == (41, 0) and not a[2]["flags"] & HOSPITALIZED
    assert a[3]["state"] == I and pos(a[3]) == (37, 5)
    assert a[4]["state"] == I and pos(a[4]) == (38, 5)
    assert a[5]["state"] == I and a[6]["state"] == I


def test_hour_23_every_recovered_agent_tries_home(cfg):
    # an old Recovered agent, isolated and in its office, is still sent home (no can_move check)
    h = at_hod(23)
    a = run_hour(cfg, [agent(35, 5, area=OFFICE0, state=R, flags=ISOLATED)], h)[0]
    assert pos(a) == random_point(0, h, HOUSE0)


def test_hour_23_severe_survivor_recovers_and_leaves_hospital(cfg):
    c = copy.deepcopy(cfg)
    c["disease"]["death_rate"] = 0.0
    h = at_hod(23)
    a = run_hour(c, [agent(41, 0, state=I, severity=SEVERE, day=23, flags=HOSPITALIZED)], h)[0]
    assert a["state"] == R and not a["flags"] & HOSPITALIZED and pos(a) == random_point(0, h, HOUSE0)


# ---------------------------------------------------------------- collisions and infection

def test_housemates_collide_lowest_id_wins_loser_keeps_updates(cfg):
    # three housemates with one free cell: each one's only candidate is (1,1); the lowest id
    # takes it (deterministic mode's merge rule), the others stay and keep their disease update
    h = at_hod(7)
    ag = [agent(1, 0), agent(0, 0, state=E, at=h - 49), agent(0, 1)]
    a = run_hour(cfg, ag, h)
    assert pos(a[0]) == (1, 1) and pos(a[1]) == (0, 0) and pos(a[2]) == (0, 1)
    assert a[1]["state"] == I


def test_infection_at_chosen_cell_from_start_of_hour_neighbours(cfg):
    c = infectious_cfg(cfg)
    h = at_hod(10)
    # infector at (32,2) in the regular window (day 6 in (5, 20]); the susceptible worker at
    # (30,0) has (30,1) and (31,0) blocked, so it moves to (31,1), a neighbour of (32,2)
    infector = agent(32, 2, area=OFFICE0, state=I, severity=ASYM, day=6, home=HOUSE1)
    ag = [agent(30, 0, area=OFFICE0), agent(31, 0, **DEAD), agent(30, 1, **DEAD), infector]
    a = run_hour(c, ag, h)
    assert pos(a[0]) == (31, 1) and a[0]["state"] == E and a[0]["at_hour"] == h


def infectious_cfg(cfg):
    c = copy.deepcopy(cfg)
    c["disease"]["regular_transmission_rate"] = 1.0  # gen_bool(1.0) always fires
    c["disease"]["high_transmission_rate"] = 1.0
    return c


# Isolated agents cannot move but can still be infected (allocation_map.rs:349-356,
# default_disease_handler.rs:72): they pin a susceptible to its cell for the infection cases.
INFECTIOUS = dict(state=I, severity=ASYM, day=6, home=HOUSE1)  # day 6 is in the regular window (5, 20]


def test_isolated_susceptible_is_infected_in_place(cfg):
    h = at_hod(10)
    a = run_hour(infectious_cfg(cfg), [agent(33, 5, area=OFFICE0, flags=ISOLATED),
                                       agent(34, 6, area=OFFICE0, flags=ISOLATED, **INFECTIOUS)], h)
    assert pos(a[0]) == (33, 5) and a[0]["state"] == E and a[0]["at_hour"] == h


def test_no_infection_from_outside_current_area(cfg):
    # (29,5) is in the transport strip: a Moore neighbour of (30,5) but outside office 0
    h = at_hod(10)
    a = run_hour(infectious_cfg(cfg), [agent(30, 5, area=OFFICE0, flags=ISOLATED),
                                       agent(29, 5, area=TRANSPORT, flags=ISOLATED, **INFECTIOUS)], h)
    assert a[0]["state"] == S


def test_no_infection_from_hospitalized(cfg):
    h = at_hod(10)
    a = run_hour(infectious_cfg(cfg), [agent(35, 5, area=OFFICE0, flags=ISOLATED),
                                       agent(36, 5, area=OFFICE0, flags=HOSPITALIZED, **INFECTIOUS)], h)
    assert a[0]["state"] == S


def test_no_infection_from_outside_the_grid(cfg):
    # row y = g = 50 belongs to the housing area but not to the grid (allocation_map.rs:156-159)
    h = at_hod(10)
    a = run_hour(infectious_cfg(cfg), [agent(10, 49, ws=NA, area=HOUSING, flags=ISOLATED),
                                       agent(10, 50, ws=NA, area=HOUSING, flags=ISOLATED, **INFECTIOUS)], h)
    assert a[0]["state"] == S


def test_rate_window_zero_does_not_infect(cfg):
    # day 5 is not in (5, 20]; day 7 with immunity -2 is day 5 as well
    h = at_hod(10)
    inf = dict(INFECTIOUS)
    inf["day"] = 5
    a = run_hour(infectious_cfg(cfg), [agent(30, 5, area=OFFICE0, flags=ISOLATED),
                                       agent(31, 5, area=OFFICE0, flags=ISOLATED, **inf),
                                       agent(30, 6, area=OFFICE0, flags=ISOLATED, immunity=-2,
                                             **dict(inf, day=7))], h)
    assert a[0]["state"] == S


@pytest.mark.parametrize("flag", [VACCINATED, QUARANTINED])
def test_vaccinated_or_quarantined_not_infected(cfg, flag):
    h = at_hod(10)
    a = run_hour(infectious_cfg(cfg), [agent(30, 5, area=OFFICE0, flags=ISOLATED | flag),
                                       agent(31, 5, area=OFFICE0, flags=ISOLATED, **INFECTIOUS)], h)
    assert a[0]["state"] == S


def test_no_infection_while_asleep_or_at_hours_0_and_23(cfg):
    for hd in (0, 3, 23):
        a = run_hour(infectious_cfg(cfg), [agent(33, 5, area=OFFICE0, flags=ISOLATED),
                                           agent(34, 6, area=OFFICE0, flags=ISOLATED, **INFECTIOUS)], at_hod(hd))
        assert a[0]["state"] == S, hd


# ---------------------------------------------------------------- disease progression

@pytest.mark.parametrize("elapsed,becomes", [(46, E), (49, I)])
def test_exposed_to_infected_window(cfg, elapsed, becomes):
    # exposed_duration 48 with a random factor in {-1, 0, 1}: 46 < 47 never, 49 >= 49 always
    h = at_hod(10)
    a = run_hour(cfg, [agent(0, 0, state=E, at=h - elapsed)], h)[0]
    assert a["state"] == becomes


def test_exposed_becomes_asymptomatic_or_pre(cfg):
    h = at_hod(10)
    for pct_asym, sev, at in ((1.0, ASYM, 0), (0.0, PRE, h)):
        c = copy.deepcopy(cfg)
        c["disease"]["percentage_asymptomatic_population"] = pct_asym
        a = run_hour(c, [agent(0, 0, state=E, at=h - 49)], h)[0]
        assert a["state"] == I and a["severity"] == sev and a["infection_day"] == 0 and a["at_hour"] == at


@pytest.mark.parametrize("elapsed,severe", [(47, None), (48, True), (48, False)])
def test_pre_symptomatic_turns_at_48_hours(cfg, elapsed, severe):
    c = copy.deepcopy(cfg)
    c["disease"]["percentage_severe_infected_population"] = 1.0 if severe else 0.0
    h = at_hod(10)
    a = run_hour(c, [agent(0, 0, state=I, severity=PRE, at=h - elapsed, day=2)], h)[0]
    want = PRE if severe is None else (SEVERE if severe else MILD)
    assert a["severity"] == want and a["infection_day"] == 2
