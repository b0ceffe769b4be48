"""Citizen-state stream (stream.hip + the JSON-lines sink) against the oracle.

Each hour's line must equal, byte for byte, serde_json's compact serialization of the
reference's CitizenStatesAtHr {hr, citizen_states: [CitizenState {citizen_id, state,
location {x, y}}]} (engine/src/models/events/citizen_state.rs:26-50, field order of the
structs) built from the oracle's agents after that hour, with the dense id as the
citizen_id string. The first line is the Grid (geography/grid.rs:35-47) and the last the
end message (listeners/events_kafka_producer.rs:75-87)."""
import copy
import json
import os
import time

import numpy as np
import pytest

import epirust_amd as G
from oracle import epicpu as O
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu
LETTERS = "seird"


def expected_line(hour, agents):
    states = [{"citizen_id": str(i), "state": LETTERS[int(s)], "location": {"x": int(x), "y": int(y)}}
              for i, (s, x, y) in enumerate(zip(agents["state"], agents["x"], agents["y"]))]
    return json.dumps({"hr": hour, "citizen_states": states}, separators=(",", ":"))


def area(r):
    return {"location_id": "0", "start_offset": {"x": r[0], "y": r[1]}, "end_offset": {"x": r[2], "y": r[3]}}


def small(cfg1, n=3000, g=120):
    c = copy.deepcopy(cfg1)
    c["population"]["Auto"]["number_of_agents"] = n
    c["geography_parameters"]["grid_size"] = g
    c["disease"]["regular_transmission_rate"] = 0.3
    c["disease"]["high_transmission_rate"] = 0.6
    c["starting_infections"]["exposed"] = 20
    return c


def test_stream_lines_match_oracle_every_hour(cfg1, tmp_path):
    cfg = small(cfg1)
    hours = 60
    say("[stream] %d hours, json lines vs oracle" % hours)
    path = str(tmp_path / "states.jsonl")
    gpu = G.Simulation(cfg)
    gpu.stream_citizen_states(path)
    cpu = O.OracleSim(cfg)
    expected = []
    for h in range(1, hours + 1):
        cpu.step(1)
        expected.append(expected_line(h, cpu.agents()))
    # two calls, so the sink carries over between epi_step calls
    gpu.step(25)
    gpu.step(hours - 25)
    hour, rec = gpu.citizen_states()
    a = cpu.agents()
    assert hour == hours
    assert np.array_equal(rec["x"], a["x"]) and np.array_equal(rec["y"], a["y"])
    assert "".join(c.decode() for c in rec["state"]) == "".join(LETTERS[s] for s in a["state"])
    gpu.stream_citizen_states(None)
    lines = open(path).read().split("\n")
    assert lines[-1] == ""
    lines = lines[:-1]
    assert len(lines) == hours + 2
    assert lines[-1] == '{"simulation_ended": true}'
    for h in range(hours):
        assert lines[1 + h] == expected[h], "hour %d differs" % (h + 1)
    grid = json.loads(lines[0])
    rects, _ = cpu.geometry()
    assert list(grid) == ["grid_size", "housing_area", "work_area", "transport_area", "hospital_area", "houses",
                          "offices"]
    assert grid["grid_size"] == 120
    assert grid["housing_area"] == area(rects[0]) and grid["transport_area"] == area(rects[1])
    assert grid["work_area"] == area(rects[2]) and grid["hospital_area"] == area(rects[3])
    # area_factory order: row-major 2x2 houses over the housing strip, 10x10 offices over work
    hx = (rects[0][2] - rects[0][0] + 1) // 2
    hy = (rects[0][3] - rects[0][1] + 1) // 2
    assert len(grid["houses"]) == hx * hy
    for k in (0, 1, hx, hx * hy - 1):
        x0, y0 = rects[0][0] + (k % hx) * 2, rects[0][1] + (k // hx) * 2
        assert grid["houses"][k] == area((x0, y0, x0 + 1, y0 + 1))
    ox = (rects[2][2] - rects[2][0] + 1) // 10
    assert len(grid["offices"]) == ox * ((rects[2][3] - rects[2][1] + 1) // 10)
    assert grid["offices"][-1]["end_offset"]["x"] == rects[2][0] + ox * 10 - 1
    gpu.close()
    cpu.close()


def test_stream_stops_with_the_simulation(cfg1, tmp_path):
    """No starting infections: the run stops after hour 1 and the stream holds one hour."""
    cfg = small(cfg1)
    cfg["starting_infections"] = {"exposed": 0, "infected_mild_asymptomatic": 0, "infected_mild_symptomatic": 0,
                                  "infected_severe": 0}
    say("[stream] early stop")
    path = str(tmp_path / "stop.jsonl")
    gpu = G.Simulation(cfg)
    gpu.stream_citizen_states(path)
    rows, stopped = gpu.step(100)
    assert stopped and len(rows) == 1
    gpu.stream_citizen_states(None)
    lines = open(path).read().split("\n")[:-1]
    assert len(lines) == 3 and lines[1].startswith('{"hr":1,"citizen_states":[{"citizen_id":"0","state":"s"')
    cpu = O.OracleSim(cfg)
    cpu.step(1)
    assert lines[1] == expected_line(1, cpu.agents())
    gpu.close()
    cpu.close()


def test_run_writes_the_stream_when_enabled(cfg1, tmp_path):
    cfg = small(cfg1, n=2000, g=100)
    cfg["hours"] = 50
    cfg["enable_citizen_state_messages"] = True
    say("[stream] epi_run with enable_citizen_state_messages")
    sim = G.Simulation(cfg)
    csv_path = sim.run(str(tmp_path))
    sim.close()
    stream_path = csv_path[:-4] + "_citizen_states.jsonl"
    lines = open(stream_path).read().split("\n")[:-1]
    n_rows = len(open(csv_path).read().split("\n")) - 2
    assert len(lines) == n_rows + 2
    last = json.loads(lines[-2])
    assert last["hr"] == n_rows and len(last["citizen_states"]) == 2000


def test_stream_at_1m_agents_matches_binary_records(tmp_path):
    """1,000,000 agents, 3 hours: the device-formatted text parses to the same records as
    epi_citizen_states reads back."""
    here = os.path.dirname(os.path.abspath(__file__))
    cfg = json.load(open(os.path.join(here, "golden", "configs", "cfg2_1m.json")))
    say("[stream] 1M agents")
    path = str(tmp_path / "big.jsonl")
    gpu = G.Simulation(cfg)
    gpu.stream_citizen_states(path)
    t0 = time.perf_counter()
    gpu.step(3)
    dt = time.perf_counter() - t0
    hour, rec = gpu.citizen_states()
    gpu.stream_citizen_states(None)
    say("[stream] 1M agents: 3 streamed hours in %.3f s, %.0f MB written" % (dt, os.path.getsize(path) / 1e6))
    with open(path) as f:
        lines = f.read().split("\n")[:-1]
    assert len(lines) == 5
    msg = json.loads(lines[3])
    assert msg["hr"] == 3 == hour
    s = msg["citizen_states"]
    assert len(s) == 1000000
    assert all(s[i]["citizen_id"] == str(i) for i in (0, 1, 999999))
    xs = np.fromiter((r["location"]["x"] for r in s), dtype=np.int32, count=len(s))
    ys = np.fromiter((r["location"]["y"] for r in s), dtype=np.int32, count=len(s))
    assert np.array_equal(xs, rec["x"]) and np.array_equal(ys, rec["y"])
    assert "".join(r["state"] for r in s) == b"".join(rec["state"]).decode()
    gpu.close()
