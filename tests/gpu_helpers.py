"""Shared helpers for the GPU parity tests (comparison against the CPU oracle)."""
import copy
import sys

import numpy as np

FIELDS = ["x", "y", "state", "severity", "infection_day", "at_hour", "flags", "work_status", "work_start_at",
          "immunity", "current_area"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_one_agent_per_cell(agents, where=""):
    """No two agents share a cell (the occupancy map holds one citizen per key,
    allocation_map.rs:44-48): the invariant the housemate arbitration and the block claim
    filter must keep."""
    key = np.asarray(agents["x"]).astype(np.int64) * 65536 + np.asarray(agents["y"])
    if len(np.unique(key)) != len(key):
        raise AssertionError("%s: %d agents share a cell" % (where, len(key) - len(np.unique(key))))


def assert_agents_equal(gpu, cpu, where=""):
    assert_one_agent_per_cell(gpu, where)
    for f in FIELDS:
        a, b = np.asarray(gpu[f]), np.asarray(cpu[f])
        if a.shape != b.shape or not np.array_equal(a, b):
            bad = np.nonzero((a != b).reshape(len(a), -1).any(axis=1))[0] if a.shape == b.shape else []
            i = int(bad[0]) if len(bad) else -1
            detail = ""
            if i >= 0:
                detail = "; agent %d: gpu=%s cpu=%s" % (i, {k: np.asarray(gpu[k])[i].tolist() for k in FIELDS},
                                                        {k: np.asarray(cpu[k])[i].tolist() for k in FIELDS})
            raise AssertionError("%s field %s differs for %d agents%s" % (where, f, len(bad), detail))


def interventions_config(cfg1, n=4000, g=160):
    cfg = copy.deepcopy(cfg1)
    cfg["population"]["Auto"]["number_of_agents"] = n
    cfg["geography_parameters"]["grid_size"] = g
    cfg["disease"]["regular_transmission_rate"] = 0.3
    cfg["disease"]["high_transmission_rate"] = 0.6
    cfg["starting_infections"] = {"exposed": 10, "infected_mild_asymptomatic": 3, "infected_mild_symptomatic": 3,
                                  "infected_severe": 3}
    cfg["interventions"] = [{"Lockdown": {"at_number_of_infections": 40, "essential_workers_population": 0.2}},
                            {"BuildNewHospital": {"spread_rate_threshold": 5}},
                            {"Vaccinate": {"at_hour": 100, "percent": 0.3}},
                            {"Vaccinate": {"at_hour": 300, "percent": 0.1}},
                            {"Vaccinate": {"at_hour": 300, "percent": 0.5}}]
    return cfg
