"""The measurement entry points on the GPU: epi_last_run_ms times exactly the hour loop of
the last step call (device time from HIP events, hours enqueued), and epi_profile's
per-kernel launch counts follow the hour structure (step at hours of day 0 and 7..23,
commit launch every hour, hospital scan at hour of day 0)."""
import pytest

import epirust_amd as G
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


def test_last_run_ms_covers_the_step_calls_hours(cfg1):
    say("[measure] last_run_ms")
    s = G.Simulation(cfg1)
    assert s.last_run_ms() == (0.0, 0)  # nothing ran yet
    rows, _ = s.step(50)  # two whole days (graph replays) and two single hours
    ms, hours = s.last_run_ms()
    assert hours == len(rows) == 50
    assert 0.0 < ms < 1000.0
    rows, _ = s.step(1)
    ms1, hours1 = s.last_run_ms()
    assert hours1 == 1 and 0.0 < ms1 < ms
    s.close()


def test_profile_launch_counts_follow_the_hour_structure(cfg1):
    say("[measure] profile launch counts")
    s = G.Simulation(cfg1)
    p = s.profile(48)  # hours 1..48: hours of day 0 are 24 and 48
    assert p["hours"] == 48
    assert p["hospital_scan"][1] == 2
    assert p["step"][1] == 2 * 18
    assert p["commit"][1] == 48
    assert all(p[k][0] > 0.0 for k in ("hospital_scan", "step", "commit"))
    s.close()
