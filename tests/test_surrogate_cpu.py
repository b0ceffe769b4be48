"""Host logic of the surrogate pipeline (epirust_amd/surrogate.py), no GPU: daily sampling,
the peak / final-size definitions, the parameter-point config and the calibrated gate."""
import numpy as np
import pytest

from epirust_amd import surrogate as S


def rows_for(hours, n=1000):
    """Hourly rows (hour, S, E, I, H, R, D) with I = hour, R = 2*hour, the rest in S."""
    out = []
    for h in range(1, hours + 1):
        out.append([h, n - 3 * h, 0, h, 0, 2 * h, 0])
    return np.array(out, dtype=np.uint32)


def test_daily_curves_sample_hour_1_of_each_day_and_carry_forward():
    # the reference's plots read hour % 24 == 1 (engine/plot/plot.py:33); a run that
    # stopped at hour 50 repeats its last row
    c = S.daily_curves(rows_for(50), 1000, 4)
    assert c.shape == (4, 6)
    assert np.allclose(c[:, 2] * 1000, [1, 25, 49, 50])
    assert np.allclose(c[:, 4] * 1000, [2, 50, 98, 100])
    assert np.allclose(c.sum(axis=1), 1.0)


def test_peak_and_final_follow_the_reference_definitions():
    # peak = max infected; final = infected + hospitalized + recovered + deceased on the
    # last day (engine/plot/update_total_infections.py:19)
    c = np.zeros((3, 6))
    c[:, 2] = [0.1, 0.3, 0.2]
    c[-1] = [0.4, 0.05, 0.2, 0.05, 0.25, 0.05]
    peak, final = S.peak_and_final(c)
    assert peak == pytest.approx(0.3)
    assert final == pytest.approx(0.2 + 0.05 + 0.25 + 0.05)
    batch = np.stack([c, c * 0.5])
    p2, f2 = S.peak_and_final(batch)
    assert p2.shape == (2,) and p2[1] == pytest.approx(0.15)


def test_point_config_maps_the_nine_inputs():
    spec = S.load_spec()
    p = S.to_params(np.full(9, 0.5), spec)
    c = S.point_config(p, spec, seed=7)
    n = spec["agents"]
    assert c["population"]["Auto"]["number_of_agents"] == n
    assert c["geography_parameters"]["grid_size"] == spec["grid_size"]
    assert c["disease"]["regular_transmission_rate"] == pytest.approx(0.03)
    assert c["disease"]["high_transmission_rate"] == pytest.approx(0.25)
    lock = c["interventions"][0]["Lockdown"]
    assert lock["at_number_of_infections"] == round(0.0125 * n)
    vacc = [iv["Vaccinate"] for iv in c["interventions"][1:]]
    assert len(vacc) == (spec["hours"] - spec["vaccination_start_hour"]) // 24
    assert vacc[0]["at_hour"] == spec["vaccination_start_hour"] and vacc[1]["at_hour"] - vacc[0]["at_hour"] == 24
    assert c["starting_infections"]["exposed"] == round(52.5)
    assert c["seed"] == 7
    # no vaccination at the low end of its range
    c0 = S.point_config(S.to_params(np.zeros(9), spec), spec)
    assert len(c0["interventions"]) == 1


def test_calibrated_threshold():
    e = np.array([0.01, 0.02, 0.03, 0.2, 0.05])
    ok = np.array([1, 1, 0, 0, 1], dtype=bool)
    assert S.calibrate_threshold(e, ok, 0.95) == pytest.approx(0.02)  # simulates the 0.03, 0.05, 0.2 points
    assert S.calibrate_threshold(e, ok, 0.6) == pytest.approx(0.2)    # answers everything
    assert S.calibrate_threshold(e, np.zeros(5, bool), 0.95) == 0.0   # simulates everything


def test_relative_errors():
    ref = np.zeros((2, 3, 6))
    ref[:, :, 2] = 0.2
    ref[:, -1, 4] = 0.3
    pred = ref.copy()
    pred[1, :, 2] = 0.22
    ep, ef = S.relative_errors(pred, ref)
    assert ep[0] == 0 and ep[1] == pytest.approx(0.1)
    assert ef[1] == pytest.approx(0.02 / 0.5)
