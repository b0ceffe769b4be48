"""Host logic of the surrogate pipeline (epirust_amd/surrogate.py), no GPU: daily sampling,
the peak / final-size definitions, the parameter-point config, the calibrated gate and the
projection of the judging ensemble's noise to other ensemble sizes."""
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


def test_noise_model_recovers_a_local_spread():
    # residuals of spread 0.01 where x0 < 0.5 and 0.04 elsewhere (random signs): the RMS over
    # the nearest training points recovers each side's spread away from the boundary
    g = np.random.default_rng(3)
    x = g.random((4000, 2))
    sd = np.where(x[:, 0] < 0.5, 0.01, 0.04)
    sign = g.choice([-1.0, 1.0], size=4000)
    nm = S.NoiseModel(x, sign * sd, 2 * sign * sd, k=64)
    sp = nm.spread(np.array([[0.1, 0.5], [0.9, 0.5]]))
    assert sp[0] == pytest.approx([0.01, 0.02])
    assert sp[1] == pytest.approx([0.04, 0.08])
    assert S.NoiseModel(x[:10], sd[:10], sd[:10], k=64).k == 10  # k is capped at the training size


class _Member:
    """Stands in for a network: returns fixed standardized outputs."""

    def __init__(self, y):
        self.y = y

    def forward(self, x):
        return np.repeat(self.y[None], len(x), axis=0)


def test_predicted_error_adds_disagreement_and_ensemble_standard_error():
    spec = dict(S.load_spec(), days=2)
    # curves: infected 0.1 then 0.2 (peak 0.2), recovered 0.3 on the last day (final 0.5)
    c = np.zeros((2, 6))
    c[:, 2] = [0.1, 0.2]
    c[1, 4] = 0.3
    c[:, 0] = 1 - c.sum(axis=1)
    y = c.reshape(-1)
    x = np.full((1, 9), 0.5)
    # identical members, residual spread 0.02 (peak) and 0.04 (final), 16 seeds: standard
    # errors 0.005 and 0.01, relative 0.025 and 0.02 -> predicted error 0.025
    nm = S.NoiseModel(np.random.default_rng(0).random((100, 9)), np.full(100, 0.02), np.full(100, -0.04))
    sur = S.Surrogate(spec, [_Member(np.zeros(12)), _Member(np.zeros(12))], y, np.ones(12), nm)
    mean, err = sur.predict(x, seeds=16)
    assert np.allclose(mean[0], c)
    assert err[0] == pytest.approx(0.025)
    # members at 0.2 -/+ 0.03 on the peak day: 0.06 apart, relative 0.3 of the mean peak 0.2
    d = np.zeros(12)
    d[1 * 6 + 2] = 0.03  # day 1 infected
    sur2 = S.Surrogate(spec, [_Member(d), _Member(-d)], y, np.ones(12), nm)
    _, err2 = sur2.predict(x, seeds=16)
    assert err2[0] == pytest.approx(np.hypot(0.06 / 0.2, 0.025))
    # without a noise model the error is the disagreement alone
    _, err3 = S.Surrogate(spec, [_Member(d), _Member(-d)], y, np.ones(12)).predict(x)
    assert err3[0] == pytest.approx(0.06 / 0.2)


def test_mean_within_projection_closed_form():
    """One point whose single-run values alternate m - d, m + d (sample CV set exactly) and
    a final size with no spread: the projected fraction is P(|Z| < bound * sqrt(N) / cv)
    of a standard normal, here 0.95 at the N where that ratio is 1.959964."""
    from scipy.stats import norm
    bound, k, m = 0.05, 32, 0.2
    target_cv = bound * np.sqrt(128) / norm.ppf(0.975)  # makes N = 128 the 95% point
    d = target_cv * m / np.sqrt(k / (k - 1))             # std(ddof=1) of +-d is d * sqrt(k / (k - 1))
    peaks = np.array([[m - d, m + d] * (k // 2)])
    finals = np.full((1, k), 0.5)
    proj = S.mean_within_projection(peaks, finals, bound, [32, 128, 512])
    assert proj["128"] == pytest.approx(0.95, abs=1e-9)
    for n in (32, 512):
        assert proj[str(n)] == pytest.approx(2 * norm.cdf(bound * np.sqrt(n) / target_cv) - 1, abs=1e-9)
    assert proj["32"] < proj["128"] < proj["512"]


def test_mean_within_projection_matches_simulated_ensembles():
    """Points with log-spread coefficients of variation: the projection at N agrees with
    the fraction of simulated N-seed means (of normal single runs around a known mean)
    that land within the bound of that mean on both statistics."""
    rng = np.random.default_rng(7)
    pts, k, bound = 400, 64, 0.05
    mean_p, mean_f = rng.uniform(0.05, 0.3, pts), rng.uniform(0.2, 0.9, pts)
    cv_p, cv_f = np.exp(rng.uniform(np.log(0.02), np.log(0.6), pts)), np.exp(rng.uniform(np.log(0.01), np.log(0.3), pts))
    # the projection from k sampled runs per point
    peaks = mean_p[:, None] * (1 + cv_p[:, None] * rng.standard_normal((pts, k)))
    finals = mean_f[:, None] * (1 + cv_f[:, None] * rng.standard_normal((pts, k)))
    proj = S.mean_within_projection(peaks, finals, bound, [32, 256])
    for n in (32, 256):
        trials = 200
        ep = np.abs(cv_p[:, None] * rng.standard_normal((pts, trials)) / np.sqrt(n))
        ef = np.abs(cv_f[:, None] * rng.standard_normal((pts, trials)) / np.sqrt(n))
        sim = ((ep <= bound) & (ef <= bound)).mean()
        assert proj[str(n)] == pytest.approx(sim, abs=0.03), (n, proj[str(n)], sim)
