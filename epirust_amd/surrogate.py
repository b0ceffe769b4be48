"""Learned surrogate for parameter sweeps of the per-hour agent step.

An MLP (9 -> 4 x 512 GELU -> 100 days x 6 compartments) maps transmission, mobility and
intervention parameters to the daily S/E/I/H/R/D fractions of a 100,000-agent region
(632 x 632 grid, 2,400 hours). It is trained on ensembles that the batched GPU simulator
generates, with the library's own bf16-MFMA forward/backward kernels (epi_mlp_*), and the
plain simulation stays the fallback and the measure of its error:

* inputs are Sobol points over the ranges in tests/golden/configs/cfg5_surrogate.json;
* a daily sample is the row with hour % 24 == 1 (the reference's plotting rule,
  engine/plot/plot.py:33); a run that stops early carries its last row forward;
* peak infected = max over days of the infected fraction; final size = infected +
  hospitalized + recovered + deceased on the last day (the reference's `totalinfected`,
  engine/plot/update_total_infections.py:19);
* two independently initialised networks are trained on linear (standardized) fractions,
  so that the fit estimates the mean curve without the downward bias a log or square-root
  target would give;
* the predicted error of a point combines the two networks' disagreement on peak and final
  size with the standard error of the 32-seed ensemble mean it is judged against, estimated
  from the single-simulation residuals of the nearest training points; a point whose
  predicted error exceeds the threshold is simulated instead (32 seeds, batched).
"""
import copy
import ctypes
import json
import math
import os
import time

import numpy as np

import epirust_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SPEC_PATH = os.path.join(ROOT, "tests", "golden", "configs", "cfg5_surrogate.json")
BASE_PATH = os.path.join(ROOT, "tests", "golden", "configs", "cfg1_10k.json")
WIDTHS = [9, 512, 512, 512, 512, 600]


def load_spec(path=SPEC_PATH):
    with open(path) as f:
        return json.load(f)


def sobol_points(n, seed=0, dims=9):
    """Scrambled Sobol points in [0, 1)^dims."""
    from scipy.stats import qmc
    s = qmc.Sobol(d=dims, scramble=True, seed=seed)
    m = int(math.ceil(math.log2(max(n, 2))))
    return s.random_base2(m)[:n]


def to_params(unit, spec):
    lo = np.array([i["low"] for i in spec["inputs"]], dtype=np.float64)
    hi = np.array([i["high"] for i in spec["inputs"]], dtype=np.float64)
    return lo + np.asarray(unit, dtype=np.float64) * (hi - lo)


def point_config(p, spec, base=None, seed=0):
    """Reference-schema config for one parameter point (unscaled values)."""
    if base is None:
        with open(BASE_PATH) as f:
            base = json.load(f)
    c = copy.deepcopy(base)
    n = spec["agents"]
    c["population"]["Auto"] = {"number_of_agents": n, "working_percentage": float(p[3]),
                               "public_transport_percentage": float(p[4])}
    c["geography_parameters"]["grid_size"] = spec["grid_size"]
    c["hours"] = spec["hours"]
    c["disease"]["regular_transmission_rate"] = float(p[0])
    c["disease"]["high_transmission_rate"] = float(p[1])
    c["disease"]["death_rate"] = float(p[2])
    c["starting_infections"] = {"exposed": int(round(p[8])), "infected_mild_asymptomatic": 0,
                                "infected_mild_symptomatic": 0, "infected_severe": 0}
    iv = [{"Lockdown": {"at_number_of_infections": int(round(p[5] * n)), "essential_workers_population": float(p[6])}}]
    if p[7] > 0:
        start = spec["vaccination_start_hour"]
        iv += [{"Vaccinate": {"at_hour": start + 24 * k, "percent": float(p[7])}}
               for k in range((spec["hours"] - start) // 24)]
    c["interventions"] = iv
    c["seed"] = int(seed)
    return c


def daily_curves(rows, n, days):
    """(days, 6) compartment fractions from hourly rows (hour % 24 == 1), carried forward."""
    out = np.zeros((days, 6), dtype=np.float32)
    last = None
    by_hour = {int(r[0]): r for r in rows}
    for d in range(days):
        h = 24 * d + 1
        r = by_hour.get(h)
        if r is None:
            r = rows[-1] if len(rows) and rows[-1][0] < h else last
        if r is None:
            continue
        last = r
        out[d] = np.asarray(r[1:7], dtype=np.float64) / n
    return out


def peak_and_final(curves):
    """Peak infected fraction and final size from (..., days, 6) curves."""
    c = np.asarray(curves)
    peak = c[..., :, 2].max(axis=-1)
    final = c[..., -1, 2] + c[..., -1, 3] + c[..., -1, 4] + c[..., -1, 5]
    return peak, final


def mean_within_projection(peaks, finals, bound, seeds_list):
    """For each ensemble size N in `seeds_list`: the projected fraction of points at which
    the true mean is within `bound` (relative) of an N-seed ensemble mean on both peak
    infected and final size. `peaks` and `finals` are (points, seeds) single-run values.
    Per point, an N-seed mean's relative standard error is the single-run coefficient of
    variation / sqrt(N) (normal approximation); the two statistics are taken as
    independent. This is the score of a predictor that knew the true mean curve, and so an
    upper bound on the share of points a surrogate can answer without simulating them."""
    erf = np.frompyfunc(math.erf, 1, 1)

    def p_within(x, n):
        cv = x.std(axis=1, ddof=1) / np.maximum(x.mean(axis=1), 1e-12)
        z = bound * math.sqrt(n) / np.maximum(cv, 1e-12)
        return erf(z / math.sqrt(2.0)).astype(np.float64)

    return {str(n): float(np.mean(p_within(peaks, n) * p_within(finals, n))) for n in seeds_list}


def simulate(params, seeds, spec, batch=128, block=None, progress=None):
    """Daily curves (len(params), days, 6) from the batched GPU simulator."""
    days = spec["days"]
    n = spec["agents"]
    with open(BASE_PATH) as f:
        base = json.load(f)
    out = np.zeros((len(params), days, 6), dtype=np.float32)
    for s in range(0, len(params), batch):
        cfgs = [point_config(params[i], spec, base, seeds[i]) for i in range(s, min(s + batch, len(params)))]
        b = G.Batch(cfgs)
        if block:
            b.set_block_size(block)
        rows, _ = b.step(spec["hours"] - 1)
        b.close()
        for j, r in enumerate(rows):
            out[s + j] = daily_curves(r, n, days)
        if progress:
            progress(min(s + batch, len(params)), len(params))
    return out


class MLP:
    """The epi_mlp handle (bf16-MFMA MLP with Adam) over ctypes."""

    def __init__(self, widths=WIDTHS, seed=0, max_batch=1024):
        self.widths = list(widths)
        self.h = ctypes.c_void_p()
        w = (ctypes.c_uint32 * len(widths))(*widths)
        rc = G.lib().epi_mlp_create(w, len(widths), seed, max_batch, ctypes.byref(self.h))
        if rc:
            raise G.EpiError(rc, G.lib().epi_mlp_last_error().decode())
        self.max_batch = max_batch

    def close(self):
        if self.h:
            G.lib().epi_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        n = G.lib().epi_mlp_param_count(self.h)
        p = np.zeros(n, dtype=np.float32)
        G.lib().epi_mlp_get_params(self.h, p.ctypes.data, n)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        rc = G.lib().epi_mlp_set_params(self.h, p.ctypes.data, len(p))
        if rc:
            raise G.EpiError(rc, G.lib().epi_mlp_last_error().decode())

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        rc = G.lib().epi_mlp_train_step(self.h, x.ctypes.data, t.ctypes.data, len(x), lr, ctypes.byref(loss))
        if rc:
            raise G.EpiError(rc, G.lib().epi_mlp_last_error().decode())
        return loss.value

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), self.widths[-1]), dtype=np.float32)
        for s in range(0, len(x), self.max_batch):
            k = min(self.max_batch, len(x) - s)
            rc = G.lib().epi_mlp_forward(self.h, x[s:].ctypes.data, k, y[s:].ctypes.data)
            if rc:
                raise G.EpiError(rc, G.lib().epi_mlp_last_error().decode())
        return y


def train_mlp(X, Y, seed, epochs=60, batch=512, lr=1e-3, widths=WIDTHS, log=None):
    """Fits one network on standardized targets; returns (mlp, loss history)."""
    rng = np.random.default_rng(seed)
    m = MLP(widths, seed=seed, max_batch=max(batch, 1024))
    hist = []
    steps_per_epoch = max(1, len(X) // batch)
    total = epochs * steps_per_epoch
    step = 0
    for ep in range(epochs):
        perm = rng.permutation(len(X))
        acc = 0.0
        for k in range(steps_per_epoch):
            idx = perm[k * batch:(k + 1) * batch]
            # cosine decay of the learning rate
            cur = lr * 0.5 * (1 + math.cos(math.pi * step / total))
            acc += m.train_step(X[idx], Y[idx], max(cur, lr * 0.02))
            step += 1
        hist.append(acc / steps_per_epoch)
        if log:
            log(ep, hist[-1])
    return m, hist


class NoiseModel:
    """Seed-to-seed spread of peak and final size around the network's answer, as a function
    of the parameters: the root mean square of the training simulations' residuals (single
    seed minus prediction) over the k nearest training points. Divided by sqrt(seeds) it is
    the standard error of a seeds-member ensemble mean, the quantity a prediction is judged
    against."""

    def __init__(self, unit_x, resid_peak, resid_final, k=64):
        from scipy.spatial import cKDTree
        self.x = np.asarray(unit_x, dtype=np.float64)
        self.r = np.stack([np.asarray(resid_peak, np.float64), np.asarray(resid_final, np.float64)], axis=1)
        self.k = int(min(k, len(self.x)))
        self.tree = cKDTree(self.x)

    def spread(self, unit_x):
        """(n, 2) RMS single-simulation residual of peak and final size near each point."""
        _, nb = self.tree.query(np.asarray(unit_x, dtype=np.float64), k=self.k)
        nb = nb.reshape(len(unit_x), self.k)
        return np.sqrt(np.mean(self.r[nb] ** 2, axis=1))


class Surrogate:
    """Two-network ensemble with standardized I/O, a noise model and a simulation fallback."""

    def __init__(self, spec, members, y_mean, y_std, noise=None):
        self.spec = spec
        self.members = members
        self.y_mean = y_mean
        self.y_std = y_std
        self.noise = noise

    @classmethod
    def fit(cls, unit_x, curves, spec, epochs=100, seeds=(1, 2), log=None):
        days = spec["days"]
        Y = curves.reshape(len(curves), days * 6).astype(np.float64)
        y_mean = Y.mean(axis=0)
        y_std = Y.std(axis=0) + 1e-4
        Yn = ((Y - y_mean) / y_std).astype(np.float32)
        X = np.asarray(unit_x, dtype=np.float32)
        members = []
        for s in seeds:
            m, hist = train_mlp(X, Yn, seed=s, epochs=epochs, log=log)
            members.append((m, hist))
        sur = cls(spec, [m for m, _ in members], y_mean, y_std)
        fit_p, fit_f = peak_and_final(np.mean(sur.predict_members(X), axis=0))
        lab_p, lab_f = peak_and_final(curves)
        sur.noise = NoiseModel(X, lab_p - fit_p, lab_f - fit_f)
        return sur, [h for _, h in members]

    def predict_members(self, unit_x):
        days = self.spec["days"]
        out = []
        for m in self.members:
            yn = m.forward(np.asarray(unit_x, dtype=np.float32)).astype(np.float64)
            y = yn * self.y_std + self.y_mean
            out.append(np.clip(y, 0.0, 1.0).reshape(len(unit_x), days, 6))
        return out

    def predict(self, unit_x, seeds=None):
        """Mean curves and the predicted relative error of peak and final size against a
        `seeds`-member simulation ensemble (default: the evaluation ensemble size): the
        members' disagreement and the ensemble mean's standard error, added in quadrature."""
        ys = self.predict_members(unit_x)
        mean = np.mean(ys, axis=0)
        pk, fs = zip(*[peak_and_final(y) for y in ys])
        pk, fs = np.array(pk), np.array(fs)
        mp, mf = peak_and_final(mean)
        mp, mf = np.maximum(mp, 1e-9), np.maximum(mf, 1e-9)
        dis_p, dis_f = np.abs(pk[0] - pk[1]) / mp, np.abs(fs[0] - fs[1]) / mf
        if self.noise is None:
            return mean, np.maximum(dis_p, dis_f)
        seeds = self.spec["eval_seeds"] if seeds is None else seeds
        sp = self.noise.spread(unit_x) / math.sqrt(seeds)
        err = np.maximum(np.hypot(dis_p, sp[:, 0] / mp), np.hypot(dis_f, sp[:, 1] / mf))
        return mean, err

    def sweep(self, unit_x, seeds_per_point, threshold=None, block=None):
        """Curves for every point: the surrogate where its predicted error is within the
        threshold, else the mean of a seeds_per_point simulation ensemble. Returns
        (curves, fallback mask, seconds)."""
        threshold = self.spec["max_relative_error"] if threshold is None else threshold
        t0 = time.perf_counter()
        curves, err = self.predict(unit_x, seeds_per_point)
        fb = err > threshold
        idx = np.nonzero(fb)[0]
        if len(idx):
            params = to_params(unit_x[idx], self.spec)
            rep = np.repeat(params, seeds_per_point, axis=0)
            seeds = np.tile(np.arange(1, seeds_per_point + 1), len(idx))
            sims = simulate(rep, seeds, self.spec, block=block)
            curves[idx] = sims.reshape(len(idx), seeds_per_point, self.spec["days"], 6).mean(axis=1)
        return curves, fb, time.perf_counter() - t0


def calibrate_threshold(predicted_err, within, target=0.95):
    """Largest disagreement threshold t such that, simulating every point whose predicted
    error exceeds t, at least `target` of the points end within the bound (points not
    simulated count as within only if their surrogate answer is)."""
    order = np.sort(np.unique(np.concatenate([predicted_err, [0.0]])))[::-1]
    for t in order:
        answered = predicted_err <= t
        if np.mean(~answered | within) >= target:
            return float(t)
    return 0.0


def nearest_neighbour(train_x, train_curves, query_x):
    """The no-network method: each query takes the curve of its nearest training point."""
    tx = np.asarray(train_x, dtype=np.float64)
    out = np.zeros((len(query_x),) + train_curves.shape[1:], dtype=np.float32)
    for i, q in enumerate(np.asarray(query_x, dtype=np.float64)):
        out[i] = train_curves[np.argmin(((tx - q) ** 2).sum(axis=1))]
    return out


def relative_errors(pred_curves, ref_curves):
    pp, pf = peak_and_final(pred_curves)
    rp, rf = peak_and_final(ref_curves)
    return np.abs(pp - rp) / np.maximum(rp, 1e-9), np.abs(pf - rf) / np.maximum(rf, 1e-9)


def run_pipeline(out_dir, n_train=None, n_eval=None, eval_seeds=None, epochs=100, block=None, log=print):
    """Generate data, train, evaluate; writes surrogate_weights.npz and surrogate_metrics.json."""
    spec = load_spec()
    n_train = n_train or spec["train_points"]
    n_eval = n_eval or spec["eval_points"]
    eval_seeds = eval_seeds or spec["eval_seeds"]
    os.makedirs(out_dir, exist_ok=True)
    metrics = {"train_points": n_train, "eval_points": n_eval, "eval_seeds": eval_seeds, "epochs": epochs,
               "agents": spec["agents"], "grid_size": spec["grid_size"], "hours": spec["hours"]}

    def prog(label):
        last = [0.0]

        def f(done, total):
            if time.perf_counter() - last[0] > 20 or done == total:
                last[0] = time.perf_counter()
                log("[surrogate] %s %d/%d" % (label, done, total))
        return f

    ux = sobol_points(n_train, seed=0)
    t0 = time.perf_counter()
    curves = simulate(to_params(ux, spec), 1_000_000 + np.arange(n_train), spec, block=block, progress=prog("train sims"))
    metrics["train_sim_seconds"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    sur, hists = Surrogate.fit(ux, curves, spec, epochs=epochs,
                               log=lambda ep, l: log("[surrogate] epoch %d loss %.5f" % (ep, l)) if ep % 10 == 0 else None)
    metrics["train_seconds"] = time.perf_counter() - t0
    metrics["final_train_loss"] = [h[-1] for h in hists]

    ex = sobol_points(n_eval, seed=12345)
    params = to_params(ex, spec)
    rep = np.repeat(params, eval_seeds, axis=0)
    seeds = np.tile(np.arange(1, eval_seeds + 1), n_eval)
    t0 = time.perf_counter()
    ens = simulate(rep, seeds, spec, block=block, progress=prog("eval sims")).reshape(n_eval, eval_seeds, spec["days"], 6)
    plain_seconds = time.perf_counter() - t0
    ref = ens.mean(axis=1)
    metrics["plain_sweep_seconds"] = plain_seconds
    metrics["plain_points_per_s"] = n_eval / plain_seconds

    t0 = time.perf_counter()
    pred, err_hat = sur.predict(ex, eval_seeds)
    metrics["surrogate_only_seconds"] = time.perf_counter() - t0
    ep_, ef_ = relative_errors(pred, ref)
    nn = nearest_neighbour(ux, curves, ex)
    np_, nf_ = relative_errors(nn, ref)
    ok = (ep_ <= spec["max_relative_error"]) & (ef_ <= spec["max_relative_error"])
    ok_nn = (np_ <= spec["max_relative_error"]) & (nf_ <= spec["max_relative_error"])
    metrics.update({
        "surrogate_median_rel_err_peak": float(np.median(ep_)), "surrogate_median_rel_err_final": float(np.median(ef_)),
        "surrogate_fraction_within": float(ok.mean()),
        "nn_median_rel_err_peak": float(np.median(np_)), "nn_median_rel_err_final": float(np.median(nf_)),
        "nn_fraction_within": float(ok_nn.mean()),
    })
    # seed-to-seed spread of the reference ensembles (how noisy a single simulation is)
    sp, sf = peak_and_final(ens)
    metrics["ensemble_cv_peak_median"] = float(np.median(sp.std(axis=1) / np.maximum(sp.mean(axis=1), 1e-9)))
    metrics["ensemble_cv_final_median"] = float(np.median(sf.std(axis=1) / np.maximum(sf.mean(axis=1), 1e-9)))
    # What the ensemble's own noise allows: half the difference of the two half-ensemble
    # means has the variance of the full ensemble mean's error around the true mean, so the
    # fraction of points where it is within the bound estimates the score of a predictor
    # that knew the true mean curve exactly.
    h = eval_seeds // 2
    if h:
        hp = np.abs(sp[:, :h].mean(axis=1) - sp[:, h:2 * h].mean(axis=1)) / 2 / np.maximum(sp.mean(axis=1), 1e-9)
        hf = np.abs(sf[:, :h].mean(axis=1) - sf[:, h:2 * h].mean(axis=1)) / 2 / np.maximum(sf.mean(axis=1), 1e-9)
        metrics["exact_mean_fraction_within_estimate"] = float(
            ((hp <= spec["max_relative_error"]) & (hf <= spec["max_relative_error"])).mean())
    # The same bound projected to larger judging ensembles from each point's single-run
    # spread: how many seeds per point would let a perfect predictor answer 95% / 99% of
    # the points within the error bound (a sweep that simulates the rest is at most
    # 1 / (simulated fraction) times faster than the plain sweep).
    if eval_seeds >= 2:
        proj = mean_within_projection(sp, sf, spec["max_relative_error"], [32, 64, 128, 256, 512, 1024, 2048, 4096])
        metrics["exact_mean_fraction_within_projected"] = proj
        for target in (0.95, 0.99):
            ns = [int(n) for n, v in proj.items() if v >= target]
            metrics["seeds_for_exact_mean_within_%d" % round(100 * target)] = min(ns) if ns else None

    curves_fb, fb, sweep_seconds = sur.sweep(ex, eval_seeds, block=block)
    efp, eff = relative_errors(curves_fb, ref)
    okf = (efp <= spec["max_relative_error"]) & (eff <= spec["max_relative_error"])
    metrics.update({
        "fallback_fraction": float(fb.mean()),
        "surrogate_sweep_seconds": sweep_seconds,
        "surrogate_points_per_s": n_eval / sweep_seconds,
        "speedup_vs_plain_sweep": (n_eval / sweep_seconds) / (n_eval / plain_seconds),
        "with_fallback_fraction_within": float(okf.mean()),
        "with_fallback_max_rel_err_surrogate_points": float(np.max(np.maximum(efp, eff)[~fb])) if (~fb).any() else None,
    })
    # Calibrated gate: the disagreement threshold is chosen on the even-indexed evaluation
    # points (the largest threshold that keeps >= 95% of them within the error bound once
    # the points above it are simulated) and then measured on the odd-indexed points.
    calib, test = np.arange(0, n_eval, 2), np.arange(1, n_eval, 2)
    t_star = calibrate_threshold(err_hat[calib], ok[calib], target=0.95)
    sub = ex[test]
    curves_t, fb_t, sweep_t = sur.sweep(sub, eval_seeds, threshold=t_star, block=block)
    etp, etf = relative_errors(curves_t, ref[test])
    ok_t = (etp <= spec["max_relative_error"]) & (etf <= spec["max_relative_error"])
    plain_rate = n_eval / plain_seconds
    metrics["calibrated"] = {
        "target_fraction_within": 0.95, "threshold": t_star, "calibration_points": int(len(calib)),
        "test_points": int(len(test)), "test_fallback_fraction": float(fb_t.mean()),
        "test_fraction_within": float(ok_t.mean()),
        "test_median_rel_err_peak": float(np.median(etp)), "test_median_rel_err_final": float(np.median(etf)),
        "test_points_per_s": len(test) / sweep_t, "speedup_vs_plain_sweep": (len(test) / sweep_t) / plain_rate,
    }
    np.savez_compressed(os.path.join(out_dir, "surrogate_eval_points.npz"), unit_x=ex, predicted_err=err_hat,
                        rel_err_peak=ep_, rel_err_final=ef_, nn_rel_err_peak=np_, nn_rel_err_final=nf_,
                        ensemble_peak=sp, ensemble_final=sf)
    flat = [m.params() for m in sur.members]
    np.savez_compressed(os.path.join(out_dir, "surrogate_weights.npz"), member0=flat[0], member1=flat[1],
                        y_mean=sur.y_mean, y_std=sur.y_std, noise_x=sur.noise.x, noise_resid=sur.noise.r)
    with open(os.path.join(out_dir, "surrogate_metrics.json"), "w") as f:
        json.dump(metrics, f, indent=1)
    log(json.dumps(metrics))
    return metrics


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="surrogate_out")
    ap.add_argument("--train", type=int, default=0)
    ap.add_argument("--eval", type=int, default=0)
    ap.add_argument("--seeds", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=100)
    a = ap.parse_args()
    run_pipeline(a.out, a.train or None, a.eval or None, a.seeds or None, a.epochs,
                 log=lambda s: print(s, flush=True))
