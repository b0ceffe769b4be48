"""The surrogate pipeline end to end at a small scale, on the GPU: batched simulations of
Sobol points (100,000 agents, 2,400 hours each), two bf16-MFMA networks trained on them,
and the answers compared with 8-seed simulation ensembles of held-out points.

The network must beat the no-network method (each query takes its nearest training
point's curve) on median relative error of both peak infected and final size, and the
sweep's fallback must return exactly the simulation ensemble for the points it simulates.
"""
import numpy as np
import pytest

from epirust_amd import surrogate as S
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


def test_surrogate_beats_nearest_neighbour_and_falls_back_exactly():
    spec = S.load_spec()
    n_train, n_eval, seeds = 512, 32, 8
    say("[surrogate] %d training simulations" % n_train)
    ux = S.sobol_points(n_train, seed=0)
    curves = S.simulate(S.to_params(ux, spec), 1_000_000 + np.arange(n_train), spec)
    assert np.all(np.abs(curves.sum(axis=2) - 1.0) < 1e-4)  # compartments sum to the population
    say("[surrogate] training")
    sur, hists = S.Surrogate.fit(ux, curves, spec, epochs=60)
    for h in hists:
        assert h[-1] < 0.5 * h[0]
    say("[surrogate] %d evaluation points x %d seeds" % (n_eval, seeds))
    ex = S.sobol_points(n_eval, seed=12345)
    rep = np.repeat(S.to_params(ex, spec), seeds, axis=0)
    ens = S.simulate(rep, np.tile(np.arange(1, seeds + 1), n_eval), spec).reshape(n_eval, seeds, spec["days"], 6)
    ref = ens.mean(axis=1)
    pred, err_hat = sur.predict(ex)
    ep, ef = S.relative_errors(pred, ref)
    np_, nf = S.relative_errors(S.nearest_neighbour(ux, curves, ex), ref)
    say("[surrogate] median rel. err peak %.3f final %.3f; nearest neighbour %.3f %.3f"
        % (np.median(ep), np.median(ef), np.median(np_), np.median(nf)))
    assert np.median(ep) < np.median(np_)
    assert np.median(ef) < np.median(nf)
    # fallback: force a threshold that sends some points to the simulator
    t = float(np.median(err_hat))
    out, fb, _ = sur.sweep(ex, seeds, threshold=t)
    assert fb.any() and (~fb).any()
    assert np.allclose(out[fb], ref[fb], atol=1e-6)
    assert np.allclose(out[~fb], pred[~fb])
    for m in sur.members:
        m.close()
