"""Pins the fp64 reference MLP (oracle/mlp_ref.py) by central finite differences, and its
GELU against the closed form: the checker of the surrogate's MFMA forward/backward."""
import numpy as np

from oracle import mlp_ref as M


def test_gelu_closed_form_points():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    # odd/even structure: gelu(x) - gelu(-x) = x
    x = np.linspace(-4, 4, 41)
    assert np.allclose(M.gelu(x) - M.gelu(-x), x, atol=1e-12)
    # derivative against a central difference
    h = 1e-6
    assert np.allclose(M.gelu_grad(x), (M.gelu(x + h) - M.gelu(x - h)) / (2 * h), atol=1e-7)


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    widths = [5, 7, 6, 4]
    n_params = sum(k * n + n for k, n in zip(widths[:-1], widths[1:]))
    p = rng.normal(0, 0.5, n_params)
    x = rng.normal(0, 1, (9, widths[0]))
    t = rng.normal(0, 1, (9, widths[-1]))
    loss, g = M.loss_and_grad(p, widths, x, t)
    h = 1e-6
    num = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (M.loss_and_grad(p + e, widths, x, t)[0] - M.loss_and_grad(p - e, widths, x, t)[0]) / (2 * h)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-8)


def test_bf16_rounding_known_values():
    # bf16 spacing at 1.0 is 2^-7
    vals = np.array([1.0, -2.0, 3.140625, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 1.0 + 2 ** -8 + 2 ** -10,
                     1.0 + 2 ** -7 + 2 ** -8], dtype=np.float32)
    r = M.bf16_round(vals)
    assert r[0] == 1.0 and r[1] == -2.0 and r[2] == 3.140625
    assert r[3] == 1.0                                 # tie between 1 and 1+2^-7 rounds to even
    assert r[4] == np.float32(1.0 + 2 ** -7)           # below the tie rounds down
    assert r[5] == np.float32(1.0 + 2 ** -7)           # above the tie rounds up
    assert r[6] == np.float32(1.0 + 2 ** -6)           # tie between odd 1+2^-7 and 1+2^-6 rounds to even
