"""The surrogate's bf16-MFMA kernels against the fp64 reference MLP (oracle/mlp_ref.py).

Tolerances: the GEMM takes bf16 inputs whose products are exact in fp32, so against an
fp64 product of the same bf16 values the error is fp32 accumulation only (rtol 1e-5 of
sum |a||b|). The MLP stores weights and activations in bf16 (8-bit significand), so the
network output and gradients are compared by relative L2 norm: 3e-2 for the forward pass
and 8e-2 for each layer's gradients.
"""
import ctypes

import numpy as np
import pytest

import epirust_amd as G
from oracle import mlp_ref as M
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


def gemm(a16, b16, m, n, k, ta, tb):
    c = np.zeros((m, n), dtype=np.float32)
    rc = G.lib().epi_gemm_bf16(a16.ctypes.data, b16.ctypes.data, c.ctypes.data, m, n, k, ta, tb)
    assert rc == 0, G.lib().epi_mlp_last_error()
    return c


@pytest.mark.parametrize("m,n,k,ta,tb", [(64, 64, 32, 0, 0), (100, 70, 50, 0, 0), (256, 512, 512, 0, 0),
                                         (33, 129, 65, 1, 0), (130, 31, 96, 0, 1), (77, 91, 200, 1, 1),
                                         (1, 600, 512, 0, 0), (512, 512, 1000, 1, 0)])
def test_mfma_gemm_matches_fp64(m, n, k, ta, tb):
    say("[mlp] gemm %dx%dx%d ta=%d tb=%d" % (m, n, k, ta, tb))
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    a = rng.normal(0, 1, (k, m) if ta else (m, k)).astype(np.float32)
    b = rng.normal(0, 1, (n, k) if tb else (k, n)).astype(np.float32)
    c = gemm(M.bf16_bits(a), M.bf16_bits(b), m, n, k, ta, tb)
    A = M.bf16_round(a).astype(np.float64)
    B = M.bf16_round(b).astype(np.float64)
    A = A.T if ta else A
    B = B.T if tb else B
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    assert np.all(np.abs(c - ref) <= 1e-5 * scale + 1e-30)


def make_mlp(widths, batch, seed=5):
    h = ctypes.c_void_p()
    w = (ctypes.c_uint32 * len(widths))(*widths)
    rc = G.lib().epi_mlp_create(w, len(widths), seed, batch, ctypes.byref(h))
    assert rc == 0, G.lib().epi_mlp_last_error()
    return h


def params_of(h):
    n = G.lib().epi_mlp_param_count(h)
    p = np.zeros(n, dtype=np.float32)
    assert G.lib().epi_mlp_get_params(h, p.ctypes.data, n) == 0
    return p


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_and_gradients_match_reference():
    widths = [9, 512, 512, 512, 512, 600]
    say("[mlp] forward/gradients %s" % widths)
    rng = np.random.default_rng(11)
    B = 256
    h = make_mlp(widths, B)
    p = params_of(h)
    x = rng.uniform(0, 1, (B, 9)).astype(np.float32)
    t = rng.uniform(0, 1, (B, 600)).astype(np.float32)
    y = np.zeros((B, 600), dtype=np.float32)
    assert G.lib().epi_mlp_forward(h, x.ctypes.data, B, y.ctypes.data) == 0
    y_ref, _ = M.forward(p, widths, x)
    assert rel(y, y_ref) < 3e-2
    g = np.zeros_like(p)
    loss = ctypes.c_float()
    assert G.lib().epi_mlp_gradients(h, x.ctypes.data, t.ctypes.data, B, g.ctypes.data, ctypes.byref(loss)) == 0
    l_ref, g_ref = M.loss_and_grad(p, widths, x, t)
    assert abs(loss.value - l_ref) / l_ref < 3e-2
    o = 0
    for k, n in zip(widths[:-1], widths[1:]):
        for size in (k * n, n):
            assert rel(g[o:o + size], g_ref[o:o + size]) < 8e-2, "layer block at %d" % o
            o += size
    G.lib().epi_mlp_destroy(h)


def test_training_reduces_loss_on_a_smooth_target():
    widths = [9, 128, 128, 16]
    say("[mlp] training %s" % widths)
    rng = np.random.default_rng(2)
    X = rng.uniform(0, 1, (2048, 9)).astype(np.float32)
    W = rng.normal(0, 1, (9, 16))
    T = np.tanh(X @ W).astype(np.float32)
    h = make_mlp(widths, 256)
    losses = []
    for step in range(400):
        idx = rng.integers(0, len(X), 256)
        xb = np.ascontiguousarray(X[idx])  # named: the arrays must outlive the call
        tb = np.ascontiguousarray(T[idx])
        loss = ctypes.c_float()
        rc = G.lib().epi_mlp_train_step(h, xb.ctypes.data, tb.ctypes.data, 256, 1e-3, ctypes.byref(loss))
        assert rc == 0
        losses.append(loss.value)
    assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:20])
    G.lib().epi_mlp_destroy(h)
