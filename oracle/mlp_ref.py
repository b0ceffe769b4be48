"""fp64 numpy restatement of the surrogate MLP (test infrastructure only: the checker
for the bf16-MFMA kernels in epirust_amd/csrc/mlp.hip).

Network: widths w0 -> ... -> wL, GELU (tanh form) after every hidden layer, identity
output, loss = mean((y - t)^2). Parameters are flattened as W0 (K x N row-major), b0,
W1, b1, ... exactly like the library's epi_mlp_get_params / epi_mlp_set_params.
"""
import numpy as np

C = np.sqrt(2.0 / np.pi)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C * (1.0 + 3.0 * 0.044715 * z * z)


def unflatten(params, widths):
    out, o = [], 0
    for k, n in zip(widths[:-1], widths[1:]):
        W = params[o:o + k * n].reshape(k, n)
        o += k * n
        b = params[o:o + n]
        o += n
        out.append((W, b))
    assert o == len(params)
    return out


def forward(params, widths, x):
    layers = unflatten(np.asarray(params, dtype=np.float64), widths)
    a = np.asarray(x, dtype=np.float64)
    cache = []
    for i, (W, b) in enumerate(layers):
        z = a @ W + b
        cache.append((a, z))
        a = z if i == len(layers) - 1 else gelu(z)
    return a, cache


def loss_and_grad(params, widths, x, t):
    y, cache = forward(params, widths, x)
    t = np.asarray(t, dtype=np.float64)
    loss = np.mean((y - t) ** 2)
    dz = 2.0 * (y - t) / y.size
    layers = unflatten(np.asarray(params, dtype=np.float64), widths)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        a, z = cache[i]
        W, _ = layers[i]
        grads[i] = (a.T @ dz, dz.sum(axis=0))
        if i > 0:
            da = dz @ W.T
            dz = da * gelu_grad(cache[i - 1][1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
