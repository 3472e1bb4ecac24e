"""NumPy restatement of the wave-term network, used as the checker of the device kernels.

Layout of the flat parameter vector (as libmhx stores it): for each of the five
layers, W (out x in, row-major) followed by b (out); layer widths
64 -> 128 -> 128 -> 128 -> 128 -> 6.  Inputs u (m x 2) map to features
[sin(2 pi u B^T), cos(2 pi u B^T)] with the 32 x 2 Fourier matrix B; hidden
layers use SiLU (or ReLU); the loss is the mean of squared errors over m x 6 outputs.
"""
import numpy as np

IN = (64, 128, 128, 128, 128)
OUT = (128, 128, 128, 128, 6)
NPARAMS = sum(o * i + o for i, o in zip(IN, OUT))


def unpack(params):
    out, o = [], 0
    for i, n in zip(IN, OUT):
        W = params[o:o + n * i].reshape(n, i)
        o += n * i
        b = params[o:o + n]
        o += n
        out.append((W, b))
    assert o == len(params)
    return out


def inputs(r, v):
    """(r, v) -> u in [0, 1]^2: log(r + 1e-3) over [log 1e-3, log 100.001], log(max(v, 5e-4)) over
    [log 5e-4, log 50]."""
    r = np.asarray(r, dtype=np.float64)
    v = np.maximum(np.asarray(v, dtype=np.float64), 5e-4)
    u1 = (np.log(r + 1e-3) - np.log(1e-3)) / (np.log(100.0 + 1e-3) - np.log(1e-3))
    u2 = (np.log(v) - np.log(5e-4)) / (np.log(50.0) - np.log(5e-4))
    return np.stack([u1, u2], 1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def relu(z):
    return np.maximum(z, 0.0)


def drelu(z):
    return (z > 0.0).astype(z.dtype)


ACT = {"silu": (silu, dsilu), "relu": (relu, drelu)}


def forward(params, B, U, keep=False, act="silu"):
    layers = unpack(params)
    p = 2.0 * np.pi * (U @ B.T)
    h = np.concatenate([np.sin(p), np.cos(p)], 1)
    hs, zs = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ W.T + b
        zs.append(z)
        h = ACT[act][0](z) if l < 4 else z
        hs.append(h)
    return (h, hs, zs) if keep else h


def loss_grad(params, B, U, Y, act="silu"):
    """MSE loss and its gradient wrt the flat parameters (reverse mode by hand)."""
    out, hs, zs = forward(params, B, U, keep=True, act=act)
    m = U.shape[0]
    d = out - Y
    loss = float(np.mean(d * d))
    dz = 2.0 * d / (m * 6)
    layers = unpack(params)
    grads = [None] * 5
    for l in range(4, -1, -1):
        W, _ = layers[l]
        grads[l] = (dz.T @ hs[l], dz.sum(0))
        if l > 0:
            dz = (dz @ W) * ACT[act][1](zs[l - 1])
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, g


def init_params(rng):
    """Uniform(+-1/sqrt(fan_in)) weights and biases, N(0, 4^2) Fourier matrix."""
    parts = []
    for i, n in zip(IN, OUT):
        bound = 1.0 / np.sqrt(i)
        parts.append(rng.uniform(-bound, bound, n * i + n))
    return np.concatenate(parts), 4.0 * rng.standard_normal((32, 2))
