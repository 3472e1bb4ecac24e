"""Host-side checks of the wave-term surrogate: the NumPy checker's reverse mode against
central differences, the input map's end points, and the shipped weights file's format."""
import json
import os

import numpy as np
import pytest

import surrogate_np as snp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "marinehydro.jl_amd", "weights", "delhommeau_surrogate.mhxs")


def test_param_count():
    assert snp.NPARAMS == 128 * 64 + 128 + 3 * (128 * 128 + 128) + 6 * 128 + 6 == 58630


@pytest.mark.parametrize("act", ["silu", "relu"])
def test_reverse_mode_matches_central_differences(act):
    """A wrong transpose, a dropped activation derivative or a bias gradient summed over the wrong
    axis all fail this.  (ReLU: rows with a hidden pre-activation within 1e-5 of its kink are left
    out and the step is 1e-7, so that no difference step crosses it.)"""
    rng = np.random.default_rng(5)
    params, B = snp.init_params(rng)
    U = rng.uniform(0, 1, (40, 2))
    Y = rng.standard_normal((40, 6))
    if act == "relu":
        _, _, zs = snp.forward(params, B, U, keep=True, act=act)
        keep = np.all([np.min(np.abs(z), axis=1) > 1e-5 for z in zs[:4]], axis=0)
        U, Y = U[keep][:12], Y[keep][:12]
        assert U.shape[0] >= 6
    else:
        U, Y = U[:12], Y[:12]
    _, g = snp.loss_grad(params, B, U, Y, act=act)
    offs = np.cumsum([0] + [o * i + o for i, o in zip(snp.IN, snp.OUT)])
    picks = []
    for l in range(5):   # weights and biases of every layer
        nW = snp.OUT[l] * snp.IN[l]
        picks += list(offs[l] + rng.integers(0, nW, 4)) + list(offs[l] + nW + rng.integers(0, snp.OUT[l], 2))
    h = 1e-6 if act == "silu" else 1e-7
    for q in picks:
        p1 = params.copy(); p1[q] += h
        p2 = params.copy(); p2[q] -= h
        fd = (snp.loss_grad(p1, B, U, Y, act=act)[0] - snp.loss_grad(p2, B, U, Y, act=act)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])) + 1e-9, (q, fd, g[q])


def test_input_map_end_points():
    u = snp.inputs([0.0, 100.0], [5e-4, 50.0])
    np.testing.assert_allclose(u, [[0.0, 0.0], [1.0, 1.0]], atol=1e-12)
    assert snp.inputs([1.0], [1e-6])[0, 1] == 0.0   # v clamped at 5e-4


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no shipped weights file")
def test_shipped_weights_format():
    with open(WEIGHTS, "rb") as f:
        assert f.readline() == b"MHXSURR1\n"
        hdr = json.loads(f.readline().decode())
        body = f.read()
    assert hdr["format"] == "mhx-surrogate" and hdr["params"] == snp.NPARAMS and hdr["outputs"] == 6
    assert hdr["version"] == 2 and hdr["activation"] == "silu"   # the 4 x 128 SiLU network
    assert len(body) == 4 * (64 + snp.NPARAMS)
    vals = np.frombuffer(body, dtype=np.float32)
    assert np.all(np.isfinite(vals))
    assert hdr["val_mse"] < 1e-3
