"""The learned wave-term surrogate on the device, through the C ABI.

  * training kernels: loss and gradient of one FP32 forward/backward against the NumPy
    checker (itself pinned by central differences on the host);
  * training: a short device run drives the held-out loss down; save/load round trip;
  * inference: the BF16-MFMA network against the FP32 NumPy network on the same weights,
    and (Wu + network) against the exact term of the oracle;
  * assembly: pairs outside the trained box take the exact FP64 path (bit-level parity with
    the exact assembly when every pair falls back); coefficients within 1e-3 of the exact
    FP64 path, and closer than the method without the network (plain GFWu).
"""
import os

import numpy as np
import pytest

import surrogate_np as snp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    return M


@pytest.fixture(scope="module")
def trained(mhx):
    """The network the library ships (weights file), else one trained here with the defaults."""
    from mhx import api
    if os.path.exists(api.SURROGATE_WEIGHTS):
        print("loading shipped weights", flush=True)
        mhx.load_surrogate()
    else:
        print("training the surrogate (defaults)", flush=True)
        mhx.train_surrogate()
    return mhx.context().surrogate_params()


def _away_from_kinks(params, B, U, act, margin=1e-4):
    """Rows whose hidden pre-activations all stay |z| > margin (for ReLU, whose derivative jumps at
    0: FP32 and FP64 may put a pre-activation this close to 0 on different sides)."""
    if act != "relu":
        return np.ones(U.shape[0], dtype=bool)
    _, _, zs = snp.forward(params, B, U.astype(np.float32).astype(np.float64), keep=True, act=act)
    return np.all([np.min(np.abs(z), axis=1) > margin for z in zs[:4]], axis=0)


@pytest.mark.parametrize("act", ["silu", "relu"])
def test_loss_and_gradient_match_numpy(mhx, act):
    print(f"surrogate loss/grad ({act})", flush=True)
    rng = np.random.default_rng(3)
    params, B = snp.init_params(rng)
    ctx = mhx.context()
    assert ctx.surrogate_param_count() == snp.NPARAMS
    ctx.set_surrogate_activation(act)
    try:
        for m in (64, 333, 4096):
            U = rng.uniform(0, 1, (m, 2))
            Y = rng.standard_normal((m, 6))
            p32, B32 = params.astype(np.float32), B.astype(np.float32)
            keep = _away_from_kinks(p32.astype(np.float64), B32.astype(np.float64), U, act)
            U, Y = U[keep], Y[keep]
            loss, g = ctx.surrogate_loss_grad(U, Y, p32, B32)
            lref, gref = snp.loss_grad(p32.astype(np.float64), B32.astype(np.float64), U.astype(np.float32).astype(
                np.float64), Y.astype(np.float32).astype(np.float64), act=act)
            assert loss == pytest.approx(lref, rel=1e-5)
            offs = np.cumsum([0] + [o * i + o for i, o in zip(snp.IN, snp.OUT)])
            for l in range(5):
                a, b = offs[l], offs[l + 1]
                rel = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
                assert rel < 2e-4, (m, l, rel)
    finally:
        ctx.set_surrogate_activation("silu")


def test_short_training_and_round_trip(mhx, tmp_path):
    print("surrogate short training", flush=True)
    ctx = mhx.context()
    val = ctx.surrogate_train(ngrid=256, steps=600, batch=2048, lr=2e-3, seed=7)
    assert np.isfinite(val) and val < 2e-2, val
    p, B = ctx.surrogate_params()
    r = np.geomspace(1e-3, 90, 50)
    v = np.geomspace(2e-3, 40, 50)
    out1 = ctx.surrogate_eval(r, v)
    path = str(tmp_path / "s.mhxs")
    ctx.surrogate_save(path)
    ctx.surrogate_load(path)
    p2, B2 = ctx.surrogate_params()
    assert np.array_equal(p, p2) and np.array_equal(B, B2)
    assert np.array_equal(out1, ctx.surrogate_eval(r, v))
    with pytest.raises(ValueError):
        ctx.surrogate_load(str(tmp_path / "missing.mhxs"))


def _points(rng, m):
    r = np.exp(rng.uniform(np.log(1e-3), np.log(100.0), m))
    v = np.exp(rng.uniform(np.log(1e-3), np.log(50.0), m))
    return r, v


def test_bf16_inference_matches_fp32_network(mhx, trained):
    """The MFMA path (BF16 weights/activations, FP32 accumulation) against the FP32 network."""
    print("surrogate bf16 vs fp32", flush=True)
    params, B = trained
    rng = np.random.default_rng(9)
    r, v = _points(rng, 3000)
    got = mhx.context().surrogate_eval(r, v)
    ref = snp.forward(params.astype(np.float64), B.astype(np.float64), snp.inputs(r, v).astype(np.float32))
    err = np.abs(got - ref)
    # BF16 rounding of weights and activations (8-bit significand) acts at the scale of the
    # largest outputs; the imaginary residuals are three orders smaller than the real ones, so
    # their bound is set by the overall scale.  A wrong operand or accumulator layout gives
    # errors of the order of the outputs themselves.
    scale = np.max(np.abs(ref), axis=0)
    top = scale.max()
    assert np.all(err.max(0) <= 0.03 * scale + 0.005 * top), (err.max(0), scale)
    assert np.sqrt(np.mean(err ** 2)) < 2e-3 * top, np.sqrt(np.mean(err ** 2))


def test_surrogate_against_exact_term(mhx, oracle, trained):
    """Wu + network / 100 against the oracle's exact term, and the gain over Wu alone."""
    print("surrogate vs exact", flush=True)
    rng = np.random.default_rng(13)
    r, v = _points(rng, 1500)
    net = mhx.context().surrogate_residual(r, v)
    e_net, e_wu, mag = [], [], []
    for q in range(len(r)):
        G, Gr, Gz = oracle.delhommeau_term(r[q], -v[q])
        W, Wr, Wz = oracle.wu_term(r[q], -v[q])
        ex = np.array([G.real, G.imag, Gr.real, Gr.imag, Gz.real, Gz.imag])
        wu = np.array([W.real, W.imag, Wr.real, Wr.imag, Wz.real, Wz.imag])
        e_net.append(wu + net[q] - ex)
        e_wu.append(wu - ex)
        mag.append(np.abs(ex))
    e_net, e_wu, mag = map(np.asarray, (e_net, e_wu, mag))
    # G and dG/dr (the assembly's inputs): absolute error against the term's magnitude
    for c in (0, 2):
        rms_net = np.sqrt(np.mean(e_net[:, c] ** 2))
        rms_wu = np.sqrt(np.mean(e_wu[:, c] ** 2))
        print(f"component {c}: rms err net {rms_net:.2e}  wu {rms_wu:.2e}  max net {np.abs(e_net[:, c]).max():.2e}",
              flush=True)
        assert rms_net < 0.1 * rms_wu, (c, rms_net, rms_wu)
        assert np.abs(e_net[:, c]).max() < 2e-3 * max(1.0, np.max(mag[:, c]))
    # imaginary parts: GFWu's are the closed forms 2 pi e^z J0(r), 2 pi e^z J1(r); the correction
    # must stay below 1e-5 of the envelope exp(-v)
    for c in (1, 3):
        rel = np.abs(e_net[:, c]) * np.exp(np.minimum(v, 5.0))
        print(f"component {c}: max err / exp(-min(v,5)) {rel.max():.2e}", flush=True)
        assert rel.max() < 1e-4


def test_all_pairs_fall_back_to_exact(mhx, trained):
    """At k = 1e-4 every pair of a small body near the surface has k|z + zeta| < 1e-3: the surrogate
    assembly is then the exact one (same FP64 kernel), to rounding."""
    print("surrogate fallback", flush=True)
    from mhx import meshes
    m = meshes.sphere(1.0, 8, 16, immersed=True)
    for direct in (True, False):
        S1, D1 = mhx.assemble_matrices(mhx.default_greens_functions("Surrogate"), m, 1e-4, direct=direct)
        S0, D0 = mhx.assemble_matrices(mhx.default_greens_functions("ExactGuevelDelhommeau"), m, 1e-4,
                                       direct=direct)
        assert np.linalg.norm(S1 - S0) <= 1e-13 * np.linalg.norm(S0)
        assert np.linalg.norm(D1 - D0) <= 1e-13 * np.linalg.norm(D0)


def _coeffs(mhx, body, names, ws, gf, schedule):
    return mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=ws, radiating_dofs=names, wave_directions=[0.0]),
                                         mhx.DirectBEM(gf), schedule=schedule)


def test_coefficients_within_1e3_of_exact(mhx, trained):
    """Capped cylinder (R=4, draft 12, 1,024 panels), 6 dofs, 8 frequencies in 0.1-3.0 rad/s: surrogate
    + batched schedule against the exact FP64 per-frequency path; and the gain over plain GFWu."""
    print("surrogate coefficients", flush=True)
    from mhx import meshes
    mesh = meshes.capped_cylinder(4.0, 8.0, 32, 20, 12)
    mhx.set_rho(1025.0); mhx.set_g(9.81)
    names = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
    body = mhx.FloatingBody(mesh, names)
    ws = np.linspace(0.1, 3.0, 8)
    ex = _coeffs(mhx, body, names, ws, "ExactGuevelDelhommeau", "per_omega")
    su = _coeffs(mhx, body, names, ws, "Surrogate", "batched")
    wu = _coeffs(mhx, body, names, ws, "Wu", "per_omega")
    gains = {}
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        for q in range(len(ws)):
            ref = ex[key][..., q, :] if key != "excitation_force" else ex[key][:, q]
            s = su[key][..., q, :] if key != "excitation_force" else su[key][:, q]
            w = wu[key][..., q, :] if key != "excitation_force" else wu[key][:, q]
            es = np.linalg.norm(s - ref) / np.linalg.norm(ref)
            ew = np.linalg.norm(w - ref) / np.linalg.norm(ref)
            print(f"{key} omega={ws[q]:.3f}: surrogate {es:.2e}  Wu {ew:.2e}", flush=True)
            assert es <= 1e-3, (key, ws[q], es)
        # over the sweep (normwise): the network must beat the fit it corrects
        es_all = np.linalg.norm(su[key] - ex[key]) / np.linalg.norm(ex[key])
        ew_all = np.linalg.norm(wu[key] - ex[key]) / np.linalg.norm(ex[key])
        gains[key] = ew_all / es_all
        print(f"{key} sweep: surrogate {es_all:.2e}  Wu {ew_all:.2e}  gain {gains[key]:.1f}x", flush=True)
    assert gains["added_mass"] > 10 and gains["excitation_force"] > 10, gains
    assert gains["radiation_damping"] > 1, gains


RELU_WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "marinehydro.jl_amd",
                            "weights", "delhommeau_surrogate_relu.mhxs")


@pytest.mark.skipif(not os.path.exists(RELU_WEIGHTS), reason="no ReLU weights file")
def test_relu_network_inference_and_accuracy(mhx, oracle):
    """The shipped ReLU variant of the network: the BF16-MFMA inference against the FP32 NumPy
    network (same bounds as the SiLU network), and (Wu + network) against the oracle's exact term
    at least ten times closer than Wu alone on G and dG/dr."""
    print("surrogate relu variant", flush=True)
    ctx = mhx.context()
    mhx.load_surrogate(RELU_WEIGHTS)
    try:
        assert ctx.surrogate_activation() == "relu"
        params, B = ctx.surrogate_params()
        rng = np.random.default_rng(21)
        r, v = _points(rng, 3000)
        got = ctx.surrogate_eval(r, v)
        ref = snp.forward(params.astype(np.float64), B.astype(np.float64), snp.inputs(r, v).astype(np.float32),
                          act="relu")
        err = np.abs(got - ref)
        scale = np.max(np.abs(ref), axis=0)
        top = scale.max()
        assert np.all(err.max(0) <= 0.03 * scale + 0.005 * top), (err.max(0), scale)
        r, v = r[:800], v[:800]
        net = ctx.surrogate_residual(r, v)
        e_net, e_wu = [], []
        for q in range(len(r)):
            G, Gr, _ = oracle.delhommeau_term(r[q], -v[q])
            W, Wr, _ = oracle.wu_term(r[q], -v[q])
            e_net.append([W.real + net[q, 0] - G.real, Wr.real + net[q, 2] - Gr.real])
            e_wu.append([W.real - G.real, Wr.real - Gr.real])
        e_net, e_wu = np.asarray(e_net), np.asarray(e_wu)
        for c in range(2):
            assert np.sqrt(np.mean(e_net[:, c] ** 2)) < 0.1 * np.sqrt(np.mean(e_wu[:, c] ** 2)), c
    finally:
        mhx.load_surrogate()
