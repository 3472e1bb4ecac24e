"""Mixed-precision dense solver (FP32 LU with trailing updates on FP32 MFMA or on split-bf16
MFMA, FP64 residuals and iterative refinement, FP64 fallback) against LAPACK and the oracle,
through the C ABI."""
import math

import numpy as np
import pytest

from test_gpu_parity import _coeff_parity, sphere64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    return M


@pytest.fixture(params=["fp32", "bf16"])
def mixed(mhx, request):
    """The mixed-precision solver with FP32 trailing updates (LU_MIXED) and with split-bf16 ones
    (LU_MIXED_BF16)."""
    from mhx import _lib
    ctx = mhx.context()
    ctx.set_lu_mode(_lib.LU_MIXED if request.param == "fp32" else _lib.LU_MIXED_BF16)
    ctx.reset_timing()
    yield ctx
    ctx.set_lu_mode(_lib.LU_AUTO)


@pytest.mark.parametrize("n", [5, 64, 130, 300, 1000, 1100])
@pytest.mark.parametrize("direct", [True, False])
def test_mixed_solve_matches_lapack(mhx, mixed, n, direct):
    """Diagonally dominated complex systems (BEM-like): refined solution within 1e-12 of LAPACK."""
    print(f"mixed solve n={n} direct={direct}", flush=True)
    rng = np.random.default_rng(n)
    D = 0.5 * np.eye(n) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / (2 * math.sqrt(n))
    S = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    bc = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    phi, src = mhx.solve_linear(D, S, bc, direct=direct)
    ref = np.linalg.solve(D, S @ bc) if direct else S @ np.linalg.solve(D, bc)
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-12
    if not direct:
        assert np.linalg.norm(src - np.linalg.solve(D, bc)) / np.linalg.norm(src) < 1e-12
    solved, steps, fallbacks, resid = mixed.solver_stats()
    assert solved == 1 and fallbacks == 0 and resid <= 1e-12 and steps >= 1


def test_ill_conditioned_falls_back_to_fp64(mhx, mixed):
    """cond(D) = 1e10: the FP32 factor cannot be refined to 1e-12, the item is redone in FP64."""
    print("mixed fallback", flush=True)
    rng = np.random.default_rng(1)
    n = 200
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    D = (Q1 * np.geomspace(1.0, 1e-10, n)) @ Q2.conj().T
    S = np.eye(n, dtype=np.complex128)
    bc = D @ (rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))
    phi, _ = mhx.solve_linear(D, S, bc, direct=True)
    ref = np.linalg.solve(D, bc)
    assert np.linalg.norm(D @ phi - bc) / np.linalg.norm(bc) < 1e-12
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-4   # cond 1e10: both are FP64 LU solutions
    solved, steps, fallbacks, resid = mixed.solver_stats()
    assert fallbacks == 1 and solved == 0


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
def test_mixed_coefficients_match_oracle(mhx, oracle, mixed, gf, direct):
    """The 64-panel sphere (radiation + diffraction, two headings) through the mixed solver in the
    batched schedule: coefficients within 1e-9 of the oracle's FP64 LAPACK path."""
    print(f"mixed coefficients {gf} direct={direct}", flush=True)
    _coeff_parity(mhx, oracle, sphere64(), [0.5, 1.03, 2.2], [0.0, math.pi / 3], gf, direct, mode="batched")
    solved, steps, fallbacks, resid = mixed.solver_stats()
    assert solved == 3 and fallbacks == 0 and resid <= 1e-12


def test_two_stream_batch_matches_fp64(mhx, mixed):
    """A batch of 9 frequencies on a 768-panel cylinder (order >= 512, 8 or more items): the FP32
    factorization runs as two halves (4 and 5 matrices) on two streams.  Every coefficient matches
    the FP64 LU path to 1e-9, every system refined to 1e-12, none fell back."""
    print("mixed two-stream batch", flush=True)
    from mhx import _lib, meshes
    body = mhx.FloatingBody(meshes.cylinder(1.0, 2.0, 32, 16, 8), ["Surge", "Heave", "Pitch"])
    par = dict(wave_frequencies=list(np.linspace(0.3, 2.7, 9)), radiating_dofs=["Surge", "Heave", "Pitch"],
               wave_directions=[0.0, 1.0])
    got = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
    solved, steps, fallbacks, resid = mixed.solver_stats()
    assert solved == 9 and fallbacks == 0 and resid <= 1e-12
    mixed.set_lu_mode(_lib.LU_FP64)
    ref = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        err = np.linalg.norm(got[key] - ref[key]) / np.linalg.norm(ref[key])
        assert err <= 1e-9, (key, err)


def test_tuned_factorization_matches_fp64(mhx, mixed):
    """The LU tuner: the first batched FP32 factorization of a signature (12 frequencies, 768
    panels) times 1-3 streams and 2-8 panels per split update on the batch itself and keeps the
    fastest; the coefficients match the FP64 LU path to 1e-9, and a second call of the same
    signature reuses the choice and reproduces the first call's results exactly."""
    print("mixed tuned batch", flush=True)
    from mhx import _lib, meshes
    body = mhx.FloatingBody(meshes.cylinder(1.0, 2.0, 32, 16, 8), ["Surge", "Heave", "Pitch"])
    par = dict(wave_frequencies=list(np.linspace(0.35, 2.6, 12)), radiating_dofs=["Surge", "Heave", "Pitch"],
               wave_directions=[0.0])
    mixed.set_autotune(True)
    got = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
    streams, panels = mixed.lu_config()
    assert 1 <= streams <= 3 and 1 <= panels <= 8, (streams, panels)
    again = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
    assert mixed.lu_config() == (streams, panels)
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        assert np.array_equal(got[key], again[key]), key
    solved, steps, fallbacks, resid = mixed.solver_stats()
    assert solved == 24 and fallbacks == 0 and resid <= 1e-12
    mixed.set_lu_mode(_lib.LU_FP64)
    ref = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        err = np.linalg.norm(got[key] - ref[key]) / np.linalg.norm(ref[key])
        assert err <= 1e-9, (key, err)


def test_phase_timing_switch(mhx):
    """mhx_set_timing: a call made with the phase events off records no phase, one made with them
    on records its LU and solve phases, and the coefficients are the same either way."""
    print("phase timing switch", flush=True)
    from mhx import meshes
    ctx = mhx.context()
    body = mhx.FloatingBody(meshes.sphere(1.0, 8, 16, immersed=True), ["Heave"])
    par = dict(wave_frequencies=[0.8, 1.4], radiating_dofs=["Heave"], wave_directions=[0.0])
    try:
        ctx.set_timing(False)
        ctx.reset_timing()
        off = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
        assert ctx.phase_time("lu")[1] == 0 and ctx.phase_time("solve")[1] == 0
        ctx.set_timing(True)
        ctx.reset_timing()
        on = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Wu"), schedule="batched")
        assert ctx.phase_time("lu")[1] > 0 and ctx.phase_time("solve")[1] > 0
        assert ctx.phase_time("lu")[0] > 0.0
        for key in ("added_mass", "radiation_damping", "excitation_force"):
            assert np.array_equal(off[key], on[key]), key
    finally:
        ctx.set_timing(False)
        ctx.reset_timing()
