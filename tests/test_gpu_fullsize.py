"""Parity at the workloads' own sizes, through the C ABI.

- The dense LU (FP64, FP32 and split-bf16 mixed precision) against numpy LAPACK at orders
  2049, 3073 and 4096 (the leaf tile of more than 2048 rows), eight systems per batch (the
  batch is factored as two halves on two streams).
- The 4,096-panel capped cylinder (R = 4 m, draft 12 m) of the headline workload on the plain
  FP64 exact path: 64 seeded rows of S and D against the oracle, and the full coefficients at
  one frequency against the oracle (its assembly, boundary conditions and pressure
  integration, with LAPACK for the dense solve).
- The same body on the surrogate + batched path over all 64 frequencies against the exact
  FP64 path: every added-mass, damping and excitation block within 1e-3 (normwise).
- The 1,024-panel cylinder over its full 32-frequency sweep, direct and indirect, and the
  seeded 8-body workload with the exact term, against the oracle.
"""
import math
import os

import numpy as np
import pytest

from test_gpu_parity import close

pytestmark = pytest.mark.gpu

DOFS6 = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
KEYS = ("added_mass", "radiation_damping", "excitation_force", "Froude_Krylov_force")
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1025.0)
    M.set_g(9.81)
    yield M
    M.context().set_lu_mode(0)
    M.set_rho(1000.0)


@pytest.fixture(scope="module")
def capcyl():
    from mhx import meshes
    m = meshes.capped_cylinder(4.0, 8.0, 64, 40, 24)
    assert m.nfaces == 4096
    return m


def _systems(n, batch, seed):
    """Half BEM-like (1/2 I + a small dense part, no pivoting needed), half complex Gaussian
    (pivoting needed): (batch, n, n) and right-hand sides (batch, n, 7)."""
    rng = np.random.default_rng(seed)
    A = np.empty((batch, n, n), dtype=np.complex128)
    for q in range(batch):
        G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        A[q] = 0.5 * np.eye(n) + G / n if q % 2 == 0 else G
    B = rng.standard_normal((batch, n, 7)) + 1j * rng.standard_normal((batch, n, 7))
    return A, B


@pytest.mark.parametrize("n", [2049, 3073, 4096])
def test_lu_modes_against_lapack(mhx, n):
    from mhx import _lib
    A, B = _systems(n, 8, n)
    print(f"lu batch of 8, n={n}: numpy LAPACK reference", flush=True)
    ref = np.linalg.solve(A, B)
    refH = np.linalg.solve(np.conj(np.transpose(A[:2], (0, 2, 1))), B[:2])
    ctx = mhx.context()
    try:
        for name, mode in (("fp64", _lib.LU_FP64), ("mixed", _lib.LU_MIXED), ("mixed_bf16", _lib.LU_MIXED_BF16)):
            print(f"lu batch of 8, n={n}, {name}", flush=True)
            ctx.set_lu_mode(mode)
            X = mhx.lu_solve_batched(A, B)
            for q in range(8):
                err = np.linalg.norm(X[q] - ref[q]) / np.linalg.norm(ref[q])
                assert err < 1e-10, (name, q, err)
        ctx.set_lu_mode(_lib.LU_FP64)
        XH = mhx.lu_solve_batched(A[:2], B[:2], trans=2)
        for q in range(2):
            assert np.linalg.norm(XH[q] - refH[q]) / np.linalg.norm(refH[q]) < 1e-10
    finally:
        ctx.set_lu_mode(_lib.LU_AUTO)


def test_config4_plain_exact_sampled_rows(mhx, oracle, capcyl):
    """64 seeded rows (8 blocks of 8) of S and D (direct) and of K (indirect) at two frequencies
    of the sweep, plain FP64 exact path against the oracle, 1e-9."""
    oracle.set_threads(THREADS)
    om = oracle.Mesh.from_mhx(capcyl)
    rng = np.random.default_rng(4096)
    starts = np.sort(rng.choice(np.arange(0, 4096 - 8, 8), size=8, replace=False))
    ws = np.linspace(0.1, 3.0, 64)
    for w, direct in ((ws[22], True), (ws[63], True), (ws[22], False)):
        k = w * w / 9.81
        print(f"config4 rows: omega={w:.4f} direct={direct}", flush=True)
        S, D = mhx.assemble_matrices(mhx.default_greens_functions("ExactGuevelDelhommeau"), capcyl, k, direct=direct)
        for i0 in starts:
            So, Do = oracle.assemble_rows(om, k, (int(i0), int(i0) + 8), gf=oracle.GF_DELHOMMEAU, direct=direct)
            for got, ref in ((S[i0:i0 + 8], So), (D[i0:i0 + 8], Do)):
                rel, elem, ok = close(got, ref)
                assert ok, (w, direct, int(i0), rel, elem)


def test_config4_plain_exact_coefficients(mhx, oracle, capcyl):
    """Full coefficients of the 4,096-panel body at one frequency (6 dofs, beta = 0, rho = 1025),
    plain FP64 exact path (per frequency, FP64 LU) against the oracle, 1e-9."""
    oracle.set_threads(THREADS)
    w = float(np.linspace(0.1, 3.0, 64)[22])
    body = mhx.FloatingBody(capcyl, DOFS6)
    print(f"config4 coefficients: device, omega={w:.4f}", flush=True)
    out = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=[w], radiating_dofs=DOFS6, wave_directions=[0.0]),
                                        mhx.DirectBEM("ExactGuevelDelhommeau"))
    print("config4 coefficients: oracle", flush=True)
    om = oracle.Mesh.from_mhx(capcyl)
    ref = oracle.hydro_coeffs_lapack(om, [body.dofs[nm] for nm in DOFS6], [w], [0.0], rho=1025.0, g=9.81,
                                     gf=oracle.GF_DELHOMMEAU, direct=True)
    for key in KEYS:
        rel, elem, ok = close(out[key][..., 0], ref[key])
        assert ok, (key, rel, elem)


def _blocks(res, q):
    """Coefficient blocks of frequency q: A, B (6 x 6) and F_ex (6,)."""
    return (res["added_mass"][:, :, q, 0], res["radiation_damping"][:, :, q, 0], res["excitation_force"][:, q, 0, 0])


def test_config4_surrogate_batched_all_frequencies(mhx, capcyl):
    """The headline path (learned wave term, all 64 frequencies as one batch, mixed-precision
    LU) against the plain exact FP64 path on every frequency: each A, B and F_ex block within
    1e-3 normwise; the elementwise error floored at 1e-3 max|X| is reported."""
    mhx.load_surrogate()
    ws = np.linspace(0.1, 3.0, 64)
    body = mhx.FloatingBody(capcyl, DOFS6)
    par = dict(wave_frequencies=ws, radiating_dofs=DOFS6, wave_directions=[0.0])
    print("config4 surrogate batched: 64 frequencies", flush=True)
    su = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("Surrogate"), schedule="batched")
    print("config4 exact per frequency: 64 frequencies", flush=True)
    ex = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM("ExactGuevelDelhommeau"))
    worst_norm, worst_elem = 0.0, 0.0
    for q in range(64):
        for name, a, b in zip(("A", "B", "F_ex"), _blocks(su, q), _blocks(ex, q)):
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            elem = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * np.max(np.abs(b))))
            worst_norm = max(worst_norm, rel)
            worst_elem = max(worst_elem, elem)
            assert rel <= 1e-3, (q, ws[q], name, rel)
    print(f"largest normwise block error {worst_norm:.3e}, floored elementwise {worst_elem:.3e}", flush=True)


@pytest.mark.parametrize("direct", [True, False])
def test_config2_full_sweep(mhx, oracle, direct):
    """The 1,024-panel truncated cylinder (R = 5, draft 10) over all 32 frequencies in 0.2-2.0,
    6 dofs, beta = 0, exact term, batched schedule, against the oracle (1e-9)."""
    from mhx import meshes
    oracle.set_threads(THREADS)
    m = meshes.cylinder(5.0, 10.0, 32, 24, 8)
    assert m.nfaces == 1024
    ws = np.linspace(0.2, 2.0, 32)
    form = mhx.DirectBEM("ExactGuevelDelhommeau") if direct else mhx.IndirectBEM("ExactGuevelDelhommeau")
    body = mhx.FloatingBody(m, DOFS6)
    print(f"config2 sweep direct={direct}: device", flush=True)
    out = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=ws, radiating_dofs=DOFS6, wave_directions=[0.0]),
                                        form, schedule="batched")
    print(f"config2 sweep direct={direct}: oracle", flush=True)
    om = oracle.Mesh.from_mhx(m)
    ref = oracle.hydro_coeffs_lapack(om, [body.dofs[nm] for nm in DOFS6], ws, [0.0], rho=1025.0, g=9.81,
                                     gf=oracle.GF_DELHOMMEAU, direct=direct)
    for q in range(32):
        for key in KEYS:
            got = out[key][..., q, 0] if key in ("added_mass", "radiation_damping") else out[key][:, q, :, 0]
            exp = ref[key][..., q] if key in ("added_mass", "radiation_damping") else ref[key][:, q, :]
            rel, elem, ok = close(got, exp)
            assert ok, (q, key, rel, elem)


def test_config3_exact_full_workload(mhx, oracle):
    """The seeded 8-body workload (256-4,096 panels, 16-64 frequencies each, 6 dofs, 4 headings)
    with the exact term: the multi-body sweep (eager and replayed graph) equals each body's
    batched single-body path to 1e-10, and one frequency of every body equals the oracle (1e-9)."""
    from mhx import meshes
    oracle.set_threads(THREADS)
    betas = [0.0, math.pi / 4, math.pi / 2, math.pi]
    data = meshes.random_bodies()
    jobs = [(mhx.FloatingBody(m, DOFS6), {"wave_frequencies": w, "radiating_dofs": DOFS6, "wave_directions": betas})
            for m, w, _ in data]
    form = mhx.DirectBEM("ExactGuevelDelhommeau")
    print("config3 exact: eager sweep", flush=True)
    res = mhx.hydrodynamic_coefficients_many(jobs, form)
    print("config3 exact: graph sweep", flush=True)
    res_g = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
    for q, (body, par) in enumerate(jobs):
        print(f"  body {q}: {body.mesh.nfaces} panels, {len(par['wave_frequencies'])} frequencies", flush=True)
        ref = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
        for key in KEYS:
            rel, elem, ok = close(res[q][key], ref[key], tol=1e-10)
            assert ok, (q, key, rel, elem)
            rel, elem, ok = close(res_g[q][key], res[q][key], tol=1e-12)
            assert ok, (q, "graph", key, rel, elem)
        iw = len(par["wave_frequencies"]) // 2
        w = float(par["wave_frequencies"][iw])
        om = oracle.Mesh.from_mhx(body.mesh)
        ora = oracle.hydro_coeffs_lapack(om, [body.dofs[nm] for nm in DOFS6], [w], betas, rho=1025.0, g=9.81,
                                         gf=oracle.GF_DELHOMMEAU, direct=True)
        for key in KEYS:
            got = res[q][key][..., iw, 0] if key in ("added_mass", "radiation_damping") else res[q][key][:, iw, :, 0]
            exp = ora[key][..., 0] if key in ("added_mass", "radiation_damping") else ora[key][:, 0, :]
            rel, elem, ok = close(got, exp)
            assert ok, (q, "oracle", key, rel, elem)
