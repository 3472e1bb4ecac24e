"""The learned surrogate inside the multi-body sweep (mhx_sweep_coeffs with MHX_GF_SURROGATE): the
fused pair kernel over the (body, frequency, tile) queue of every body writes S, D and the FP32
copy of D in one launch per chunk, the LU is the FP32 split-bf16 one with a refinement step on
the device, and the whole sweep is captured into a HIP graph and replayed.

- Four small bodies: the sweep equals each body's single-body surrogate path (same pair values,
  same mixed-precision solver) to 1e-10, eager and replayed.
- The seeded 8-body workload (256-4,096 panels, 16-64 frequencies each, 6 dofs, 4 headings): the
  captured surrogate sweep, replayed twice, is within 1e-3 of the exact FP64 results on every
  added-mass, damping and excitation block of every body and frequency.
"""
import math

import numpy as np
import pytest

from test_gpu_parity import close

pytestmark = pytest.mark.gpu

DOFS6 = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1025.0)
    M.set_g(9.81)
    M.load_surrogate()
    yield M
    M.set_rho(1000.0)


def _small_jobs(mhx):
    from mhx import meshes
    bodies = [
        (meshes.sphere(1.0, 8, 16), ["Surge", "Heave", "Pitch"], [0.3, 0.8, 1.7]),
        (meshes.sphere_with_count(1.5, 256), ["Heave", "Surge"], [0.2, 0.5, 1.1, 2.7]),
        (meshes.cylinder(2.0, 3.0, 24, 8, 6), ["Heave", "Roll"], [0.45, 1.2]),
        (meshes.box(3.0, 2.0, 1.0, 512), ["Heave", "Roll", "Yaw"], [1.4]),
    ]
    return [(mhx.FloatingBody(m, dofs), {"wave_frequencies": w, "radiating_dofs": dofs,
                                         "wave_directions": [0.0, math.pi / 4]}) for m, dofs, w in bodies]


def test_small_sweep_matches_single_body_surrogate(mhx):
    form = mhx.DirectBEM("Surrogate")
    jobs = _small_jobs(mhx)
    keys = ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force")
    results = []
    for graph in (False, True, True):
        print(f"surrogate sweep, graph={graph}", flush=True)
        results.append(mhx.hydrodynamic_coefficients_many(jobs, form, graph=graph))
    for (body, par), got in zip(jobs, results[0]):
        ref = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
        for key in keys:
            rel, elem, ok = close(got[key], ref[key], tol=1e-10)
            assert ok, (body.mesh.nfaces, key, rel, elem)
    for res in results[1:]:
        for got, ref in zip(res, results[0]):
            for key in keys:
                rel, elem, ok = close(got[key], ref[key], tol=1e-12)
                assert ok, ("graph", key, rel, elem)


def _blocks(res, q):
    A = res["added_mass"][:, :, q, 0]
    B = res["radiation_damping"][:, :, q, 0]
    F = res["excitation_force"][:, q, :, 0]
    return A, B, F


def test_config3_surrogate_sweep_graph_against_exact(mhx):
    from mhx import meshes
    betas = [0.0, math.pi / 4, math.pi / 2, math.pi]
    data = meshes.random_bodies()
    jobs = [(mhx.FloatingBody(m, DOFS6), {"wave_frequencies": w, "radiating_dofs": DOFS6, "wave_directions": betas})
            for m, w, _ in data]
    ctx = mhx.context()
    print("config3: exact FP64 sweep", flush=True)
    ex = mhx.hydrodynamic_coefficients_many(jobs, mhx.DirectBEM("ExactGuevelDelhommeau"))
    form = mhx.DirectBEM("Surrogate")
    ctx.reset_timing()
    print("config3: surrogate sweep, captured", flush=True)
    first = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
    for rep in range(2):
        print(f"config3: surrogate sweep, replay {rep}", flush=True)
        res = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
        for got, ref in zip(res, first):
            for key in ("added_mass", "radiation_damping", "excitation_force"):
                rel, elem, ok = close(got[key], ref[key], tol=1e-12)
                assert ok, ("replay", key, rel, elem)
    solved, steps, fallbacks, resid = ctx.solver_stats()
    print(f"mixed-precision items {solved}, FP64 redos {fallbacks}, largest residual {resid:.2e}", flush=True)
    worst = 0.0
    for q, ((body, par), got, ref) in enumerate(zip(jobs, res, ex)):
        for w in range(len(par["wave_frequencies"])):
            for name, a, b in zip(("A", "B", "F_ex"), _blocks(got, w), _blocks(ref, w)):
                rel = np.linalg.norm(a - b) / np.linalg.norm(b)
                worst = max(worst, rel)
                assert rel <= 1e-3, (q, body.mesh.nfaces, par["wave_frequencies"][w], name, rel)
    print(f"largest normwise block error against the exact path {worst:.3e}", flush=True)


@pytest.mark.parametrize("gf", ["Surrogate", "Wu"])
def test_graph_schedule_equals_batched(mhx, gf):
    """schedule="graph" (MHX_MODE_GRAPH: the batched schedule captured on the first call, replayed
    on the next with new frequencies) equals schedule="batched" to 1e-10."""
    from mhx import meshes
    body = mhx.FloatingBody(meshes.capped_cylinder(2.0, 3.0, 32, 12, 8), ["Surge", "Heave", "Pitch"])
    form = mhx.DirectBEM(gf)
    for ws in ([0.4, 0.9, 1.3, 2.1, 2.6], [0.5, 1.0, 1.4, 2.2, 2.7]):
        par = dict(wave_frequencies=ws, radiating_dofs=body.dof_names, wave_directions=[0.0, math.pi / 2])
        print(f"graph schedule {gf} omegas {ws[0]}..", flush=True)
        g = mhx.hydrodynamic_coefficients(body, par, form, schedule="graph")
        b = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
        for key in ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force"):
            rel, elem, ok = close(g[key], b[key], tol=1e-10)
            assert ok, (gf, key, rel, elem)


def test_fp64_path_of_the_surrogate_storage(mhx):
    """Refinement limit 0 sends every system down the FP64 path.  The batched schedule then widens
    the FP32 S and D the pair kernel stored and factors them in FP64: the same matrices the mixed
    solver refines to 1e-12, so the coefficients agree to 1e-9.  The sweep redoes every item
    through the per-frequency FP64 path, which assembles S and D again in FP64: within FP32
    rounding (1e-5) of the sweep's own results."""
    form = mhx.DirectBEM("Surrogate")
    body, par = _small_jobs(mhx)[1]
    nw = len(par["wave_frequencies"])
    ctx = mhx.context()
    keys = ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force")
    print("refine limit: reference runs", flush=True)
    ref = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
    ref_sweep = mhx.hydrodynamic_coefficients_many([(body, par)], form)[0]
    try:
        ctx.set_refine_limit(0)
        ctx.reset_timing()
        print("refine limit 0: batched", flush=True)
        got = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
        solved, _, fallbacks, _ = ctx.solver_stats()
        assert solved == 0 and fallbacks == nw, (solved, fallbacks)
        ctx.reset_timing()
        print("refine limit 0: sweep", flush=True)
        sweep = mhx.hydrodynamic_coefficients_many([(body, par)], form)[0]
        solved, _, fallbacks, _ = ctx.solver_stats()
        assert solved == 0 and fallbacks == nw, (solved, fallbacks)
    finally:
        ctx.set_refine_limit(-1)
    for key in keys:
        rel, elem, ok = close(got[key], ref[key], tol=1e-9, abs_floor=1e-10)
        assert ok, ("batched", key, rel, elem)
        rel, elem, ok = close(sweep[key], ref_sweep[key], tol=1e-5, abs_floor=1e-6)
        assert ok, ("sweep", key, rel, elem)
    with pytest.raises(ValueError):
        ctx.set_refine_limit(9)
