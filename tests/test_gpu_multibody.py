"""Multi-body sweep (mhx_sweep_coeffs): several bodies of uneven panel counts, each with its own
frequencies, through one assembly queue launch per stage, batched LU per body, and the captured
graph replay.  Checked against the single-body FP64 path (itself within 1e-9 of the oracle) and
against the oracle directly on the small bodies."""
import math

import numpy as np
import pytest

from test_gpu_parity import close

pytestmark = pytest.mark.gpu

KEYS = ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force", "excitation_force")


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1025.0)
    M.set_g(9.81)
    yield M
    M.set_rho(1000.0)


def _small_jobs(mhx):
    from mhx import meshes
    bodies = [
        (meshes.sphere(1.0, 6, 6), ["Heave", "Pitch"], [0.6, 1.03, 1.9]),
        (meshes.sphere(1.0, 8, 16), ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"], [0.3, 0.8]),
        (meshes.sphere_with_count(1.5, 256), ["Heave", "Surge"], [0.2, 0.5, 1.1, 2.7]),
        (meshes.box(3.0, 2.0, 1.0, 512), ["Heave", "Roll", "Yaw"], [1.4]),
    ]
    return [(mhx.FloatingBody(m, dofs), {"wave_frequencies": w, "radiating_dofs": dofs,
                                         "wave_directions": [0.0, math.pi / 4]}) for m, dofs, w in bodies]


def _check(got, ref, tol):
    for key in KEYS:
        rel, elem, ok = close(got[key], ref[key], tol=tol)
        assert ok, (key, rel, elem)


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
def test_multibody_matches_single_body_path(mhx, gf, direct):
    """Four bodies (18, 64, 256, 512 panels), 1-4 frequencies each, two headings: every body's
    coefficients equal the single-body FP64 path's to 1e-11."""
    print(f"multibody small {gf} direct={direct}", flush=True)
    form = mhx.DirectBEM(gf) if direct else mhx.IndirectBEM(gf)
    jobs = _small_jobs(mhx)
    res = mhx.hydrodynamic_coefficients_many(jobs, form)
    for (body, par), got in zip(jobs, res):
        ref = mhx.hydrodynamic_coefficients(body, par, form)
        _check(got, ref, 1e-11)


def test_multibody_matches_oracle(mhx, oracle):
    """The two smallest bodies of the sweep against the oracle directly (1e-9)."""
    print("multibody oracle", flush=True)
    jobs = _small_jobs(mhx)
    res = mhx.hydrodynamic_coefficients_many(jobs, mhx.DirectBEM("ExactGuevelDelhommeau"))
    for (body, par), got in list(zip(jobs, res))[:2]:
        om = oracle.Mesh.from_mhx(body.mesh)
        names = par["radiating_dofs"]
        r = oracle.hydro_coeffs(om, [body.dofs[nm] for nm in names], par["wave_frequencies"], par["wave_directions"],
                                rho=1025.0, g=9.81, gf=oracle.GF_DELHOMMEAU, direct=True)
        for key in KEYS:
            rel, elem, ok = close(np.asarray(got[key]).reshape(np.asarray(r[key]).shape), r[key])
            assert ok, (key, rel, elem)


def test_graph_replays_match_eager(mhx):
    """The captured sweep replayed with new frequency values (same counts), with other device work
    in between, equals the eager sweep every time; a change of frequency count rebuilds it."""
    print("multibody graph replay", flush=True)
    jobs = _small_jobs(mhx)
    form = mhx.DirectBEM()
    rng = np.random.default_rng(3)
    for rep in range(4):
        jr = []
        for body, par in jobs:
            w = np.sort(rng.uniform(0.2, 2.5, size=len(par["wave_frequencies"])))
            jr.append((body, dict(par, wave_frequencies=w)))
        got = mhx.hydrodynamic_coefficients_many(jr, form, graph=True)
        # eager work on the same bodies between replays (single-body path, then the eager sweep)
        single = mhx.hydrodynamic_coefficients(*jr[1], form)
        ref = mhx.hydrodynamic_coefficients_many(jr, form, graph=False)
        _check(single, ref[1], 1e-11)
        for g, r in zip(got, ref):
            _check(g, r, 1e-12)
    # different frequency counts: a new graph
    jr = [(b, dict(p, wave_frequencies=[0.7, 1.3])) for b, p in jobs]
    got = mhx.hydrodynamic_coefficients_many(jr, form, graph=True)
    ref = mhx.hydrodynamic_coefficients_many(jr, form)
    for g, r in zip(got, ref):
        _check(g, r, 1e-12)


def test_config3_seeded_uneven_workload(mhx):
    """The seeded 8-body workload (spheres, cylinders, boxes of 256-4096 panels, 16-64
    log-uniform frequencies in 0.1-3.0 each, 6 dofs, headings 0, pi/4, pi/2, pi), FP64 Wu, direct:
    the multi-body sweep (eager and replayed graph) equals each body's batched single-body
    sweep to 1e-10."""
    print("multibody config 3", flush=True)
    from mhx import meshes
    dofs6 = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
    betas = [0.0, math.pi / 4, math.pi / 2, math.pi]
    jobs = [(mhx.FloatingBody(m, dofs6), {"wave_frequencies": w, "radiating_dofs": dofs6, "wave_directions": betas})
            for m, w, _ in meshes.random_bodies()]
    form = mhx.DirectBEM()
    res = mhx.hydrodynamic_coefficients_many(jobs, form)
    res_g = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
    res_g2 = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
    for q, (body, par) in enumerate(jobs):
        print(f"  body {q}: {body.mesh.nfaces} panels, {len(par['wave_frequencies'])} frequencies", flush=True)
        ref = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
        _check(res[q], ref, 1e-10)
        _check(res_g[q], res[q], 1e-12)
        _check(res_g2[q], res[q], 1e-12)


def test_graph_replay_after_constant_and_buffer_changes(mhx):
    """A captured sweep is not replayed with stale state: (i) rho and g changed between two graph
    calls (the force and boundary-condition kernels carry them by value) give the eager sweep's
    coefficients at the new constants; (ii) a single-body call with 10 headings between two graph
    calls reallocates that body's work buffers (more headings than the buffers hold), and the next
    graph call equals the eager sweep."""
    print("multibody graph: constants and buffer generations", flush=True)
    jobs = _small_jobs(mhx)
    form = mhx.DirectBEM()
    try:
        mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
        mhx.set_rho(1000.0)
        mhx.set_g(9.80665)
        got = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
        ref = mhx.hydrodynamic_coefficients_many(jobs, form)
        for g, r in zip(got, ref):
            _check(g, r, 1e-12)
        body, par = jobs[2]
        mhx.hydrodynamic_coefficients(body, dict(par, wave_directions=list(np.linspace(0.0, 3.0, 10))), form)
        got = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
        ref = mhx.hydrodynamic_coefficients_many(jobs, form)
        for g, r in zip(got, ref):
            _check(g, r, 1e-12)
    finally:
        mhx.set_rho(1025.0)
        mhx.set_g(9.81)
