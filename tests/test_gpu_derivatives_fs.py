"""Frequency derivatives of a body with a lid and of a body at forward speed (mhx_hydro_coeffs_fs_jvp).

The reference differentiates both in omega: the Heave added mass of a lid body with ForwardDiff
against central differences (test/solve_problem_differentiation.jl:94-98) and the Surge added
mass of a moving sphere (U = 0.2, beta = 2.0, IndirectBEM) with ForwardDiff and Enzyme
(test/forward_speed_tests.jl:280-303).  Here every coefficient's tangent along omega, from the
device's tangent kernels, is checked against fourth-order central differences of the oracle's
coefficients (GFWu: smooth in omega, so the differences are accurate to ~1e-10), and with the
exact Delhommeau term against differences of the device's own coefficients.  At U = 0 without a
lid the new entry point equals the plain tangent (mhx_hydro_coeffs_jvp)."""
import math

import numpy as np
import pytest

from test_gpu_parity import close

pytestmark = pytest.mark.gpu

KEYS = ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force")


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1025.0)
    M.set_g(9.81)
    yield M
    M.set_rho(1000.0)


def _d4(f, w, h):
    """Fourth-order central difference of the dict-valued f at w (arrays over frequencies)."""
    p2, p1, m1, m2 = f(w + 2 * h), f(w + h), f(w - h), f(w - 2 * h)
    return {k: (-p2[k] + 8 * p1[k] - 8 * m1[k] + m2[k]) / (12 * h) for k in KEYS}


def _check(got, ref, tol, what):
    for key in KEYS:
        g, r = np.asarray(got[key]), np.asarray(ref[key]).reshape(np.asarray(got[key]).shape)
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"{what} {key}: {rel:.2e}", flush=True)
        assert rel <= tol, (what, key, rel)


def _lid_body(mhx):
    from mhx import meshes
    hull = meshes.cylinder(1.0, 1.0, 16, 4, 3)
    lid = meshes.disk_lid(1.0, 3, 16, z=-0.1)
    return mhx.FloatingBody(hull, ["Heave", "Surge", "Pitch"], lid_mesh=lid), hull, lid


def test_lid_body_domega_against_oracle_differences(mhx, oracle):
    from mhx import meshes
    body, hull, lid = _lid_body(mhx)
    names = body.dof_names
    ws, betas = np.array([0.9, 1.7]), [0.0, math.pi / 3]
    par = dict(wave_frequencies=ws, radiating_dofs=names, wave_directions=betas)
    print("lid body: tangent", flush=True)
    val, tan = mhx.hydrodynamic_coefficients_jvp(body, par, {"wave_frequencies": np.ones(2)}, mhx.DirectBEM())
    full = meshes.combine(hull, lid)
    om = oracle.Mesh.from_mhx(full)
    dl = [np.vstack([body.dofs[nm], np.zeros((lid.nfaces, 3))]) for nm in names]

    def ora(w):
        return oracle.hydro_coeffs(om, dl, w, betas, rho=1025.0, g=9.81, gf=oracle.GF_WU, direct=True,
                                   nhull=hull.nfaces)
    # the values equal the oracle's; the tangents its differences
    _check(val, ora(ws), 1e-9, "lid value")
    _check(tan, _d4(ora, ws, 1e-3), 1e-7, "lid d/domega")


def test_lid_body_exact_domega_against_device_differences(mhx):
    body, _, _ = _lid_body(mhx)
    names = body.dof_names
    ws, betas = np.array([1.2]), [0.0]
    form = mhx.DirectBEM("ExactGuevelDelhommeau")
    par = dict(wave_frequencies=ws, radiating_dofs=names, wave_directions=betas)
    print("lid body exact: tangent", flush=True)
    _, tan = mhx.hydrodynamic_coefficients_jvp(body, par, {"wave_frequencies": np.ones(1)}, form)

    def dev(w):
        return mhx.hydrodynamic_coefficients(body, dict(par, wave_frequencies=w), form)
    _check(tan, _d4(dev, ws, 1e-2), 1e-5, "lid exact d/domega")


def _fs_body(mhx):
    from mhx import meshes
    return mhx.FloatingBody(meshes.sphere(1.0, 8, 16), ["Surge", "Heave", "Pitch"], rotation_center=(1.0, 1.0, 0.0))


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
def test_forward_speed_domega(mhx, oracle, gf):
    """Sphere R = 1 (64 panels), Surge/Heave/Pitch about (1, 1, 0), U in {0, 0.2, 1.5}, beta in
    {2.0, 0}, IndirectBEM: GFWu against the oracle's differences, the exact term against the
    device's own."""
    body = _fs_body(mhx)
    names = body.dof_names
    ws, betas, us = np.array([0.8, 1.3]), [2.0, 0.0], [0.0, 0.2, 1.5]
    form = mhx.IndirectBEM(gf)
    par = dict(wave_frequencies=ws, radiating_dofs=names, wave_directions=betas, forward_speeds=us)
    print(f"forward speed {gf}: tangent", flush=True)
    val, tan = mhx.hydrodynamic_coefficients_jvp(body, par, {"wave_frequencies": np.ones(2)}, form)
    assert tan["added_mass"].shape == (3, 3, 2, 3, 2)
    if gf == "Wu":
        om = oracle.Mesh.from_mhx(body.mesh)

        def ora(w):
            return oracle.hydro_coeffs_fs(om, [body.dofs[nm] for nm in names],
                                          [oracle.gradient_of_motion(nm) for nm in names], w, betas, us, rho=1025.0,
                                          g=9.81, gf=oracle.GF_WU)
        _check(val, ora(ws), 1e-9, "fs value")
        _check(tan, _d4(ora, ws, 1e-3), 1e-7, "fs d/domega")
    else:
        def dev(w):
            return mhx.hydrodynamic_coefficients(body, dict(par, wave_frequencies=w), form)
        _check(val, dev(ws), 1e-10, "fs exact value")
        _check(tan, _d4(dev, ws, 1e-2), 1e-5, "fs exact d/domega")


def test_at_rest_without_lid_equals_plain_tangent(mhx):
    body = _fs_body(mhx)
    names = body.dof_names
    ws, betas = np.array([0.7, 1.9]), [0.0, math.pi / 4]
    par = dict(wave_frequencies=ws, radiating_dofs=names, wave_directions=betas)
    dW = {"wave_frequencies": np.array([1.0, -0.5])}
    print("at rest: both tangents", flush=True)
    v0, t0 = mhx.hydrodynamic_coefficients_jvp(body, par, dW, mhx.DirectBEM())
    v1, t1 = mhx.hydrodynamic_coefficients_jvp(body, dict(par, forward_speeds=[0.0]), dW, mhx.DirectBEM())
    for key in KEYS:
        a0, a1 = np.asarray(t0[key]), np.asarray(t1[key])
        a1 = a1.reshape(a0.shape) if key in ("diffraction_force", "Froude_Krylov_force") else a1[:, :, :, 0, 0][..., None]
        rel, elem, ok = close(a1, a0, tol=1e-12)
        assert ok, (key, rel, elem)
