"""Forward-mode derivatives of the coefficients (tangent assembly, tangent right-hand sides and
forces, tangent solve on the FP64 LU) through the C ABI.

Checks: the reference notebook's ForwardDiff/Enzyme derivatives (mh_tutorial.ipynb cells 5, 9,
13: d/dr of the sphere radius and d/domega) and central differences of the oracle along random
directions of every input field (vertices, centers, normals, areas, dof motions, frequencies).
"""
import json
import math
import os

import numpy as np
import pytest

from test_gpu_parity import sphere18, sphere64

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))["notebook_sphere"]


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1000.0)
    M.set_g(9.81)
    return M


def _forms(mhx):
    return {"direct_Wu": mhx.DirectBEM("Wu"), "indirect_Wu": mhx.IndirectBEM("Wu"),
            "direct_Delhommeau": mhx.DirectBEM("ExactGuevelDelhommeau"),
            "indirect_Delhommeau": mhx.IndirectBEM("ExactGuevelDelhommeau")}


def _radius_and_omega(mhx, form, omegas):
    """(A, dA/dr, dA/domega, Re F_ex, dRe F_ex/dr, dRe F_ex/domega) of the 18-panel sphere, heave."""
    m = sphere18()
    body = mhx.FloatingBody(m, ["Heave"])
    par = {"wave_frequencies": omegas, "radiating_dofs": ["Heave"], "wave_directions": [0.0]}
    p, tr = mhx.hydrodynamic_coefficients_jvp(body, par, mhx.mesh_scaling_tangent(m), form)
    _, tw = mhx.hydrodynamic_coefficients_jvp(body, par, {"wave_frequencies": np.ones(len(omegas))}, form)
    return (p["added_mass"].ravel(), tr["added_mass"].ravel(), tw["added_mass"].ravel(),
            p["excitation_force"].ravel().real, tr["excitation_force"].ravel().real,
            tw["excitation_force"].ravel().real)


def test_notebook_derivatives_direct_wu(mhx):
    """Printed to 6 decimals in the notebook; the reference's own ForwardDiff and Enzyme agree to
    1e-12 relative, so the bar is the print precision."""
    print("tangent notebook direct Wu", flush=True)
    d = REF["derivatives_direct_Wu"]
    A, dAr, dAw, F, dFr, dFw = _radius_and_omega(mhx, mhx.DirectBEM(), [1.03])
    assert A[0] == pytest.approx(REF["A33"], abs=1e-6)
    assert dAr[0] == pytest.approx(d["dA_dr"], abs=2e-6)
    assert dAw[0] == pytest.approx(d["dA_domega"], abs=2e-6)
    assert dFr[0] == pytest.approx(d["dReFex_dr"], abs=2e-6)
    assert dFw[0] == pytest.approx(d["dReFex_domega"], abs=2e-6)
    # several frequencies in one call: d A(omega_i) / dr and the gradient of sum_i A(omega_i)
    # with respect to the frequency vector (each A depends on its own omega only)
    A, dAr, dAw, *_ = _radius_and_omega(mhx, mhx.DirectBEM(), REF["omegas"])
    np.testing.assert_allclose(dAr, d["dA_dr_of_omega"], rtol=1e-9)
    np.testing.assert_allclose(dAw, d["dsumA_domegas"], rtol=1e-9)
    assert dAr.sum() == pytest.approx(d["dsumA_dr"], abs=2e-6)


@pytest.mark.parametrize("key", ["indirect_Wu", "direct_Delhommeau", "indirect_Delhommeau"])
def test_notebook_derivative_table(mhx, key):
    """(dA/dr, dA/domega, dRe F_ex/dr) printed to 4 decimals (mh_tutorial.ipynb cell 13)."""
    print("tangent notebook", key, flush=True)
    _, dAr, dAw, _, dFr, _ = _radius_and_omega(mhx, _forms(mhx)[key], [1.03])
    ref = REF["derivatives_table"][key]
    np.testing.assert_allclose([dAr[0], dAw[0], dFr[0]], ref, rtol=0, atol=1e-4)


def _direction(mesh, rng):
    """A direction of the mesh fields along which the coefficients are smooth: an affine map
    x -> x + h (M x + t) of the whole body (each centre stays on its panel's plane; a centre
    moved off its own plane makes the self-influence double-layer term jump by 1/2, for the
    reference as for any panel method: the in-plane test has a 1e-14 threshold), plus in-plane
    moves of the centres and
    independent perturbations of the normals and areas.  Returns the perturbed mesh at step h
    and the exact tangent at h = 0."""
    from mhx.meshes import Mesh
    n = mesh.nfaces
    M = 0.2 * rng.standard_normal((3, 3))
    t = 0.1 * rng.standard_normal(3)
    jit = 0.05 * rng.standard_normal((n, 3))
    jit -= np.sum(jit * mesh.normals, 1)[:, None] * mesh.normals
    dN = 0.05 * rng.standard_normal((n, 3))
    dA = 0.05 * mesh.areas * rng.standard_normal(n)

    def at(h):
        F = np.eye(3) + h * M
        nn = mesh.normals @ np.linalg.inv(F)
        s = np.linalg.norm(nn, axis=1)
        # the in-plane move is applied before the map, so each centre stays on its mapped panel
        return Mesh(mesh.vertices @ F.T + h * t, mesh.faces, (mesh.centers + h * jit) @ F.T + h * t,
                    nn / s[:, None] + h * dN, mesh.areas * abs(np.linalg.det(F)) * s + h * dA, mesh.radii)

    N = mesh.normals
    nMn = np.einsum("ij,jk,ik->i", N, M, N)
    tangent = {"vertices": mesh.vertices @ M.T + t, "centers": mesh.centers @ M.T + t + jit,
               "normals": -(N @ M) + nMn[:, None] * N + dN, "areas": mesh.areas * (np.trace(M) - nMn) + dA}
    return at, tangent


def _oracle_coeffs(oracle, mesh, dofs, omegas, betas, gf, direct):
    om = oracle.Mesh.from_mhx(mesh)
    return oracle.hydro_coeffs(om, dofs, omegas, betas, rho=1000.0, g=9.81,
                               gf=oracle.GF_WU if gf == "Wu" else oracle.GF_DELHOMMEAU, direct=direct)


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("meshname", ["sphere18", "sphere64"])
def test_random_direction_vs_oracle_central_difference(mhx, oracle, gf, direct, meshname):
    """Every input field perturbed at once along a seeded random direction (vertices, centers,
    normals, areas, a non-rigid dof field, the frequencies): the tangent matches the oracle's
    fourth-order central difference to 1e-6 of each block's scale, and the primal outputs match
    the oracle to 1e-9."""
    print(f"tangent random {meshname} {gf} direct={direct}", flush=True)
    m = sphere18() if meshname == "sphere18" else sphere64()
    n = m.nfaces
    rng = np.random.default_rng(7 + n + 2 * direct)
    at, t = _direction(m, rng)
    surge = np.zeros((n, 3)); surge[:, 0] = 1.0
    heave = np.zeros((n, 3)); heave[:, 2] = 1.0
    flex = np.stack([np.zeros(n), m.centers[:, 0], m.centers[:, 0] ** 2], 1)
    dflex = 0.3 * rng.standard_normal((n, 3))
    dofs = {"Surge": surge, "Heave": heave, "Flex": flex}
    t["dofs"] = {"Flex": dflex}
    omegas = np.array([0.7, 1.6])
    t["wave_frequencies"] = np.array([0.4, -0.9])
    betas = [0.0, 0.9]
    form = mhx.DirectBEM(gf) if direct else mhx.IndirectBEM(gf)
    body = mhx.FloatingBody(m, dofs)
    names = ["Surge", "Heave", "Flex"]
    par = {"wave_frequencies": omegas, "radiating_dofs": names, "wave_directions": betas}
    p, tg = mhx.hydrodynamic_coefficients_jvp(body, par, t, form)

    def orc(h):
        return _oracle_coeffs(oracle, at(h), [surge, heave, flex + h * dflex], omegas + h * t["wave_frequencies"],
                              betas, gf, direct)

    r0 = orc(0.0)
    h = 5e-5
    rp, rm, rp2, rm2 = orc(h), orc(-h), orc(2 * h), orc(-2 * h)
    for key in ("added_mass", "radiation_damping", "excitation_force", "Froude_Krylov_force"):
        ref0 = np.asarray(r0[key])
        got0 = np.asarray(p[key]).reshape(ref0.shape)
        assert np.linalg.norm(got0 - ref0) <= 1e-9 * np.linalg.norm(ref0), key
        fd = (8 * (np.asarray(rp[key]) - np.asarray(rm[key])) - (np.asarray(rp2[key]) - np.asarray(rm2[key]))) / (12 * h)
        got = np.asarray(tg[key]).reshape(fd.shape)
        err = np.max(np.abs(got - fd)) / np.max(np.abs(fd))
        assert err < 1e-6, (key, err)


def test_zero_direction_gives_zero_tangent(mhx):
    print("tangent zero direction", flush=True)
    m = sphere18()
    body = mhx.FloatingBody(m, ["Heave", "Pitch"])
    par = {"wave_frequencies": [0.9, 1.2], "radiating_dofs": ["Heave", "Pitch"], "wave_directions": [0.3]}
    p, tg = mhx.hydrodynamic_coefficients_jvp(body, par, {}, mhx.DirectBEM())
    ref = mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM())
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        np.testing.assert_allclose(p[key], ref[key], rtol=1e-12, atol=1e-9)
        assert np.all(tg[key] == 0)


def test_rigid_rotation_follows_centers(mhx, oracle):
    """A body built from dof names: the pitch motion c x e_y moves with the centers, so the
    tangent of a translated mesh includes the dof tangent (checked against the oracle's central
    difference of a translated sphere)."""
    print("tangent rigid pitch", flush=True)
    m = sphere18()
    shift = np.array([0.3, -0.2, 0.0])
    t = {"vertices": np.tile(shift, (m.nvertices, 1)), "centers": np.tile(shift, (m.nfaces, 1))}
    body = mhx.FloatingBody(m, ["Heave", "Pitch"])
    par = {"wave_frequencies": [1.1], "radiating_dofs": ["Heave", "Pitch"], "wave_directions": [0.5]}
    _, tg = mhx.hydrodynamic_coefficients_jvp(body, par, t, mhx.DirectBEM())

    def orc(h):
        from mhx.meshes import Mesh
        mm = Mesh(m.vertices + h * shift, m.faces, m.centers + h * shift, m.normals, m.areas, m.radii)
        om = oracle.Mesh.from_mhx(mm)
        return oracle.hydro_coeffs(om, [oracle.rigid_dof(om, 2), oracle.rigid_dof(om, 4)], [1.1], [0.5],
                                   rho=1000.0, g=9.81)

    h = 1e-5
    rp, rm = orc(h), orc(-h)
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        fd = (rp[key] - rm[key]) / (2 * h)
        got = np.asarray(tg[key]).reshape(fd.shape)
        assert np.max(np.abs(got - fd)) <= 1e-6 * np.max(np.abs(fd)) + 1e-6, key


# ---------------------------------------------------------------------------- reverse mode
def _dot_tangent(grads, t, names):
    s = 0.0
    for key in ("vertices", "centers", "normals", "areas", "wave_frequencies"):
        if key in t:
            s += float(np.sum(grads[key] * t[key]))
    for nm, d in t.get("dofs", {}).items():
        s += float(np.sum(grads["dofs"][nm] * d))
    return s


def _dot_cotangent(cot, out):
    s = 0.0
    for key in ("added_mass", "radiation_damping"):
        s += float(np.sum(np.asarray(cot[key]).ravel() * np.asarray(out[key]).ravel()))
    for key in ("diffraction_force", "Froude_Krylov_force"):
        s += float(np.sum((np.conj(np.asarray(cot[key]).ravel()) * np.asarray(out[key]).ravel()).real))
    return s


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("meshname", ["sphere18", "sphere64"])
def test_vjp_transposes_jvp(mhx, gf, direct, meshname):
    """<VJP(ybar), t> = <ybar, JVP(t)> for a seeded random cotangent of every output and a random
    direction of every input (the reverse mode is the transpose of the forward mode, which is
    pinned against the notebook and the oracle above), to 1e-10 of the terms' scale."""
    print(f"vjp dot test {meshname} {gf} direct={direct}", flush=True)
    m = sphere18() if meshname == "sphere18" else sphere64()
    n = m.nfaces
    rng = np.random.default_rng(11 + n + 2 * direct)
    _, t = _direction(m, rng)
    surge = np.zeros((n, 3)); surge[:, 0] = 1.0
    heave = np.zeros((n, 3)); heave[:, 2] = 1.0
    flex = np.stack([np.zeros(n), m.centers[:, 0], m.centers[:, 0] ** 2], 1)
    t["dofs"] = {"Flex": 0.3 * rng.standard_normal((n, 3)), "Surge": 0.1 * rng.standard_normal((n, 3))}
    omegas = np.array([0.7, 1.6])
    t["wave_frequencies"] = np.array([0.4, -0.9])
    betas = [0.0, 0.9]
    form = mhx.DirectBEM(gf) if direct else mhx.IndirectBEM(gf)
    body = mhx.FloatingBody(m, {"Surge": surge, "Heave": heave, "Flex": flex})
    names = ["Surge", "Heave", "Flex"]
    par = {"wave_frequencies": omegas, "radiating_dofs": names, "wave_directions": betas}
    cot = {"added_mass": rng.standard_normal((3, 3, 2, 1)), "radiation_damping": rng.standard_normal((3, 3, 2, 1)),
           "diffraction_force": rng.standard_normal((3, 2, 2, 1)) + 1j * rng.standard_normal((3, 2, 2, 1)),
           "Froude_Krylov_force": rng.standard_normal((3, 2, 2, 1)) + 1j * rng.standard_normal((3, 2, 2, 1))}
    p1, tg = mhx.hydrodynamic_coefficients_jvp(body, par, t, form)
    p2, gr = mhx.hydrodynamic_coefficients_vjp(body, par, cot, form)
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        np.testing.assert_allclose(p2[key], p1[key], rtol=1e-12, atol=1e-9)
    lhs = _dot_tangent(gr, t, names)
    rhs = _dot_cotangent(cot, tg)
    scale = sum(abs(float(np.sum(gr[k] * t[k]))) for k in ("vertices", "centers", "normals", "areas"))
    assert abs(lhs - rhs) <= 1e-10 * max(scale, abs(rhs)), (lhs, rhs, scale)


def test_vjp_notebook_derivatives(mhx):
    """Enzyme reverse in the notebook (mh_tutorial.ipynb cells 5, 9, 13): dA/dr, dA/domega,
    dRe F_ex/dr, dRe F_ex/domega at omega = 1.03, and the gradient of sum_i A(omega_i) with
    respect to the frequencies and r."""
    print("vjp notebook", flush=True)
    d = REF["derivatives_direct_Wu"]
    m = sphere18()
    body = mhx.FloatingBody(m, ["Heave"])
    sc = mhx.mesh_scaling_tangent(m)
    par = {"wave_frequencies": [1.03], "radiating_dofs": ["Heave"], "wave_directions": [0.0]}
    _, g = mhx.hydrodynamic_coefficients_vjp(body, par, {"added_mass": np.ones(1)})
    assert _dot_tangent(g, sc, []) == pytest.approx(d["dA_dr"], abs=2e-6)
    assert g["wave_frequencies"][0] == pytest.approx(d["dA_domega"], abs=2e-6)
    # chained through the face properties, the radius direction is the vertices alone
    gv = mhx.chain_mesh_gradient(m, g)
    assert float(np.sum(gv * m.vertices)) == pytest.approx(d["dA_dr"], abs=2e-6)
    _, g = mhx.hydrodynamic_coefficients_vjp(body, par, {"excitation_force": np.ones(1, dtype=complex)})
    assert _dot_tangent(g, sc, []) == pytest.approx(d["dReFex_dr"], abs=2e-6)
    assert g["wave_frequencies"][0] == pytest.approx(d["dReFex_domega"], abs=2e-6)
    par = {"wave_frequencies": REF["omegas"], "radiating_dofs": ["Heave"]}
    _, g = mhx.hydrodynamic_coefficients_vjp(body, par, {"added_mass": np.ones(3)})
    np.testing.assert_allclose(g["wave_frequencies"], d["dsumA_domegas"], rtol=1e-9)
    assert _dot_tangent(g, sc, []) == pytest.approx(d["dsumA_dr"], abs=5e-6)


def test_vjp_rigid_rotation_folds_into_centers(mhx):
    """Body built from dof names (heave, pitch): the pitch motion follows the centres, so the
    gradient with respect to the centres includes it; checked by the transpose identity along a
    translation of the mesh."""
    print("vjp rigid pitch", flush=True)
    m = sphere18()
    shift = np.array([0.3, -0.2, 0.1])
    t = {"vertices": np.tile(shift, (m.nvertices, 1)), "centers": np.tile(shift, (m.nfaces, 1))}
    body = mhx.FloatingBody(m, ["Heave", "Pitch"])
    par = {"wave_frequencies": [1.1], "radiating_dofs": ["Heave", "Pitch"], "wave_directions": [0.5]}
    rng = np.random.default_rng(5)
    cot = {"added_mass": rng.standard_normal((2, 2, 1)), "radiation_damping": rng.standard_normal((2, 2, 1)),
           "diffraction_force": rng.standard_normal((2, 1, 1)) + 1j * rng.standard_normal((2, 1, 1)),
           "Froude_Krylov_force": rng.standard_normal((2, 1, 1)) + 1j * rng.standard_normal((2, 1, 1))}
    _, tg = mhx.hydrodynamic_coefficients_jvp(body, par, t, mhx.DirectBEM())
    _, gr = mhx.hydrodynamic_coefficients_vjp(body, par, cot, mhx.DirectBEM())
    lhs = float(np.sum(gr["vertices"] * t["vertices"]) + np.sum(gr["centers"] * t["centers"]))
    rhs = _dot_cotangent(cot, tg)
    assert lhs == pytest.approx(rhs, rel=1e-10)


def test_adjoint_sensitivities_hemisphere_2048(mhx):
    """Hemisphere of radius 1 m, 2,048 panels, 16 frequencies in 0.5-2.0 rad/s, FP64: gradients
    of heave added mass, damping and Re(heave excitation) at beta = 0 with respect to the radius,
    the frequencies and the vertex coordinates, against central differences (step 1e-6) of the
    FP64 omega-solve (itself within 1e-9 of the oracle).  Vertex coordinates are checked along
    smooth directions (an affine map of the body): moving one vertex of a quad alone bends the
    panel, the centre leaves the panel's plane and the self-influence double-layer term jumps,
    for the reference as here, so a single-coordinate difference has no derivative to match."""
    print("adjoint hemisphere 2048 x 16", flush=True)
    from mhx import meshes
    mhx.set_rho(1025.0)
    try:
        m = meshes.sphere_with_count(1.0, 2048)
        ws = np.linspace(0.5, 2.0, 16)
        par = {"wave_frequencies": ws, "radiating_dofs": ["Heave"], "wave_directions": [0.0]}
        form = mhx.DirectBEM()

        def coeffs(mesh, w=ws):
            body = mhx.FloatingBody(mesh, ["Heave"])
            r = mhx.hydrodynamic_coefficients(body, dict(par, wave_frequencies=w), form)
            return np.stack([r["added_mass"].ravel(), r["radiation_damping"].ravel(), r["excitation_force"].ravel().real])

        h = 1e-6
        fd_r = (coeffs(m.scaled(1 + h)) - coeffs(m.scaled(1 - h))) / (2 * h)
        fd_w = (coeffs(m, ws + h) - coeffs(m, ws - h)) / (2 * h)
        rng = np.random.default_rng(2048)
        at, tdir = _direction(m, rng)
        fd_v = (coeffs(at(h)) - coeffs(at(-h))) / (2 * h)
        body = mhx.FloatingBody(m, ["Heave"])
        sc = mhx.mesh_scaling_tangent(m)
        for q, key in enumerate(("added_mass", "radiation_damping", "excitation_force")):
            cot = {key: np.ones(16, dtype=complex if key == "excitation_force" else float)}
            p, g = mhx.hydrodynamic_coefficients_vjp(body, par, cot, form)
            dr = _dot_tangent(g, sc, [])
            assert dr == pytest.approx(fd_r[q].sum(), rel=1e-6), key
            np.testing.assert_allclose(g["wave_frequencies"], fd_w[q], rtol=1e-6, atol=1e-6 * np.max(np.abs(fd_w[q])))
            dv = _dot_tangent(g, tdir, [])
            assert dv == pytest.approx(fd_v[q].sum(), rel=1e-6), key
        # the forward mode gives the per-frequency radius derivatives
        _, tr = mhx.hydrodynamic_coefficients_jvp(body, par, sc, form)
        got = np.stack([tr["added_mass"].ravel(), tr["radiation_damping"].ravel(), tr["excitation_force"].ravel().real])
        for q in range(3):
            err = np.max(np.abs(got[q] - fd_r[q])) / np.max(np.abs(fd_r[q]))
            print(f"  output {q}: max |jvp - fd| / max |fd| = {err:.2e}", flush=True)
            assert err < 1e-6
    finally:
        mhx.set_rho(1000.0)
