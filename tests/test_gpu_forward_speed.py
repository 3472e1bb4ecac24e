"""Forward speed (forward_speeds: problems_and_results.jl:236-380, waves.jl:136-151, 183-199,
266-303, solve.jl:13-19, 35-55): the device path against the oracle's forward-speed restatement,
the reduction to the plain path at U = 0 (any formulation), single problems against the
coefficient grid."""
import math

import numpy as np
import pytest

from test_gpu_parity import close

pytestmark = pytest.mark.gpu

KEYS = ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force", "excitation_force")
DOFS = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1025.0)
    M.set_g(9.81)
    yield M
    M.set_rho(1000.0)


def _sphere_body(mhx):
    from mhx import meshes
    # rotation centre off the axis, as the reference's forward-speed test, for nonzero couplings
    return mhx.FloatingBody(meshes.sphere(1.0, 8, 16), DOFS, rotation_center=(1.0, 1.0, 0.0))


def _oracle_fs(oracle, body, names, ws, betas, us, gf):
    return oracle.hydro_coeffs_fs(oracle.Mesh.from_mhx(body.mesh), [body.dofs[nm] for nm in names],
                                  [oracle.gradient_of_motion(nm) for nm in names], ws, betas, us, rho=1025.0, g=9.81,
                                  gf=oracle.GF_WU if gf == "Wu" else oracle.GF_DELHOMMEAU)


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
def test_forward_speed_matches_oracle(mhx, oracle, gf):
    """Sphere R = 1 (64 panels), six rigid dofs about (1, 1, 0), speeds 0, 0.2, -0.3, 2 and 8 m/s,
    frequencies 0.5, 1.0, 1.7, headings 0, pi/6, pi/2, 2: every coefficient within 1e-9 of the
    oracle.  U = 8 at omega = 1.7 and beta = 0 or pi/6 has a negative Doppler-shifted frequency
    (beta_e = beta + pi); beta = pi/2 keeps omega_e = omega with the U terms present."""
    print(f"forward speed parity {gf}", flush=True)
    body = _sphere_body(mhx)
    ws, betas, us = [0.5, 1.0, 1.7], [0.0, math.pi / 6, math.pi / 2, 2.0], [0.0, 0.2, -0.3, 2.0, 8.0]
    assert mhx.compute_encountered_values(1.7, 0.0, 8.0, 9.81)[2] == math.pi
    par = {"wave_frequencies": ws, "radiating_dofs": DOFS, "wave_directions": betas, "forward_speeds": us}
    got = mhx.hydrodynamic_coefficients(body, par, mhx.IndirectBEM(gf))
    ref = _oracle_fs(oracle, body, DOFS, ws, betas, us, gf)
    assert ref["info"] == 0
    assert got["added_mass"].shape == (6, 6, 3, 5, 4) and got["diffraction_force"].shape == (6, 3, 4, 5)
    for key in KEYS:
        rel, elem, ok = close(got[key], ref[key])
        assert ok, (key, rel, elem)
    # the speed terms matter: U = 2 differs from U = 0 by far more than the tolerance
    a = got["added_mass"]
    assert np.linalg.norm(a[:, :, :, 3] - a[:, :, :, 0]) > 1e-3 * np.linalg.norm(a[:, :, :, 0])


def test_forward_speed_influenced_subset_and_larger_mesh(mhx, oracle):
    """Cylinder R = 1, draft 2 (768 panels), radiating Surge and Pitch, influenced Heave, Surge and
    Pitch, U = 2 at omega = 1.2 and heading pi/6, GFWu: within 1e-9 of the oracle."""
    print("forward speed 768 panels", flush=True)
    from mhx import meshes
    body = mhx.FloatingBody(meshes.cylinder(1.0, 2.0, 32, 16, 8), ["Surge", "Heave", "Pitch"],
                            rotation_center=(0.0, 0.0, -0.5))
    inf, rad = ["Heave", "Surge", "Pitch"], ["Surge", "Pitch"]
    par = {"wave_frequencies": [1.2], "radiating_dofs": rad, "influenced_dofs": inf, "wave_directions": [math.pi / 6],
           "forward_speeds": [2.0]}
    got = mhx.hydrodynamic_coefficients(body, par, mhx.IndirectBEM())
    ref = _oracle_fs(oracle, body, inf, [1.2], [math.pi / 6], [2.0], "Wu")
    ii = [0, 1, 2]
    rr = [inf.index(r) for r in rad]
    pairs = {"added_mass": ref["added_mass"][np.ix_(ii, rr)], "radiation_damping": ref["radiation_damping"][np.ix_(ii, rr)],
             "diffraction_force": ref["diffraction_force"], "Froude_Krylov_force": ref["Froude_Krylov_force"]}
    for key, r in pairs.items():
        rel, elem, ok = close(got[key], r)
        assert ok, (key, rel, elem)


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
def test_forward_speeds_at_rest_equal_the_plain_path(mhx, gf):
    """forward_speeds = [0] with DirectBEM (allowed at rest): every heading's radiation block equals
    the plain path's, and the forces equal the plain forces, to 1e-11; twelve headings with six
    dofs take two heading chunks."""
    print(f"forward speed at rest {gf}", flush=True)
    body = _sphere_body(mhx)
    ws = [0.4, 1.1, 2.3]
    betas = list(np.linspace(0.0, 2 * math.pi, 12, endpoint=False))
    form = mhx.DirectBEM(gf)
    fs = mhx.hydrodynamic_coefficients(body, {"wave_frequencies": ws, "radiating_dofs": DOFS, "wave_directions": betas,
                                              "forward_speeds": [0.0]}, form)
    plain = mhx.hydrodynamic_coefficients(body, {"wave_frequencies": ws, "radiating_dofs": DOFS,
                                                 "wave_directions": betas}, form)
    for ib in range(len(betas)):
        for key in ("added_mass", "radiation_damping"):
            rel, elem, ok = close(fs[key][..., ib], plain[key], tol=1e-11)
            assert ok, (key, ib, rel, elem)
    for key in ("diffraction_force", "Froude_Krylov_force", "excitation_force"):
        rel, elem, ok = close(fs[key], plain[key], tol=1e-11)
        assert ok, (key, rel, elem)


def test_single_problems_match_the_grid(mhx):
    """solve(RadiationProblem(...; beta, forward_speed)) and solve(DiffractionProblem(...;
    forward_speed)) equal the corresponding entries of the forward-speed grid (1e-11);
    added_mass / radiation_damping of a moving body divide by the encountered frequency."""
    print("forward speed single problems", flush=True)
    body = _sphere_body(mhx)
    w, beta, U = 1.3, 2.0, 0.6
    form = mhx.IndirectBEM("ExactGuevelDelhommeau")
    grid = mhx.hydrodynamic_coefficients(body, {"wave_frequencies": [w], "radiating_dofs": DOFS,
                                                "wave_directions": [beta], "forward_speeds": [U]}, form)
    for r, dof in enumerate(DOFS):
        prob = mhx.RadiationProblem(body, w, dof, beta=beta, forward_speed=U)
        we, ke, be = mhx.compute_encountered_values(w, beta, U, 9.81)
        assert (prob.encountered_omega, prob.encountered_wavenumber, prob.encountered_beta) == (we, ke, be)
        res = mhx.solve(prob, form)
        am, bd = mhx.added_mass(res), mhx.radiation_damping(res)
        got = np.array([[am[nm] for nm in DOFS], [bd[nm] for nm in DOFS]])
        ref = np.array([grid["added_mass"][:, r, 0, 0, 0], grid["radiation_damping"][:, r, 0, 0, 0]])
        rel, elem, ok = close(got, ref, tol=1e-11)
        assert ok, (dof, rel, elem)
    res = mhx.solve(mhx.DiffractionProblem(body, w, beta, forward_speed=U), form)
    fd, fk = mhx.diffraction_force(res), mhx.froude_krylov_force(res)
    got = np.array([[fd[nm] for nm in DOFS], [fk[nm] for nm in DOFS]])
    ref = np.array([grid["diffraction_force"][:, 0, 0, 0], grid["Froude_Krylov_force"][:, 0, 0, 0]])
    rel, elem, ok = close(got, ref, tol=1e-11)
    assert ok, (rel, elem)


def test_forward_speed_surrogate_within_1e3_of_exact(mhx):
    """The learned wave term with forward speed: sphere R = 1 (64 panels) and cylinder R = 1, draft 2
    (768 panels), six dofs, speeds 0, 1.5 and 8 m/s, three frequencies, headings 0 and pi/3.  S, D
    and K_x come from the surrogate's (G, dG/dr) table, the solves are the default mixed precision
    of the surrogate (split-bf16 LU refined to 1e-12): every coefficient block within 1e-3
    (normwise) of the exact FP64 path; the solves ran mixed, with no FP64 fallback."""
    from mhx import meshes
    mhx.load_surrogate()
    ctx = mhx.context()
    for name, mesh in (("sphere64", meshes.sphere(1.0, 8, 16)), ("cylinder768", meshes.cylinder(1.0, 2.0, 32, 16, 8))):
        print(f"forward speed surrogate {name}", flush=True)
        body = mhx.FloatingBody(mesh, DOFS, rotation_center=(0.0, 0.0, -0.5))
        par = {"wave_frequencies": [0.6, 1.2, 2.0], "radiating_dofs": DOFS, "wave_directions": [0.0, math.pi / 3],
               "forward_speeds": [0.0, 1.5, 8.0]}
        ex = mhx.hydrodynamic_coefficients(body, par, mhx.IndirectBEM("ExactGuevelDelhommeau"))
        ctx.reset_timing()
        su = mhx.hydrodynamic_coefficients(body, par, mhx.IndirectBEM("Surrogate"))
        mixed, _, fallbacks, _ = ctx.solver_stats()
        assert mixed > 0 and fallbacks == 0, (mixed, fallbacks)
        for key in KEYS:
            for iu in range(3):
                r = ex[key][..., iu, :] if key in ("added_mass", "radiation_damping") else ex[key][..., iu]
                s = su[key][..., iu, :] if key in ("added_mass", "radiation_damping") else su[key][..., iu]
                err = np.linalg.norm(s - r) / np.linalg.norm(r)
                print(f"  {key} U-index {iu}: {err:.2e}", flush=True)
                assert err <= 1e-3, (name, key, iu, err)
