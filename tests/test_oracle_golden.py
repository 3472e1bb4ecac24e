"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

Checks use the reference's own golden vectors (tests/golden/reference_values.json,
each block cites the reference file:line), closed forms, library routines
(scipy / mpmath / numpy LAPACK) and invariants, chosen so that a dropped term,
sign or index error fails one of them.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def _birk_eval(O, corners, p):
    T, q, lc, A = O.birk_panel_geometry(np.array(corners))
    fl = T.T @ (np.array(p) - q)
    phi = O.velocity_potential(*fl, lc)
    v = T @ O.velocity_derivatives(*fl, lc)
    H = T @ O.velocity_hessian(*fl, lc) @ T.T
    return phi, v, fl, H, (T, q, lc, A)


@pytest.mark.parametrize("panel", ["birk_square", "birk_twisted", "birk_triangle"])
def test_birk_tables(oracle, panel):
    """Birk (2021) tables, test/rankine_vectorized.jl:26-153 (atol 1e-6, Laplace trace 1e-8)."""
    g = REF[panel]
    _, _, _, _, (T, q, lc, A) = _birk_eval(oracle, g["corners"], [0, 0, 5])
    assert abs(A - g["area"]) < 1e-6
    if "T" in g:
        np.testing.assert_allclose(T, np.array(g["T"]), atol=1e-6)
        np.testing.assert_allclose(q, g["qgc"], atol=1e-6)
    if "local_corners" in g:
        np.testing.assert_allclose(lc, np.array(g["local_corners"]), atol=1e-6)
    for case in g["cases"]:
        p, phi_e, v_e = case[0], case[1], case[2]
        phi, v, fl, H, _ = _birk_eval(oracle, g["corners"], p)
        assert abs(phi - phi_e) < 1e-6
        np.testing.assert_allclose(v, v_e, atol=1e-5)
        assert abs(np.trace(H)) < 1e-8
        if panel == "birk_twisted":
            np.testing.assert_allclose(fl, case[3], atol=1e-6)
        elif panel == "birk_square":
            np.testing.assert_allclose(H, np.array(case[3]), atol=1e-6)
    for p, H_e in g.get("hessian_cases", []):
        _, _, _, H, _ = _birk_eval(oracle, g["corners"], p)
        np.testing.assert_allclose(H, np.array(H_e), atol=1e-6)


def test_birk_fused_matches_separate(oracle):
    """test/rankine_vectorized.jl:155-175 (1e-14), including a degenerate edge."""
    lc = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]])
    for x, y, z in ((0.1, 0.2, 0.4), (0.0, 0.0, 0.0), (0.5, 1.0, 0.0), (-0.25, 0.25, 0.5)):
        phi2, v2 = oracle.birk_phi_and_velocity(x, y, z, lc)
        assert abs(oracle.velocity_potential(x, y, z, lc) - phi2) <= 1e-14
        np.testing.assert_allclose(oracle.velocity_derivatives(x, y, z, lc), v2, atol=1e-14, rtol=1e-14)
    tri = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [0.5, 0.5]])
    phi2, v2 = oracle.birk_phi_and_velocity(0.1, 0.2, 0.3, tri)
    assert abs(oracle.velocity_potential(0.1, 0.2, 0.3, tri) - phi2) <= 1e-14
    np.testing.assert_allclose(oracle.velocity_derivatives(0.1, 0.2, 0.3, tri), v2, atol=1e-14)


def test_birk_against_brute_force_quadrature(oracle):
    """Independent pin: 2-D Gauss-Legendre of 1/|x - xi| over a planar quad for |z| > 0."""
    corners = np.array([[0.0, 0.0, 0.0], [1.3, 0.1, 0.0], [1.1, 0.9, 0.0], [-0.2, 0.8, 0.0]])
    T, q, lc, A = oracle.birk_panel_geometry(corners)
    xg, wg = np.polynomial.legendre.leggauss(96)
    u = 0.5 * (xg + 1); wu = 0.5 * wg
    for p in ([0.4, 0.3, 0.35], [2.0, -1.0, -0.7], [0.5, 0.5, -0.05]):
        # bilinear map of the unit square onto the quad
        P0, P1, P2, P3 = corners
        tot = 0.0; grad = np.zeros(3)
        for a, wa in zip(u, wu):
            for b, wb in zip(u, wu):
                X = (1 - a) * (1 - b) * P0 + a * (1 - b) * P1 + a * b * P2 + (1 - a) * b * P3
                Xa = (1 - b) * (P1 - P0) + b * (P2 - P3)
                Xb = (1 - a) * (P3 - P0) + a * (P2 - P1)
                J = np.linalg.norm(np.cross(Xa, Xb))
                d = np.array(p) - X
                r = np.linalg.norm(d)
                tot += wa * wb * J / r
                grad += wa * wb * J * d / r ** 3
        phi, gsrc = oracle.vrankine_from_geom(np.array(p), q, 10.0, A, T, q, lc)
        assert abs(phi - tot) < 2e-5 * abs(tot)
        np.testing.assert_allclose(gsrc, grad, rtol=5e-4, atol=1e-5)


def test_rankine_point_source_far_field(oracle):
    """Far field (r > 7 radius) is area/r and area (x - xi)/r^3: test/greens_function.jl:40-47, 68-73."""
    corners = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.5, 0.0]]) + [1.0, 1.0, -2.0]
    T, q, lc, A = oracle.birk_panel_geometry(corners)
    radius = math.sqrt(2) / 2
    c2 = np.array([1.0, 1.0, -2.0])
    p = np.array([20.0, 15.0, -9.0])
    phi, g = oracle.vrankine_from_geom(p, c2, radius, A, T, q, lc)
    r = np.linalg.norm(p - c2)
    assert phi == pytest.approx(A / r, abs=1e-15)
    np.testing.assert_allclose(g, (p - c2) / r ** 3 * A, atol=1e-15)
    # just inside 7 radius: the panel integral differs from the point source by O((radius/r)^2)
    p = c2 + np.array([0.0, 0.0, 4.9 * radius])
    phi, g = oracle.vrankine_from_geom(p, c2, radius, A, T, q, lc)
    r = np.linalg.norm(p - c2)
    assert 1e-4 < abs(phi - A / r) / (A / r) < 1e-2
    # image value 1/sqrt(1+1+9) of test/greens_function.jl:68
    e1 = np.array([0.0, 0.0, -1.0])
    phi_img, _ = oracle.vrankine_from_geom(e1 * [1, 1, -1], c2, radius, A, T, q, lc)
    assert phi_img == pytest.approx(1 / math.sqrt(11), abs=1e-1)


def test_struve_fits_against_scipy(oracle):
    from scipy.special import struve
    for x in (0.1, 0.7, 2.0, 2.999, 3.001, 5.0, 12.0, 40.0):
        assert abs(oracle.struve_h0(x) - struve(0, x)) < 5e-7 * max(1, abs(struve(0, x))) + 2e-7
        assert abs(oracle.struve_h1(x) - struve(1, x)) < 5e-7 * max(1, abs(struve(1, x))) + 2e-7


def test_expint_against_mpmath(oracle):
    import mpmath as mp
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(400):
        R = 10 ** rng.uniform(-3, 2.05)
        ang = rng.uniform(-math.pi * 0.999, math.pi * 0.999)
        z = complex(R * math.cos(ang), R * math.sin(ang))
        ref = complex(mp.exp(z) * mp.e1(z))
        got = oracle.expE1(z)
        worst = max(worst, abs(got - ref) / abs(ref))
    assert worst < 5e-14
    # branch: approached from above on the negative real axis, E1(-x + i0) = -Ei(x) - i pi
    e = oracle.expint_e1(complex(-3.0, 0.0))
    assert e.imag == pytest.approx(-math.pi, abs=1e-14)
    assert e.real == pytest.approx(-float(mp.ei(3.0)), rel=1e-14)


def test_gk15_rule_is_exact_to_degree_22(oracle):
    xk, wk, wg = oracle.gk15_rule()
    X = np.concatenate([xk[:7], [0.0], -xk[:7][::-1]])
    WK = np.concatenate([wk[:7], [wk[7]], wk[:7][::-1]])
    for d in range(23):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.sum(WK * X ** d) - exact) < 1e-14
    gx, gw = np.polynomial.legendre.leggauss(7)
    np.testing.assert_allclose(np.sort(np.abs(gx))[::-1][::2][:3], xk[[1, 3, 5]], atol=1e-15)
    np.testing.assert_allclose(sorted(set(np.round(gw, 15))), sorted(set(np.round(wg, 15))), atol=1e-14)


def test_delhommeau_capytaine_values(oracle):
    """test/greens_function.jl:80-92 (atol 1e-4) plus the exact imaginary parts."""
    g = REF["delhommeau_capytaine"]
    G, Gr, Gz = oracle.delhommeau_term(1.0, -1.0)
    assert abs(G - complex(*g["G_1_m1"])) < 1e-4
    assert abs(Gr - complex(*g["Gr_1_m1"])) < 1e-4
    # Capytaine's greens(e1, e2, k=1): k * G(r=sqrt 2, z=-3)
    G2, Gr2, Gz2 = oracle.delhommeau_term(math.sqrt(2), -3.0)
    assert abs(G2 - complex(*g["greens_e1_e2_k1"])) < 1e-4
    # gradient w.r.t. the first point: k (k^2/r (x - xi) Gr, ..., k Gz) with x - xi = (-1, -1, 1)
    grad = [Gr2 / math.sqrt(2) * -1, Gr2 / math.sqrt(2) * -1, Gz2]
    for got, exp in zip(grad, g["gradient_wrt_first"]):
        assert abs(got - complex(*exp)) < 1e-4


@pytest.mark.parametrize("r,z", [(0.0, -0.3), (0.5, -0.1), (1.0, -1.0), (3.0, -0.5), (7.5, -2.0), (20.0, -4.0),
                                 (0.05, -8.0), (45.0, -20.0)])
def test_imaginary_parts_closed_form(oracle, r, z):
    """Im G = 2 pi e^z J0(r), Im dG/dr = -2 pi e^z J1(r), dG/dz = G + 2/rho for both wave terms."""
    from scipy.special import j0, j1
    ref = 2 * math.pi * math.exp(z)
    # the reference's adaptive tolerance (rtol = atol = 1e-8)
    G, Gr, Gz = oracle.delhommeau_term(r, z)
    assert abs(G.imag - ref * j0(r)) < 1e-8 * max(1, abs(G))
    assert abs(Gr.imag + ref * j1(r)) < 1e-8 * max(1, abs(Gr))
    assert abs(Gz - (G + 2 / math.hypot(r, z))) < 1e-13 * max(1, abs(Gz))
    # converged
    G, Gr, Gz = oracle.delhommeau_term(r, z, rtol=1e-14, atol=1e-16)
    assert abs(G.imag - ref * j0(r)) < 1e-12 * max(ref, 1e-3)
    assert abs(Gr.imag + ref * j1(r)) < 1e-12 * max(ref, 1e-3)
    W, Wh, Wv = oracle.wu_term(max(r, 1e-12), z)
    assert abs(W.imag - ref * j0(r)) < 1e-12 * max(1, ref)
    assert abs(Wh.imag + ref * j1(r)) < 1e-12 * max(1, ref)


def test_delhommeau_converged_against_mpmath(oracle):
    """The 1e-8 adaptive quadrature is within tolerance of the converged integral (mpmath, 20 digits)."""
    import mpmath as mp
    mp.mp.dps = 20
    for r, z in ((1.0, -1.0), (4.0, -0.4), (0.3, -2.5)):
        f = lambda t: mp.re(mp.exp(mp.mpc(z, r * mp.cos(t))) * (mp.e1(mp.mpc(z, r * mp.cos(t))) / mp.pi + 1j))
        ref = 4 * mp.quad(f, [0, mp.pi / 2])
        G, _, _ = oracle.delhommeau_term(r, z)
        assert abs(G.real - float(ref)) < 1e-8 * max(1.0, abs(float(ref)))
        Gc, _, _ = oracle.delhommeau_term(r, z, rtol=1e-13, atol=1e-14)
        assert abs(Gc.real - float(ref)) < 1e-12 * max(1.0, abs(float(ref)))


@pytest.mark.parametrize("k", [0.1, 1.0, 10.0])
def test_wu_close_to_delhommeau(oracle, k):
    """test/greens_function.jl:96-99: Wu and exact Delhommeau agree to 2e-2 on the element pair."""
    r, z = k * math.sqrt(2), k * -3.0
    G, Gr, Gz = oracle.delhommeau_term(r, z)
    W, Wh, Wv = oracle.wu_term(r, z)
    assert abs(k * G - k * W) < 2e-2
    g_del = k * np.array([k * k / r * 1 * Gr, k * k / r * 1 * Gr, k * Gz])
    g_wu = k * np.array([k * k / r * 1 * Wh, k * k / r * 1 * Wh, k * Wv])
    assert np.max(np.abs(g_del - g_wu)) < 2e-2


def _notebook_sphere(O):
    from mhx import meshes
    m = meshes.sphere(1.0, 6, 6, immersed=True)
    return m, O.Mesh.from_mhx(m)


def test_notebook_end_to_end(oracle):
    """mh_tutorial.ipynb: full solve on the notebook's 18-panel sphere, all four formulations."""
    nb = REF["notebook_sphere"]
    m, om = _notebook_sphere(oracle)
    assert m.nfaces == 18
    heave = oracle.rigid_dof(om, 2)
    r = oracle.hydro_coeffs(om, [heave], [nb["omega0"]], [0.0], rho=1000.0, g=9.81)
    assert r["added_mass"][0, 0, 0] == pytest.approx(nb["A33"], abs=1e-6)
    assert r["radiation_damping"][0, 0, 0] == pytest.approx(nb["B33"], abs=1e-6)
    assert r["diffraction_force"][0, 0, 0] == pytest.approx(complex(*nb["F_D"]), abs=1e-6)
    assert r["Froude_Krylov_force"][0, 0, 0] == pytest.approx(complex(*nb["F_FK"]), abs=1e-6)
    assert r["excitation_force"][0, 0, 0] == pytest.approx(complex(*nb["F_ex"]), abs=1e-6)
    for key, vals in nb["A_of_omega"].items():
        direct = key.startswith("direct")
        gf = oracle.GF_WU if key.endswith("Wu") else oracle.GF_DELHOMMEAU
        rr = oracle.hydro_coeffs(om, [heave], nb["omegas"], [0.0], rho=1000.0, g=9.81, gf=gf, direct=direct)
        np.testing.assert_allclose(rr["added_mass"].ravel(), vals, rtol=2e-15)
        tab = nb["table_omega0"][key]
        assert rr["radiation_damping"][0, 0, 1] == pytest.approx(tab[1], abs=6e-5)
        assert rr["excitation_force"][0, 1, 0].real == pytest.approx(tab[2], abs=6e-5)
        if key == "direct_Wu":
            np.testing.assert_allclose(rr["radiation_damping"].ravel(), nb["B_of_omega_direct_Wu"], rtol=2e-15)
            ex = np.array([complex(*c) for c in nb["Fex_of_omega_direct_Wu"]])
            np.testing.assert_allclose(rr["excitation_force"].ravel(), ex, rtol=2e-15)


def test_reference_semantics_equal_per_omega(oracle):
    m, om = _notebook_sphere(oracle)
    dofs = [oracle.rigid_dof(om, d) for d in range(6)]
    a = oracle.hydro_coeffs(om, dofs, [0.9], [0.0, 0.7], rho=1025.0, mode=0)
    b = oracle.hydro_coeffs(om, dofs, [0.9], [0.0, 0.7], rho=1025.0, mode=1)
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        np.testing.assert_allclose(a[key], b[key], rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("gf", [0, 1])
@pytest.mark.parametrize("direct", [True, False])
def test_lapack_orchestration_equals_oracle(oracle, gf, direct):
    """hydro_coeffs_lapack (the oracle's assembly, BCs and pressure integration with LAPACK for the
    dense solve, used for orders >= 1024) equals the C oracle's own hydro_coeffs to rounding."""
    from mhx import meshes
    m = meshes.sphere(1.0, 8, 16, immersed=True)
    om = oracle.Mesh.from_mhx(m)
    dofs = [oracle.rigid_dof(om, d) for d in range(6)]
    a = oracle.hydro_coeffs(om, dofs, [0.5, 1.3], [0.0, 0.7], rho=1025.0, gf=gf, direct=direct)
    b = oracle.hydro_coeffs_lapack(om, dofs, [0.5, 1.3], [0.0, 0.7], rho=1025.0, gf=gf, direct=direct)
    for key in ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force"):
        assert np.linalg.norm(a[key] - b[key]) <= 1e-13 * np.linalg.norm(a[key]), key


def test_hulme_hemisphere(oracle):
    """test/consistency_with_analytical_solutions.jl:54-79: Wu on a finer sphere, rtol 0.1."""
    from mhx import meshes
    h = REF["hulme_heave"]
    m = meshes.sphere(1.0, 10, 10, immersed=True)
    om = oracle.Mesh.from_mhx(m)
    heave = oracle.rigid_dof(om, 2)
    K = np.array(h["K"])
    w = np.sqrt(K * 9.8)
    r = oracle.hydro_coeffs(om, [heave], w, [], rho=1025.0, g=9.81)
    A = r["added_mass"].ravel() / (2 / 3 * math.pi * 1025.0)
    B = r["radiation_damping"].ravel() / (2 / 3 * math.pi * 1025.0) / w

    def isapprox(x, y, rtol):  # Julia's array isapprox is normwise
        return np.linalg.norm(x - y) <= rtol * max(np.linalg.norm(x), np.linalg.norm(y))

    assert isapprox(A, np.array(h["A"]), 0.1)
    assert isapprox(B, np.array(h["B"]), 0.1)
    assert not isapprox(A, np.array(h["A"]), 0.02)  # a 50-panel sphere is not that close


def test_assembly_invariants(oracle):
    """test/matrix_assembly.jl:10-124: S identical for direct/indirect, D != K,
    S = S_static + S_wave with a wavenumber-independent static part."""
    from mhx import meshes
    m = meshes.sphere(1.0, 8, 16, immersed=True)
    om = oracle.Mesh.from_mhx(m)
    for gf in (oracle.GF_WU, oracle.GF_DELHOMMEAU):
        Sd, D = oracle.assemble(om, 0.7, gf=gf, direct=True)
        Si, K = oracle.assemble(om, 0.7, gf=gf, direct=False)
        np.testing.assert_array_equal(Sd, Si)
        assert np.max(np.abs(D - K)) > 1e-3
        SR, DR = oracle.assemble_rankine(om, direct=True)
        SR2, _ = oracle.assemble_rankine(om, direct=True)
        np.testing.assert_array_equal(SR, SR2)
        Sw, Dw = oracle.assemble_wave(om, 0.7, gf=gf, direct=True)
        np.testing.assert_allclose(Sd, SR + Sw, rtol=0, atol=1e-15)
        np.testing.assert_allclose(D, DR + Dw, rtol=0, atol=1e-15)
    assert np.all(np.isreal(SR))


def test_lu_against_lapack(oracle):
    rng = np.random.default_rng(3)
    n = 120
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    LU, ipiv, info = oracle.zgetrf(A)
    assert info == 0
    X = oracle.zgetrs(LU, ipiv, B)
    np.testing.assert_allclose(X, np.linalg.solve(A, B), rtol=1e-11, atol=1e-11)
    XH = oracle.zgetrs(LU, ipiv, B, trans=2)
    np.testing.assert_allclose(XH, np.linalg.solve(A.conj().T, B), rtol=1e-11, atol=1e-11)
    import scipy.linalg as sl
    lu_ref, piv_ref = sl.lu_factor(A)
    np.testing.assert_array_equal(ipiv, piv_ref)  # same pivot rule as LAPACK izamax
    np.testing.assert_allclose(LU, lu_ref, rtol=1e-12, atol=1e-12)


def test_planted_mistake_is_caught(oracle):
    """A dropped image term or a wrong sign in the pressure changes A33 well beyond 1e-6."""
    nb = REF["notebook_sphere"]
    m, om = _notebook_sphere(oracle)
    heave = oracle.rigid_dof(om, 2)
    r = oracle.hydro_coeffs(om, [-heave], [nb["omega0"]], [], rho=1000.0)
    # flipping the dof sign leaves A invariant (quadratic), so the golden still holds...
    assert r["added_mass"][0, 0, 0] == pytest.approx(nb["A33"], abs=1e-6)
    # ...while flipping the normals (outward -> inward) breaks it
    m2 = meshes_flip(m)
    om2 = oracle.Mesh.from_mhx(m2)
    r2 = oracle.hydro_coeffs(om2, [oracle.rigid_dof(om2, 2)], [nb["omega0"]], [], rho=1000.0)
    assert abs(r2["added_mass"][0, 0, 0] - nb["A33"]) > 1.0


def meshes_flip(m):
    from mhx.meshes import Mesh
    return Mesh(m.vertices, m.faces[:, ::-1], m.centers, -m.normals, m.areas, m.radii)


def test_notebook_derivatives_are_radius_scaling(oracle):
    """mh_tutorial.ipynb cells 5, 9, 13: the notebook's ForwardDiff/Enzyme derivatives with respect
    to the sphere radius (mesh tangent seeded by central differences of the mesher) and to omega
    are reproduced by central differences of the oracle on the uniformly scaled sphere.  This pins
    what the derivative tests compare against: d/dr is the uniform-scaling direction."""
    from mhx import meshes
    nb = REF["notebook_sphere"]
    d = nb["derivatives_direct_Wu"]
    base = meshes.sphere(1.0, 6, 6, immersed=True)

    def coeffs(r, w, gf=oracle.GF_WU, direct=True):
        om = oracle.Mesh.from_mhx(base.scaled(r))
        res = oracle.hydro_coeffs(om, [oracle.rigid_dof(om, 2)], [w], [0.0], rho=1000.0, g=9.81, gf=gf, direct=direct)
        return res["added_mass"][0, 0, 0], res["excitation_force"][0, 0, 0].real

    h = 1e-4

    def fd(f, x):  # fourth-order central difference
        return (8 * (np.array(f(x + h)) - np.array(f(x - h))) - (np.array(f(x + 2 * h)) - np.array(f(x - 2 * h)))) / (12 * h)

    dr = fd(lambda r: coeffs(r, 1.03), 1.0)
    dw = fd(lambda w: coeffs(1.0, w), 1.03)
    assert dr[0] == pytest.approx(d["dA_dr"], abs=2e-6)
    assert dr[1] == pytest.approx(d["dReFex_dr"], abs=2e-5)
    assert dw[0] == pytest.approx(d["dA_domega"], abs=2e-6)
    assert dw[1] == pytest.approx(d["dReFex_domega"], abs=2e-6)
    for key, (gf, direct) in {"indirect_Wu": (oracle.GF_WU, False), "direct_Delhommeau": (oracle.GF_DELHOMMEAU, True),
                              "indirect_Delhommeau": (oracle.GF_DELHOMMEAU, False)}.items():
        dr = fd(lambda r: coeffs(r, 1.03, gf, direct), 1.0)
        dw = fd(lambda w: coeffs(1.0, w, gf, direct), 1.03)
        np.testing.assert_allclose([dr[0], dw[0], dr[1]], nb["derivatives_table"][key], rtol=0, atol=1e-4)


def _lid_cylinder(oracle, zlid=-0.25):
    from mhx import meshes
    hull = meshes.cylinder(1.0, 2.0, 16, 8, 4)          # R = 1, draft 2: side rows of 0.25 m
    lid = meshes.disk_lid(1.0, 4, 16, z=zlid)            # at the first vertex ring below the waterline
    full = meshes.combine(hull, lid)
    nh = hull.nfaces

    def dofs(m, pad):
        heave = np.zeros((m.nfaces, 3)); heave[:, 2] = 1.0
        heave[pad:] = 0.0
        return [heave]
    return (oracle.Mesh.from_mhx(hull), dofs(hull, hull.nfaces), oracle.Mesh.from_mhx(full), dofs(full, nh), nh)


def test_lid_removes_irregular_frequency(oracle):
    """Lid-based irregular-frequency removal (solve.jl:4-11, 22-28; test/irregular_frequency_removal.jl):
    a vertical cylinder R = 1, draft 2 has its first irregular frequency where J0(kR) = 0,
    omega^2 = g k coth(k T), k = 2.405 (omega = 4.858).  Without the lid the heave damping and
    excitation jump there (damping goes negative); with a lid they are smooth, and away from the
    irregular frequency both agree to the discretisation level."""
    oh, dh, of, df, nh = _lid_cylinder(oracle)
    w_irr = math.sqrt(9.81 * 2.405 / math.tanh(2.405 * 2.0))
    ws = np.linspace(w_irr - 0.4, w_irr + 0.4, 17)
    r0 = oracle.hydro_coeffs(oh, dh, ws, [0.0], rho=1025.0)
    r1 = oracle.hydro_coeffs(of, df, ws, [0.0], rho=1025.0, nhull=nh)
    b0, b1 = r0["radiation_damping"][0, 0], r1["radiation_damping"][0, 0]
    f0, f1 = np.abs(r0["excitation_force"][0, :, 0]), np.abs(r1["excitation_force"][0, :, 0])
    assert b0.min() < 0 and b1.min() > 0                   # the spike drives the lid-less damping negative
    jump = lambda y: np.max(np.abs(np.diff(y, 2)))           # second differences
    assert jump(b1) < 0.05 * jump(b0) and jump(f1) < 0.05 * jump(f0)
    assert np.all(np.diff(b1) < 0) and np.all(np.diff(f1) < 0)   # monotone through the irregular frequency
    ws2 = [0.8, 1.5, 2.5]
    r0 = oracle.hydro_coeffs(oh, dh, ws2, [0.0], rho=1025.0)
    r1 = oracle.hydro_coeffs(of, df, ws2, [0.0], rho=1025.0, nhull=nh)
    for key, tol in (("added_mass", 1e-4), ("radiation_damping", 3e-3), ("excitation_force", 1e-3)):
        assert np.max(np.abs(r1[key] - r0[key])) <= tol * np.max(np.abs(r0[key])), key


def test_lid_path_without_lid_panels_is_the_plain_path(oracle):
    """nhull = n (no lid panel, no panel in the free surface): the lid entry reduces to the plain path."""
    oh, dh, _, _, _ = _lid_cylinder(oracle)
    a = oracle.hydro_coeffs(oh, dh, [1.1, 2.0], [0.0, 0.5], rho=1025.0)
    b = oracle.hydro_coeffs(oh, dh, [1.1, 2.0], [0.0, 0.5], rho=1025.0, nhull=oh.nf)
    for key in ("added_mass", "radiation_damping", "excitation_force"):
        np.testing.assert_array_equal(a[key], b[key])


def test_fixed_normal_assembly_combines_to_K(oracle):
    """assemble_matrices(...; all_normals=e_c) (the forward-speed x-derivative operator): by
    linearity sum_c n_i,c K_c[i, :] = K[i, :] for the indirect K, identity jump included, for
    both wave terms -- a wrong reflection sign or identity scaling breaks it."""
    from mhx import meshes
    m = meshes.sphere(1.0, 8, 16)
    om = oracle.Mesh.from_mhx(m)
    for gf in (oracle.GF_WU, oracle.GF_DELHOMMEAU):
        _, K = oracle.assemble(om, 0.7, gf=gf, direct=False)
        Ks = [oracle.assemble_fixed_normal(om, 0.7, e, gf=gf)[1] for e in np.eye(3)]
        comb = sum(m.normals[:, c][:, None] * Ks[c] for c in range(3))
        assert np.max(np.abs(comb - K)) <= 1e-14 * np.max(np.abs(K))


def test_encountered_values(oracle):
    """compute_encountered_values (waves.jl:136-151): Doppler shift omega - k U cos(beta), its
    modulus, and the heading turned by pi when it is negative."""
    w, beta, U, g = 2.0, 0.3, 6.0, 9.81
    k = w * w / g
    dop = w - k * U * math.cos(beta)
    we, ke, be = oracle.encountered(w, beta, U, g)
    assert dop < 0 and we == pytest.approx(-dop, rel=1e-15) and ke == pytest.approx(dop * dop / g, rel=1e-14)
    assert be == pytest.approx(beta + math.pi)
    we, ke, be = oracle.encountered(1.0, 2.0, 0.5, g)
    assert we == pytest.approx(1.0 - 0.5 / g * math.cos(2.0)) and be == 2.0
    assert oracle.encountered(1.3, 0.4, 0.0, g) == (1.3, 1.3 ** 2 / g, 0.4)


def test_forward_speed_at_rest_is_the_indirect_path(oracle):
    """forward_speeds = [0]: every heading's radiation problems and the diffraction problem equal
    the plain IndirectBEM results bit for bit."""
    from mhx import meshes
    m = meshes.sphere(1.0, 8, 16)
    om = oracle.Mesh.from_mhx(m)
    heave = oracle.rigid_dof(om, 2)
    pitch = oracle.rigid_dof(om, 4)
    r0 = oracle.hydro_coeffs(om, [heave, pitch], [0.8, 1.5], [0.0, 1.0], rho=1025.0, direct=False)
    r1 = oracle.hydro_coeffs_fs(om, [heave, pitch], [oracle.gradient_of_motion("Heave"),
                                                      oracle.gradient_of_motion("Pitch")],
                                [0.8, 1.5], [0.0, 1.0], [0.0], rho=1025.0)
    for ib in range(2):
        np.testing.assert_array_equal(r1["added_mass"][:, :, :, 0, ib], r0["added_mass"])
        np.testing.assert_array_equal(r1["radiation_damping"][:, :, :, 0, ib], r0["radiation_damping"])
    np.testing.assert_array_equal(r1["excitation_force"][:, :, :, 0], r0["excitation_force"])


# ---------------------------------------------------------------------------- paper: two-sphere array
PAPER = json.load(open(os.path.join(HERE, "golden", "paper_array_sensitivities.json")))


def _paper_pair(r, dx):
    """paper/meshGradients_pair.jl:9-15: mesh_sphere(r, resolution=(6,6)).immersed_part() and the
    same moved by r + dx along x, joined."""
    from mhx import meshes
    s = meshes.mesh_sphere(r, resolution=(6, 6)).immersed_part()
    return s + s.translated_x(r + dx), s.nfaces


def _paper_coeffs(O, r, dx, w):
    """A, B of the heave of each sphere on its own panels (paper/ArraySensitivities.jl:16-50):
    GFWu, direct, rho = 1000, k = omega^2 / 9.81."""
    m, n1 = _paper_pair(r, dx)
    d1 = np.zeros((m.nfaces, 3)); d1[:n1, 2] = 1.0
    d2 = np.zeros((m.nfaces, 3)); d2[n1:, 2] = 1.0
    res = O.hydro_coeffs(O.Mesh.from_mhx(m), [d1, d2], [w], [], rho=1000.0, g=9.81, gf=O.GF_WU, direct=True)
    return np.array([res["added_mass"][:, :, 0], res["radiation_damping"][:, :, 0]])


def test_paper_array_sensitivities(oracle):
    """The reference's paper data (paper/Plots/11_heuristics_dx_DELhommeau.csv,
    11_heuristics_dr_DELhommeau.csv, added_mass_/damping_data_dimensionless.csv; 336 numbers
    computed by the reference with Zygote through its AD rules): the oracle's coefficients of the
    two-sphere array, differentiated by a fourth-order central difference (h = 1e-4, inside the
    far-field rule's branch), agree to 1e-6 (relative to max(|value|, 1)).  This pins the oracle on
    coupled bodies at separations 1.5-8 radii and kr up to 9.5."""
    h = 1e-4

    def d4(f):
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    def check(got, want):
        assert abs(got - want) <= 1e-6 * max(abs(want), 1.0), (got, want)

    g = PAPER["dA11_ddx"]
    for dx, a, b in zip(g["dx"], g["dA11"], g["dB11"]):
        d = d4(lambda t: _paper_coeffs(oracle, 1.0, dx + t, g["omega"]))
        check(d[0, 0, 0], a)
        check(d[1, 0, 0], b)
    g = PAPER["dA11_dr"]
    for dx, a, b in zip(g["dx"], g["dA11"], g["dB11"]):
        d = d4(lambda t: _paper_coeffs(oracle, 1.0 + t, dx, g["omega"]))
        check(d[0, 0, 0], a)
        check(d[1, 0, 0], b)
    g = PAPER["dA21_dr"]
    for dxr, kr, a, b in zip(g["dx_r_ratio"], g["kr"], g["dA21"], g["dB21"]):
        w = math.sqrt(kr * 9.8)  # the script's g = 9.8 for omega (paper/ArraySensitivities.jl:54, 82)
        d = d4(lambda t: _paper_coeffs(oracle, 1.0 + t, dxr, w))
        check(d[0, 1, 0], a)
        check(d[1, 1, 0], b)
