"""Parity of libmhx (HIP, gfx950) against the CPU oracle, through the C ABI.

Bar: plain FP64 path within 1e-9 of the oracle, normwise per matrix/coefficient
block and elementwise with a floor of 1e-9 * max|X|; LU results within
rounding of LAPACK (numpy) on the same inputs.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))
TOL = 1e-9


def close(got, ref, tol=TOL, abs_floor=1e-13):
    """Normwise relative error <= tol, and elementwise |d| <= 10 tol |ref| + abs_floor max|ref|
    (entries that are zero up to rounding, e.g. the cross couplings of axisymmetric bodies,
    are compared at the block's rounding scale)."""
    got = np.asarray(got); ref = np.asarray(ref)
    nrm = np.linalg.norm(ref)
    rel = np.linalg.norm(got - ref) / nrm if nrm > 0 else np.linalg.norm(got - ref)
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    d = np.abs(got - ref)
    elem_ok = bool(np.all(d <= 10 * tol * np.abs(ref) + abs_floor * scale)) if ref.size else True
    worst = float(np.max(d / np.maximum(np.abs(ref), abs_floor * scale + 1e-300))) if ref.size else 0.0
    return rel, worst, rel <= tol and elem_ok


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    return M


def sphere18():
    from mhx import meshes
    return meshes.sphere(1.0, 6, 6, immersed=True)


def sphere64():
    from mhx import meshes
    return meshes.sphere(1.0, 8, 16, immersed=True)


@pytest.mark.parametrize("n", [1, 4, 5, 63, 64, 65, 130, 300])
@pytest.mark.parametrize("trans", [0, 2])
def test_lu_solve_against_lapack(mhx, n, trans):
    print(f"lu n={n} trans={trans}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    X = mhx.lu_solve(A, B, trans=trans)
    Aop = A if trans == 0 else A.conj().T
    ref = np.linalg.solve(Aop, B)
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 1e-11


def test_lu_large_bem_like(mhx):
    """Diagonally weighted n=1000 system (BEM-like D) and a random one needing pivoting."""
    print("lu n=1000", flush=True)
    rng = np.random.default_rng(11)
    n = 1000
    A = 0.5 * np.eye(n) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n
    B = rng.standard_normal((n, 7)) + 1j * rng.standard_normal((n, 7))
    X = mhx.lu_solve(A, B)
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) < 1e-12
    A2 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    X2 = mhx.lu_solve(A2, B)
    ref = np.linalg.solve(A2, B)
    assert np.linalg.norm(X2 - ref) / np.linalg.norm(ref) < 1e-10


def test_lu_singular_raises(mhx):
    A = np.zeros((8, 8), dtype=np.complex128)
    with pytest.raises(mhx.SingularException):
        mhx.lu_solve(A, np.ones(8))


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("meshname", ["sphere18", "sphere64"])
def test_assembly_parity(mhx, oracle, gf, direct, meshname):
    print(f"assembly {meshname} {gf} direct={direct}", flush=True)
    m = sphere18() if meshname == "sphere18" else sphere64()
    om = oracle.Mesh.from_mhx(m)
    for k in (0.05, 0.108, 1.3):
        S, D = mhx.assemble_matrices(mhx.default_greens_functions(gf), m, k, direct=direct)
        So, Do = oracle.assemble(om, k, gf=oracle.GF_WU if gf == "Wu" else oracle.GF_DELHOMMEAU, direct=direct)
        for got, ref in ((S, So), (D, Do)):
            rel, elem, ok = close(got, ref)
            assert ok, (k, rel, elem)


def test_notebook_coefficients(mhx):
    """The reference notebook's printed full-precision tables (mh_tutorial.ipynb cells 11, 17)."""
    nb = REF["notebook_sphere"]
    mhx.set_rho(1000.0); mhx.set_g(9.81)
    body = mhx.FloatingBody(sphere18(), ["Heave"])
    for key, vals in nb["A_of_omega"].items():
        print("notebook", key, flush=True)
        gf = "Wu" if key.endswith("Wu") else "ExactGuevelDelhommeau"
        form = mhx.DirectBEM(gf) if key.startswith("direct") else mhx.IndirectBEM(gf)
        out = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=nb["omegas"], radiating_dofs=["Heave"],
                                                       wave_directions=[0.0]), form)
        np.testing.assert_allclose(out["added_mass"].ravel(), vals, rtol=1e-10)
        if key == "direct_Wu":
            np.testing.assert_allclose(out["radiation_damping"].ravel(), nb["B_of_omega_direct_Wu"], rtol=1e-10)
            ex = np.array([complex(*c) for c in nb["Fex_of_omega_direct_Wu"]])
            np.testing.assert_allclose(out["excitation_force"].ravel(), ex, rtol=1e-10)
    # problem-level API
    res = mhx.solve(mhx.RadiationProblem(body, 1.03), mhx.DirectBEM())
    assert mhx.added_mass(res)["Heave"] == pytest.approx(nb["A33"], abs=1e-6)
    assert mhx.radiation_damping(res)["Heave"] == pytest.approx(nb["B33"], abs=1e-6)
    dres = mhx.solve(mhx.DiffractionProblem(body, 1.03, beta=0.0), mhx.DirectBEM())
    assert mhx.diffraction_force(dres)["Heave"] == pytest.approx(complex(*nb["F_D"]), abs=1e-6)
    assert mhx.froude_krylov_force(dres)["Heave"] == pytest.approx(complex(*nb["F_FK"]), abs=1e-6)
    assert mhx.excitation_force(dres)["Heave"] == pytest.approx(complex(*nb["F_ex"]), abs=1e-6)


def _coeff_parity(mhx, oracle, mesh, omegas, betas, gf, direct, rho=1025.0, mode="per_omega"):
    mhx.set_rho(rho); mhx.set_g(9.81)
    names = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
    body = mhx.FloatingBody(mesh, names)
    form = mhx.DirectBEM(gf) if direct else mhx.IndirectBEM(gf)
    out = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=omegas, radiating_dofs=names,
                                                   wave_directions=betas), form, schedule=mode)
    om = oracle.Mesh.from_mhx(mesh)
    ref = oracle.hydro_coeffs(om, [body.dofs[nm] for nm in names], omegas, betas, rho=rho, g=9.81,
                              gf=oracle.GF_WU if gf == "Wu" else oracle.GF_DELHOMMEAU, direct=direct)
    for key in ("added_mass", "radiation_damping", "excitation_force", "Froude_Krylov_force"):
        got = out[key][..., 0]
        rel, elem, ok = close(got, ref[key])
        assert ok, (key, rel, elem)
    return out


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
def test_config1_sphere64(mhx, oracle, gf, direct):
    """64-panel half sphere (8 x 16 UV sphere, z < 0), rho=1025: radiation + diffraction."""
    print(f"config1 {gf} direct={direct}", flush=True)
    _coeff_parity(mhx, oracle, sphere64(), [0.5, 1.03, 2.2], [0.0, math.pi / 3], gf, direct)


def test_per_problem_schedule_matches(mhx, oracle):
    print("per-problem schedule", flush=True)
    _coeff_parity(mhx, oracle, sphere64(), [1.03], [0.0], "Wu", True, mode="per_problem")


def test_config2_cylinder_1024(mhx, oracle):
    """1,024-panel truncated cylinder (R=5, draft 10), 6 dofs, exact Delhommeau, direct and indirect,
    at three of the 32 sweep frequencies (the full sweep runs in the bench)."""
    from mhx import meshes
    m = meshes.cylinder(5.0, 10.0, 32, 24, 8)
    assert m.nfaces == 1024
    oracle.set_threads(min(16, os.cpu_count() or 1))
    ws = np.linspace(0.2, 2.0, 32)[[0, 13, 31]]
    for direct in (True, False):
        print(f"config2 direct={direct}", flush=True)
        _coeff_parity(mhx, oracle, m, ws, [0.0], "ExactGuevelDelhommeau", direct)


def test_argument_errors(mhx):
    body = mhx.FloatingBody(sphere18(), ["Heave"])
    with pytest.raises(ValueError):
        mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=[-1.0], radiating_dofs=["Heave"]))
    # more headings than one solve holds (1 dof + 16 headings): split into heading chunks, each
    # heading's forces equal those of a call with that heading alone
    betas = list(np.linspace(0, 3, 16))
    many = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=[1.0], radiating_dofs=["Heave"],
                                                    wave_directions=betas))
    for q in (0, 14, 15):
        one = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=[1.0], radiating_dofs=["Heave"],
                                                       wave_directions=[betas[q]]))
        np.testing.assert_allclose(many["diffraction_force"][:, :, q], one["diffraction_force"][:, :, 0], rtol=1e-12)
    # 16 dofs leave no room for a heading column: the solver's limit is reported
    m = sphere18()
    wide = mhx.FloatingBody(m, {f"d{i}": np.tile(np.eye(3)[i % 3] * (1 + i), (m.nfaces, 1)) for i in range(16)})
    with pytest.raises(mhx.DimensionMismatch):
        mhx.hydrodynamic_coefficients(wide, dict(wave_frequencies=[1.0], radiating_dofs=wide.dof_names,
                                                 wave_directions=[0.0]))


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
def test_batched_schedule_matches_oracle(mhx, oracle, gf):
    """All frequencies through the batched LU in chunks (3 items per chunk, uneven last chunk)."""
    print(f"batched {gf}", flush=True)
    mhx.context().set_batch(3)
    try:
        _coeff_parity(mhx, oracle, sphere64(), [0.3, 0.6, 0.9, 1.2, 1.5, 1.8, 2.1], [0.0, 1.0], gf, True,
                      mode="batched")
        _coeff_parity(mhx, oracle, sphere64(), [0.45, 1.7], [0.5], gf, False, mode="batched")
    finally:
        mhx.context().set_batch(64)


@pytest.mark.parametrize("zc", [0.0, 0.37])
@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
def test_clipped_horizontal_cylinder(mhx, oracle, zc, gf):
    """The reference's horizontal cylinder (mesh_horizontal_cylinder(radius=1.5, length=2.5,
    faces_max_radius=0.5).keep_immersed_part(), test/consistency_with_Capytaine.jl:211-216), and the
    same body raised by 0.37 m so that the waterline cuts faces (quadrangles split into quadrangles
    and triangles): coefficients against the oracle on the same panels, 1e-9."""
    from mhx import meshes
    m = meshes.mesh_horizontal_cylinder(length=2.5, radius=1.5, center=(0.0, 0.0, zc),
                                        faces_max_radius=0.5).immersed_part()
    if zc:
        assert np.any(m.faces[:, 0] == m.faces[:, 3])   # cut faces became triangles
    print(f"clipped horizontal cylinder zc={zc} ({m.nfaces} panels) {gf}", flush=True)
    _coeff_parity(mhx, oracle, m, [0.5, 1.0, 2.0], [0.0, math.pi / 4], gf, True)
