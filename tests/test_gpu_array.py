"""Coupled arrays of bodies (combine_floatingbodies, src/bodies.jl:165-225): one solve on the
combined mesh with every body's dofs, against the oracle."""
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


def _array(mhx):
    from mhx import meshes
    s = meshes.sphere(1.0, 8, 16)
    b1 = mhx.FloatingBody(meshes.translated(s, (-2.0, 0.0, 0.0)), ["Heave", "Surge"], body_name="wec")
    b2 = mhx.FloatingBody(meshes.translated(s, (2.0, 0.0, 0.0)), ["Heave", "Surge"], body_name="wec")
    return b1 + b2


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
def test_array_matches_oracle(mhx, oracle, gf, direct):
    """Two 64-panel spheres 4 m apart, heave and surge each (four dofs), two frequencies, headings
    0 and pi/2: all couplings within 1e-9 of the oracle on the combined mesh."""
    print(f"array parity {gf} direct={direct}", flush=True)
    arr = _array(mhx)
    assert arr.dof_names == ["wec__Heave", "wec__Surge", "wec_2__Heave", "wec_2__Surge"]
    ws, betas = [0.7, 1.6], [0.0, math.pi / 2]
    par = {"wave_frequencies": ws, "radiating_dofs": arr.dof_names, "wave_directions": betas}
    form = mhx.DirectBEM(gf) if direct else mhx.IndirectBEM(gf)
    got = mhx.hydrodynamic_coefficients(arr, par, form)
    ref = oracle.hydro_coeffs(oracle.Mesh.from_mhx(arr.mesh), [arr.dofs[n] for n in arr.dof_names], ws, betas,
                              rho=1025.0, g=9.81, gf=oracle.GF_WU if gf == "Wu" else oracle.GF_DELHOMMEAU,
                              direct=direct)
    for key in KEYS:
        rel, elem, ok = close(np.asarray(got[key]).reshape(np.asarray(ref[key]).shape), ref[key])
        assert ok, (key, rel, elem)


def test_mirror_symmetric_array(mhx):
    """Identical bodies at x = -2 and x = +2 (the meshes are mirror images through x = 0): each
    body's own heave added mass and damping are equal, the heave-heave couplings are symmetric,
    and under beam seas (beta = pi/2) both bodies feel the same heave excitation."""
    print("array symmetry", flush=True)
    arr = _array(mhx)
    par = {"wave_frequencies": [1.1], "radiating_dofs": arr.dof_names, "wave_directions": [math.pi / 2]}
    r = mhx.hydrodynamic_coefficients(arr, par, mhx.DirectBEM())
    A, B = r["added_mass"][:, :, 0, 0], r["radiation_damping"][:, :, 0, 0]
    for M in (A, B):
        assert M[0, 0] == pytest.approx(M[2, 2], rel=1e-9)
        assert M[0, 2] == pytest.approx(M[2, 0], rel=1e-6)   # reciprocity, to the discretisation
    fe = r["excitation_force"][:, 0, 0, 0]
    assert fe[0] == pytest.approx(fe[2], rel=1e-9)
