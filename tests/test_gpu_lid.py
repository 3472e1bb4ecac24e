"""Bodies with a lid (irregular-frequency removal, solve.jl:4-11, 22-28): the device path against
the oracle's lid path, in every schedule."""
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


def _body(mhx, zlid):
    from mhx import meshes
    hull = meshes.cylinder(1.0, 2.0, 16, 8, 4)
    lid = meshes.disk_lid(1.0, 4, 16, z=zlid)
    return mhx.FloatingBody(hull, ["Heave", "Surge", "Pitch"], lid_mesh=lid), hull, lid


@pytest.mark.parametrize("gf", ["Wu", "ExactGuevelDelhommeau"])
@pytest.mark.parametrize("direct", [True, False])
def test_lid_body_matches_oracle(mhx, oracle, gf, direct):
    """Cylinder R = 1, draft 2 (192 panels) with a 64-panel lid at z = -0.25, three dofs, two
    headings, frequencies below, at and above the first irregular frequency (4.858): every
    coefficient within 1e-9 of the oracle's lid path."""
    print(f"lid parity {gf} direct={direct}", flush=True)
    from mhx import meshes
    body, hull, lid = _body(mhx, -0.25)
    ws = [1.2, 4.858, 5.3]
    par = {"wave_frequencies": ws, "radiating_dofs": ["Heave", "Surge", "Pitch"],
           "wave_directions": [0.0, math.pi / 3]}
    form = mhx.DirectBEM(gf) if direct else mhx.IndirectBEM(gf)
    got = mhx.hydrodynamic_coefficients(body, par, form)
    full = meshes.combine(hull, lid)
    nh = hull.nfaces
    dofs = [np.concatenate([body.dofs[nm], np.zeros((lid.nfaces, 3))]) for nm in par["radiating_dofs"]]
    ref = oracle.hydro_coeffs(oracle.Mesh.from_mhx(full), dofs, ws, par["wave_directions"], rho=1025.0, g=9.81,
                              gf=oracle.GF_WU if gf == "Wu" else oracle.GF_DELHOMMEAU, direct=direct, nhull=nh)
    for key in KEYS:
        rel, elem, ok = close(np.asarray(got[key]).reshape(np.asarray(ref[key]).shape), ref[key])
        assert ok, (key, rel, elem)


def test_lid_body_schedules_agree(mhx):
    """The same lid body through the batched schedule and the multi-body sweep (with a lid-less
    body beside it) equals the per-frequency path."""
    print("lid schedules", flush=True)
    from mhx import meshes
    body, _, _ = _body(mhx, -0.25)
    other = mhx.FloatingBody(meshes.sphere(1.0, 8, 16), ["Heave", "Surge", "Pitch"])
    par = {"wave_frequencies": [0.9, 2.1, 4.9], "radiating_dofs": ["Heave", "Surge", "Pitch"],
           "wave_directions": [0.0]}
    form = mhx.DirectBEM()
    ref = mhx.hydrodynamic_coefficients(body, par, form)
    bat = mhx.hydrodynamic_coefficients(body, par, form, schedule="batched")
    many = mhx.hydrodynamic_coefficients_many([(other, par), (body, par)], form)
    for got in (bat, many[1]):
        for key in KEYS:
            rel, elem, ok = close(got[key], ref[key], tol=1e-11)
            assert ok, (key, rel, elem)
