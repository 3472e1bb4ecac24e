"""Host-side forward-speed logic: encountered values (waves.jl:136-151), the gradient of motion of
rigid dofs (bodies.jl:55-66), the problems' encountered fields (problems_and_results.jl:40-53,
118-133) and the argument errors raised before any device work (solve.jl:35-40)."""
import math

import numpy as np
import pytest

import mhx
from mhx import meshes


@pytest.mark.parametrize("w,beta,U", [(1.0, 0.3, 0.0), (0.5, 2.0, 0.2), (1.7, 0.0, 8.0), (1.7, math.pi / 6, 8.0),
                                      (2.2, math.pi / 2, 3.0), (0.9, math.pi, -1.5)])
def test_encountered_values_match_the_oracle(oracle, w, beta, U):
    got = mhx.compute_encountered_values(w, beta, U, 9.81)
    ref = oracle.encountered(w, beta, U, 9.81)
    np.testing.assert_allclose(got, ref, rtol=1e-15, atol=0)


def test_encountered_values_relations():
    """omega_e = |omega - k U cos beta| and k_e = omega_e^2 / g; heading pi/2 leaves omega unchanged;
    a negative Doppler shift turns the heading by pi; U = 0 returns the incident values."""
    g = 9.81
    w, b, U = 1.7, 0.0, 8.0
    we, ke, be = mhx.compute_encountered_values(w, b, U, g)
    assert we == pytest.approx(abs(w - w * w / g * U)) and ke == pytest.approx(we * we / g) and be == math.pi
    we, ke, be = mhx.compute_encountered_values(1.3, math.pi / 2, 5.0, g)
    assert we == pytest.approx(1.3, abs=1e-15) and be == math.pi / 2
    assert mhx.compute_encountered_values(1.3, 0.4, 0.0, g) == (1.3, 1.3 * 1.3 / g, 0.4)


def test_gradient_of_motion(oracle):
    m = meshes.sphere(1.0, 4, 8)
    for name in mhx.api.RIGID_DOFS:
        gm = mhx.evaluate_gradient_of_motion(m, name)
        assert gm.shape == (m.nfaces, 3)
        np.testing.assert_array_equal(gm, np.tile(oracle.gradient_of_motion(name), (m.nfaces, 1)))
    with pytest.raises(ValueError, match="rigid body dofs"):
        mhx.evaluate_gradient_of_motion(m, "Flex")


def test_problem_encountered_fields():
    body = mhx.FloatingBody(meshes.sphere(1.0, 4, 8), ["Heave", "Pitch"])
    p = mhx.RadiationProblem(body, 1.7, "Pitch", beta=0.0, forward_speed=8.0, g=9.81)
    assert (p.encountered_omega, p.encountered_wavenumber, p.encountered_beta) == \
        mhx.compute_encountered_values(1.7, 0.0, 8.0, 9.81)
    d = mhx.DiffractionProblem(body, 1.0, 0.5, forward_speed=0.0)
    assert (d.encountered_omega, d.encountered_beta) == (1.0, 0.5)
    with pytest.raises(ValueError, match="wave direction"):
        mhx.RadiationProblem(body, 1.0, "Heave", forward_speed=0.3)


def test_forward_speed_argument_errors():
    """Raised on the host before any device call, with the reference's messages."""
    m = meshes.sphere(1.0, 4, 8)
    body = mhx.FloatingBody(m, ["Heave", "Pitch"])
    par = {"wave_frequencies": [1.0], "radiating_dofs": ["Heave"], "wave_directions": [0.0], "forward_speeds": [0.5]}
    with pytest.raises(ValueError, match="not supported with the direct method"):
        mhx.hydrodynamic_coefficients(body, par, mhx.DirectBEM())
    nobeta = {"wave_frequencies": [1.0], "radiating_dofs": ["Heave"], "forward_speeds": [0.5]}
    with pytest.raises(ValueError, match="wave_directions"):
        mhx.hydrodynamic_coefficients(body, nobeta, mhx.IndirectBEM())
    flex = mhx.FloatingBody(m, {"flex": np.ones((m.nfaces, 3))})
    par_flex = dict(par, radiating_dofs=["flex"])
    with pytest.raises(ValueError, match="rigid body dofs"):
        mhx.hydrodynamic_coefficients(flex, par_flex, mhx.IndirectBEM())
