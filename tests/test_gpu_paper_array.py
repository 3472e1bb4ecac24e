"""Derivatives of a coupled two-sphere array against the reference's paper data (paper/Plots,
336 numbers computed by the reference with Zygote through its AD rules; setup of
paper/ArraySensitivities.jl and paper/meshGradients_pair.jl): the forward mode
(mhx_hydro_coeffs_jvp) along the mesh tangents of a separation change and of a radius change,
and the reverse mode (mhx_hydro_coeffs_vjp) chained with the same tangents."""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PAPER = json.load(open(os.path.join(HERE, "golden", "paper_array_sensitivities.json")))
TOL = 1e-8  # relative to max(|value|, 1)


@pytest.fixture(scope="module")
def mhx():
    import mhx as M
    print("device context", flush=True)
    M.context()
    M.set_rho(1000.0)
    M.set_g(9.81)
    return M


def _pair(mhx, dx):
    """Sphere 1 = mesh_sphere(1, resolution=(6,6)).immersed_part(), sphere 2 = the same moved by
    1 + dx along x; each sphere's heave on its own panels.  Tangents of the mesh fields for
    d/d(dx) (sphere 2 moves along x) and d/dr at r = 1 (both spheres scale about their centres'
    origin, sphere 2's shift r + dx grows with r): vertices, centres (x r for r), areas (2A/r)."""
    from mhx import meshes
    s = meshes.mesh_sphere(1.0, resolution=(6, 6)).immersed_part()
    m = s + s.translated_x(1.0 + dx)
    n1, v1 = s.nfaces, s.nvertices
    d1 = np.zeros((m.nfaces, 3)); d1[:n1, 2] = 1.0
    d2 = np.zeros((m.nfaces, 3)); d2[n1:, 2] = 1.0
    body = mhx.FloatingBody(m, {"h1": d1, "h2": d2})
    ex = np.array([1.0, 0.0, 0.0])
    t_dx = {"vertices": np.zeros_like(m.vertices), "centers": np.zeros_like(m.centers)}
    t_dx["vertices"][v1:] = ex
    t_dx["centers"][n1:] = ex
    t_dr = {"vertices": np.concatenate([s.vertices, s.vertices + ex]),
            "centers": np.concatenate([s.centers, s.centers + ex]), "areas": 2.0 * m.areas}
    return body, t_dx, t_dr


def _par(w):
    return {"wave_frequencies": [w], "radiating_dofs": ["h1", "h2"], "influenced_dofs": ["h1", "h2"]}


def _check(got, want):
    assert abs(got - want) <= TOL * max(abs(want), 1.0), (got, want)


def test_forward_mode_separation_and_radius(mhx):
    """d A11/d(dx), d B11/d(dx), d A11/dr, d B11/dr at omega = 1.03 for separations 1.5-8
    (56 numbers) and d A21/dr, d B21/dr over x/r = 1.5-8 and kr = 0.5-9.5 (280 numbers)."""
    print("paper array: forward mode", flush=True)
    form = mhx.DirectBEM()
    worst = 0.0
    for key, tkey in (("dA11_ddx", 0), ("dA11_dr", 1)):
        g = PAPER[key]
        for dx, a, b in zip(g["dx"], g["dA11"], g["dB11"]):
            body, t_dx, t_dr = _pair(mhx, dx)
            _, tan = mhx.hydrodynamic_coefficients_jvp(body, _par(g["omega"]), (t_dx, t_dr)[tkey], form)
            for got, want in ((tan["added_mass"][0, 0, 0, 0], a), (tan["radiation_damping"][0, 0, 0, 0], b)):
                _check(got, want)
                worst = max(worst, abs(got - want) / max(abs(want), 1.0))
    g = PAPER["dA21_dr"]
    for dxr, kr, a, b in zip(g["dx_r_ratio"], g["kr"], g["dA21"], g["dB21"]):
        body, _, t_dr = _pair(mhx, dxr)
        _, tan = mhx.hydrodynamic_coefficients_jvp(body, _par(math.sqrt(kr * 9.8)), t_dr, form)
        for got, want in ((tan["added_mass"][1, 0, 0, 0], a), (tan["radiation_damping"][1, 0, 0, 0], b)):
            _check(got, want)
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
    print(f"  worst relative difference {worst:.3e}", flush=True)


def test_reverse_mode_radius(mhx):
    """The reverse mode chained with the radius tangent: <grad A11, t_dr> and <grad B11, t_dr> at
    omega = 1.03 for every separation, and <grad A21, t_dr> on a spread of the (x/r, kr) grid."""
    print("paper array: reverse mode", flush=True)
    form = mhx.DirectBEM()

    def along(grads, t):
        return sum(float(np.sum(grads[k] * t[k])) for k in t)

    g = PAPER["dA11_dr"]
    for dx, a, b in zip(g["dx"], g["dA11"], g["dB11"]):
        body, _, t_dr = _pair(mhx, dx)
        for key, want in (("added_mass", a), ("radiation_damping", b)):
            cot = np.zeros((2, 2, 1, 1)); cot[0, 0] = 1.0
            _, grads = mhx.hydrodynamic_coefficients_vjp(body, _par(g["omega"]), {key: cot}, form)
            _check(along(grads, t_dr), want)
    g = PAPER["dA21_dr"]
    for q in range(0, len(g["kr"]), 7):
        body, _, t_dr = _pair(mhx, g["dx_r_ratio"][q])
        for key, want in (("added_mass", g["dA21"][q]), ("radiation_damping", g["dB21"][q])):
            cot = np.zeros((2, 2, 1, 1)); cot[1, 0] = 1.0
            _, grads = mhx.hydrodynamic_coefficients_vjp(body, _par(math.sqrt(g["kr"][q] * 9.8)), {key: cot}, form)
            _check(along(grads, t_dr), want)
