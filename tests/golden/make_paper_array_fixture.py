"""Writes paper_array_sensitivities.json from the data files of the reference's paper
(paper/Plots/*.csv of MarineHydro.jl, read as CSV text; nothing of the reference is run).

The numbers are derivatives of the two-sphere array of paper/ArraySensitivities.jl and
paper/meshGradients_pair.jl: sphere 1 = Capytaine mesh_sphere(radius=r, center=(0,0,0),
resolution=(6,6)).immersed_part(), sphere 2 = the same translated by r + dx along x, heave of
each sphere restricted to its own panels, omega fixed, rho = 1000, k = omega^2 / 9.81,
GFWu, direct method (the files' "DELhommeau" suffix notwithstanding: the GFWu numbers match,
the exact-term ones do not):
  11_heuristics_dx_DELhommeau.csv   d/d(dx) of A11, B11 at omega = 1.03, r = 1
  11_heuristics_dr_DELhommeau.csv   d/dr of A11, B11 at omega = 1.03, r = 1 (both radii and the
                                    shift r + dx move with r)
  added_mass_data_dimensionless.csv d/dr of A21 (= A12) at r = 1, dx = dx_r_ratio (held fixed),
  damping_data_dimensionless.csv    omega = sqrt(kr * 9.8 / r) (g = 9.8 in that script)
heuristics_dx*.csv are left out: no setup tried reproduces them.

Run from the repository root:  python tests/golden/make_paper_array_fixture.py
"""
import csv
import json
import os

SRC = "/root/reference/paper/Plots"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "paper_array_sensitivities.json")


def read(name):
    with open(os.path.join(SRC, name), newline="") as f:
        return [{k: float(v) for k, v in row.items()} for row in csv.DictReader(f)]


def main():
    dx = read("11_heuristics_dx_DELhommeau.csv")
    dr = read("11_heuristics_dr_DELhommeau.csv")
    am = read("added_mass_data_dimensionless.csv")
    bd = read("damping_data_dimensionless.csv")
    assert [(a["dx_r_ratio"], a["kr"]) for a in am] == [(b["dx_r_ratio"], b["kr"]) for b in bd]
    out = {
        "_source": "MarineHydro.jl paper/Plots: 11_heuristics_dx_DELhommeau.csv, 11_heuristics_dr_DELhommeau.csv, "
                   "added_mass_data_dimensionless.csv, damping_data_dimensionless.csv (generated by "
                   "paper/ArraySensitivities.jl with the meshes of paper/meshGradients_pair.jl)",
        "setup": {"sphere": "mesh_sphere(radius=r, resolution=(6,6)).immersed_part(), 18 panels",
                  "second_sphere_shift_x": "r + dx", "rho": 1000.0, "g_wavenumber": 9.81,
                  "g_omega_dimensionless": 9.8, "green_function": "GFWu", "formulation": "direct"},
        "dA11_ddx": {"omega": 1.03, "r": 1.0, "dx": [row["dx"] for row in dx],
                     "dA11": [row["A11_grad_r"] for row in dx], "dB11": [row["B11_grad_r"] for row in dx]},
        "dA11_dr": {"omega": 1.03, "r": 1.0, "dx": [row["dx"] for row in dr],
                    "dA11": [row["A11_grad_r"] for row in dr], "dB11": [row["B11_grad_r"] for row in dr]},
        "dA21_dr": {"r": 1.0, "dx_r_ratio": [row["dx_r_ratio"] for row in am], "kr": [row["kr"] for row in am],
                    "dA21": [row["grad_r"] for row in am], "dB21": [row["grad_r"] for row in bd]},
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
