"""mhx — MI355X (gfx950) implementation of MarineHydro.jl's boundary-element hot path.

The computation lives in libmhx.so (HIP kernels behind the C ABI of
include/mhx.h); this package is the host-side mirror of the reference's Julia
interface plus the mesh generators used by the workloads.
"""
from . import meshes  # noqa: F401
from ._lib import DimensionMismatch, MHXError, SingularException  # noqa: F401
from .api import (SETTINGS, DiffractionProblem, DirectBEM, ExactGuevelDelhommeau, FloatingBody, GFWu,  # noqa: F401
                  IndirectBEM, RadiationProblem, Rankine, RankineReflected, added_mass, assemble_matrices,
                  compute_wavenumber, context, default_greens_functions, diffraction_force, excitation_force,
                  froude_krylov_force, hydrodynamic_coefficients, is_direct, lu_solve, lu_solve_batched,
                  radiation_damping,
                  rotational_dofs, set_g, set_rho, solve, solve_linear, translational_dofs,
                  SurrogateGuevelDelhommeau, train_surrogate, load_surrogate,
                  hydrodynamic_coefficients_jvp, hydrodynamic_coefficients_vjp, mesh_scaling_tangent,
                  chain_mesh_gradient, hydrodynamic_coefficients_many,
                  combine_floatingbodies, make_body_name_list_unique, compute_encountered_values,
                  evaluate_gradient_of_motion)
from .meshes import Mesh  # noqa: F401
