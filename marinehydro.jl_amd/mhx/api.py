"""Host-side mirror of MarineHydro.jl's user interface for the BEM hot path.

Names, argument meaning and error behaviour follow the reference so that a
caller (or a test ported from the reference's test suite) reads the same:

  FloatingBody, translational_dofs, rotational_dofs        src/bodies.jl:9-118
  DirectBEM, IndirectBEM, default_greens_functions         src/formulations.jl:7-37
  RadiationProblem, DiffractionProblem                     src/problems_and_results.jl:14-150
  solve(problem, formulation)                              src/solve.jl:59-78
  added_mass, radiation_damping, diffraction_force,
  froude_krylov_force, excitation_force                    src/problems_and_results.jl:202-219
  hydrodynamic_coefficients                                src/problems_and_results.jl:464-495
  assemble_matrices(gfs, mesh, k; direct, include_identity) src/matrix_assembly.jl:164-208
  solve_linear(D, S, bc; direct)  (the reference's solve(D, S, bc))  src/matrix_assembly.jl:436-449
  SETTINGS, set_rho, set_g                                 src/constants.jl:10-29
  compute_encountered_values, forward_speed= / forward_speeds  src/waves.jl:136-151,
                                                           src/problems_and_results.jl:236-380

Every computation runs in libmhx on the GPU; this module only validates
arguments, keeps device contexts and packs results.  ArgumentError maps to
ValueError, DimensionMismatch to mhx.DimensionMismatch.
"""
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DimensionMismatch, SingularException  # noqa: F401
from .meshes import Mesh

RIGID_DOFS = ("Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw")


# --------------------------------------------------------------------------- constants
@dataclass
class PhysicalConstants:
    rho: float = 1000.0
    g: float = 9.81


SETTINGS = PhysicalConstants()


def set_rho(val):
    SETTINGS.rho = float(val)


def set_g(val):
    SETTINGS.g = float(val)


def compute_wavenumber(omega, g=None):
    """Deep-water dispersion k = omega^2 / g (src/waves.jl:133-134)."""
    g = SETTINGS.g if g is None else g
    return omega * omega / g


def compute_encountered_values(omega, beta, forward_speed, g=None):
    """(omega_e, k_e, beta_e) seen by a body moving at forward_speed along +x in waves of
    frequency omega and heading beta (src/waves.jl:136-151): omega_e = |omega - k U cos beta|,
    k_e = omega_e^2 / g, beta_e = beta, or beta + pi when the Doppler-shifted frequency is
    negative."""
    g = SETTINGS.g if g is None else g
    k = compute_wavenumber(omega, g)
    if forward_speed == 0:
        return omega, k, beta
    doppler = omega - k * forward_speed * np.cos(beta)
    we = abs(doppler)
    return we, compute_wavenumber(we, g), (beta if doppler >= 0 else beta + np.pi)


def evaluate_gradient_of_motion(mesh, rigid_dof_name):
    """d(dof)/dx on every panel (nf x 3) of a rigid dof, used by the forward-speed radiation
    condition (src/bodies.jl:55-66): Pitch (0, 0, 1), Yaw (0, -1, 0), the others 0."""
    out = np.zeros((mesh.nfaces, 3))
    out[:] = _motion_gradient(rigid_dof_name)
    return out


def _motion_gradient(name):
    """The constant row of evaluate_gradient_of_motion for a rigid dof name."""
    if name not in RIGID_DOFS:
        raise ValueError("forward speed problems are only developed for rigid body dofs")
    return {"Pitch": (0.0, 0.0, 1.0), "Yaw": (0.0, -1.0, 0.0)}.get(name, (0.0, 0.0, 0.0))


# --------------------------------------------------------------------------- Green functions
class GreensFunction:
    wavenumber_independent = False


class Rankine(GreensFunction):
    wavenumber_independent = True


class RankineReflected(GreensFunction):
    wavenumber_independent = True


class GFWu(GreensFunction):
    code = _lib.GF_WU


class ExactGuevelDelhommeau(GreensFunction):
    code = _lib.GF_DELHOMMEAU


class SurrogateGuevelDelhommeau(GreensFunction):
    """The exact wave term through the learned surrogate: GFWu plus a network residual
    (BF16 MFMA inference), with the exact FP64 term for pairs outside the trained box
    (k R_h > 100, k|z + zeta| > 50) and near the free surface (k|z + zeta| < 1e-3).
    No counterpart in the reference; it stands in for ExactGuevelDelhommeau()."""
    code = _lib.GF_SURROGATE


GREENS = {"Wu": GFWu, "ExactGuevelDelhommeau": ExactGuevelDelhommeau, "Surrogate": SurrogateGuevelDelhommeau}


def default_greens_functions(gf="Wu"):
    if gf not in GREENS:
        raise ValueError(f'unknown Green\'s function "{gf}"; use "Wu", "ExactGuevelDelhommeau" or "Surrogate"')
    return (Rankine(), RankineReflected(), GREENS[gf]())


def _wave_code(gfs):
    gfs = tuple(gfs)
    kinds = [type(g) for g in gfs]
    if len(gfs) != 3 or kinds[0] is not Rankine or kinds[1] is not RankineReflected or \
            kinds[2] not in GREENS.values():
        raise ValueError("the GPU path supports (Rankine(), RankineReflected(), GFWu() | ExactGuevelDelhommeau() | "
                         "SurrogateGuevelDelhommeau())")
    return gfs[2].code


class BEMFormulation:
    direct = True

    def __init__(self, greens="Wu"):
        self.greens = default_greens_functions(greens) if isinstance(greens, str) else tuple(greens)
        self.gf = _wave_code(self.greens)


class DirectBEM(BEMFormulation):
    direct = True


class IndirectBEM(BEMFormulation):
    direct = False


def is_direct(f):
    return f.direct


# --------------------------------------------------------------------------- contexts
_contexts = {}


def context(device=0):
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = _contexts[device] = _lib.Context(SETTINGS.rho, SETTINGS.g, device)
    return ctx


# --------------------------------------------------------------------------- surrogate
SURROGATE_WEIGHTS = os.path.join(_lib.PKG_ROOT, "weights", "delhommeau_surrogate.mhxs")        # 4 x 128 SiLU
SURROGATE_WEIGHTS_RELU = os.path.join(_lib.PKG_ROOT, "weights", "delhommeau_surrogate_relu.mhxs")  # ReLU variant
_surrogate_ready = set()


def train_surrogate(device=0, save_to=None, **kw):
    """Train the wave-term network on the device (exact FP64 samples, Adam); keeps it in the
    device context and optionally writes it to `save_to`.  Returns the validation MSE."""
    ctx = context(device)
    val = ctx.surrogate_train(**kw)
    _surrogate_ready.add(device)
    if save_to:
        ctx.surrogate_save(save_to)
    return val


def load_surrogate(path=None, device=0):
    """Load trained weights (default: the weights shipped with the package)."""
    context(device).surrogate_load(path or SURROGATE_WEIGHTS)
    _surrogate_ready.add(device)


def _ensure_surrogate(gf, device=0):
    if gf == _lib.GF_SURROGATE and device not in _surrogate_ready:
        if not os.path.exists(SURROGATE_WEIGHTS):
            raise ValueError("SurrogateGuevelDelhommeau needs a trained network: call mhx.train_surrogate() or "
                             "mhx.load_surrogate(path)")
        load_surrogate(device=device)

# --------------------------------------------------------------------------- bodies
def translational_dofs(mesh, name):
    dof = np.zeros((mesh.nfaces, 3), order="F")
    dof[:, ("Surge", "Sway", "Heave").index(name)] = 1.0
    return dof


def rotational_dofs(mesh, name, rotation_center=(0.0, 0.0, 0.0)):
    if name not in ("Roll", "Pitch", "Yaw"):
        raise ValueError(f"unknown rotational dof {name}")
    axis = np.zeros(3)
    axis[("Roll", "Pitch", "Yaw").index(name)] = 1.0
    pos = mesh.centers - np.asarray(rotation_center, dtype=np.float64)[None, :]
    return np.asfortranarray(np.cross(axis[None, :], pos))


class FloatingBody:
    """FloatingBody(mesh, dofs, name) with dofs a dict name -> (n, 3) motion, or
    FloatingBody(mesh, ["Heave", ...], rotation_center, name) for rigid dofs."""

    def __init__(self, mesh, dofs, rotation_center=(0.0, 0.0, 0.0), body_name="body", lid_mesh=None):
        if not isinstance(mesh, Mesh):
            raise TypeError("mesh must be an mhx.meshes.Mesh")
        if lid_mesh is not None and not isinstance(lid_mesh, Mesh):
            raise TypeError("lid_mesh must be an mhx.meshes.Mesh")
        self.mesh = mesh
        self.lid_mesh = lid_mesh  # irregular-frequency removal (bodies.jl:9-28); dofs live on the hull
        self.body_name = body_name
        self._rotation = {}  # rigid rotational dofs: name -> axis (their motion follows the centers)
        if isinstance(dofs, dict):
            items = dofs.items()
        else:
            items = []
            for name in dofs:
                name = str(name)
                if name in ("Surge", "Sway", "Heave"):
                    items.append((name, translational_dofs(mesh, name)))
                else:
                    items.append((name, rotational_dofs(mesh, name, rotation_center)))
                    self._rotation[name] = np.eye(3)[("Roll", "Pitch", "Yaw").index(name)]
        self.dofs = {}
        for name, d in items:
            d = np.asfortranarray(d, dtype=np.float64)
            if d.shape != (mesh.nfaces, 3):
                raise DimensionMismatch(f"dof {name} must have shape ({mesh.nfaces}, 3)")
            self.dofs[str(name)] = d
        self._handle = None
        self._handle_key = None

    @property
    def dof_names(self):
        return list(self.dofs)

    def _dof_stack(self, names):
        return np.ascontiguousarray(np.concatenate([self.dofs[nm].ravel(order="F") for nm in names]))


def make_body_name_list_unique(names):
    """make_body_name_list_unique (src/bodies.jl:138-161): repeated names get _2, _3, ..."""
    seen, out = {}, []
    for s in names:
        if s not in seen:
            out.append(s)
            seen[s] = 2
        else:
            count = seen[s]
            new = f"{s}_{count}"
            while new in seen or new in names:
                count += 1
                new = f"{s}_{count}"
            out.append(new)
            seen[s] = count + 1
            seen[new] = 1
    return out


def combine_floatingbodies(bodies, new_body_name=None):
    """combine_floatingbodies (src/bodies.jl:165-217): one body whose mesh is the bodies' meshes
    in order, whose dofs are every body's dofs renamed "body__dof" and padded with zero rows on
    the other bodies' panels, and whose lid is the combined lids.  The coupled radiation and
    diffraction problem of the array is then one solve on the combined mesh."""
    from .meshes import combine_meshes
    bodies = list(bodies)
    names = make_body_name_list_unique([b.body_name.replace(" ", "_") for b in bodies])
    mesh = combine_meshes([b.mesh for b in bodies])
    ntot = mesh.nfaces
    dofs, nbf = {}, 0
    for b, bn in zip(bodies, names):
        for dn, d in b.dofs.items():
            full = np.zeros((ntot, 3))
            full[nbf:nbf + b.mesh.nfaces] = d
            dofs[f"{bn}__{dn}"] = full
        nbf += b.mesh.nfaces
    lids = [b.lid_mesh for b in bodies if getattr(b, "lid_mesh", None) is not None]
    lid = combine_meshes(lids) if lids else None
    return FloatingBody(mesh, dofs, body_name=new_body_name or "+".join(names), lid_mesh=lid)


FloatingBody.__add__ = lambda self, other: combine_floatingbodies(
    [self] + (list(other) if isinstance(other, (list, tuple)) else [other]))


# --------------------------------------------------------------------------- problems
@dataclass
class RadiationProblem:
    floatingbody: FloatingBody
    omega: float
    radiating_dof: str = None
    influenced_dofs: list = None
    rho: float = None
    g: float = None
    beta: float = None
    forward_speed: float = 0.0

    def __post_init__(self):
        inf = list(self.influenced_dofs) if self.influenced_dofs is not None else self.floatingbody.dof_names
        rad = self.radiating_dof
        if rad is None:
            if len(inf) != 1:
                raise ValueError("radiating_dof must be given when the body has several dofs")
            rad = inf[0]
        if rad not in self.floatingbody.dofs:
            raise ValueError("the radiating_dof Symbol must be a key of floatingbody.dofs")
        if not set(inf) <= set(self.floatingbody.dofs):
            raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
        self.influenced_dofs, self.radiating_dof = inf, rad
        self.rho = SETTINGS.rho if self.rho is None else float(self.rho)
        self.g = SETTINGS.g if self.g is None else float(self.g)
        self.wavenumber = compute_wavenumber(self.omega, self.g)
        _encounter(self)


@dataclass
class DiffractionProblem:
    floatingbody: FloatingBody
    omega: float
    beta: float = 0.0
    influenced_dofs: list = None
    rho: float = None
    g: float = None
    forward_speed: float = 0.0

    def __post_init__(self):
        inf = list(self.influenced_dofs) if self.influenced_dofs is not None else self.floatingbody.dof_names
        if not set(inf) <= set(self.floatingbody.dofs):
            raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
        self.influenced_dofs = inf
        self.rho = SETTINGS.rho if self.rho is None else float(self.rho)
        self.g = SETTINGS.g if self.g is None else float(self.g)
        self.wavenumber = compute_wavenumber(self.omega, self.g)
        _encounter(self)


def _encounter(problem):
    """encountered_omega / _wavenumber / _beta of a problem (problems_and_results.jl:40-53,
    118-133): the incident values at rest, compute_encountered_values otherwise (which needs a
    heading)."""
    U = float(problem.forward_speed)
    problem.forward_speed = U
    if U == 0:
        problem.encountered_omega, problem.encountered_wavenumber = problem.omega, problem.wavenumber
        problem.encountered_beta = problem.beta
        return
    if problem.beta is None:
        raise ValueError("a problem with nonzero forward speed needs a wave direction beta")
    problem.encountered_omega, problem.encountered_wavenumber, problem.encountered_beta = \
        compute_encountered_values(problem.omega, problem.beta, U, problem.g)


@dataclass
class RadiationResult:
    problem: RadiationProblem
    forces: dict


@dataclass
class DiffractionResult:
    problem: DiffractionProblem
    forces: dict
    fk: dict = field(default_factory=dict)


def _result_omega(res):
    """problems_and_results.jl:202-203: the encountered frequency of a moving body."""
    p = res.problem
    return p.omega if p.forward_speed == 0 else p.encountered_omega


def added_mass(res):
    w = _result_omega(res)
    return {k: v.real / w ** 2 for k, v in res.forces.items()}


def radiation_damping(res):
    w = _result_omega(res)
    return {k: v.imag / w for k, v in res.forces.items()}


def diffraction_force(res):
    return dict(res.forces)


def froude_krylov_force(res):
    return dict(res.fk)


def excitation_force(res):
    return {k: res.forces[k] + res.fk[k] for k in res.forces}


# --------------------------------------------------------------------------- solves
def _ensure_handle(body, names, ctx):
    """The body's device-resident copy (geometry, dofs, Rankine cache, work buffers) for this
    list of dofs, created on first use."""
    key = tuple(names)
    if body._handle is None or body._handle_key != key:
        _release(body)
        h = _lib.C.c_void_p()
        lid = getattr(body, "lid_mesh", None)
        if lid is None:
            cm = _lib.cmesh(body.mesh)
            dofs = body._dof_stack(names)
            ctx.check(_lib.lib().mhx_body_create(ctx.h, _lib.C.byref(cm), len(names), _lib.ptr(dofs),
                                                 _lib.C.byref(h)))
        else:
            # hull + lid (solve.jl:22-28), the dofs padded with zero rows on the lid (_pad_hull_bc)
            from .meshes import combine
            full = combine(body.mesh, lid)
            nh, nl = body.mesh.nfaces, lid.nfaces
            dofs = np.ascontiguousarray(np.concatenate(
                [np.concatenate([body.dofs[nm], np.zeros((nl, 3))]).ravel(order="F") for nm in names]))
            cm = _lib.cmesh(full)
            ctx.check(_lib.lib().mhx_body_create_lid(ctx.h, _lib.C.byref(cm), nh, len(names), _lib.ptr(dofs),
                                                     _lib.C.byref(h)))
        body._handle, body._handle_key, body._ctx = h, key, ctx
    return body._handle


def _body_coeffs(body, names, omegas, betas, gf, direct, mode, rho, g, device=0):
    ctx = context(device)
    _ensure_surrogate(gf, device)
    ctx.set_constants(rho, g)
    L = _lib.lib()
    _ensure_handle(body, names, ctx)
    nd = len(names)
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    be = np.ascontiguousarray(betas, dtype=np.float64)
    nw, nb = om.size, be.size
    cb = 16 - nd  # right-hand sides of one solve: nd radiation + at most cb diffraction columns
    if 0 < cb < nb:
        parts = [_body_coeffs(body, names, om, be[i:i + cb], gf, direct, mode, rho, g, device)
                 for i in range(0, nb, cb)]
        return (parts[0][0], parts[0][1], np.concatenate([p[2] for p in parts], axis=2),
                np.concatenate([p[3] for p in parts], axis=2))
    A = np.zeros((nd, nd, nw), order="F")
    B = np.zeros((nd, nd, nw), order="F")
    Fd = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    Fk = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    ctx.check(L.mhx_body_coeffs(ctx.h, body._handle, nw, _lib.ptr(om), nb, _lib.ptr(be) if nb else None, gf,
                                int(direct), mode, _lib.ptr(A), _lib.ptr(B), _lib.ptr(Fd), _lib.ptr(Fk)))
    return A, B, Fd[:, :, :nb], Fk[:, :, :nb]


def _body_coeffs_fs(body, names, omegas, betas, speeds, rad, gf, direct, rho, g, device=0):
    """mhx_body_coeffs_fs: A, B (nd, nd, nw, nU, nb); Fd, Fk (nd, nw, nb, nU).  The gradient of
    motion is needed for the radiating dofs only (rad) and only when some speed is nonzero."""
    us = np.ascontiguousarray(speeds, dtype=np.float64)
    nd = len(names)
    dgrad = np.zeros((nd, 3))
    if np.any(us != 0):
        if direct:
            raise ValueError("nonzero forward speed is not supported with the direct method; "
                             "use IndirectBEM (direct = false)")
        for i, nm in enumerate(names):
            if nm in rad:
                dgrad[i] = _motion_gradient(nm)
    dgrad = np.ascontiguousarray(dgrad)
    ctx = context(device)
    _ensure_surrogate(gf, device)
    ctx.set_constants(rho, g)
    _ensure_handle(body, names, ctx)
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    be = np.ascontiguousarray(betas, dtype=np.float64)
    nw, nb, nu = om.size, be.size, us.size
    A = np.zeros((nd, nd, nw, nu, nb), order="F")
    B = np.zeros((nd, nd, nw, nu, nb), order="F")
    Fd = np.zeros((nd, nw, nb, nu), dtype=np.complex128, order="F")
    Fk = np.zeros((nd, nw, nb, nu), dtype=np.complex128, order="F")
    ctx.check(_lib.lib().mhx_body_coeffs_fs(ctx.h, body._handle, nw, _lib.ptr(om), nb, _lib.ptr(be), nu, _lib.ptr(us),
                                            _lib.ptr(dgrad), gf, int(direct), _lib.ptr(A), _lib.ptr(B), _lib.ptr(Fd),
                                            _lib.ptr(Fk)))
    return A, B, Fd, Fk


def _release(body):
    if body._handle is not None:
        _lib.lib().mhx_body_destroy(body._handle)
        body._handle = None


def solve(problem, formulation=None):
    """solve(problem, DirectBEM()/IndirectBEM()) (src/solve.jl:59-65)."""
    formulation = DirectBEM() if formulation is None else formulation
    body = problem.floatingbody
    names = list(problem.influenced_dofs)
    if problem.forward_speed != 0:
        # one (U, omega, beta) group of mhx_body_coeffs_fs (solve.jl:59-65 with _problem_omega_k)
        rad = [problem.radiating_dof] if isinstance(problem, RadiationProblem) else []
        if rad and rad[0] not in names:
            names = names + rad
        A, B, Fd, Fk = _body_coeffs_fs(body, names, [problem.omega], [problem.beta], [problem.forward_speed], rad,
                                       formulation.gf, formulation.direct, problem.rho, problem.g)
        ninf = len(problem.influenced_dofs)
        if rad:
            r, w = names.index(rad[0]), problem.encountered_omega
            return RadiationResult(problem, {nm: complex(A[i, r, 0, 0, 0] * w * w, B[i, r, 0, 0, 0] * w)
                                             for i, nm in enumerate(names[:ninf])})
        return DiffractionResult(problem, {nm: complex(Fd[i, 0, 0, 0]) for i, nm in enumerate(names)},
                                 {nm: complex(Fk[i, 0, 0, 0]) for i, nm in enumerate(names)})
    if isinstance(problem, RadiationProblem):
        if problem.radiating_dof not in names:
            names = names + [problem.radiating_dof]
        A, B, _, _ = _body_coeffs(body, names, [problem.omega], [], formulation.gf, formulation.direct,
                                  _lib.MODE_PER_OMEGA, problem.rho, problem.g)
        r = names.index(problem.radiating_dof)
        w = problem.omega
        forces = {nm: complex(A[i, r, 0] * w * w, B[i, r, 0] * w) for i, nm in enumerate(problem.influenced_dofs)}
        return RadiationResult(problem, forces)
    if isinstance(problem, DiffractionProblem):
        _, _, Fd, Fk = _body_coeffs(body, names, [problem.omega], [problem.beta], formulation.gf, formulation.direct,
                                    _lib.MODE_PER_OMEGA, problem.rho, problem.g)
        return DiffractionResult(problem, {nm: complex(Fd[i, 0, 0]) for i, nm in enumerate(names)},
                                 {nm: complex(Fk[i, 0, 0]) for i, nm in enumerate(names)})
    raise TypeError("problem must be a RadiationProblem or a DiffractionProblem")


def hydrodynamic_coefficients(floatingbody, parameters, formulation=None, schedule="per_omega"):
    """hydrodynamic_coefficients(body, (wave_frequencies, radiating_dofs, wave_directions), formulation)
    (src/problems_and_results.jl:464-495).  Layouts as the reference:
    added_mass / radiation_damping (n_inf, n_rad, n_omega, 1); forces (n_inf, n_omega, n_beta, 1).
    schedule="per_problem" reproduces the reference's one assembly + LU per problem;
    schedule="batched" processes the frequencies in chunks with a batched LU; schedule="graph" is
    the batched schedule captured into a HIP graph on the first call and replayed after."""
    formulation = DirectBEM() if formulation is None else formulation
    ws = np.atleast_1d(np.asarray(parameters["wave_frequencies"], dtype=np.float64))
    inf = list(parameters.get("influenced_dofs", floatingbody.dof_names))
    rad = list(parameters.get("radiating_dofs", []))
    betas = list(parameters.get("wave_directions", []))
    for nm in inf + rad:
        if nm not in floatingbody.dofs:
            raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
    names = inf + [r for r in rad if r not in inf]
    if "forward_speeds" in parameters:
        return _coefficients_fs(floatingbody, parameters, formulation, ws, inf, rad, betas, names)
    mode = {"per_omega": _lib.MODE_PER_OMEGA, "per_problem": _lib.MODE_PER_PROBLEM,
            "batched": _lib.MODE_BATCHED, "graph": _lib.MODE_GRAPH}[schedule]
    A, B, Fd, Fk = _body_coeffs(floatingbody, names, ws, betas, formulation.gf, formulation.direct, mode,
                                SETTINGS.rho, SETTINGS.g)
    ii = [names.index(nm) for nm in inf]
    rr = [names.index(nm) for nm in rad]
    out = {}
    if rad:
        out["added_mass"] = A[np.ix_(ii, rr)][..., None]
        out["radiation_damping"] = B[np.ix_(ii, rr)][..., None]
    else:
        out["added_mass"] = out["radiation_damping"] = np.zeros(0)
    if betas:
        fd = Fd[ii][..., None]
        fk = Fk[ii][..., None]
        out.update(diffraction_force=fd, Froude_Krylov_force=fk, excitation_force=fd + fk)
    else:
        out.update(diffraction_force=np.zeros(0), Froude_Krylov_force=np.zeros(0), excitation_force=np.zeros(0))
    return out


def _coefficients_fs(body, parameters, formulation, ws, inf, rad, betas, names):
    """hydrodynamic_coefficients with forward_speeds (problems_and_results.jl:236-380): radiation
    problems over (beta, dof, omega, U), diffraction over (beta, omega, U).  Layouts as the
    reference: added_mass / radiation_damping (n_inf, n_rad, n_omega, n_U, n_beta); forces
    (n_inf, n_omega, n_beta, n_U)."""
    us = np.atleast_1d(np.asarray(parameters["forward_speeds"], dtype=np.float64))
    if not betas:
        raise ValueError("forward_speeds needs wave_directions (the encountered frequency depends on them)")
    A, B, Fd, Fk = _body_coeffs_fs(body, names, ws, betas, us, rad, formulation.gf, formulation.direct,
                                   SETTINGS.rho, SETTINGS.g)
    ii = [names.index(nm) for nm in inf]
    rr = [names.index(nm) for nm in rad]
    out = {}
    if rad:
        out["added_mass"] = A[np.ix_(ii, rr)]
        out["radiation_damping"] = B[np.ix_(ii, rr)]
    else:
        out["added_mass"] = out["radiation_damping"] = np.zeros(0)
    fd, fk = Fd[ii], Fk[ii]
    out.update(diffraction_force=fd, Froude_Krylov_force=fk, excitation_force=fd + fk)
    return out


def mesh_scaling_tangent(mesh, center=(0.0, 0.0, 0.0), scale=1.0):
    """Tangent of the mesh fields along a uniform scaling about `center`: for a body of size r
    (scale = r) built as r times a unit body, d/dr of (vertices, centers, normals, areas) is
    ((V - c) / r, (C - c) / r, 0, 2 A / r).  This is what the reference's fd_mesh_rules! seeds for
    the notebook's mesh_sphere(radius=r) (ad_rules_geometry.jl:174-230), up to its difference
    step."""
    c = np.asarray(center, dtype=np.float64)[None, :]
    return {"vertices": (mesh.vertices - c) / scale, "centers": (mesh.centers - c) / scale,
            "normals": np.zeros_like(mesh.normals), "areas": 2.0 * mesh.areas / scale}


def hydrodynamic_coefficients_jvp(floatingbody, parameters, tangent, formulation=None):
    """Forward-mode derivative of hydrodynamic_coefficients (src/problems_and_results.jl:464-495)
    along `tangent`, a dict with any of: "vertices" (nv, 3), "centers", "normals" (nf, 3),
    "areas" (nf,), "dofs" {name: (nf, 3)}, "wave_frequencies" (nw,).  Missing entries are zero;
    the rigid rotational dofs of a body built from dof names follow the centers' tangent.
    Returns (coefficients, tangents), two dicts with the layout of hydrodynamic_coefficients.
    The reference obtains the same numbers with ForwardDiff or Enzyme through its AD rules
    (ad_rules.jl:79-414); here the tangent assembly, right-hand sides and forces are device
    kernels and the tangent solve reuses each frequency's FP64 LU."""
    formulation = DirectBEM() if formulation is None else formulation
    if formulation.gf not in (_lib.GF_WU, _lib.GF_DELHOMMEAU):
        raise ValueError("derivatives are available for GFWu and ExactGuevelDelhommeau")
    if getattr(floatingbody, "lid_mesh", None) is not None or "forward_speeds" in parameters:
        return _coefficients_domega(floatingbody, parameters, tangent, formulation)
    body = floatingbody
    mesh = body.mesh
    ws = np.ascontiguousarray(np.atleast_1d(np.asarray(parameters["wave_frequencies"], dtype=np.float64)))
    inf = list(parameters.get("influenced_dofs", body.dof_names))
    rad = list(parameters.get("radiating_dofs", []))
    betas = np.ascontiguousarray(parameters.get("wave_directions", []), dtype=np.float64)
    for nm in inf + rad:
        if nm not in body.dofs:
            raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
    names = inf + [r for r in rad if r not in inf]
    nd, nw, nb = len(names), ws.size, betas.size

    def field(key, shape):
        a = tangent.get(key)
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != shape:
            raise DimensionMismatch(f"tangent {key} must have shape {shape}")
        return a

    dV = field("vertices", mesh.vertices.shape)
    dC = field("centers", mesh.centers.shape)
    dN = field("normals", mesh.normals.shape)
    dA = field("areas", mesh.areas.shape)
    dW = field("wave_frequencies", ws.shape)
    tdofs = dict(tangent.get("dofs", {}))
    rot = getattr(body, "_rotation", {})
    dd = []
    for nm in names:
        if nm in tdofs:
            d = np.asarray(tdofs[nm], dtype=np.float64)
            if d.shape != (mesh.nfaces, 3):
                raise DimensionMismatch(f"tangent dof {nm} must have shape ({mesh.nfaces}, 3)")
        elif nm in rot and dC is not None:
            d = np.cross(rot[nm][None, :], dC)
        else:
            d = np.zeros((mesh.nfaces, 3))
        dd.append(np.asarray(d).ravel(order="F"))
    ddofs = np.ascontiguousarray(np.concatenate(dd)) if nd else None
    dofs = body._dof_stack(names) if nd else None
    ctx = context()
    ctx.set_constants(SETTINGS.rho, SETTINGS.g)
    A = np.zeros((nd, nd, nw), order="F")
    B = np.zeros((nd, nd, nw), order="F")
    Fd = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    Fk = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    tA, tB, tFd, tFk = (np.zeros_like(x) for x in (A, B, Fd, Fk))
    cm = _lib.cmesh(mesh)
    tm = _lib.CMeshTangent(_lib.ptr(dV), _lib.ptr(dC), _lib.ptr(dN), _lib.ptr(dA))
    ctx.check(_lib.lib().mhx_hydro_coeffs_jvp(
        ctx.h, _lib.C.byref(cm), _lib.C.byref(tm), nd, _lib.ptr(dofs), _lib.ptr(ddofs), nw, _lib.ptr(ws),
        _lib.ptr(dW), nb, _lib.ptr(betas) if nb else None, formulation.gf, int(formulation.direct),
        *(_lib.ptr(x) for x in (A, B, Fd, Fk, tA, tB, tFd, tFk))))
    ii = [names.index(nm) for nm in inf]
    rr = [names.index(nm) for nm in rad]

    def label(A, B, Fd, Fk):
        out = {}
        if rad:
            out["added_mass"] = A[np.ix_(ii, rr)][..., None]
            out["radiation_damping"] = B[np.ix_(ii, rr)][..., None]
        else:
            out["added_mass"] = out["radiation_damping"] = np.zeros(0)
        if nb:
            fd, fk = Fd[ii][:, :, :nb, None], Fk[ii][:, :, :nb, None]
            out.update(diffraction_force=fd, Froude_Krylov_force=fk, excitation_force=fd + fk)
        else:
            out.update(diffraction_force=np.zeros(0), Froude_Krylov_force=np.zeros(0),
                       excitation_force=np.zeros(0))
        return out

    return label(A, B, Fd, Fk), label(tA, tB, tFd, tFk)


def _coefficients_domega(body, parameters, tangent, formulation):
    """Frequency tangent of hydrodynamic_coefficients for a body with a lid and/or forward speeds
    (mhx_hydro_coeffs_fs_jvp): the reference differentiates both in omega (ForwardDiff/Enzyme,
    test/solve_problem_differentiation.jl:94-98, test/forward_speed_tests.jl:280-303).  Only the
    "wave_frequencies" entry of the tangent is taken; a mesh or dof tangent is an error here.
    Layouts as hydrodynamic_coefficients (with forward_speeds as its forward-speed layout)."""
    if set(tangent) - {"wave_frequencies"}:
        raise ValueError("for a body with a lid or forward speeds only the wave_frequencies tangent is available")
    ws = np.ascontiguousarray(np.atleast_1d(np.asarray(parameters["wave_frequencies"], dtype=np.float64)))
    dW = tangent.get("wave_frequencies")
    dW = None if dW is None else np.ascontiguousarray(np.broadcast_to(np.asarray(dW, dtype=np.float64), ws.shape))
    inf = list(parameters.get("influenced_dofs", body.dof_names))
    rad = list(parameters.get("radiating_dofs", []))
    betas = list(parameters.get("wave_directions", []))
    if not betas:
        raise ValueError("the frequency tangent needs wave_directions")
    fs = "forward_speeds" in parameters
    us = np.ascontiguousarray(np.atleast_1d(np.asarray(parameters.get("forward_speeds", [0.0]), dtype=np.float64)))
    for nm in inf + rad:
        if nm not in body.dofs:
            raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
    names = inf + [r for r in rad if r not in inf]
    nd, nw, nb, nu = len(names), ws.size, len(betas), us.size
    dgrad = np.zeros((nd, 3))
    if np.any(us != 0):
        if formulation.direct:
            raise ValueError("nonzero forward speed is not supported with the direct method; use IndirectBEM (direct = false)")
        for i, nm in enumerate(names):
            if nm in rad:
                dgrad[i] = _motion_gradient(nm)
    lid = getattr(body, "lid_mesh", None)
    if lid is None:
        mesh, nhull = body.mesh, body.mesh.nfaces
        dofs = body._dof_stack(names)
    else:
        from .meshes import combine
        mesh, nhull = combine(body.mesh, lid), body.mesh.nfaces
        dofs = np.ascontiguousarray(np.concatenate(
            [np.concatenate([body.dofs[nm], np.zeros((lid.nfaces, 3))]).ravel(order="F") for nm in names]))
    be = np.ascontiguousarray(betas, dtype=np.float64)
    dgrad = np.ascontiguousarray(dgrad)
    ctx = context()
    ctx.set_constants(SETTINGS.rho, SETTINGS.g)
    A = np.zeros((nd, nd, nw, nu, nb), order="F")
    B = np.zeros_like(A)
    Fd = np.zeros((nd, nw, nb, nu), dtype=np.complex128, order="F")
    Fk = np.zeros_like(Fd)
    tA, tB, tFd, tFk = (np.zeros_like(x) for x in (A, B, Fd, Fk))
    cm = _lib.cmesh(mesh)
    ctx.check(_lib.lib().mhx_hydro_coeffs_fs_jvp(
        ctx.h, _lib.C.byref(cm), nhull, nd, _lib.ptr(dofs), nw, _lib.ptr(ws), _lib.ptr(dW), nb, _lib.ptr(be), nu,
        _lib.ptr(us), _lib.ptr(dgrad), formulation.gf, int(formulation.direct),
        *(_lib.ptr(x) for x in (A, B, Fd, Fk, tA, tB, tFd, tFk))))
    ii = [names.index(nm) for nm in inf]
    rr = [names.index(nm) for nm in rad]

    def label(A, B, Fd, Fk):
        out = {}
        if rad:
            a, b = A[np.ix_(ii, rr)], B[np.ix_(ii, rr)]
            if not fs:   # (n_inf, n_rad, n_omega, 1): the plain layout, one heading's radiation results
                a, b = a[:, :, :, 0, 0][..., None], b[:, :, :, 0, 0][..., None]
            out["added_mass"], out["radiation_damping"] = a, b
        else:
            out["added_mass"] = out["radiation_damping"] = np.zeros(0)
        fd, fk = Fd[ii], Fk[ii]
        if not fs:       # (n_inf, n_omega, n_beta, 1)
            fd, fk = fd[..., 0][..., None], fk[..., 0][..., None]
        out.update(diffraction_force=fd, Froude_Krylov_force=fk, excitation_force=fd + fk)
        return out

    return label(A, B, Fd, Fk), label(tA, tB, tFd, tFk)


def hydrodynamic_coefficients_vjp(floatingbody, parameters, cotangents, formulation=None):
    """Reverse-mode gradient of J = sum over the outputs of hydrodynamic_coefficients of
    <cotangent, output> (real outputs: elementwise product; complex forces: Re(conj(c) * F)),
    with `cotangents` a dict keyed like the outputs ("added_mass", "radiation_damping",
    "diffraction_force", "Froude_Krylov_force", "excitation_force"; missing = zero; an
    excitation cotangent applies to both of its terms).  Returns (coefficients, gradients):
    gradients holds dJ/d"vertices" (nv, 3), "centers", "normals" (nf, 3), "areas" (nf,),
    "dofs" {name: (nf, 3)} and "wave_frequencies" (nw,).  For a body built from dof names the
    rotational dofs follow the centers, and their gradient is folded into "centers".  The
    reference's Enzyme / Zygote reverse mode gives the same numbers through its rules
    (ad_rules.jl:79-414); here the pullbacks are device kernels and the transposed solve reuses
    each frequency's FP64 LU."""
    formulation = DirectBEM() if formulation is None else formulation
    if formulation.gf not in (_lib.GF_WU, _lib.GF_DELHOMMEAU):
        raise ValueError("derivatives are available for GFWu and ExactGuevelDelhommeau")
    if getattr(floatingbody, "lid_mesh", None) is not None:
        raise ValueError("derivatives of a body with a lid are not supported")
    body = floatingbody
    mesh = body.mesh
    ws = np.ascontiguousarray(np.atleast_1d(np.asarray(parameters["wave_frequencies"], dtype=np.float64)))
    inf = list(parameters.get("influenced_dofs", body.dof_names))
    rad = list(parameters.get("radiating_dofs", []))
    betas = np.ascontiguousarray(parameters.get("wave_directions", []), dtype=np.float64)
    for nm in inf + rad:
        if nm not in body.dofs:
            raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
    names = inf + [r for r in rad if r not in inf]
    nd, nw, nb = len(names), ws.size, betas.size
    ii = [names.index(nm) for nm in inf]
    rr = [names.index(nm) for nm in rad]
    Ab = np.zeros((nd, nd, nw), order="F")
    Bb = np.zeros((nd, nd, nw), order="F")
    Fdb = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    Fkb = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")

    def cot(key, shape, dtype):
        c = cotangents.get(key)
        if c is None:
            return None
        c = np.asarray(c, dtype=dtype)
        if c.size != int(np.prod(shape)):
            raise DimensionMismatch(f"cotangent {key} must have {int(np.prod(shape))} elements (shape {shape})")
        return c.reshape(shape)

    for key, dst in (("added_mass", Ab), ("radiation_damping", Bb)):
        c = cot(key, (len(inf), len(rad), nw), np.float64)
        if c is not None:
            dst[np.ix_(ii, rr)] += c
    for key, dsts in (("diffraction_force", (Fdb,)), ("Froude_Krylov_force", (Fkb,)),
                      ("excitation_force", (Fdb, Fkb))):
        c = cot(key, (len(inf), nw, nb), np.complex128)
        if c is not None:
            for dst in dsts:
                dst[ii, :, :nb] += c
    dofs = body._dof_stack(names) if nd else None
    ctx = context()
    ctx.set_constants(SETTINGS.rho, SETTINGS.g)
    A = np.zeros((nd, nd, nw), order="F")
    B = np.zeros((nd, nd, nw), order="F")
    Fd = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    Fk = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
    gV = np.zeros_like(mesh.vertices)
    gC = np.zeros_like(mesh.centers)
    gN = np.zeros_like(mesh.normals)
    gA = np.zeros_like(mesh.areas)
    gD = np.zeros(3 * mesh.nfaces * max(nd, 1))
    gW = np.zeros(nw)
    cg = _lib.CMeshGrad(*(_lib.ptr(x) for x in (gV, gC, gN, gA, gD, gW)))
    cm = _lib.cmesh(mesh)
    ctx.check(_lib.lib().mhx_hydro_coeffs_vjp(
        ctx.h, _lib.C.byref(cm), nd, _lib.ptr(dofs), nw, _lib.ptr(ws), nb, _lib.ptr(betas) if nb else None,
        formulation.gf, int(formulation.direct), _lib.ptr(Ab), _lib.ptr(Bb), _lib.ptr(Fdb), _lib.ptr(Fkb),
        _lib.ptr(A), _lib.ptr(B), _lib.ptr(Fd), _lib.ptr(Fk), _lib.C.byref(cg)))
    grads = {"vertices": gV, "centers": gC, "normals": gN, "areas": gA, "wave_frequencies": gW, "dofs": {}}
    rot = getattr(body, "_rotation", {})
    for q, nm in enumerate(names):
        gd = gD[3 * mesh.nfaces * q:3 * mesh.nfaces * (q + 1)].reshape(mesh.nfaces, 3, order="F").copy()
        grads["dofs"][nm] = gd
        if nm in rot:  # dof = axis x (c - rotation centre): cbar += dofbar x axis
            gC += np.cross(gd, rot[nm][None, :])
    out = {}
    if rad:
        out["added_mass"] = A[np.ix_(ii, rr)][..., None]
        out["radiation_damping"] = B[np.ix_(ii, rr)][..., None]
    else:
        out["added_mass"] = out["radiation_damping"] = np.zeros(0)
    if nb:
        fd, fk = Fd[ii][:, :, :nb, None], Fk[ii][:, :, :nb, None]
        out.update(diffraction_force=fd, Froude_Krylov_force=fk, excitation_force=fd + fk)
    else:
        out.update(diffraction_force=np.zeros(0), Froude_Krylov_force=np.zeros(0), excitation_force=np.zeros(0))
    return out, grads


def chain_mesh_gradient(mesh, grads):
    """Field-wise gradient (vertices, centers, normals, areas held independent, as the
    reference's StaticArraysMesh and its AD treat them) -> gradient with respect to the vertices
    when the centres, normals and areas are the face properties of the vertices (Capytaine's
    formulas, mhx.meshes.face_properties)."""
    from .meshes import face_properties_vjp
    return grads["vertices"] + face_properties_vjp(mesh.vertices, mesh.faces, grads["centers"], grads["normals"],
                                                   grads["areas"])


def _label(A, B, Fd, Fk, ii, rr, nb):
    out = {}
    if rr:
        out["added_mass"] = A[np.ix_(ii, rr)][..., None]
        out["radiation_damping"] = B[np.ix_(ii, rr)][..., None]
    else:
        out["added_mass"] = out["radiation_damping"] = np.zeros(0)
    if nb:
        fd, fk = Fd[ii][:, :, :nb, None], Fk[ii][:, :, :nb, None]
        out.update(diffraction_force=fd, Froude_Krylov_force=fk, excitation_force=fd + fk)
    else:
        out.update(diffraction_force=np.zeros(0), Froude_Krylov_force=np.zeros(0), excitation_force=np.zeros(0))
    return out


def hydrodynamic_coefficients_many(jobs, formulation=None, graph=False):
    """hydrodynamic_coefficients over several bodies at once (the reference's
    solve(problems::Vector), src/solve.jl:75-78, over every body's problems): `jobs` is a list
    of (floatingbody, parameters) with each body's own wave_frequencies and dof lists and the
    same wave_directions for all.  The bodies may have different panel counts; their
    (body, frequency) items share one assembly queue launch per stage on the device, the LU is
    batched per body.  graph=True captures the whole sweep into a HIP graph on the first call
    and replays it when the same bodies and frequency counts come again.  Returns one result
    dict per job, laid out as hydrodynamic_coefficients."""
    formulation = DirectBEM() if formulation is None else formulation
    if formulation.gf not in (_lib.GF_WU, _lib.GF_DELHOMMEAU, _lib.GF_SURROGATE):
        raise ValueError("the multi-body sweep runs GFWu, ExactGuevelDelhommeau or the surrogate")
    _ensure_surrogate(formulation.gf)
    if not jobs:
        return []
    betas = None
    ctx = context()
    ctx.set_constants(SETTINGS.rho, SETTINGS.g)
    arr = (_lib.CSweepBody * len(jobs))()
    keep, meta = [], []
    for q, (body, par) in enumerate(jobs):
        ws = np.ascontiguousarray(np.atleast_1d(np.asarray(par["wave_frequencies"], dtype=np.float64)))
        inf = list(par.get("influenced_dofs", body.dof_names))
        rad = list(par.get("radiating_dofs", []))
        be = np.ascontiguousarray(par.get("wave_directions", []), dtype=np.float64)
        if betas is None:
            betas = be
        elif be.shape != betas.shape or np.any(be != betas):
            raise ValueError("every body of a multi-body sweep takes the same wave_directions")
        for nm in inf + rad:
            if nm not in body.dofs:
                raise ValueError("each influenced_dofs Symbol must be a key of floatingbody.dofs")
        names = inf + [r for r in rad if r not in inf]
        h = _ensure_handle(body, names, ctx)
        nd, nw, nb = len(names), ws.size, betas.size
        A = np.zeros((nd, nd, nw), order="F")
        B = np.zeros((nd, nd, nw), order="F")
        Fd = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
        Fk = np.zeros((nd, nw, max(nb, 1)), dtype=np.complex128, order="F")
        arr[q] = _lib.CSweepBody(h, nw, _lib.ptr(ws), _lib.ptr(A), _lib.ptr(B), _lib.ptr(Fd), _lib.ptr(Fk))
        keep.append((ws, A, B, Fd, Fk))
        meta.append(([names.index(nm) for nm in inf], [names.index(nm) for nm in rad]))
    nb = betas.size
    ctx.check(_lib.lib().mhx_sweep_coeffs(ctx.h, arr, len(jobs), nb, _lib.ptr(betas) if nb else None, formulation.gf,
                                          int(formulation.direct), _lib.SWEEP_GRAPH if graph else 0))
    return [_label(A, B, Fd, Fk, ii, rr, nb) for (_, A, B, Fd, Fk), (ii, rr) in zip(keep, meta)]


def assemble_matrices(green_functions, mesh, wavenumber, direct=True, include_identity=True, device=0):
    """S, D (or K for the indirect method) as n x n complex column-major arrays."""
    gf = _wave_code(green_functions)
    ctx = context(device)
    _ensure_surrogate(gf, device)
    n = mesh.nfaces
    S = _lib.complex_out((n, n))
    D = _lib.complex_out((n, n))
    cm = _lib.cmesh(mesh)
    ctx.check(_lib.lib().mhx_assemble(ctx.h, _lib.C.byref(cm), float(wavenumber), gf, int(direct),
                                      int(include_identity), _lib.ptr(S), _lib.ptr(D)))
    return S, D


def solve_linear(D, S, bc, direct=True, device=0):
    """The reference's solve(D, S, bc; direct): returns (phi, sources)."""
    D = _lib.as_c128(D)
    S = _lib.as_c128(S)
    bc = _lib.as_c128(bc)
    vec = bc.ndim == 1
    bc2 = bc.reshape(-1, 1, order="F") if vec else bc
    n = D.shape[0]
    if D.shape != (n, n) or S.shape != (n, n) or bc2.shape[0] != n:
        raise DimensionMismatch("D, S must be n x n and bc must have n rows")
    R = bc2.shape[1]
    phi = _lib.complex_out((n, R))
    src = _lib.complex_out((n, R))
    ctx = context(device)
    ctx.check(_lib.lib().mhx_solve(ctx.h, n, R, _lib.ptr(D), _lib.ptr(S), _lib.ptr(bc2), int(direct),
                                   _lib.ptr(phi), _lib.ptr(src)))
    if vec:
        return phi[:, 0], (None if direct else src[:, 0])
    return phi, (None if direct else src)


def lu_solve(A, B, trans=0, device=0):
    """X = A \\ B (trans=0) or A^H \\ B (trans=2) through the device LU."""
    A = _lib.as_c128(A)
    B = np.array(B, dtype=np.complex128, order="F", copy=True)
    vec = B.ndim == 1
    B2 = B.reshape(-1, 1, order="F") if vec else B
    n = A.shape[0]
    if A.shape != (n, n) or B2.shape[0] != n:
        raise DimensionMismatch("A must be n x n and B must have n rows")
    ctx = context(device)
    ctx.check(_lib.lib().mhx_lu_solve(ctx.h, n, B2.shape[1], _lib.ptr(A), int(trans), _lib.ptr(B2)))
    return B2[:, 0] if vec else B2


def lu_solve_batched(A, B, trans=0, device=0):
    """X_q = A_q \\ B_q (trans=0) or A_q^H \\ B_q (trans=2) for a batch of systems of one order,
    through the device's batched LU: A (batch, n, n), B (batch, n, R) or (batch, n).  The
    precision follows the context's LU mode (context().set_lu_mode)."""
    A = np.asarray(A, dtype=np.complex128)
    B = np.asarray(B, dtype=np.complex128)
    if A.ndim != 3 or A.shape[1] != A.shape[2]:
        raise DimensionMismatch("A must be (batch, n, n)")
    vec = B.ndim == 2
    B3 = B[:, :, None] if vec else B
    nb, n = A.shape[0], A.shape[1]
    if B3.shape[0] != nb or B3.shape[1] != n:
        raise DimensionMismatch("B must be (batch, n, R) with the batch and order of A")
    R = B3.shape[2]
    Ac = np.empty((nb, n, n), dtype=np.complex128)
    Ac[...] = np.transpose(A, (0, 2, 1))   # each matrix column-major in its slot
    Bc = np.empty((nb, R, n), dtype=np.complex128)
    Bc[...] = np.transpose(B3, (0, 2, 1))
    ctx = context(device)
    ctx.check(_lib.lib().mhx_lu_solve_batched(ctx.h, n, R, nb, _lib.ptr(Ac), int(trans), _lib.ptr(Bc)))
    X = np.transpose(Bc, (0, 2, 1))
    return X[:, :, 0] if vec else X
