"""ctypes wrapper over oracle/liboracle.so — the CPU FP64 restatement of the
reference's BEM hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (libmhx.so and the
mhx package) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

GF_WU = 0
GF_DELHOMMEAU = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class OraMesh(C.Structure):
    _fields_ = [("nv", C.c_int64), ("nf", C.c_int64), ("vertices", _dp), ("faces", _ip),
                ("centers", _dp), ("normals", _dp), ("areas", _dp), ("radii", _dp)]


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_delhommeau_term.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.POINTER(C.c_int)]
        L.ora_velocity_potential.restype = C.c_double
        L.ora_velocity_potential.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.ora_velocity_derivatives.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.ora_velocity_hessian.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.ora_birk_phi_and_velocity.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.ora_wu_wave_term_and_derivs.argtypes = [C.c_double, C.c_double, _dp]
        L.ora_struve_h0.restype = C.c_double
        L.ora_struve_h0.argtypes = [C.c_double]
        L.ora_struve_h1.restype = C.c_double
        L.ora_struve_h1.argtypes = [C.c_double]
        L.ora_expint_e1.argtypes = [C.c_double, C.c_double, _dp]
        L.ora_expE1.argtypes = [C.c_double, C.c_double, _dp]
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_radiation_bc.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ora_airy_bc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.ora_airy_pressure.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.ora_integrate_pressure.argtypes = [C.c_void_p, _dp, _dp, _dp]
        for name in ("ora_assemble", "ora_assemble_rows", "ora_assemble_rankine", "ora_assemble_wave",
                     "ora_hydro_coeffs", "ora_hydro_coeffs_lid", "ora_zgetrf", "ora_zgetrs", "ora_solve",
                     "ora_assemble_fixed_normal", "ora_hydro_coeffs_fs"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def set_threads(n):
    lib().ora_set_threads(int(n))


class Mesh:
    """Holds contiguous arrays in the reference's StaticArraysMesh layout."""

    def __init__(self, vertices, faces, centers, normals, areas, radii):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.faces = np.ascontiguousarray(faces, dtype=np.int64)
        self.centers = np.ascontiguousarray(centers, dtype=np.float64)
        self.normals = np.ascontiguousarray(normals, dtype=np.float64)
        self.areas = np.ascontiguousarray(areas, dtype=np.float64)
        self.radii = np.ascontiguousarray(radii, dtype=np.float64)
        self.nf = self.faces.shape[0]
        self.s = OraMesh(self.vertices.shape[0], self.nf, _p(self.vertices), self.faces.ctypes.data_as(_ip),
                         _p(self.centers), _p(self.normals), _p(self.areas), _p(self.radii))

    @classmethod
    def from_mhx(cls, m):
        return cls(m.vertices, m.faces, m.centers, m.normals, m.areas, m.radii)


def birk_panel_geometry(V):
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(4, 3)
    T = np.zeros(9); q = np.zeros(3); lc = np.zeros(8); a = C.c_double()
    lib().ora_birk_panel_geometry(_p(V), _p(T), _p(q), _p(lc), C.byref(a))
    return T.reshape(3, 3).T.copy(), q, lc.reshape(4, 2), a.value


def velocity_potential(x, y, z, lc):
    lc = np.ascontiguousarray(lc, dtype=np.float64).reshape(8)
    return lib().ora_velocity_potential(x, y, z, _p(lc))


def velocity_derivatives(x, y, z, lc):
    lc = np.ascontiguousarray(lc, dtype=np.float64).reshape(8)
    v = np.zeros(3)
    lib().ora_velocity_derivatives(x, y, z, _p(lc), _p(v))
    return v


def velocity_hessian(x, y, z, lc):
    lc = np.ascontiguousarray(lc, dtype=np.float64).reshape(8)
    H = np.zeros(9)
    lib().ora_velocity_hessian(x, y, z, _p(lc), _p(H))
    return H.reshape(3, 3).T.copy()


def birk_phi_and_velocity(x, y, z, lc):
    lc = np.ascontiguousarray(lc, dtype=np.float64).reshape(8)
    phi = np.zeros(1); v = np.zeros(3)
    lib().ora_birk_phi_and_velocity(x, y, z, _p(lc), _p(phi), _p(v))
    return phi[0], v


def vrankine_from_geom(p, c, radius, area, T, qgc, lc):
    p = np.ascontiguousarray(p, dtype=np.float64); c = np.ascontiguousarray(c, dtype=np.float64)
    Tc = np.ascontiguousarray(np.asarray(T).T.reshape(9)); q = np.ascontiguousarray(qgc, dtype=np.float64)
    l8 = np.ascontiguousarray(lc, dtype=np.float64).reshape(8)
    phi = C.c_double(); g = np.zeros(3)
    lib().ora_vrankine_from_geom(_p(p), _p(c), C.c_double(radius), C.c_double(area), _p(Tc), _p(q), _p(l8),
                                 C.byref(phi), _p(g))
    return phi.value, g


def wu_term(hh, vv):
    o = np.zeros(6)
    lib().ora_wu_wave_term_and_derivs(hh, vv, _p(o))
    return o[0] + 1j * o[1], o[2] + 1j * o[3], o[4] + 1j * o[5]


def struve_h0(x):
    return lib().ora_struve_h0(x)


def struve_h1(x):
    return lib().ora_struve_h1(x)


def expint_e1(z):
    o = np.zeros(2)
    lib().ora_expint_e1(z.real, z.imag, _p(o))
    return o[0] + 1j * o[1]


def expE1(z):
    o = np.zeros(2)
    lib().ora_expE1(z.real, z.imag, _p(o))
    return o[0] + 1j * o[1]


def delhommeau_term(r, z, rtol=1e-8, atol=1e-8, return_boxes=False):
    o = np.zeros(6)
    nb = (C.c_int * 2)()
    lib().ora_delhommeau_term(r, z, rtol, atol, _p(o), nb)
    G, Gr, Gz = o[0] + 1j * o[1], o[2] + 1j * o[3], o[4] + 1j * o[5]
    if return_boxes:
        return G, Gr, Gz, (nb[0], nb[1])
    return G, Gr, Gz


def gk15_rule():
    xk = np.zeros(8); wk = np.zeros(8); wg = np.zeros(4)
    lib().ora_gk15_rule(_p(xk), _p(wk), _p(wg))
    return xk, wk, wg


def assemble(mesh, k, gf=GF_WU, direct=True, include_identity=True):
    n = mesh.nf
    S = np.zeros((n, n), dtype=np.complex128, order="F")
    D = np.zeros((n, n), dtype=np.complex128, order="F")
    rc = lib().ora_assemble(C.byref(mesh.s), C.c_double(k), gf, int(direct), int(include_identity),
                            _p(S), _p(D))
    assert rc == 0
    return S, D


def assemble_fixed_normal(mesh, k, fixed, gf=GF_WU, direct=False):
    """assemble_matrices(...; all_normals=fixed) (matrix_assembly.jl:183-245)."""
    n = mesh.nf
    S = np.zeros((n, n), dtype=np.complex128, order="F")
    D = np.zeros((n, n), dtype=np.complex128, order="F")
    fx = np.ascontiguousarray(fixed, dtype=np.float64)
    rc = lib().ora_assemble_fixed_normal(C.byref(mesh.s), C.c_double(k), gf, int(direct), _p(fx), _p(S), _p(D))
    assert rc == 0
    return S, D


def encountered(omega, beta, U, g=9.81):
    """compute_encountered_values (waves.jl:136-151): (omega_e, k_e, beta_e)."""
    out = np.zeros(3)
    lib().ora_encountered(C.c_double(omega), C.c_double(beta), C.c_double(U), C.c_double(g), _p(out))
    return tuple(out)


def gradient_of_motion(name):
    """evaluate_gradient_of_motion (bodies.jl:55-66) as the constant per-panel vector."""
    return {"Pitch": (0.0, 0.0, 1.0), "Yaw": (0.0, -1.0, 0.0)}.get(name, (0.0, 0.0, 0.0))


def hydro_coeffs_fs(mesh, dofs, dgrads, omegas, betas, Us, rho=1000.0, g=9.81, gf=GF_WU):
    """Forward speed, IndirectBEM.  dofs: list of (n,3); dgrads: list of 3-vectors.  Returns
    A, B (ndof, ndof, nw, nU, nb), Fd, Fk (ndof, nw, nb, nU)."""
    n = mesh.nf
    ndof = len(dofs)
    dd = np.zeros((n, 3, ndof), order="F")
    for d, m in enumerate(dofs):
        dd[:, :, d] = m
    dg = np.ascontiguousarray(np.asarray(dgrads, dtype=np.float64).reshape(ndof, 3))
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    be = np.ascontiguousarray(betas, dtype=np.float64)
    us = np.ascontiguousarray(Us, dtype=np.float64)
    nw, nb, nu = len(om), len(be), len(us)
    A = np.zeros((ndof, ndof, nw, nu, nb), order="F"); B = np.zeros_like(A)
    Fd = np.zeros((ndof, nw, nb, nu), dtype=np.complex128, order="F"); Fk = np.zeros_like(Fd)
    info = lib().ora_hydro_coeffs_fs(C.byref(mesh.s), C.c_int64(ndof), _p(dd), _p(dg), C.c_int64(nw), _p(om),
                                     C.c_int64(nb), _p(be), C.c_int64(nu), _p(us), C.c_double(rho), C.c_double(g),
                                     gf, _p(A), _p(B), _p(Fd), _p(Fk))
    return dict(added_mass=A, radiation_damping=B, diffraction_force=Fd, Froude_Krylov_force=Fk,
                excitation_force=Fd + Fk, info=info)


def assemble_rows(mesh, k, rows, gf=GF_WU, direct=True, include_identity=True):
    i0, i1 = rows
    n = mesh.nf
    S = np.zeros((i1 - i0, n), dtype=np.complex128, order="F")
    D = np.zeros((i1 - i0, n), dtype=np.complex128, order="F")
    rc = lib().ora_assemble_rows(C.byref(mesh.s), C.c_double(k), gf, int(direct), int(include_identity),
                                 C.c_int64(i0), C.c_int64(i1), _p(S), _p(D))
    assert rc == 0
    return S, D


def assemble_rankine(mesh, direct=True, include_identity=True):
    n = mesh.nf
    S = np.zeros((n, n), dtype=np.complex128, order="F")
    D = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().ora_assemble_rankine(C.byref(mesh.s), int(direct), int(include_identity), _p(S), _p(D))
    return S, D


def assemble_wave(mesh, k, gf=GF_WU, direct=True):
    n = mesh.nf
    S = np.zeros((n, n), dtype=np.complex128, order="F")
    D = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().ora_assemble_wave(C.byref(mesh.s), C.c_double(k), gf, int(direct), _p(S), _p(D))
    return S, D


def zgetrf(A):
    A = np.array(A, dtype=np.complex128, order="F")
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    info = lib().ora_zgetrf(C.c_int64(n), _p(A), C.c_int64(n), ipiv.ctypes.data_as(_ip))
    return A, ipiv, info


def zgetrs(LU, ipiv, B, trans=0):
    B = np.array(B, dtype=np.complex128, order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = LU.shape[0]
    lib().ora_zgetrs(trans, C.c_int64(n), C.c_int64(B.shape[1]), _p(LU), C.c_int64(n),
                     ipiv.ctypes.data_as(_ip), _p(B), C.c_int64(n))
    return B


def solve(D, S, bc, direct=True):
    D = np.asfortranarray(D, dtype=np.complex128); S = np.asfortranarray(S, dtype=np.complex128)
    bc = np.asfortranarray(bc, dtype=np.complex128)
    vec = bc.ndim == 1
    bc2 = bc.reshape(-1, 1, order="F") if vec else bc
    n, R = bc2.shape
    phi = np.zeros((n, R), dtype=np.complex128, order="F")
    src = np.zeros((n, R), dtype=np.complex128, order="F")
    info = lib().ora_solve(C.c_int64(n), C.c_int64(R), _p(D), _p(S), _p(bc2), int(direct), _p(phi), _p(src))
    if vec:
        return phi[:, 0], (None if direct else src[:, 0]), info
    return phi, (None if direct else src), info


def rigid_dof(mesh, which, rc=(0.0, 0.0, 0.0)):
    dof = np.zeros((mesh.nf, 3), order="F")
    rcv = np.ascontiguousarray(rc, dtype=np.float64)
    lib().ora_rigid_dof(C.byref(mesh.s), int(which), _p(rcv), _p(dof))
    return dof


def hydro_coeffs(mesh, dofs, omegas, betas, rho=1000.0, g=9.81, gf=GF_WU, direct=True, mode=0, nhull=None):
    """dofs: list of (n,3) arrays.  Returns A, B (ndof, ndof, nw), Fd, Fk (ndof, nw, nb).
    nhull: with a lid, the mesh is hull panels [0, nhull) followed by the lid panels and the dofs
    are zero on the lid (solve.jl:4-11, 22-28)."""
    n = mesh.nf
    ndof = len(dofs)
    dd = np.zeros((n, 3, ndof), order="F")
    for d, m in enumerate(dofs):
        dd[:, :, d] = m
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    be = np.ascontiguousarray(betas, dtype=np.float64)
    nw, nb = len(om), len(be)
    A = np.zeros((ndof, ndof, nw), order="F"); B = np.zeros((ndof, ndof, nw), order="F")
    Fd = np.zeros((ndof, nw, max(nb, 1)), dtype=np.complex128, order="F")
    Fk = np.zeros((ndof, nw, max(nb, 1)), dtype=np.complex128, order="F")
    if nhull is None:
        info = lib().ora_hydro_coeffs(C.byref(mesh.s), C.c_int64(ndof), _p(dd), C.c_int64(nw), _p(om), C.c_int64(nb),
                                      _p(be), C.c_double(rho), C.c_double(g), gf, int(direct), int(mode),
                                      _p(A), _p(B), _p(Fd), _p(Fk))
    else:
        info = lib().ora_hydro_coeffs_lid(C.byref(mesh.s), C.c_int64(nhull), C.c_int64(ndof), _p(dd), C.c_int64(nw),
                                          _p(om), C.c_int64(nb), _p(be), C.c_double(rho), C.c_double(g), gf,
                                          int(direct), int(mode), _p(A), _p(B), _p(Fd), _p(Fk))
    return dict(added_mass=A, radiation_damping=B, diffraction_force=Fd[:, :, :nb],
                Froude_Krylov_force=Fk[:, :, :nb], excitation_force=(Fd + Fk)[:, :, :nb], info=info)


def hydro_coeffs_lapack(mesh, dofs, omegas, betas, rho=1000.0, g=9.81, gf=GF_WU, direct=True):
    """hydro_coeffs (mode 0, no lid) with the dense solve done by LAPACK (numpy's OpenBLAS,
    multi-threaded) instead of the oracle's own single-threaded LU, for orders where that LU is
    slow (n >= 1024).  The same steps as hydro_coeffs_impl in mh_oracle.c, one frequency at a time:
    ora_assemble, ora_radiation_bc / ora_airy_bc, phi = D^-1 (S bc) (direct) or phi = S (D^-1 bc)
    (indirect) (matrix_assembly.jl:436-449), p = i rho omega phi (solve.jl:50),
    ora_integrate_pressure, A = Re F / omega^2, B = Im F / omega (problems_and_results.jl:207-208),
    Froude-Krylov from ora_airy_pressure (waves.jl:294-303)."""
    import scipy.linalg as sl
    n = mesh.nf
    ndof = len(dofs)
    dd = [np.asfortranarray(np.asarray(m, dtype=np.float64).reshape(n, 3)) for m in dofs]  # n x 3 column-major
    om = np.asarray(omegas, dtype=np.float64)
    be = np.asarray(betas, dtype=np.float64)
    nw, nb = len(om), len(be)
    R = ndof + nb
    A = np.zeros((ndof, ndof, nw), order="F"); B = np.zeros((ndof, ndof, nw), order="F")
    Fd = np.zeros((ndof, nw, nb), dtype=np.complex128, order="F")
    Fk = np.zeros((ndof, nw, nb), dtype=np.complex128, order="F")
    L = lib()
    ms = C.byref(mesh.s)
    F2 = np.zeros(2)
    for iw, w in enumerate(om):
        k = w * w / g  # compute_wavenumber (waves.jl:133-134)
        bc = np.zeros((n, R), dtype=np.complex128, order="F")
        col = np.zeros(n, dtype=np.complex128)
        for r in range(ndof):
            L.ora_radiation_bc(ms, _p(dd[r]), C.c_double(w), _p(col))
            bc[:, r] = col
        for ib in range(nb):
            L.ora_airy_bc(ms, C.c_double(w), C.c_double(k), C.c_double(be[ib]), C.c_double(g), _p(col))
            bc[:, ndof + ib] = col
        S, D = assemble(mesh, k, gf=gf, direct=direct)
        lu = sl.lu_factor(D, check_finite=False)
        if direct:
            phi = sl.lu_solve(lu, S @ bc, check_finite=False)
        else:
            phi = S @ sl.lu_solve(lu, bc, check_finite=False)
        for r in range(R):
            p = np.ascontiguousarray(1j * rho * w * phi[:, r])  # solve.jl:50
            for d in range(ndof):
                L.ora_integrate_pressure(ms, _p(dd[d]), _p(p), _p(F2))
                if r < ndof:
                    A[d, r, iw] = F2[0] / (w * w)
                    B[d, r, iw] = F2[1] / w
                else:
                    Fd[d, iw, r - ndof] = F2[0] + 1j * F2[1]
        for ib in range(nb):
            p = np.zeros(n, dtype=np.complex128)
            L.ora_airy_pressure(ms, C.c_double(w), C.c_double(k), C.c_double(be[ib]), C.c_double(g), C.c_double(rho),
                                _p(p))
            for d in range(ndof):
                L.ora_integrate_pressure(ms, _p(dd[d]), _p(p), _p(F2))
                Fk[d, iw, ib] = F2[0] + 1j * F2[1]
    return dict(added_mass=A, radiation_damping=B, diffraction_force=Fd, Froude_Krylov_force=Fk,
                excitation_force=Fd + Fk)
