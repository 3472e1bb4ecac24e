"""The Julia glue's libmhx calls replayed through ctypes with Julia's memory layouts.

The Julia glue that binds libmhx into the reference's package hands libmhx the reference's own arrays: a StaticArraysMesh's fields
(contiguous xyz triples, 1-based faces), dofs packed as reduce(vcat, vec.(...)) of each dof's
column-major n x 3 matrix (with zero rows over a lid's panels), column-major output arrays, and
MHX_MODE_BATCHED as the mode.  Here the same buffers are built in NumPy and passed to the same
entry points in the same order, and the results are checked against the oracle (1e-9):

- hydrodynamic_coefficients(body, parameters, formulation): mhx_body_create + mhx_body_coeffs,
  a plain body and a body with a lid (mhx_body_create_lid);
- solve(prob, DirectBEM()): mhx_assemble (S, D) then mhx_solve(D, S, bc) for one radiation problem;
- _backend_ldiv: mhx_lu_solve;
- the forward rule: mhx_hydro_coeffs_jvp along d/domega against central differences;
- forward speed: mhx_body_coeffs_fs with the ndof x 3 gradient-of-motion rows.
"""
import ctypes as C
import math

import numpy as np
import pytest

from test_gpu_parity import close

pytestmark = pytest.mark.gpu

RHO, G = 1025.0, 9.81
MODE_BATCHED = 2
KEYS = ("added_mass", "radiation_damping", "diffraction_force", "Froude_Krylov_force")


@pytest.fixture(scope="module")
def lib():
    import mhx
    from mhx import _lib
    print("device context", flush=True)
    ctx = mhx.context()
    L = _lib.lib()
    # the Python mirror (used for the forward-speed comparison) sets the context from its settings
    mhx.set_rho(RHO)
    mhx.set_g(G)
    # ctx() of the glue: mhx_set_constants(SETTINGS.rho, SETTINGS.g) before every call
    ctx.check(L.mhx_set_constants(ctx.h, RHO, G))
    yield L, ctx, _lib
    mhx.set_rho(1000.0)
    ctx.check(L.mhx_set_constants(ctx.h, 1000.0, 9.81))


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def julia_mesh(m):
    """The StaticArraysMesh{Float64} fields as Julia stores them."""
    return dict(vertices=np.ascontiguousarray(m.vertices, dtype=np.float64),
                faces=np.ascontiguousarray(m.faces, dtype=np.int64),          # 1-based SVector{4,Int}
                centers=np.ascontiguousarray(m.centers, dtype=np.float64),
                normals=np.ascontiguousarray(m.normals, dtype=np.float64),
                areas=np.ascontiguousarray(m.areas, dtype=np.float64),
                radii=np.ascontiguousarray(m.radii, dtype=np.float64))


def mhxmesh(_lib, jm):
    return _lib.CMesh(jm["vertices"].shape[0], jm["faces"].shape[0], dp(jm["vertices"]),
                      jm["faces"].ctypes.data_as(C.POINTER(C.c_int64)), dp(jm["centers"]), dp(jm["normals"]),
                      dp(jm["areas"]), dp(jm["radii"]))


def packed_dofs(dofs, names, nlid=0):
    """reduce(vcat, [vec(vcat(Matrix(body.dofs[k]), zeros(nlid, 3))) for k in names])"""
    return np.concatenate([np.vstack([dofs[k], np.zeros((nlid, 3))]).ravel(order="F") for k in names])


def body_coeffs(L, ctx, h, ws, bs, gf, direct, nd):
    nw, nb = len(ws), len(bs)
    A = np.zeros((nd, nd, nw), order="F"); B = np.zeros((nd, nd, nw), order="F")
    Fd = np.zeros((nd, nw, nb), dtype=np.complex128, order="F"); Fk = np.zeros_like(Fd)
    om = np.ascontiguousarray(ws, dtype=np.float64); be = np.ascontiguousarray(bs, dtype=np.float64)
    ctx.check(L.mhx_body_coeffs(ctx.h, h, nw, dp(om), nb, dp(be), gf, direct, MODE_BATCHED, dp(A), dp(B), dp(Fd),
                                dp(Fk)))
    return dict(added_mass=A, radiation_damping=B, diffraction_force=Fd, Froude_Krylov_force=Fk)


@pytest.mark.parametrize("gf", [0, 1])
def test_hydrodynamic_coefficients_plain_body(lib, oracle, gf):
    L, ctx, _lib = lib
    import mhx
    m = mhx.meshes.cylinder(1.0, 1.5, 16, 6, 3)
    body = mhx.FloatingBody(m, ["Heave", "Surge", "Pitch"])
    names = body.dof_names
    jm = julia_mesh(m)
    dofs = packed_dofs(body.dofs, names)
    h = C.c_void_p()
    ws, bs = [0.7, 1.4, 2.2], [0.0, math.pi / 3]
    print(f"glue replay: plain body, gf={gf}", flush=True)
    ctx.check(L.mhx_body_create(ctx.h, C.byref(mhxmesh(_lib, jm)), len(names), dp(dofs), C.byref(h)))
    try:
        got = body_coeffs(L, ctx, h, ws, bs, gf, 1, len(names))
    finally:
        L.mhx_body_destroy(h)
    ref = oracle.hydro_coeffs(oracle.Mesh.from_mhx(m), [body.dofs[k] for k in names], ws, bs, rho=RHO, g=G,
                              gf=oracle.GF_WU if gf == 0 else oracle.GF_DELHOMMEAU, direct=True)
    ref = dict(added_mass=ref["added_mass"], radiation_damping=ref["radiation_damping"],
               diffraction_force=ref["diffraction_force"], Froude_Krylov_force=ref["Froude_Krylov_force"])
    for key in KEYS:
        rel, elem, ok = close(got[key], ref[key])
        assert ok, (key, rel, elem)


def test_more_headings_than_one_solve_takes(lib, oracle):
    """Six dofs and twelve headings through mhx_body_coeffs (the glue passes every heading at
    once): 18 right-hand sides, more than one solve takes, so the library solves the headings in
    chunks of ten and two; every block equals the oracle's."""
    L, ctx, _lib = lib
    import mhx
    m = mhx.meshes.cylinder(1.0, 1.5, 16, 6, 3)
    body = mhx.FloatingBody(m, ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"])
    names = body.dof_names
    jm = julia_mesh(m)
    dofs = packed_dofs(body.dofs, names)
    h = C.c_void_p()
    ws, bs = [0.8, 1.9], list(np.linspace(0.0, 2 * math.pi, 12, endpoint=False))
    print("glue replay: 6 dofs, 12 headings", flush=True)
    ctx.check(L.mhx_body_create(ctx.h, C.byref(mhxmesh(_lib, jm)), len(names), dp(dofs), C.byref(h)))
    try:
        got = body_coeffs(L, ctx, h, ws, bs, 0, 1, len(names))
    finally:
        L.mhx_body_destroy(h)
    ref = oracle.hydro_coeffs(oracle.Mesh.from_mhx(m), [body.dofs[k] for k in names], ws, bs, rho=RHO, g=G,
                              gf=oracle.GF_WU, direct=True)
    for key in KEYS:
        rel, elem, ok = close(got[key], ref[key])
        assert ok, (key, rel, elem)


def test_backend_ldiv_many_columns(lib):
    """_backend_ldiv with 40 right-hand sides (more than one solve takes): mhx_lu_solve solves
    them in column chunks of 16, 16 and 8; the result equals LAPACK's (A and A^H)."""
    L, ctx, _lib = lib
    rng = np.random.default_rng(7)
    n, k = 300, 40
    A = 0.5 * np.eye(n) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / (2 * math.sqrt(n))
    Bm = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    print("glue replay: backend_ldiv, 40 columns", flush=True)
    for trans in (0, 2):
        Af = np.asfortranarray(A)
        X = np.asfortranarray(Bm.copy())
        ctx.check(L.mhx_lu_solve(ctx.h, n, k, dp(Af), trans, dp(X)))
        ref = np.linalg.solve(A if trans == 0 else A.conj().T, Bm)
        assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 1e-12


def test_hydrodynamic_coefficients_lid_body(lib, oracle):
    """FloatingBody(HIPMesh(hull), HIPMesh(lid), dofs): hull + lid panels, dofs padded with zero
    rows on the lid, mhx_body_create_lid with the hull's panel count."""
    L, ctx, _lib = lib
    import mhx
    hull = mhx.meshes.cylinder(1.0, 1.0, 16, 4, 3)
    lid = mhx.meshes.disk_lid(1.0, 3, 16, z=-0.1)
    body = mhx.FloatingBody(hull, ["Heave", "Surge"])
    names = body.dof_names
    full = mhx.meshes.combine(hull, lid)
    jm = julia_mesh(full)
    dofs = packed_dofs(body.dofs, names, lid.nfaces)
    h = C.c_void_p()
    ws, bs = [1.1, 2.5], [0.0]
    print("glue replay: lid body", flush=True)
    ctx.check(L.mhx_body_create_lid(ctx.h, C.byref(mhxmesh(_lib, jm)), hull.nfaces, len(names), dp(dofs), C.byref(h)))
    try:
        got = body_coeffs(L, ctx, h, ws, bs, 0, 1, len(names))
    finally:
        L.mhx_body_destroy(h)
    om = oracle.Mesh.from_mhx(full)
    dl = [np.vstack([body.dofs[k], np.zeros((lid.nfaces, 3))]) for k in names]
    ref = oracle.hydro_coeffs(om, dl, ws, bs, rho=RHO, g=G, gf=oracle.GF_WU, direct=True, nhull=hull.nfaces)
    for key in KEYS:
        rel, elem, ok = close(got[key], ref[key])
        assert ok, (key, rel, elem)


def test_solve_problem_and_backend_ldiv(lib, oracle):
    """solve(prob, DirectBEM()): S, D from mhx_assemble into column-major ComplexF64 matrices, then
    mhx_solve(D, S, bc) (direct and indirect); _backend_ldiv: mhx_lu_solve."""
    L, ctx, _lib = lib
    import mhx
    m = mhx.meshes.sphere(1.0, 8, 8)
    jm = julia_mesh(m)
    n = m.nfaces
    k = 1.3 ** 2 / G
    S = np.zeros((n, n), dtype=np.complex128, order="F"); D = np.zeros_like(S)
    print("glue replay: assemble + solve", flush=True)
    ctx.check(L.mhx_assemble(ctx.h, C.byref(mhxmesh(_lib, jm)), k, 0, 1, 1, dp(S), dp(D)))
    So, Do = oracle.assemble_rows(oracle.Mesh.from_mhx(m), k, (0, n), gf=oracle.GF_WU, direct=True)
    for got, ref in ((S, So), (D, Do)):
        rel, elem, ok = close(got, ref)
        assert ok, (rel, elem)
    body = mhx.FloatingBody(m, ["Heave"])
    bc = np.asfortranarray((-1j * 1.3 * (m.normals * body.dofs["Heave"]).sum(1)).astype(np.complex128))
    phi = np.zeros(n, dtype=np.complex128); src = np.zeros(n, dtype=np.complex128)
    ctx.check(L.mhx_solve(ctx.h, n, 1, dp(D), dp(S), dp(bc), 1, dp(phi), dp(src)))
    ref = np.linalg.solve(Do, So @ bc)
    assert np.linalg.norm(phi - ref) <= 1e-10 * np.linalg.norm(ref)
    # indirect: sources = K \ bc, phi = S sources
    Kd = np.zeros_like(S); Si = np.zeros_like(S)
    ctx.check(L.mhx_assemble(ctx.h, C.byref(mhxmesh(_lib, jm)), k, 0, 0, 1, dp(Si), dp(Kd)))
    ctx.check(L.mhx_solve(ctx.h, n, 1, dp(Kd), dp(Si), dp(bc), 0, dp(phi), dp(src)))
    sref = np.linalg.solve(Kd, bc)
    assert np.linalg.norm(src - sref) <= 1e-10 * np.linalg.norm(sref)
    assert np.linalg.norm(phi - Si @ sref) <= 1e-10 * np.linalg.norm(Si @ sref)
    # _backend_ldiv(A::MhxMatrix, b): the solution overwrites the right-hand sides
    rng = np.random.default_rng(5)
    b = np.asfortranarray(rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3)))
    x = b.copy(order="F")
    ctx.check(L.mhx_lu_solve(ctx.h, n, 3, dp(D), 0, dp(x)))
    xr = np.linalg.solve(D, b)
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_frule_omega_direction(lib):
    """The frule's call: mhx_hydro_coeffs_jvp with only the frequency tangent set (the mesh tangent
    struct of NULL fields, ddofs NULL) against central differences of mhx_hydro_coeffs."""
    L, ctx, _lib = lib
    import mhx
    m = mhx.meshes.sphere(1.0, 8, 8)
    body = mhx.FloatingBody(m, ["Heave", "Surge"])
    names = body.dof_names
    jm = julia_mesh(m)
    dofs = packed_dofs(body.dofs, names)
    nd, ws, bs = len(names), np.array([0.9, 1.6]), np.array([0.0, math.pi / 4])
    nw, nb = len(ws), len(bs)

    def coeffs(w):
        A = np.zeros((nd, nd, nw), order="F"); B = np.zeros_like(A)
        Fd = np.zeros((nd, nw, nb), dtype=np.complex128, order="F"); Fk = np.zeros_like(Fd)
        ctx.check(L.mhx_hydro_coeffs(ctx.h, C.byref(mhxmesh(_lib, jm)), nd, dp(dofs), nw, dp(w), nb, dp(bs), 0, 1,
                                     MODE_BATCHED, dp(A), dp(B), dp(Fd), dp(Fk)))
        return A, B, Fd, Fk

    tan = _lib.CMeshTangent(None, None, None, None)
    dw = np.ones(nw)
    outs = [np.zeros((nd, nd, nw), order="F") for _ in range(4)]
    couts = [np.zeros((nd, nw, nb), dtype=np.complex128, order="F") for _ in range(4)]
    A, B, dA, dB = outs
    Fd, Fk, dFd, dFk = couts
    print("glue replay: frule", flush=True)
    ctx.check(L.mhx_hydro_coeffs_jvp(ctx.h, C.byref(mhxmesh(_lib, jm)), C.byref(tan), nd, dp(dofs), None, nw, dp(ws),
                                     dp(dw), nb, dp(bs), 0, 1, dp(A), dp(B), dp(Fd), dp(Fk), dp(dA), dp(dB), dp(dFd),
                                     dp(dFk)))
    h = 1e-5
    p = coeffs(ws + h)
    q = coeffs(ws - h)
    for got, a, b in zip((dA, dB, dFd, dFk), p, q):
        fd = (a - b) / (2 * h)
        assert np.linalg.norm(got - fd) <= 1e-6 * np.linalg.norm(fd), (np.linalg.norm(got - fd), np.linalg.norm(fd))


def test_forward_speed_call(lib):
    """hydrodynamic_coefficients with forward_speeds: mhx_body_coeffs_fs with dgrad = zeros(3, nd)
    in Julia (C: ndof x 3 row-major) filled for the radiating dofs; equals the Python mirror's call
    (itself checked against the oracle in the forward-speed tests)."""
    L, ctx, _lib = lib
    import mhx
    m = mhx.meshes.cylinder(1.0, 1.5, 16, 6, 3)
    body = mhx.FloatingBody(m, ["Heave", "Pitch"])
    names = body.dof_names
    jm = julia_mesh(m)
    dofs = packed_dofs(body.dofs, names)
    ws, bs, us = np.array([0.8, 1.5]), np.array([0.0, math.pi]), np.array([0.0, 1.5])
    nd, nw, nb, nu = len(names), len(ws), len(bs), len(us)
    dgrad_julia = np.zeros((3, nd), order="F")
    dgrad_julia[:, names.index("Pitch")] = [0.0, 0.0, 1.0]   # evaluate_gradient_of_motion(mesh, "Pitch")[1, :]
    h = C.c_void_p()
    print("glue replay: forward speed", flush=True)
    ctx.check(L.mhx_body_create(ctx.h, C.byref(mhxmesh(_lib, jm)), nd, dp(dofs), C.byref(h)))
    try:
        A = np.zeros((nd, nd, nw, nu, nb), order="F"); B = np.zeros_like(A)
        Fd = np.zeros((nd, nw, nb, nu), dtype=np.complex128, order="F"); Fk = np.zeros_like(Fd)
        ctx.check(L.mhx_body_coeffs_fs(ctx.h, h, nw, dp(ws), nb, dp(bs), nu, dp(us), dp(dgrad_julia), 0, 0, dp(A), dp(B),
                                       dp(Fd), dp(Fk)))
    finally:
        L.mhx_body_destroy(h)
    ref = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=ws, radiating_dofs=names, wave_directions=bs,
                                                   forward_speeds=us), mhx.IndirectBEM())
    for key, got in (("added_mass", A), ("radiation_damping", B), ("diffraction_force", Fd),
                     ("Froude_Krylov_force", Fk)):
        r = np.asarray(ref[key]).reshape(got.shape, order="A")
        rel, elem, ok = close(got, r, tol=1e-12)
        assert ok, (key, rel, elem)
