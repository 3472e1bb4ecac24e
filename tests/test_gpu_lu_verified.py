"""The dense LU with verified diagonal pivots (the default pivoting): matrices that keep LAPACK's
diagonal pivots are factored without the pivot search; a matrix that breaks LAPACK's rule
anywhere, here at a late panel, or that has an exactly zero pivot, is factored again with row
interchanges; the results equal LAPACK's either way.  Also: orders above 4096 (the pivoted
leaves' larger row tiles), a lid body whose BEM matrix needs row interchanges near its irregular
frequency, and the multi-body sweep's redo of such a (body, frequency) item."""
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
    ctx = M.context()
    M.set_rho(1025.0)
    M.set_g(9.81)
    yield M
    ctx.set_pivoting(0)
    ctx.set_lu_mode(0)
    M.set_rho(1000.0)


def _bem_like(rng, n):
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * np.eye(n) + G / n


def test_stats_and_late_pivot(mhx):
    """A batch of four order-2048 systems: two BEM-like (diagonal pivots kept), one whose rows
    1500 and 1900 are exchanged (LAPACK interchanges them back at column 1500, panel 23), one
    complex Gaussian: the two last ones take the pivot search; every solution equals LAPACK's."""
    from mhx import _lib
    ctx = mhx.context()
    rng = np.random.default_rng(7)
    n = 2048
    A = np.stack([_bem_like(rng, n) for _ in range(3)] + [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))])
    A[2][[1500, 1900]] = A[2][[1900, 1500]]
    B = rng.standard_normal((4, n, 5)) + 1j * rng.standard_normal((4, n, 5))
    ref = np.linalg.solve(A, B)
    for mode in (_lib.LU_FP64, _lib.LU_MIXED, _lib.LU_MIXED_BF16):
        print(f"verified pivots, lu mode {mode}", flush=True)
        ctx.set_lu_mode(mode)
        ctx.reset_timing()
        X = mhx.lu_solve_batched(A, B)
        diag, piv = ctx.pivot_stats()
        # FP64: exactly two of each; mixed precision: a system whose refinement does not converge is
        # solved again in FP64 (counted again)
        assert (diag, piv) == (2, 2) if mode == _lib.LU_FP64 else (diag >= 2 and piv >= 2), (mode, diag, piv)
        for q in range(4):
            assert np.linalg.norm(X[q] - ref[q]) <= 1e-10 * np.linalg.norm(ref[q]), (mode, q)
    ctx.set_lu_mode(_lib.LU_AUTO)
    # the same batch with the pivot search on every matrix
    ctx.set_pivoting(_lib.PIVOT_ALWAYS)
    ctx.reset_timing()
    X = mhx.lu_solve_batched(A, B)
    assert ctx.pivot_stats() == (0, 4)
    for q in range(4):
        assert np.linalg.norm(X[q] - ref[q]) <= 1e-10 * np.linalg.norm(ref[q])
    ctx.set_pivoting(_lib.PIVOT_VERIFIED)


def test_zero_pivot_falls_back(mhx):
    """A zero in the first diagonal position (the matrix is regular with the interchange): the
    verified pass flags it and the pivoted factorization solves it; an exactly singular matrix is
    reported as such."""
    rng = np.random.default_rng(8)
    n = 300
    A = _bem_like(rng, n)
    A[0, 0] = 0.0
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    X = mhx.lu_solve(A, B)
    assert np.linalg.norm(A @ X - B) <= 1e-12 * np.linalg.norm(B)
    Z = np.zeros((70, 70), dtype=np.complex128)
    with pytest.raises(mhx.SingularException):
        mhx.lu_solve(Z, np.ones(70))


@pytest.mark.parametrize("n", [4097, 6000])
def test_orders_above_4096(mhx, n):
    """Orders above 4096 in FP64 (verified and pivoted) and both mixed precisions against LAPACK."""
    from mhx import _lib
    ctx = mhx.context()
    rng = np.random.default_rng(n)
    A = np.stack([_bem_like(rng, n), rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))])
    B = rng.standard_normal((2, n, 3)) + 1j * rng.standard_normal((2, n, 3))
    print(f"order {n}: LAPACK", flush=True)
    ref = np.linalg.solve(A, B)
    try:
        for mode in (_lib.LU_FP64, _lib.LU_MIXED, _lib.LU_MIXED_BF16):
            print(f"order {n}: lu mode {mode}", flush=True)
            ctx.set_lu_mode(mode)
            X = mhx.lu_solve_batched(A, B)
            for q in range(2):
                assert np.linalg.norm(X[q] - ref[q]) <= 1e-10 * np.linalg.norm(ref[q]), (mode, q)
    finally:
        ctx.set_lu_mode(_lib.LU_AUTO)


def _lid_body(mhx):
    from mhx import meshes
    hull = meshes.cylinder(1.0, 2.0, 32, 8, 4)
    lid = meshes.disk_lid(1.0, 4, 32, z=-0.25)
    return mhx.FloatingBody(hull, ["Heave", "Surge", "Pitch"], lid_mesh=lid), hull, lid


def test_lid_body_needing_interchanges(mhx, oracle):
    """A cylinder (R = 1, draft 2, 256 panels) with a 256-panel lid at z = -0.25: near its irregular
    frequency LAPACK interchanges rows (omega = 4.6, 4.7); the device path takes the pivot search for
    those frequencies and matches the oracle's lid path (1e-9), and keeps the diagonal pivots at
    omega = 1."""
    body, hull, lid = _lid_body(mhx)
    from mhx import meshes
    ws = [1.0, 4.6, 4.7]
    ctx = mhx.context()
    ctx.reset_timing()
    out = mhx.hydrodynamic_coefficients(body, dict(wave_frequencies=ws, radiating_dofs=body.dof_names,
                                                   wave_directions=[0.0, math.pi / 3]), mhx.DirectBEM())
    diag, piv = ctx.pivot_stats()
    assert piv >= 2 and diag >= 1, (diag, piv)
    full = meshes.combine(hull, lid)
    om = oracle.Mesh.from_mhx(full)
    dofs = [np.vstack([body.dofs[nm], np.zeros((lid.nfaces, 3))]) for nm in body.dof_names]
    ref = oracle.hydro_coeffs(om, dofs, ws, [0.0, math.pi / 3], rho=1025.0, g=9.81, gf=oracle.GF_WU, direct=True,
                              nhull=hull.nfaces)
    for key in KEYS:
        rel, elem, ok = close(np.asarray(out[key]).reshape(np.asarray(ref[key]).shape), ref[key])
        assert ok, (key, rel, elem)


def test_sweep_redoes_items_needing_interchanges(mhx):
    """The lid body in a multi-body sweep (eager and captured graph) beside a plain sphere: the
    frequencies whose factorization breaks LAPACK's rule are solved again with row interchanges
    after the sweep, and every body's coefficients equal the single-body path's (1e-11)."""
    from mhx import meshes
    body, hull, lid = _lid_body(mhx)
    sph = mhx.FloatingBody(meshes.sphere(1.0, 8, 16), ["Heave", "Surge"])
    jobs = [(body, dict(wave_frequencies=[1.0, 4.6, 4.7, 2.0], radiating_dofs=body.dof_names, wave_directions=[0.0])),
            (sph, dict(wave_frequencies=[0.5, 1.5], radiating_dofs=sph.dof_names, wave_directions=[0.0]))]
    form = mhx.DirectBEM()
    ctx = mhx.context()
    for graph in (False, True, True):
        print(f"sweep redo, graph={graph}", flush=True)
        ctx.reset_timing()
        res = mhx.hydrodynamic_coefficients_many(jobs, form, graph=graph)
        diag, piv = ctx.pivot_stats()
        assert piv >= 2 and diag >= 4, (diag, piv)
        for (b, par), got in zip(jobs, res):
            ref = mhx.hydrodynamic_coefficients(b, par, form)
            for key in KEYS:
                rel, elem, ok = close(got[key], ref[key], tol=1e-11)
                assert ok, (graph, key, rel, elem)
