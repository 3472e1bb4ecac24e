"""Host-side mesh logic: the reverse mode of the face properties (centres, normals, areas as
functions of the vertices) used to chain field-wise gradients through the geometry."""
import numpy as np

from mhx import meshes


def _fd(mesh, dV, h=1e-6):
    p = meshes.face_properties(mesh.vertices + h * dV, mesh.faces)
    m = meshes.face_properties(mesh.vertices - h * dV, mesh.faces)
    return [(a - b) / (2 * h) for a, b in zip(p[:3], m[:3])]


def test_face_properties_vjp_is_transpose_of_central_difference():
    """<vjp(gC, gN, gA), dV> = <(gC, gN, gA), d(face_properties)[dV]> on a mesh with triangles
    (poles) and quads, for random cotangents and a random vertex direction."""
    for mesh in (meshes.sphere(1.0, 6, 6), meshes.sphere(1.3, 8, 16), meshes.sphere_with_count(1.0, 256)):
        rng = np.random.default_rng(mesh.nfaces)
        gC, gN, gA = rng.standard_normal((mesh.nfaces, 3)), rng.standard_normal((mesh.nfaces, 3)), rng.standard_normal(mesh.nfaces)
        dV = rng.standard_normal(mesh.vertices.shape)
        dC, dN, dA = _fd(mesh, dV)
        rhs = np.sum(gC * dC) + np.sum(gN * dN) + np.sum(gA * dA)
        lhs = np.sum(meshes.face_properties_vjp(mesh.vertices, mesh.faces, gC, gN, gA) * dV)
        assert abs(lhs - rhs) <= 1e-8 * (abs(rhs) + np.sum(np.abs(gC * dC)) + np.sum(np.abs(gA * dA)))


def test_face_properties_scale_with_the_body():
    """Uniform scaling: centres scale by s, normals are unchanged, areas scale by s^2, so the
    tangent of the face properties along V is (C, 0, 2A) -- what the radius derivative uses."""
    m = meshes.sphere(1.0, 6, 6)
    dC, dN, dA = _fd(m, m.vertices)
    np.testing.assert_allclose(dC, m.centers, atol=1e-8)
    np.testing.assert_allclose(dN, 0.0, atol=1e-8)
    np.testing.assert_allclose(dA, 2 * m.areas, atol=1e-8)
