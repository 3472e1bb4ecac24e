"""Host-side mesh logic: the reverse mode of the face properties (centres, normals, areas as
functions of the vertices) used to chain field-wise gradients through the geometry."""
import numpy as np
import pytest

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


# ---------------------------------------------------------------------------- Capytaine-compatible I/O
def test_capytaine_fields_round_trip(tmp_path):
    """to_capytaine / from_capytaine (Mesh(PyObject), src/meshes.jl:122-143): 0-based faces under
    Capytaine's names, every field preserved; the .npz file round trip likewise."""
    m = meshes.sphere(1.3, 8, 16)
    d = m.to_capytaine()
    assert d["faces"].min() == 0 and d["nb_faces"] == m.nfaces and d["nb_vertices"] == m.nvertices
    np.testing.assert_array_equal(d["faces"] + 1, m.faces)
    for back in (meshes.Mesh.from_capytaine(d), meshes.Mesh.load(_saved(m, tmp_path))):
        for f in ("vertices", "faces", "centers", "normals", "areas", "radii"):
            np.testing.assert_array_equal(getattr(back, f), getattr(m, f))


def _saved(m, tmp_path):
    p = tmp_path / "mesh.npz"
    m.save(p)
    return p


def test_capytaine_object_fields_taken_as_given(tmp_path):
    """An object with Capytaine's attributes (what PyCall hands the reference): its stored face
    properties are used, not recomputed, as the reference's Mesh(PyObject) does; a file with only
    vertices and 0-based faces gets the face properties computed by Capytaine's rules."""
    from types import SimpleNamespace
    m = meshes.sphere(1.0, 6, 6)
    d = m.to_capytaine()
    d["faces_radiuses"] = d["faces_radiuses"] * 2.0  # a value no recomputation would give
    obj = SimpleNamespace(**d)
    back = meshes.Mesh.from_capytaine(obj)
    np.testing.assert_array_equal(back.radii, 2.0 * m.radii)
    p = tmp_path / "vf.npz"
    np.savez(p, vertices=m.vertices, faces=m.faces - 1)
    vf = meshes.Mesh.load(p)
    for f in ("centers", "normals", "areas", "radii"):
        np.testing.assert_allclose(getattr(vf, f), getattr(m, f), rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        meshes.Mesh.from_capytaine(dict(d, faces=d["faces"][:, :3]))


def test_capytaine_meshers():
    """mesh_sphere(...).immersed_part() is the generator's notebook sphere (same vertices, faces and
    face properties); translated_x moves positions only; + joins (combine_meshes); a face across
    z = 0 is refused."""
    a = meshes.mesh_sphere(1.0, resolution=(6, 6)).immersed_part()
    b = meshes.sphere(1.0, 6, 6)
    for f in ("vertices", "faces", "centers", "normals", "areas", "radii"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    c = meshes.mesh_sphere(2.0, center=(0.5, -1.0, 0.0), resolution=(10, 12)).immersed_part()
    assert c.nfaces == 60 and np.max(c.vertices[:, 2]) == 0.0
    np.testing.assert_allclose(c.centers.mean(axis=0)[:2], (0.5, -1.0), atol=1e-12)
    t = a.translated_x(3.0)
    np.testing.assert_array_equal(t.centers[:, 0], a.centers[:, 0] + 3.0)
    np.testing.assert_array_equal(t.normals, a.normals)
    j = a + t
    assert j.nfaces == 2 * a.nfaces and j.nvertices == 2 * a.nvertices
    np.testing.assert_array_equal(j.faces[a.nfaces:], a.faces + a.nvertices)
    with pytest.raises(ValueError, match="crosses"):
        meshes.mesh_sphere(1.0, center=(0.0, 0.0, 0.1), resolution=(6, 6)).immersed_part()
