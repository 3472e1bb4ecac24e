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
    face properties); translated_x moves positions only; + joins (combine_meshes)."""
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


def _polygon_below(ys, zs):
    """Area of the planar polygon (ys, zs) below z = 0 and its chord length on z = 0, in the
    plane, by the shoelace formula on the polygon clipped against z <= 0."""
    pts = []
    n = len(ys)
    for e in range(n):
        p, q = (ys[e], zs[e]), (ys[(e + 1) % n], zs[(e + 1) % n])
        if p[1] <= 0:
            pts.append(p)
        if (p[1] < 0 < q[1]) or (q[1] < 0 < p[1]):
            t = p[1] / (p[1] - q[1])
            pts.append((p[0] + t * (q[0] - p[0]), 0.0))
    y = np.array([a for a, _ in pts]); z = np.array([b for _, b in pts])
    area = 0.5 * abs(np.dot(y, np.roll(z, -1)) - np.dot(z, np.roll(y, -1)))
    on = y[np.abs(z) == 0.0]
    return area, (on.max() - on.min()) if on.size >= 2 else 0.0


def test_horizontal_cylinder_is_a_closed_outward_prism():
    """mesh_horizontal_cylinder: the closed polygonal prism (regular ntheta-gon cross-section):
    sum n a = 0, volume (divergence theorem, sum z n_z a) = L x polygon area to rounding, every
    face radius within faces_max_radius, counts from the square-face rule."""
    R, L, rmax = 1.5, 2.5, 0.5
    m = meshes.mesh_horizontal_cylinder(length=L, radius=R, faces_max_radius=rmax)
    s = np.sqrt(2.0) * rmax
    nr, nth, nx = int(np.ceil(R / s)), int(np.ceil(2 * np.pi * R / s)), int(np.ceil(L / s))
    assert m.nfaces == nth * nx + 2 * nr * nth
    assert np.max(m.radii) <= rmax
    np.testing.assert_allclose((m.normals * m.areas[:, None]).sum(0), 0.0, atol=1e-12)
    poly = 0.5 * nth * R * R * np.sin(2 * np.pi / nth)
    assert np.sum(m.centers[:, 2] * m.normals[:, 2] * m.areas) == pytest.approx(L * poly, rel=1e-13)
    np.testing.assert_allclose(m.centers.mean(axis=0), 0.0, atol=1e-12)
    c = meshes.mesh_horizontal_cylinder(length=L, radius=R, center=(1.0, -2.0, -3.0), resolution=(2, 8, 3))
    assert c.nfaces == 8 * 3 + 2 * 2 * 8
    np.testing.assert_allclose(c.centers.mean(axis=0), (1.0, -2.0, -3.0), atol=1e-12)


@pytest.mark.parametrize("zc", [0.0, 0.37, -0.9, 0.8])
def test_waterline_clipping(zc):
    """immersed_part cuts the faces that cross z = 0: nothing is left above z = 0; the cut surface
    plus the waterplane closes (sum n a = -A_wp e_z, with A_wp = L x chord of the cross-section
    polygon at z = 0); the submerged volume is L x the polygon's area below z = 0 (both computed
    in the cross-section plane by the shoelace formula); the intersection points are shared by
    the faces on either side of a cut edge (no duplicate vertices)."""
    R, L, nth = 1.0, 3.0, 36
    m = meshes.mesh_horizontal_cylinder(length=L, radius=R, center=(0.0, 0.0, zc), resolution=(3, nth, 5))
    im = m.immersed_part()
    assert np.max(im.vertices[:, 2]) <= 0.0
    th = np.arange(nth) * (2 * np.pi / nth)
    area, chord = _polygon_below(R * np.cos(th), zc - R * np.sin(th))
    vol = np.sum(im.centers[:, 2] * im.normals[:, 2] * im.areas)
    assert vol == pytest.approx(L * area, rel=1e-12)
    na = (im.normals * im.areas[:, None]).sum(0)
    np.testing.assert_allclose(na[:2], 0.0, atol=1e-12)
    assert na[2] == pytest.approx(-L * chord, rel=1e-12, abs=1e-12)
    assert len(np.unique(np.round(im.vertices, 12), axis=0)) == im.nvertices
    assert im.nfaces > 0 and np.all(im.areas > 0)


def test_clipping_one_quad():
    """A unit square in the vertical plane x = 0 standing on one vertex, cut at the height of its
    centre: the part below is the lower half of the diamond (area 1/2) as one quadrangle-turned-
    triangle; with the cut lower, a small triangle; with one vertex above, a pentagon as a quad plus
    a triangle."""
    h = np.sqrt(0.5)
    V = np.array([[0, 0, -h], [0, h, 0.0], [0, 0, h], [0, -h, 0.0]], dtype=float)   # diamond, centre z = 0
    F = np.array([[1, 2, 3, 4]])
    for dz, area, nf in ((0.0, 0.5, 1), (0.5 * h, 0.125, 1), (-0.5 * h, 1.0 - 0.125, 2)):
        im = meshes.Mesh(V + [0, 0, dz], F).immersed_part()
        assert im.nfaces == nf
        assert im.areas.sum() == pytest.approx(area, rel=1e-13)
    # a vertical square with its top edge above the surface: a quadrangle
    sq = meshes.Mesh(np.array([[0, 0, -1], [0, 1, -1], [0, 1, 0.5], [0, 0, 0.5]], dtype=float), F).immersed_part()
    assert sq.nfaces == 1 and sq.faces[0, 0] != sq.faces[0, 3]
    assert sq.areas[0] == pytest.approx(1.0)
