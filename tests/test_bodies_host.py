"""Host-side body logic: combining floating bodies into one coupled array
(src/bodies.jl:138-225) and combining meshes (src/meshes.jl:60-98)."""
import numpy as np

import mhx
from mhx import meshes


def test_unique_body_names():
    """make_body_name_list_unique: repeats get _2, _3 and skip names already taken."""
    assert mhx.make_body_name_list_unique(["a", "b", "a", "a"]) == ["a", "b", "a_2", "a_3"]
    assert mhx.make_body_name_list_unique(["a", "a", "a_2"]) == ["a", "a_3", "a_2"]


def test_combined_body_layout():
    s1 = meshes.sphere(1.0, 6, 6)
    s2 = meshes.translated(meshes.sphere(0.5, 8, 16), (5.0, 0.0, 0.0))
    b1 = mhx.FloatingBody(s1, ["Heave", "Pitch"], body_name="float")
    b2 = mhx.FloatingBody(s2, ["Heave"], body_name="float", lid_mesh=meshes.disk_lid(0.5, 2, 8, z=-0.1))
    c = b1 + b2
    n1, n2 = s1.nfaces, s2.nfaces
    assert c.mesh.nfaces == n1 + n2 and c.mesh.nvertices == s1.nvertices + s2.nvertices
    assert c.dof_names == ["float__Heave", "float__Pitch", "float_2__Heave"]
    np.testing.assert_array_equal(c.dofs["float__Pitch"][:n1], b1.dofs["Pitch"])
    assert not c.dofs["float__Pitch"][n1:].any()
    np.testing.assert_array_equal(c.dofs["float_2__Heave"][n1:], b2.dofs["Heave"])
    assert not c.dofs["float_2__Heave"][:n1].any()
    np.testing.assert_array_equal(c.mesh.faces[n1:], s2.faces + s1.nvertices)
    assert c.lid_mesh is not None and c.lid_mesh.nfaces == b2.lid_mesh.nfaces
    assert c.body_name == "float+float_2"
