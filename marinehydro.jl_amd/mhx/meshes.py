"""Panel meshes in the reference's StaticArraysMesh layout (src/meshes.jl:42-51).

A `Mesh` holds, as contiguous numpy arrays:
  vertices (nv, 3) float64, faces (nf, 4) int64 **1-based**, centers (nf, 3),
  normals (nf, 3), areas (nf,), radii (nf,)
i.e. exactly the memory of Julia's Vector{SVector{3,Float64}} / Vector{SVector{4,Int}},
so the C ABI receives the reference's own arrays without a copy.

Face properties follow the Capytaine conventions the reference imports
(src/meshes.jl:122-142, faces_centers / faces_normals / faces_areas /
faces_radiuses): a face whose first and last vertex indices coincide is a
triangle; quads are split into triangles (0,1,2) and (0,2,3) for area and
area-weighted centre, the quad normal is (v2-v0) x (v3-v1), the radius is the
largest centre-to-vertex distance.

The generators build bodies of revolution from a meridian profile the way
Capytaine's mesh_sphere does (profile from the bottom pole upwards, rotated
counter-clockwise about z, face (i, i+n, i+n+1, i+1)), so the reference's
notebook sphere (radius 1, resolution (6, 6), immersed part, 18 panels) is
reproduced vertex for vertex.
"""
import numpy as np


class Mesh:
    def __init__(self, vertices, faces, centers=None, normals=None, areas=None, radii=None):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.int64).reshape(-1, 4)
        if f.size and f.min() < 1:
            raise ValueError("faces must be 1-based (StaticArraysMesh convention)")
        if f.size and f.max() > self.vertices.shape[0]:
            raise ValueError("face index beyond the vertex count")
        self.faces = f
        if centers is None:
            centers, normals, areas, radii = face_properties(self.vertices, self.faces)
        self.centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        self.normals = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        self.areas = np.ascontiguousarray(areas, dtype=np.float64).reshape(-1)
        self.radii = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        nf = self.faces.shape[0]
        for name, arr, shape in (("centers", self.centers, (nf, 3)), ("normals", self.normals, (nf, 3)),
                                 ("areas", self.areas, (nf,)), ("radii", self.radii, (nf,))):
            if arr.shape != shape:
                raise ValueError(f"{name} must have shape {shape}, got {arr.shape}")

    @property
    def nfaces(self):
        return self.faces.shape[0]

    @property
    def nvertices(self):
        return self.vertices.shape[0]

    nf = nfaces

    def copy(self):
        return Mesh(self.vertices.copy(), self.faces.copy(), self.centers.copy(), self.normals.copy(),
                    self.areas.copy(), self.radii.copy())

    def scaled(self, s):
        """Uniform scaling of every field, as the reference's radius-derivative tests do."""
        return Mesh(self.vertices * s, self.faces, self.centers * s, self.normals, self.areas * s * s, self.radii * s)

    # ---- Capytaine-compatible I/O (src/meshes.jl:122-143): 0-based faces, Capytaine field names ----
    CAPYTAINE_FIELDS = ("vertices", "faces", "faces_centers", "faces_normals", "faces_areas", "faces_radiuses")

    @classmethod
    def from_capytaine(cls, mesh):
        """Mesh(PyObject) / StaticArraysMesh(PyObject) (src/meshes.jl:122-143): a Capytaine mesh, or
        any object or dict with its fields (vertices, 0-based faces, faces_centers, faces_normals,
        faces_areas, faces_radiuses), taken as given; only the faces move to 1-based indices."""
        get = mesh.__getitem__ if isinstance(mesh, dict) else (lambda k: getattr(mesh, k))
        f = np.asarray(get("faces"), dtype=np.int64)
        if f.ndim != 2 or f.shape[1] != 4:
            raise ValueError("faces must be (nf, 4): Capytaine stores triangles with the last index repeated")
        if f.size and f.min() < 0:
            raise ValueError("Capytaine faces are 0-based and non-negative")
        return cls(np.asarray(get("vertices"), dtype=np.float64), f + 1, get("faces_centers"), get("faces_normals"),
                   get("faces_areas"), get("faces_radiuses"))

    def to_capytaine(self):
        """The fields under Capytaine's names, faces 0-based (what Mesh(PyObject) reads)."""
        return {"vertices": self.vertices.copy(), "faces": self.faces - 1, "faces_centers": self.centers.copy(),
                "faces_normals": self.normals.copy(), "faces_areas": self.areas.copy(),
                "faces_radiuses": self.radii.copy(), "nb_vertices": self.nvertices, "nb_faces": self.nfaces}

    def save(self, path):
        """Write the Capytaine fields to an .npz (numeric arrays only; 0-based faces), e.g. for a mesh
        exported from Capytaine elsewhere with np.savez(path, **{k: getattr(m, k) for k in
        Mesh.CAPYTAINE_FIELDS})."""
        d = self.to_capytaine()
        np.savez(path, **{k: d[k] for k in self.CAPYTAINE_FIELDS})

    @classmethod
    def load(cls, path):
        """Read an .npz written by save() or by the np.savez line above (no pickled objects)."""
        with np.load(path, allow_pickle=False) as z:
            missing = [k for k in ("vertices", "faces") if k not in z.files]
            if missing:
                raise ValueError(f"{path}: missing {missing}")
            d = {k: z[k] for k in z.files}
        if all(k in d for k in cls.CAPYTAINE_FIELDS):
            return cls.from_capytaine(d)
        # vertices and faces only: face properties computed here (Capytaine's rules)
        return cls(d["vertices"], np.asarray(d["faces"], dtype=np.int64) + 1)

    # ---- Capytaine mesh methods used by the reference's tests and paper scripts ----
    def __add__(self, other):
        """mesh1 + mesh2 (src/meshes.jl:94-96, Capytaine's join): vertices and faces concatenated."""
        return combine(self, other)

    def translated(self, shift):
        return translated(self, shift)

    def translated_x(self, d):
        return translated(self, (d, 0.0, 0.0))

    def translated_y(self, d):
        return translated(self, (0.0, d, 0.0))

    def translated_z(self, d):
        return translated(self, (0.0, 0.0, d))

    def immersed_part(self, tol=1e-12):
        """Capytaine's keep_immersed_part (the reference's meshes are the immersed parts of
        Capytaine meshes: test/consistency_with_Capytaine.jl): the faces at or below the free
        surface z = 0, with every face that crosses it cut at the waterline.  Vertices within
        tol * (mesh size) of z = 0 are put on it.  A cut face keeps its part below z = 0 (polygon
        clipping against the half-space z <= 0): a quadrangle with one vertex above becomes a
        pentagon, split into a quadrangle and a triangle; with two above, a quadrangle; with
        three, a triangle (triangles stored as (a, b, c, a)).  The points where an edge meets
        z = 0 are shared by the faces on both sides of the edge.  Faces with no vertex strictly
        below z = 0 are dropped; unused vertices are dropped (the others keep their order)."""
        V = self.vertices.copy()
        scale = max(1.0, float(np.max(np.abs(V))) if V.size else 1.0)
        V[np.abs(V[:, 2]) <= tol * scale, 2] = 0.0
        verts = [tuple(v) for v in V]
        cut = {}   # (i, j) sorted original vertex pair -> new vertex (0-based)

        def cross_point(i, j):
            key = (min(i, j), max(i, j))
            if key not in cut:
                p, q = V[key[0]], V[key[1]]
                t = p[2] / (p[2] - q[2])
                x = p + t * (q - p)
                x[2] = 0.0
                cut[key] = len(verts)
                verts.append(tuple(x))
            return cut[key]

        faces = []
        for f in self.faces - 1:
            poly = [int(f[0])]
            for v in f[1:]:
                if int(v) != poly[-1]:
                    poly.append(int(v))
            if len(poly) > 1 and poly[-1] == poly[0]:
                poly.pop()                      # triangle stored as (a, b, c, a)
            z = V[poly, 2]
            if not np.any(z < 0.0):
                continue                        # at or above the surface
            if not np.any(z > 0.0):
                faces.append(list(f))           # wholly immersed: unchanged
                continue
            out = []
            for e in range(len(poly)):
                i, j = poly[e], poly[(e + 1) % len(poly)]
                if V[i, 2] <= 0.0:
                    out.append(i)
                if (V[i, 2] < 0.0 < V[j, 2]) or (V[j, 2] < 0.0 < V[i, 2]):
                    out.append(cross_point(i, j))
            if len(out) == 3:
                faces.append([out[0], out[1], out[2], out[0]])
            elif len(out) == 4:
                faces.append(out)
            elif len(out) == 5:
                faces.append(out[:4])
                faces.append([out[0], out[3], out[4], out[0]])
            else:
                raise ValueError(f"immersed_part: unexpected cut polygon with {len(out)} vertices")
        if not faces:
            return Mesh(np.zeros((0, 3)), np.zeros((0, 4), dtype=np.int64))
        F = np.asarray(faces, dtype=np.int64)
        Vn = np.asarray(verts, dtype=np.float64)
        used = np.unique(F)
        remap = np.full(Vn.shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        return Mesh(Vn[used], remap[F] + 1)

    keep_immersed_part = immersed_part


def face_properties(vertices, faces):
    """Capytaine's compute_faces_properties on 1-based faces."""
    V = vertices
    F = faces - 1
    nf = F.shape[0]
    centers = np.zeros((nf, 3))
    normals = np.zeros((nf, 3))
    areas = np.zeros(nf)
    tri = F[:, 0] == F[:, 3]
    quad = ~tri
    t = F[tri]
    if t.size:
        cr = np.cross(V[t[:, 1]] - V[t[:, 0]], V[t[:, 2]] - V[t[:, 0]])
        ta = np.linalg.norm(cr, axis=1)
        normals[tri] = cr / ta[:, None]
        areas[tri] = ta / 2.0
        centers[tri] = np.sum(V[t[:, :3]], axis=1) / 3.0
    q = F[quad]
    if q.size:
        qn = np.cross(V[q[:, 2]] - V[q[:, 0]], V[q[:, 3]] - V[q[:, 1]])
        normals[quad] = qn / np.linalg.norm(qn, axis=1)[:, None]
        a1 = np.linalg.norm(np.cross(V[q[:, 1]] - V[q[:, 0]], V[q[:, 2]] - V[q[:, 0]]), axis=1) * 0.5
        a2 = np.linalg.norm(np.cross(V[q[:, 3]] - V[q[:, 0]], V[q[:, 2]] - V[q[:, 0]]), axis=1) * 0.5
        areas[quad] = a1 + a2
        c1 = np.sum(V[q[:, :3]], axis=1) / 3.0
        c2 = (np.sum(V[q[:, 2:4]], axis=1) + V[q[:, 0]]) / 3.0
        centers[quad] = (a1[:, None] * c1 + a2[:, None] * c2) / areas[quad][:, None]
    rv = centers[None, :, :] - np.moveaxis(V[F], 0, 1)
    radii = np.max(np.linalg.norm(rv, axis=2), axis=0)
    return centers, normals, areas, radii


def face_properties_vjp(vertices, faces, g_centers=None, g_normals=None, g_areas=None):
    """Pull the cotangents of the face properties (centres, normals, areas of face_properties)
    back to the vertices: returns dJ/dV (nv, 3) for J depending on V only through them.  The
    radii only select a far-field rule (no derivative).  Reverse mode of the formulas above:
    normalisation n = x/|x| gives xbar = (nbar - (nbar . n) n)/|x|, a cross product x = u x w
    gives ubar = w x xbar and wbar = xbar x u, and the area-weighted centre
    C = (a1 c1 + a2 c2)/(a1 + a2) gives c1bar = Cbar a1/A and a1bar = Cbar . (c1 - C)/A."""
    V = np.asarray(vertices, dtype=np.float64)
    F = np.asarray(faces) - 1
    nf = F.shape[0]
    gC = np.zeros((nf, 3)) if g_centers is None else np.asarray(g_centers, dtype=np.float64)
    gN = np.zeros((nf, 3)) if g_normals is None else np.asarray(g_normals, dtype=np.float64)
    gA = np.zeros(nf) if g_areas is None else np.asarray(g_areas, dtype=np.float64)
    out = np.zeros_like(V)

    def unit_bar(x, nbar):
        nx = np.linalg.norm(x, axis=1)[:, None]
        n = x / nx
        return (nbar - np.sum(nbar * n, 1)[:, None] * n) / nx

    def cross_bar(u, w, xbar):  # x = u x w
        return np.cross(w, xbar), np.cross(xbar, u)

    tri = F[:, 0] == F[:, 3]
    t, q = F[tri], F[~tri]
    if t.size:
        e1, e2 = V[t[:, 1]] - V[t[:, 0]], V[t[:, 2]] - V[t[:, 0]]
        cr = np.cross(e1, e2)
        ta = np.linalg.norm(cr, axis=1)[:, None]
        crb = unit_bar(cr, gN[tri]) + (0.5 * gA[tri])[:, None] * cr / ta
        b1, b2 = cross_bar(e1, e2, crb)
        gc = gC[tri] / 3.0
        np.add.at(out, t[:, 0], gc - b1 - b2)
        np.add.at(out, t[:, 1], gc + b1)
        np.add.at(out, t[:, 2], gc + b2)
    if q.size:
        V0, V1, V2, V3 = (V[q[:, k]] for k in range(4))
        d1, d2 = V2 - V0, V3 - V1
        qb = unit_bar(np.cross(d1, d2), gN[~tri])
        b1, b2 = cross_bar(d1, d2, qb)
        np.add.at(out, q[:, 2], b1)
        np.add.at(out, q[:, 0], -b1)
        np.add.at(out, q[:, 3], b2)
        np.add.at(out, q[:, 1], -b2)
        x1 = np.cross(V1 - V0, V2 - V0)
        x2 = np.cross(V3 - V0, V2 - V0)
        a1 = 0.5 * np.linalg.norm(x1, axis=1)
        a2 = 0.5 * np.linalg.norm(x2, axis=1)
        A = a1 + a2
        c1 = (V0 + V1 + V2) / 3.0
        c2 = (V2 + V3 + V0) / 3.0
        C = (a1[:, None] * c1 + a2[:, None] * c2) / A[:, None]
        Cb = gC[~tri]
        c1b = Cb * (a1 / A)[:, None]
        c2b = Cb * (a2 / A)[:, None]
        a1b = gA[~tri] + np.sum(Cb * (c1 - C), 1) / A
        a2b = gA[~tri] + np.sum(Cb * (c2 - C), 1) / A
        x1b = (0.5 * a1b / np.linalg.norm(x1, axis=1))[:, None] * x1
        x2b = (0.5 * a2b / np.linalg.norm(x2, axis=1))[:, None] * x2
        u1, w1 = cross_bar(V1 - V0, V2 - V0, x1b)
        u2, w2 = cross_bar(V3 - V0, V2 - V0, x2b)
        np.add.at(out, q[:, 1], u1 + c1b / 3.0)
        np.add.at(out, q[:, 2], w1 + w2 + (c1b + c2b) / 3.0)
        np.add.at(out, q[:, 3], u2 + c2b / 3.0)
        np.add.at(out, q[:, 0], -(u1 + w1 + u2 + w2) + (c1b + c2b) / 3.0)
    return out


def combine_meshes(meshes):
    """combine_meshes (src/meshes.jl:60-98): vertices concatenated, each mesh's faces shifted by
    the vertex count before it, face properties concatenated."""
    meshes = list(meshes)
    offs = np.cumsum([0] + [m.vertices.shape[0] for m in meshes[:-1]])
    return Mesh(np.concatenate([m.vertices for m in meshes]),
                np.concatenate([m.faces + o for m, o in zip(meshes, offs)]),
                np.concatenate([m.centers for m in meshes]), np.concatenate([m.normals for m in meshes]),
                np.concatenate([m.areas for m in meshes]), np.concatenate([m.radii for m in meshes]))


def combine(mesh1, mesh2):
    """mesh1 + mesh2 (src/meshes.jl:94-96)."""
    return combine_meshes([mesh1, mesh2])


def translated(mesh, shift):
    """The mesh moved by a constant vector (every field that is a position moves)."""
    s = np.asarray(shift, dtype=np.float64)[None, :]
    return Mesh(mesh.vertices + s, mesh.faces, mesh.centers + s, mesh.normals, mesh.areas, mesh.radii)


def disk_lid(radius, nrings, nsectors, z=0.0):
    """A horizontal disk at height z (the lid of a vertical cylinder or sphere cut at z): nrings
    rings x nsectors sectors, the centre ring triangles, normals pointing up (+z), as Capytaine's
    generate_lid orients a lid."""
    rs = np.linspace(0.0, radius, nrings + 1)
    # profile from the rim inwards so that the revolution's outward orientation points up
    prof = [(r, z) for r in rs[::-1]]
    m = revolution(prof, nsectors)
    if np.mean(m.normals[:, 2]) < 0:
        m = Mesh(m.vertices, m.faces[:, ::-1])
    return m


def rect_lid(length, width, nx, ny, z=0.0):
    """A horizontal rectangle centred on the z axis at height z, nx x ny quads, normals up."""
    xs = np.linspace(-length / 2, length / 2, nx + 1)
    ys = np.linspace(-width / 2, width / 2, ny + 1)
    verts = [(x, y, z) for y in ys for x in xs]
    faces = []
    for b in range(ny):
        for a in range(nx):
            v0 = b * (nx + 1) + a + 1
            faces.append([v0, v0 + 1, v0 + nx + 2, v0 + nx + 1])
    return Mesh(np.array(verts, dtype=np.float64), np.array(faces, dtype=np.int64))


def revolution(profile_rz, nphi, keep=None):
    """Rotate a meridian profile [(r0, z0), ...] (bottom to top) about the z axis.

    Points with r == 0 are on the axis and are merged, turning the adjacent faces
    into triangles stored as (a, b, c, a).  Face (i, i+n, i+n+1, i+1) orders
    vertices so that the normal points away from the axis for a profile that
    climbs from the bottom pole.  `keep(i_profile_segment)` filters segments.
    """
    prof = np.asarray(profile_rz, dtype=np.float64)
    npr = prof.shape[0]
    dphi = 2 * np.pi / nphi
    verts = []
    index = {}

    def vid(ip, jphi):
        r, z = prof[ip]
        key = (ip, 0) if r == 0.0 else (ip, jphi % nphi)
        if key not in index:
            ang = (jphi % nphi) * dphi
            index[key] = len(verts) + 1
            verts.append((r * np.cos(ang), r * np.sin(ang), z))
        return index[key]

    faces = []
    for j in range(nphi):
        for i in range(npr - 1):
            if keep is not None and not keep(i):
                continue
            f = [vid(i, j), vid(i, j + 1), vid(i + 1, j + 1), vid(i + 1, j)]
            if f[0] == f[1]:        # bottom pole: (P, P, a, b) -> (P, a, b, P)
                f = [f[0], f[2], f[3], f[0]]
            elif f[2] == f[3]:      # top pole: (a, b, P, P) -> (P, a, b, P)
                f = [f[2], f[0], f[1], f[2]]
            faces.append(f)
    return Mesh(np.array(verts), np.array(faces, dtype=np.int64))


def sphere(radius=1.0, ntheta=10, nphi=10, immersed=True):
    """Capytaine-style UV sphere: ntheta meridian segments (pole to pole), nphi sectors.
    With `immersed` only the segments below z=0 are kept (ntheta even)."""
    theta = np.linspace(0.0, np.pi, ntheta + 1)
    prof = np.stack([radius * np.sin(theta), -radius * np.cos(theta)], axis=1)
    prof[0, 0] = 0.0
    prof[-1, 0] = 0.0
    if immersed:
        if ntheta % 2:
            raise ValueError("immersed sphere needs an even ntheta")
        prof = prof[: ntheta // 2 + 1]
        prof[-1, 1] = 0.0
    return revolution(prof, nphi)


def mesh_sphere(radius=1.0, center=(0.0, 0.0, 0.0), resolution=(10, 10)):
    """Capytaine's mesh_sphere(radius=, center=, resolution=(ntheta, nphi)) (the mesher of the
    reference's tests and notebooks): the full UV sphere; .immersed_part() keeps the wetted half."""
    ntheta, nphi = resolution
    m = sphere(radius, ntheta, nphi, immersed=False)
    return translated(m, center) if any(c != 0.0 for c in center) else m


def mesh_horizontal_cylinder(length=10.0, radius=1.0, center=(0.0, 0.0, 0.0), resolution=(2, 8, 10),
                             faces_max_radius=None):
    """Capytaine's mesh_horizontal_cylinder (the mesher of the reference's horizontal-cylinder
    tests, test/consistency_with_Capytaine.jl:211-216, 313-326): a closed cylinder along the x
    axis, resolution = (nr, ntheta, nx) panels along a radius of the end disks, around the
    circumference and along the length.  With faces_max_radius the counts are the smallest that
    keep every face's radius (largest centre-to-vertex distance) at or below it, taking square
    faces of side sqrt(2) faces_max_radius: nr = ceil(R / s), ntheta = ceil(2 pi R / s),
    nx = ceil(L / s).  The circumference starts at the +y side (theta measured from +y towards
    -z), so that an even ntheta puts a vertex row on each side of the waterline plane through
    the axis.  Normals point out of the body; each end disk's centre ring is triangles.
    .immersed_part() keeps the wetted part."""
    R, L = float(radius), float(length)
    if faces_max_radius is not None:
        s = np.sqrt(2.0) * float(faces_max_radius)
        nr, nth, nx = (max(1, int(np.ceil(R / s))), max(3, int(np.ceil(2 * np.pi * R / s))),
                       max(1, int(np.ceil(L / s))))
    else:
        nr, nth, nx = (int(r) for r in resolution)
    th = np.arange(nth) * (2 * np.pi / nth)
    ys, zs = R * np.cos(th), -R * np.sin(th)          # theta = 0 at +y, increasing towards -z
    xs = np.linspace(-L / 2, L / 2, nx + 1)
    verts, faces = [], []

    def add(v):
        verts.append(v)
        return len(verts)

    side = [[add((x, ys[t], zs[t])) for t in range(nth)] for x in xs]
    for a in range(nx):
        for t in range(nth):
            t1 = (t + 1) % nth
            # outward: (t1 - t) is along -theta x ... ordering fixed by the normal check below
            faces.append([side[a][t], side[a][t1], side[a + 1][t1], side[a + 1][t]])
    for end, x in ((0, xs[0]), (1, xs[-1])):
        rings = [side[0 if end == 0 else nx]]
        for k in range(nr - 1, 0, -1):
            rr = R * k / nr
            rings.append([add((x, rr * np.cos(t), -rr * np.sin(t))) for t in th])
        c = add((x, 0.0, 0.0))
        for k in range(len(rings)):
            outer = rings[k]
            inner = rings[k + 1] if k + 1 < len(rings) else None
            for t in range(nth):
                t1 = (t + 1) % nth
                if inner is None:
                    faces.append([c, outer[t], outer[t1], c])
                else:
                    faces.append([outer[t], outer[t1], inner[t1], inner[t]])
    m = Mesh(np.asarray(verts, dtype=np.float64), np.asarray(faces, dtype=np.int64))
    # orient every face outwards: lateral faces away from the axis, end faces along +-x
    Fo = m.faces.copy()
    cen = m.centers
    radial = np.stack([np.zeros(len(cen)), cen[:, 1], cen[:, 2]], axis=1)
    is_end = np.isclose(np.abs(cen[:, 0]), L / 2) & np.isclose(np.abs(m.normals[:, 0]), 1.0)
    out_dir = np.where(is_end[:, None], np.stack([np.sign(cen[:, 0]), np.zeros(len(cen)), np.zeros(len(cen))], 1),
                       radial)
    flip = np.sum(m.normals * out_dir, axis=1) < 0
    Fo[flip] = Fo[flip][:, ::-1]
    # a reversed triangle (c, b, a, c) keeps its repeated index last-and-first
    m = Mesh(m.vertices, Fo)
    c0 = np.asarray(center, dtype=np.float64)
    return translated(m, c0) if np.any(c0 != 0.0) else m


def hemisphere(radius=1.0, nlat=32, nlon=64):
    """Submerged hemisphere: nlat latitude rows from the bottom pole to z=0, nlon sectors."""
    theta = np.linspace(0.0, np.pi / 2, nlat + 1)
    prof = np.stack([radius * np.sin(theta), -radius * np.cos(theta)], axis=1)
    prof[0, 0] = 0.0
    prof[-1, 1] = 0.0
    return revolution(prof, nlon)


def cylinder(radius=5.0, draft=10.0, ncirc=32, nvert=24, nrings=8):
    """Vertical truncated cylinder: bottom disk of nrings rings (centre ring triangles)
    and a side of nvert rows, ncirc sectors: ncirc * (nvert + nrings) panels."""
    rr = np.linspace(0.0, radius, nrings + 1)
    bottom = np.stack([rr, np.full_like(rr, -draft)], axis=1)
    zz = np.linspace(-draft, 0.0, nvert + 1)
    side = np.stack([np.full_like(zz, radius), zz], axis=1)
    prof = np.concatenate([bottom, side[1:]], axis=0)
    return revolution(prof, ncirc)


def capped_cylinder(radius=4.0, side_depth=8.0, ncirc=64, nside=40, ncap=24):
    """Vertical cylinder (z in [-side_depth, 0]) closed below by a hemispherical cap:
    draft = side_depth + radius, ncirc * (nside + ncap) panels."""
    th = np.linspace(0.0, np.pi / 2, ncap + 1)
    cap = np.stack([radius * np.sin(th), -side_depth - radius * np.cos(th)], axis=1)
    cap[0, 0] = 0.0
    zz = np.linspace(-side_depth, 0.0, nside + 1)
    side = np.stack([np.full_like(zz, radius), zz], axis=1)
    prof = np.concatenate([cap, side[1:]], axis=0)
    return revolution(prof, ncirc)


def _box_counts(L, W, T, n):
    best = None
    for nx in range(1, 257):
        for ny in range(1, 257):
            rem = n - nx * ny
            if rem <= 0:
                break
            if rem % (2 * (nx + ny)):
                continue
            nz = rem // (2 * (nx + ny))
            h = np.array([L / nx, W / ny, T / nz])
            score = np.log(h.max() / h.min())
            if best is None or score < best[0]:
                best = (score, nx, ny, nz)
    if best is None:
        raise ValueError(f"no box grid with exactly {n} panels")
    return best[1:]


def box(length, width, draft, n):
    """Immersed part of a rectangular box (centred on the z axis, top at z=0) with
    exactly n panels: bottom nx*ny, two sides nx*nz, two ends ny*nz."""
    nx, ny, nz = _box_counts(length, width, draft, n)
    xs = np.linspace(-length / 2, length / 2, nx + 1)
    ys = np.linspace(-width / 2, width / 2, ny + 1)
    zs = np.linspace(-draft, 0.0, nz + 1)
    verts, faces, index = [], [], {}

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(verts) + 1
            verts.append(p)
        return index[key]

    def grid(u, v, f):  # quads with outward orientation given by f(u,v)
        for a in range(len(u) - 1):
            for b in range(len(v) - 1):
                p = [f(u[a], v[b]), f(u[a + 1], v[b]), f(u[a + 1], v[b + 1]), f(u[a], v[b + 1])]
                faces.append([vid(np.array(q, dtype=np.float64)) for q in p])

    grid(ys, xs, lambda y, x: (x, y, -draft))                 # bottom, normal -z
    grid(xs, zs, lambda x, z: (x, -width / 2, z))             # side y=-W/2, normal -y
    grid(zs, xs, lambda z, x: (x, width / 2, z))              # side y=+W/2, normal +y
    grid(zs, ys, lambda z, y: (-length / 2, y, z))            # end x=-L/2, normal -x
    grid(ys, zs, lambda y, z: (length / 2, y, z))             # end x=+L/2, normal +x
    m = Mesh(np.array(verts), np.array(faces, dtype=np.int64))
    assert m.nfaces == n
    return m


def sphere_with_count(radius, n):
    """Immersed hemisphere with exactly n in {256, ..., 4096} panels."""
    table = {256: (8, 32), 512: (16, 32), 1024: (16, 64), 2048: (32, 64), 4096: (32, 128)}
    nlat, nlon = table[n]
    return hemisphere(radius, nlat, nlon)


def cylinder_with_count(radius, draft, n):
    table = {256: 16, 512: 32, 1024: 32, 2048: 64, 4096: 64}
    nc = table[n]
    rows = n // nc
    nr = int(np.clip(round(rows * radius / (radius + draft)), 1, rows - 1))
    return cylinder(radius, draft, nc, rows - nr, nr)


def random_bodies(seed=20261203, nbodies=8):
    """Seeded uneven workload: spheres (R~U[0.5,3]), cylinders (R~U[1,6], draft~U[2,12]),
    boxes (sides~U[1,10], half immersed), panel counts from {256,...,4096},
    16-64 log-uniform frequencies in [0.1, 3.0] rad/s each."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for b in range(nbodies):
        kind = ["sphere", "cylinder", "box"][int(rng.integers(0, 3))]
        n = int(rng.choice([256, 512, 1024, 2048, 4096]))
        if kind == "sphere":
            R = rng.uniform(0.5, 3.0)
            mesh = sphere_with_count(R, n)
            desc = dict(kind=kind, radius=R)
        elif kind == "cylinder":
            R = rng.uniform(1.0, 6.0); T = rng.uniform(2.0, 12.0)
            mesh = cylinder_with_count(R, T, n)
            desc = dict(kind=kind, radius=R, draft=T)
        else:
            L, W, H = rng.uniform(1.0, 10.0, size=3)
            mesh = box(L, W, H / 2, n)
            desc = dict(kind=kind, sides=(L, W, H))
        nw = int(rng.integers(16, 65))
        omegas = np.sort(np.exp(rng.uniform(np.log(0.1), np.log(3.0), size=nw)))
        desc.update(n=n, nw=nw)
        out.append((mesh, omegas, desc))
    return out
