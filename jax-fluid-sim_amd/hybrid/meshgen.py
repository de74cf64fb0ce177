"""Synthetic hybrid meshes for the hot-path configurations.

The reference reads Gmsh files through meshio (hybrid/mesh.py:186-205); its
benchmark configurations are boxes, so this module generates them directly:

* ``box_mesh``: an n_x x n_y x n_z lattice of hexahedra on [origin, origin+L],
  each axis optionally periodic.  Periodic axes identify opposite lattice nodes
  through ``canonical`` (face matching) while every cell keeps its own physical
  corner coordinates, so geometry is unchanged.
* ``split_fraction``: a seeded fraction of the hexahedra is replaced by 2
  prisms (split on the z-faces along the 0-2 diagonal) or 6 tetrahedra (the
  Kuhn split around the 0-6 diagonal, hybrid/mesh.py:68-75 decomposition).
  Both splits put the same diagonal on every shared face, so split cells are
  mutually conforming; a hexahedron face next to two split triangles becomes
  two triangle sub-faces.
* ``smoke_mesh``: the 3-cell hex/prism/tet mesh of the reference's
  tools/generate_hybrid_smoke_mesh.py:12-89.

Boundary planes are tagged 1..6 = x-, x+, y-, y+, z-, z+.
"""

from dataclasses import dataclass, field

import numpy as np

PLANE_TAGS = {"xmin": 1, "xmax": 2, "ymin": 3, "ymax": 4, "zmin": 5, "zmax": 6}

# Prism split (z-faces cut along local corners 0-2) and Kuhn tetra split.
_WEDGE_SPLIT = ((0, 1, 2, 4, 5, 6), (0, 2, 3, 4, 6, 7))
_TETRA_SPLIT = ((0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6),
                (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6))


@dataclass
class MeshArrays:
    """Raw mesh input: points, corner connectivity per element type, the
    face-matching id map, the period vector (0 where not periodic), tags of
    boundary faces keyed by sorted canonical corner ids, and tag names."""
    points: np.ndarray
    cells: dict
    canonical: np.ndarray
    period: np.ndarray
    physical_names: dict = field(default_factory=dict)
    lengths: np.ndarray = None
    origin: np.ndarray = None
    default_tag: int = -1
    face_tags: dict = None     # sorted corner ids -> tag (meshes read from files)

    @property
    def number_of_cells(self):
        return sum(len(c) for c in self.cells.values())

    def tag_of(self, face_nodes):
        """Boundary tag of a face given its corner node ids (any shape
        (..., k)): the tag of the file's boundary face with those corners
        (a boundary face without one raises, hybrid/mesh.py:477-483),
        ``default_tag`` when set, else the box plane tag."""
        if self.face_tags is not None:
            nodes = np.asarray(face_nodes)
            flat = nodes.reshape(-1, nodes.shape[-1])
            out = np.array([self.face_tags.get(tuple(sorted(int(v) for v in row)), -1)
                            for row in flat], np.int32).reshape(nodes.shape[:-1])
            missing = int(np.count_nonzero(out < 0))
            if missing:
                raise ValueError(f"{missing} boundary faces lack physical tags")
            return out
        if self.default_tag >= 0:
            return np.full(np.shape(face_nodes)[:-1], self.default_tag, np.int32)
        return plane_tags(self.points, face_nodes, self.lengths, self.origin)


def _orient_tetra(conn, points):
    v = points[conn]
    det = np.einsum("ei,ei->e", v[:, 1] - v[:, 0],
                    np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))
    flip = det < 0
    conn = conn.copy()
    conn[flip, 1], conn[flip, 2] = conn[flip, 2].copy(), conn[flip, 1].copy()
    return conn


def box_mesh(n, lengths=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
             periodic=(False, False, False), split_fraction=0.0,
             tetra_share=0.5, seed=0, plane_names=None):
    """Hexahedral box, optionally periodic and with seeded prism/tet splits."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
    L = np.asarray(lengths, float)
    o = np.asarray(origin, float)
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                          indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # x fastest
    points = np.stack([o[0] + L[0] * i / nx, o[1] + L[1] * j / ny,
                       o[2] + L[2] * k / nz], -1)

    def nid(a, b, c):
        return (c * (ny + 1) + b) * (nx + 1) + a

    ci = i % nx if periodic[0] else i
    cj = j % ny if periodic[1] else j
    ck = k % nz if periodic[2] else k
    canonical = nid(ci, cj, ck).astype(np.int64)

    hi, hj, hk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi, hj, hk = (a.transpose(2, 1, 0).ravel() for a in (hi, hj, hk))
    hexes = np.stack([nid(hi, hj, hk), nid(hi + 1, hj, hk), nid(hi + 1, hj + 1, hk),
                      nid(hi, hj + 1, hk), nid(hi, hj, hk + 1), nid(hi + 1, hj, hk + 1),
                      nid(hi + 1, hj + 1, hk + 1), nid(hi, hj + 1, hk + 1)], -1)

    cells = {"hexahedron": hexes}
    if split_fraction > 0.0:
        rng = np.random.default_rng(seed)
        count = int(round(split_fraction * len(hexes)))
        chosen = rng.permutation(len(hexes))[:count]
        to_tet = rng.random(count) < tetra_share
        keep = np.ones(len(hexes), bool)
        keep[chosen] = False
        wedge_src = hexes[chosen[~to_tet]]
        tet_src = hexes[chosen[to_tet]]
        cells["hexahedron"] = hexes[keep]
        if len(wedge_src):
            cells["wedge"] = np.concatenate([wedge_src[:, list(w)] for w in _WEDGE_SPLIT])
        if len(tet_src):
            tets = np.concatenate([tet_src[:, list(t)] for t in _TETRA_SPLIT])
            cells["tetra"] = _orient_tetra(tets, points)

    names = dict(plane_names or {})
    physical_names = {names.get(name, name): tag for name, tag in PLANE_TAGS.items()}
    period = np.array([L[a] if periodic[a] else 0.0 for a in range(3)])
    return MeshArrays(points, cells, canonical, period, physical_names, L, o)


def plane_tags(points, face_nodes, lengths, origin, tol=1e-9):
    """Tags (1..6) of boundary faces from the box plane their corners lie on;
    -1 where a face lies on no box plane.  ``face_nodes``: (..., k)."""
    x = points[np.asarray(face_nodes)]
    o = np.asarray(origin, float)
    L = np.asarray(lengths, float)
    out = np.full(x.shape[:-2], -1, np.int32)
    for axis in range(3):
        t = tol * max(L[axis], 1.0)
        lo = np.all(np.abs(x[..., axis] - o[axis]) < t, -1)
        hi = np.all(np.abs(x[..., axis] - (o[axis] + L[axis])) < t, -1)
        out = np.where(lo, 2 * axis + 1, out)
        out = np.where(hi, 2 * axis + 2, out)
    return out


def smoke_mesh():
    """The reference's 3-cell hybrid smoke mesh
    (tools/generate_hybrid_smoke_mesh.py:15-38): a prism sharing one triangle
    with a tetra and one quad with a hexahedron; all boundary faces are
    tagged 13 ("walls")."""
    points = np.array([
        [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0],
        [1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [0.0, 0.0, 2.0], [1.0, -1.0, 0.0],
        [0.0, -1.0, 0.0], [1.0, -1.0, 1.0], [0.0, -1.0, 1.0]])
    cells = {
        "hexahedron": np.array([[0, 1, 7, 8, 3, 4, 9, 10]]),
        "wedge": np.array([[0, 1, 2, 3, 4, 5]]),
        "tetra": np.array([[3, 4, 5, 6]]),
    }
    return MeshArrays(points, cells, np.arange(len(points)), np.zeros(3),
                      {"fluid": 1, "walls": 13}, default_tag=13)
