"""Type-blocked mesh topology restated from hybrid/mesh.py.

Input is raw arrays (the reference reads them through meshio, which is absent
here): points, per-type corner connectivity, optional boundary tags keyed by
the sorted corner ids, and two extensions the reference lacks:

* ``canonical``: node-id map used for face matching.  Periodic boxes identify
  opposite boundary nodes through it while each cell keeps its own physical
  coordinates; with the identity map this is exactly hybrid/mesh.py:378-405.
* sub-faces: an unmatched quad face whose corner set is covered by two
  unmatched triangles becomes two triangle faces owned by the quad side.

Face orientation (normal from owner to neighbour) follows hybrid/mesh.py:361-376
and :427-432; for periodic faces the owner->neighbour displacement uses the
minimum image across the period.
"""

import itertools
from dataclasses import dataclass

import numpy as np

# Local faces of each type, hybrid/mesh.py:37-59 (== hybrid/geometry.py:12-24).
LOCAL_FACES = {
    "tetra": ((0, 2, 1), (0, 1, 3), (1, 2, 3), (2, 0, 3)),
    "wedge": ((0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5)),
    "hexahedron": ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4),
                   (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)),
}

# Tetra decompositions used for centres/volumes, hybrid/mesh.py:61-76.
TETRA_DECOMPOSITIONS = {
    "tetra": ((0, 1, 2, 3),),
    "wedge": ((0, 1, 2, 3), (1, 2, 4, 3), (2, 4, 5, 3)),
    "hexahedron": ((0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6),
                   (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)),
}

BLOCK_ORDER = ("hexahedron", "wedge", "tetra")


@dataclass
class ElementBlock:
    """hybrid/mesh.py:122-139 (linear geometry only)."""
    element_type: str
    connectivity: np.ndarray
    centers: np.ndarray
    volumes: np.ndarray


@dataclass
class FaceBlock:
    """hybrid/mesh.py:142-175 plus ``owner_corners``: owner-reference
    coordinates of the face vertices (for ordinary faces these equal the owner
    reference vertices of LOCAL_FACES[type][owner_face])."""
    face_type: str
    nodes: np.ndarray
    owner_block: np.ndarray
    owner_local: np.ndarray
    owner_face: np.ndarray
    neighbor_block: np.ndarray
    neighbor_local: np.ndarray
    neighbor_face: np.ndarray
    centers: np.ndarray
    normals: np.ndarray
    areas: np.ndarray
    boundary_tags: np.ndarray
    owner_corners: np.ndarray
    neighbor_corners: np.ndarray

    @property
    def internal(self):
        return self.neighbor_block >= 0


def _tet_volume(v):
    return abs(np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], 1))) / 6.0


def from_arrays(m, allow_untagged=False):
    """Build from a mesh generator's raw arrays."""
    return HybridMesh(m.points, m.cells, m.tag_of, m.physical_names, m.canonical,
                      m.period, allow_untagged)


class HybridMesh:
    def __init__(self, points, cells, tag_of=None, physical_names=None,
                 canonical=None, period=None, allow_untagged=False):
        self.points = np.asarray(points, dtype=np.float64)
        self.physical_names = dict(physical_names or {})
        self.canonical = (np.arange(len(self.points)) if canonical is None
                          else np.asarray(canonical))
        self.period = None if period is None else np.asarray(period, float)
        self.element_blocks = []
        for t in BLOCK_ORDER:
            if t in cells and len(cells[t]):
                conn = np.asarray(cells[t], dtype=np.int32)
                centers, volumes = self._cell_geometry(t, conn)
                self.element_blocks.append(ElementBlock(t, conn, centers, volumes))
        self.block_index = {b.element_type: i for i, b in enumerate(self.element_blocks)}
        self.face_blocks = self._build_faces(tag_of)
        self._validate(allow_untagged)

    def _cell_geometry(self, element_type, conn):
        """hybrid/mesh.py:222-238."""
        centers, volumes = [], []
        for cell in conn:
            v = self.points[cell]
            pieces, wc = [], []
            for tet in TETRA_DECOMPOSITIONS[element_type]:
                tv = v[list(tet)]
                vol = _tet_volume(tv)
                pieces.append(vol)
                wc.append(vol * tv.mean(0))
            tot = float(np.sum(pieces))
            centers.append(np.sum(wc, 0) / tot)
            volumes.append(tot)
        return np.asarray(centers), np.asarray(volumes)

    def _key(self, nodes):
        return tuple(sorted(int(self.canonical[n]) for n in nodes))

    def _displacement(self, a, b):
        d = b - a
        if self.period is not None:
            for k in range(3):
                if self.period[k] > 0:
                    d[k] -= self.period[k] * np.round(d[k] / self.period[k])
        return d

    def _face_geometry(self, nodes, owner_center):
        """hybrid/mesh.py:361-376 (outward w.r.t. the owner centre)."""
        v = self.points[nodes]
        center = v.mean(0)
        if len(nodes) == 3:
            av = 0.5 * np.cross(v[1] - v[0], v[2] - v[0])
        else:
            av = 0.5 * (np.cross(v[1] - v[0], v[2] - v[0]) + np.cross(v[2] - v[0], v[3] - v[0]))
        if np.dot(av, self._displacement(owner_center, center)) < 0.0:
            av = -av
        area = np.linalg.norm(av)
        return center, av / area, area

    def _corner_reference(self, block_id, local_cell, local_face, face_nodes):
        """Reference coordinates, in that cell, of the given face nodes, matched
        by canonical id (the lookup of hybrid/geometry.py:205-224)."""
        from .elements import reference_vertices
        block = self.element_blocks[block_id]
        ref = reference_vertices(block.element_type)
        idx = LOCAL_FACES[block.element_type][local_face]
        cell_nodes = block.connectivity[local_cell, list(idx)]
        lookup = {int(self.canonical[n]): ref[i] for n, i in zip(cell_nodes, idx)}
        return np.asarray([lookup[int(self.canonical[n])] for n in face_nodes])

    def _build_faces(self, tag_of):
        """hybrid/mesh.py:378-459 with canonical ids and sub-faces."""
        pending = {}
        finished = {"triangle": [], "quad": []}
        for block_id, block in enumerate(self.element_blocks):
            for local_cell, cell_nodes in enumerate(block.connectivity):
                for local_face, fv in enumerate(LOCAL_FACES[block.element_type]):
                    nodes = np.asarray([cell_nodes[i] for i in fv], dtype=np.int32)
                    key = self._key(nodes)
                    record = (block_id, local_cell, local_face, nodes)
                    if key in pending:
                        owner = pending.pop(key)
                        finished["triangle" if len(nodes) == 3 else "quad"].append(
                            (owner, record, -1, None))
                    else:
                        pending[key] = record

        # Sub-faces: a quad covered by two triangles (extension).
        consumed = set()
        for key, rec in list(pending.items()):
            if len(key) != 4:
                continue
            tris = [k for k in itertools.combinations(key, 3)
                    if k in pending and k not in consumed]
            if len(tris) != 2:
                continue
            pending.pop(key)
            for tk in tris:
                consumed.add(tk)
                trec = pending.pop(tk)
                # Owner = quad side; its face "nodes" are the triangle vertices,
                # taken in the triangle cell's node order.
                sub_owner = (rec[0], rec[1], rec[2], trec[3])
                finished["triangle"].append((sub_owner, trec, -1, "sub"))

        for key, owner in pending.items():
            # hybrid/mesh.py:401-405; tags come from a callable on the corner ids
            # instead of meshio boundary cells.
            tag = -1 if tag_of is None else int(tag_of(np.asarray(owner[3])[None])[0])
            finished["triangle" if len(key) == 3 else "quad"].append((owner, None, tag, None))

        result = {}
        for ftype, records in finished.items():
            if not records:
                continue
            cols = {k: [] for k in ("nodes", "ob", "ol", "of", "nb", "nl", "nf", "c",
                                    "n", "a", "t", "oc", "nc")}
            for owner, nbr, tag, kind in records:
                ob, ol, of, nodes = owner
                oblock = self.element_blocks[ob]
                center, normal, area = self._face_geometry(nodes, oblock.centers[ol])
                oc = self._corner_reference(ob, ol, of, nodes)
                if nbr is None:
                    nb = nl = nf = -1
                    nc = np.full_like(oc, np.nan)
                else:
                    nb, nl, nf, _ = nbr
                    disp = self._displacement(oblock.centers[ol],
                                              self.element_blocks[nb].centers[nl])
                    if np.dot(normal, disp) < 0.0:
                        normal = -normal
                    nc = self._corner_reference(nb, nl, nf, nodes)
                for k, v in zip(cols, (nodes, ob, ol, of, nb, nl, nf, center, normal,
                                       area, tag, oc, nc)):
                    cols[k].append(v)
            result[ftype] = FaceBlock(
                ftype, np.asarray(cols["nodes"], np.int32),
                np.asarray(cols["ob"], np.int32), np.asarray(cols["ol"], np.int32),
                np.asarray(cols["of"], np.int32), np.asarray(cols["nb"], np.int32),
                np.asarray(cols["nl"], np.int32), np.asarray(cols["nf"], np.int32),
                np.asarray(cols["c"]), np.asarray(cols["n"]), np.asarray(cols["a"]),
                np.asarray(cols["t"], np.int32), np.asarray(cols["oc"]),
                np.asarray(cols["nc"]))
        return result

    def _validate(self, allow_untagged):
        """hybrid/mesh.py:461-477."""
        vols = np.concatenate([b.volumes for b in self.element_blocks])
        if not np.all(np.isfinite(vols)) or np.any(vols <= 0):
            raise ValueError("Mesh contains non-positive or invalid cell volumes")
        for ftype, fb in self.face_blocks.items():
            if np.any(fb.areas <= 0):
                raise ValueError(f"Mesh contains invalid {ftype} face areas")
            if np.any(fb.boundary_tags[fb.internal] != -1):
                raise ValueError("Internal faces cannot have boundary tags")
            if not allow_untagged and np.any(fb.boundary_tags[~fb.internal] < 0):
                raise ValueError(f"boundary {ftype} faces lack physical tags")
