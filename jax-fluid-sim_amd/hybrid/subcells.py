"""Reference tables of the 64-subcell finite-volume fallback (MOOD).

The subdivisions, subcell numbering, the numbering of subcell faces (internal
and on the parent's boundary) and the subcell projection follow
hybrid/subcell.py:19-274 and :277-396 of the reference.  The numbering is part
of the behaviour: the reference indexes subcell arrays with boundary-subface
numbers (hybrid/subcell.py:774-776, 805-807, 842-848), so the same numbers
are produced here (checked against the oracle in tests).

For affine elements every table is element-independent: the projection
P (64, N) averages each basis function over each subcell, the reconstruction
R = (P^T V P)^-1 P^T V is its volume-weighted least-squares left inverse, and
subcell face area vectors map to physical space by |det J| J^-T (Nanson).
"""

from functools import lru_cache

import numpy as np

from .elements import LOCAL_FACES, basis, volume_quadrature

NSUB = 64


def _triangle_lattice(n):
    """Vertices (i/n, j/n), i + j <= n, and the n^2 lattice triangles (each
    'up' triangle followed by the 'down' one sharing its hypotenuse)."""
    verts, where = [], {}
    for i in range(n + 1):
        for j in range(n + 1 - i):
            where[(i, j)] = len(verts)
            verts.append((i / n, j / n))
    tris = []
    for i in range(n):
        for j in range(n - i):
            tris.append((where[i, j], where[i + 1, j], where[i, j + 1]))
            if i + j + 2 <= n:
                tris.append((where[i + 1, j], where[i + 1, j + 1], where[i, j + 1]))
    return np.asarray(verts), np.asarray(tris, np.int64)


def _tetra_cells():
    """Two levels of red refinement of the unit tetrahedron: per level each
    tetrahedron gives its four corner children, then the inner octahedron
    cut into four along one diagonal (hybrid/subcell.py:75-110)."""
    pts = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)]
    index = {p: k for k, p in enumerate(pts)}
    cells = [(0, 1, 2, 3)]

    def midpoint(a, b):
        m = tuple(0.5 * (x + y) for x, y in zip(pts[a], pts[b]))
        if m not in index:
            index[m] = len(pts)
            pts.append(m)
        return index[m]

    for _ in range(2):
        nxt = []
        for a, b, c, d in cells:
            ab, ac, ad = midpoint(a, b), midpoint(a, c), midpoint(a, d)
            bc, bd, cd = midpoint(b, c), midpoint(b, d), midpoint(c, d)
            nxt.extend([(a, ab, ac, ad), (ab, b, bc, bd), (ac, bc, c, cd), (ad, bd, cd, d),
                        (ab, ac, ad, bd), (ab, bc, ac, bd), (ac, ad, bd, cd), (ac, bd, bc, cd)])
        cells = nxt
    return np.asarray(pts), np.asarray(cells, np.int64)


def _wedge_cells(n=4):
    tv, tris = _triangle_lattice(n)
    m = len(tv)
    verts = np.asarray([(x, y, layer / n) for layer in range(n + 1) for x, y in tv])
    cells = [(layer * m + a, layer * m + b, layer * m + c,
              (layer + 1) * m + a, (layer + 1) * m + b, (layer + 1) * m + c)
             for layer in range(n) for a, b, c in tris]
    return verts, np.asarray(cells, np.int64)


def _hexahedron_cells(n=4):
    s = n + 1
    verts = np.asarray([(i / n, j / n, k / n) for k in range(s) for j in range(s) for i in range(s)])

    def v(i, j, k):
        return (k * s + j) * s + i

    cells = [(v(i, j, k), v(i + 1, j, k), v(i + 1, j + 1, k), v(i, j + 1, k),
              v(i, j, k + 1), v(i + 1, j, k + 1), v(i + 1, j + 1, k + 1), v(i, j + 1, k + 1))
             for k in range(n) for j in range(n) for i in range(n)]
    return verts, np.asarray(cells, np.int64)


# Planes of the local faces, as functions that vanish on them
# (hybrid/subcell.py:167-180 ordering).
_FACE_PLANES = {
    "tetra": (lambda x: x[:, 2], lambda x: x[:, 1], lambda x: x.sum(1) - 1.0, lambda x: x[:, 0]),
    "wedge": (lambda x: x[:, 2], lambda x: x[:, 2] - 1.0, lambda x: x[:, 1],
              lambda x: x[:, 0] + x[:, 1] - 1.0, lambda x: x[:, 0]),
    "hexahedron": (lambda x: x[:, 2], lambda x: x[:, 2] - 1.0, lambda x: x[:, 1],
                   lambda x: x[:, 0] - 1.0, lambda x: x[:, 1] - 1.0, lambda x: x[:, 0]),
}


def area_vector(v):
    """Area vector of a planar triangle/quad from its vertices (k, 3)."""
    a = 0.5 * np.cross(v[1] - v[0], v[2] - v[0])
    if len(v) == 4:
        a = a + 0.5 * np.cross(v[2] - v[0], v[3] - v[0])
    return a


class SubcellTables:
    """Element-independent subcell tables of one element type and degree."""

    def __init__(self, element_type, p):
        self.element_type = element_type
        self.p = p
        build = {"tetra": _tetra_cells, "wedge": _wedge_cells,
                 "hexahedron": _hexahedron_cells}[element_type]
        self.vertices, self.cells = build()
        assert len(self.cells) == NSUB
        lf_list = LOCAL_FACES[element_type]
        # Faces: the first time a vertex set is met it is pending; the second
        # time it becomes an internal face (owner = first cell).  Whatever is
        # left pending lies on the parent boundary, in first-met order.
        pending = {}
        io, inb, ifv = [], [], []
        for c, conn in enumerate(self.cells):
            for lf in lf_list:
                fv = tuple(int(conn[k]) for k in lf)
                key = tuple(sorted(fv))
                hit = pending.pop(key, None)
                if hit is None:
                    pending[key] = (c, fv)
                else:
                    io.append(hit[0])
                    inb.append(c)
                    ifv.append(hit[1])
        bo, bpf, bfv = [], [], []
        planes = _FACE_PLANES[element_type]
        for c, fv in pending.values():
            x = self.vertices[list(fv)]
            on = [k for k, plane in enumerate(planes) if np.all(np.abs(plane(x)) <= 1e-10)]
            if len(on) != 1:
                raise ValueError("subcell boundary face not on exactly one parent face")
            bo.append(c)
            bpf.append(on[0])
            bfv.append(fv)
        self.internal_owner = np.asarray(io, np.int32)
        self.internal_neighbor = np.asarray(inb, np.int32)
        self.boundary_owner = np.asarray(bo, np.int32)
        self.boundary_parent_face = np.asarray(bpf, np.int32)
        centers = self.vertices[self.cells].mean(1)
        self.centers = centers
        # Internal area vectors oriented owner -> neighbour, centre offsets.
        a_int, dc = [], []
        for o, nb, fv in zip(io, inb, ifv):
            a = area_vector(self.vertices[list(fv)])
            d = centers[nb] - centers[o]
            a_int.append(a if a @ d >= 0.0 else -a)
            dc.append(d)
        self.internal_area = np.asarray(a_int)
        self.internal_dc = np.asarray(dc)
        # Boundary area vectors oriented away from the owning subcell.
        a_b, cb = [], []
        for o, fv in zip(bo, bfv):
            x = self.vertices[list(fv)]
            a = area_vector(x)
            cen = x.mean(0)
            a_b.append(a if a @ (cen - centers[o]) >= 0.0 else -a)
            cb.append(cen)
        self.boundary_area = np.asarray(a_b)
        self.boundary_center = np.asarray(cb)
        self._projection()

    def _projection(self):
        """P[s, i] = mean of basis i over subcell s (exact quadrature on each
        subcell's affine image of the reference element), R its volume-weighted
        least-squares left inverse."""
        t, p = self.element_type, self.p
        pts, w = volume_quadrature(t, max(p, 2))
        cv = self.vertices[self.cells]                       # (64, nv, 3)
        if t == "hexahedron":
            J = np.stack([cv[:, 1] - cv[:, 0], cv[:, 3] - cv[:, 0], cv[:, 4] - cv[:, 0]], -1)
        else:
            J = np.stack([cv[:, 1] - cv[:, 0], cv[:, 2] - cv[:, 0], cv[:, 3] - cv[:, 0]], -1)
        det = np.abs(np.linalg.det(J))
        mapped = cv[:, :1, :] + np.einsum("sdr,qr->sqd", J, pts)
        phi = basis(t, p, mapped.reshape(-1, 3)).reshape(NSUB, len(w), -1)
        vol = det * w.sum()
        P = np.einsum("sqi,q->si", phi, w) * det[:, None] / vol[:, None]
        self.volume = vol
        self.P = P
        gram = P.T @ (vol[:, None] * P)
        self.R = np.linalg.solve(gram, P.T * vol[None, :])

    def boundary_ids(self, local_face):
        return np.flatnonzero(self.boundary_parent_face == local_face)

    def internal_csr(self):
        """Per subcell, its internal faces in increasing face order."""
        lists = [[] for _ in range(NSUB)]
        for f, (o, nb) in enumerate(zip(self.internal_owner, self.internal_neighbor)):
            lists[o].append(f)
            lists[nb].append(f)
        ptr = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int32)
        return ptr, np.asarray([f for x in lists for f in x], np.int32)

    def boundary_targets(self, indexing):
        """Per subcell, the boundary subfaces whose flux it receives, and the
        subcell each boundary subface reads its finite-volume trace from.
        'reference': subface id j is used as a subcell index (gathers clamp
        to 63, scatters drop j >= 64); 'owner': the subcell owning j."""
        nb = len(self.boundary_owner)
        if indexing == "reference":
            gather = np.minimum(np.arange(nb), NSUB - 1)
            target = np.where(np.arange(nb) < NSUB, np.arange(nb), -1)
        elif indexing == "owner":
            gather = self.boundary_owner.copy()
            target = self.boundary_owner.copy()
        else:
            raise ValueError(f"unknown subface indexing {indexing!r}")
        lists = [[] for _ in range(NSUB)]
        for j, s in enumerate(target):
            if s >= 0:
                lists[s].append(j)
        ptr = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int32)
        return ptr, np.asarray([j for x in lists for j in x], np.int32), gather.astype(np.int32)


@lru_cache(maxsize=None)
def subcell_tables(element_type, p):
    return SubcellTables(element_type, p)


def _face_map(src, dst, k):
    """Affine map taking the first k corners src (k, 3) onto dst (k, 3) along
    the face plane: x = src0 + u (src1 - src0) + v (srcL - src0) ->
    dst0 + u (dst1 - dst0) + v (dstL - dst0), L = 3 for quads, 2 for triangles."""
    last = 3 if k == 4 else 2
    A = np.stack([src[1] - src[0], src[last] - src[0]], 1)      # (3, 2)
    B = np.stack([dst[1] - dst[0], dst[last] - dst[0]], 1)

    def apply(x):
        uv, *_ = np.linalg.lstsq(A, (x - src[0]).T, rcond=None)
        return (dst[0][:, None] + B @ uv).T

    return apply


def pair_subfaces(to, lf_o, corners_o, tn, lf_n, corners_n, p, sub):
    """Pairs of boundary subfaces across one face (hybrid/subcell.py:473-533):
    returns osub, nsub (16,), owner/neighbour reference points (16, 3), the
    side whose subface geometry gives the normal and area (0 owner, 1
    neighbour) and that side's reference area vector (16, 3), oriented from
    the owner to the neighbour."""
    so = subcell_tables(to, p)
    oids = so.boundary_ids(lf_o)
    if tn is None:
        return dict(osub=oids, nsub=None, opt=so.boundary_center[oids], npt=None,
                    src=0, area=so.boundary_area[oids])
    sn = subcell_tables(tn, p)
    nids = sn.boundary_ids(lf_n)
    if not sub:
        k = 4 if len(LOCAL_FACES[to][lf_o]) == 4 else 3
        to_n = _face_map(corners_o[:k], corners_n[:k], k)
        mapped = to_n(so.boundary_center[oids])
        d = np.linalg.norm(mapped[:, None, :] - sn.boundary_center[nids][None], axis=-1)
        perm = np.argmin(d, 1)
        if len(np.unique(perm)) != len(oids) or d[np.arange(len(oids)), perm].max() > 1e-9:
            raise ValueError("non-conforming subfaces")
        nsub = nids[perm]
        return dict(osub=oids, nsub=nsub, opt=so.boundary_center[oids],
                    npt=sn.boundary_center[nsub], src=0, area=so.boundary_area[oids])
    # Sub-face: the neighbour's triangle covers half of the owner's quad face.
    to_o = _face_map(corners_n[:3], corners_o[:3], 3)
    owner_pts = to_o(sn.boundary_center[nids])
    ocen = so.boundary_center[oids]
    osel = oids[np.argmin(np.linalg.norm(owner_pts[:, None] - ocen[None], axis=-1), 1)]
    return dict(osub=osel, nsub=nids, opt=owner_pts, npt=sn.boundary_center[nids], src=1,
                area=-sn.boundary_area[nids])
