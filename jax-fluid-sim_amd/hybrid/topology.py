"""Vectorized face topology and affine element metrics for the device path.

Face pairing follows hybrid/mesh.py:378-459: faces of all cells are keyed by
their sorted corner ids (``canonical`` ids here, so periodic boxes pair
across the period), the first cell in (hexahedron, wedge, tetra) x cell x
local-face order owns a face, and the normal points from owner to neighbour
(hybrid/mesh.py:427-432, minimum-image displacement for periodic pairs).
Unmatched quads covered by two unmatched triangles become two triangle
sub-faces owned by the quad side; the remaining faces are boundary faces.

Element metrics assume affine P1 geometry (every box-mesh cell is affine);
a non-affine cell raises.  For an affine cell the reference's per-point
metrics (hybrid/geometry.py:139-203) are constant: J^-1, |det J|, and
M_phys^-1 = min(1, V/1e-4) M_ref^-1 / |det J|.
"""

from dataclasses import dataclass

import numpy as np

from .elements import LOCAL_FACES, TYPES, corners

SLIVER_VOLUME = 1.0e-4  # hybrid/geometry.py:190
_REF_VOLUME = {"hexahedron": 1.0, "wedge": 0.5, "tetra": 1.0 / 6.0}


@dataclass
class Block:
    element_type: str
    connectivity: np.ndarray   # (E, k) corner node ids
    coords: np.ndarray         # (E, k, 3)
    jac: np.ndarray            # (E, 3, 3) d x_d / d xi_r  [d, r]
    jinv: np.ndarray           # (E, 3, 3) d xi_r / d x_d  [r, d]
    det: np.ndarray            # (E,)
    volume: np.ndarray         # (E,) = |det| * reference volume

    @property
    def count(self):
        return len(self.connectivity)


@dataclass
class Faces:
    """All faces of the mesh (one row per face or sub-face)."""
    face_type: np.ndarray      # 0 quad, 1 triangle
    sub: np.ndarray            # 1 when the owner side is a quad face cut in two
    owner_block: np.ndarray
    owner_local: np.ndarray
    owner_face: np.ndarray
    neighbor_block: np.ndarray  # -1 on the boundary
    neighbor_local: np.ndarray
    neighbor_face: np.ndarray
    owner_corners: np.ndarray   # (F, 4, 3) owner reference coords (triangles: 3 used)
    neighbor_corners: np.ndarray
    normal: np.ndarray          # (F, 3) unit, owner -> neighbour
    area: np.ndarray            # (F,)
    tag: np.ndarray             # (F,) boundary tag, -1 inside


class Topology:
    def __init__(self, mesh):
        self.mesh = mesh
        self.points = np.asarray(mesh.points, float)
        self.canonical = np.asarray(mesh.canonical, np.int64)
        self.period = np.asarray(mesh.period, float)
        self.blocks = []
        for t in TYPES:
            conn = mesh.cells.get(t)
            if conn is not None and len(conn):
                self.blocks.append(self._block(t, np.asarray(conn, np.int64)))
        self.block_of_type = {b.element_type: i for i, b in enumerate(self.blocks)}
        self.faces = self._faces()

    def _block(self, t, conn):
        X = self.points[conn]
        ref = corners(t)
        # Affine map from corners 0, and three independent corners.
        pick = {"hexahedron": (1, 3, 4), "wedge": (1, 2, 3), "tetra": (1, 2, 3)}[t]
        J = np.stack([X[:, k] - X[:, 0] for k in pick], -1)   # columns = d x / d xi_r
        mapped = X[:, :1, :] + np.einsum("edr,kr->ekd", J, ref)
        scale = np.maximum(np.abs(X).max(axis=(1, 2)), 1.0)
        if np.any(np.abs(mapped - X).max(axis=(1, 2)) > 1e-10 * scale):
            raise NotImplementedError("non-affine cells need per-point metrics")
        det = np.linalg.det(J)
        if np.any(det == 0.0):
            raise ValueError(f"Singular {t} mapping")
        return Block(t, conn, X, J, np.linalg.inv(J), det, np.abs(det) * _REF_VOLUME[t])

    # -- face pairing ------------------------------------------------------
    def _records(self, nv):
        recs = []
        for bid, b in enumerate(self.blocks):
            for lf, fv in enumerate(LOCAL_FACES[b.element_type]):
                if len(fv) != nv:
                    continue
                nodes = b.connectivity[:, list(fv)]
                E = b.count
                recs.append(dict(block=np.full(E, bid), local=np.arange(E),
                                 face=np.full(E, lf), nodes=nodes))
        if not recs:
            z = np.zeros(0, np.int64)
            return dict(block=z, local=z, face=z, nodes=np.zeros((0, nv), np.int64), rank=z)
        out = {k: np.concatenate([r[k] for r in recs]) for k in ("block", "local", "face", "nodes")}
        # Reference iteration order: block, then cell, then local face.
        order = np.lexsort((out["face"], out["local"], out["block"]))
        rank = np.empty_like(order)
        rank[order] = np.arange(len(rank))
        out["rank"] = rank
        return out

    def _encode(self, keys):
        base = int(self.canonical.max()) + 1
        code = np.zeros(len(keys), np.int64)
        for k in range(keys.shape[1]):
            code = code * base + keys[:, k]
        return code

    def _pairs(self, recs):
        keys = np.sort(self.canonical[recs["nodes"]], axis=1)
        order = np.lexsort(keys.T[::-1])
        ks = keys[order]
        same = np.all(ks[1:] == ks[:-1], axis=1)
        if np.any(same[1:] & same[:-1]):
            raise ValueError("a face is shared by more than two cells")
        first = np.flatnonzero(same)
        a, b = order[first], order[first + 1]
        swap = recs["rank"][a] > recs["rank"][b]
        owner = np.where(swap, b, a)
        nbr = np.where(swap, a, b)
        paired = np.zeros(len(keys), bool)
        paired[a] = True
        paired[b] = True
        return owner, nbr, np.flatnonzero(~paired), keys

    def _ref_of(self, block_ids, locals_, faces, target_nodes):
        """Reference coordinates, inside the given cells, of ``target_nodes``
        (F, m) matched against each cell's local-face corners by canonical id."""
        F, m = target_nodes.shape
        out = np.full((F, 4, 3), np.nan)
        tc = self.canonical[target_nodes]
        for bid, b in enumerate(self.blocks):
            ref = corners(b.element_type)
            for lf, fv in enumerate(LOCAL_FACES[b.element_type]):
                sel = np.flatnonzero((block_ids == bid) & (faces == lf))
                if not len(sel):
                    continue
                cn = self.canonical[b.connectivity[locals_[sel]][:, list(fv)]]  # (S, nv)
                hit = tc[sel][:, :, None] == cn[:, None, :]                      # (S, m, nv)
                if not np.all(hit.sum(-1) == 1):
                    raise ValueError("face corner lookup failed")
                pos = np.argmax(hit, -1)
                out[sel, :m] = ref[np.asarray(fv)[pos]]
        return out

    def _phys_of(self, block_ids, locals_, target_nodes):
        """Physical coordinates of target nodes taken from the given cell's own
        corners (matched by canonical id), so periodic images stay local."""
        F, m = target_nodes.shape
        out = np.full((F, m, 3), np.nan)
        tc = self.canonical[target_nodes]
        for bid, b in enumerate(self.blocks):
            sel = np.flatnonzero(block_ids == bid)
            if not len(sel):
                continue
            cn = self.canonical[b.connectivity[locals_[sel]]]
            hit = tc[sel][:, :, None] == cn[:, None, :]
            pos = np.argmax(hit, -1)
            out[sel] = b.coords[locals_[sel][:, None], pos]
        return out

    def _faces(self):
        cols = {k: [] for k in ("face_type", "sub", "ob", "ol", "of", "nb", "nl", "nf",
                                "onodes")}
        pending = {}
        for ftype, nv in ((0, 4), (1, 3)):
            r = self._records(nv)
            if not len(r["nodes"]):
                pending[nv] = (r, np.zeros(0, np.int64), np.zeros((0, nv), np.int64))
                continue
            own, nbr, rest, keys = self._pairs(r)
            for k, v in (("ob", r["block"][own]), ("ol", r["local"][own]), ("of", r["face"][own]),
                         ("nb", r["block"][nbr]), ("nl", r["local"][nbr]), ("nf", r["face"][nbr])):
                cols[k].append(v)
            cols["face_type"].append(np.full(len(own), ftype))
            cols["sub"].append(np.zeros(len(own), np.int64))
            on = np.full((len(own), 4), -1, np.int64)
            on[:, :nv] = r["nodes"][own]
            cols["onodes"].append(on)
            pending[nv] = (r, rest, keys[rest])

        # Sub-faces: unmatched quad covered by two unmatched triangles.
        rq, rest_q, keys_q = pending[4]
        rt, rest_t, keys_t = pending[3]
        tri_left = np.ones(len(rest_t), bool)
        quad_left = np.ones(len(rest_q), bool)
        if len(rest_q) and len(rest_t):
            tcode = self._encode(keys_t)
            torder = np.argsort(tcode)
            tsorted = tcode[torder]
            found = np.full((len(rest_q), 4), -1, np.int64)
            for s, drop in enumerate(range(4)):
                sub = np.delete(keys_q, drop, axis=1)
                code = self._encode(sub)
                pos = np.searchsorted(tsorted, code)
                pos = np.minimum(pos, len(tsorted) - 1)
                ok = tsorted[pos] == code
                found[ok, s] = torder[pos[ok]]
            two = (found >= 0).sum(1) == 2
            qi = np.flatnonzero(two)
            for qidx in qi:
                tris = found[qidx][found[qidx] >= 0]
                quad_left[qidx] = False
                for t in tris:
                    tri_left[t] = False
            if len(qi):
                tri_pairs = np.stack([found[q][found[q] >= 0] for q in qi])   # (S, 2)
                qrec = np.repeat(rest_q[qi], 2)
                trec = rest_t[tri_pairs.ravel()]
                cols["ob"].append(rq["block"][qrec]); cols["ol"].append(rq["local"][qrec])
                cols["of"].append(rq["face"][qrec])
                cols["nb"].append(rt["block"][trec]); cols["nl"].append(rt["local"][trec])
                cols["nf"].append(rt["face"][trec])
                cols["face_type"].append(np.full(len(qrec), 1))
                cols["sub"].append(np.ones(len(qrec), np.int64))
                on = np.full((len(qrec), 4), -1, np.int64)
                on[:, :3] = rt["nodes"][trec]
                cols["onodes"].append(on)

        for nv, r, rest, left, ftype in ((4, rq, rest_q, quad_left, 0), (3, rt, rest_t, tri_left, 1)):
            b = rest[left] if len(rest) else rest
            if not len(b):
                continue
            cols["ob"].append(r["block"][b]); cols["ol"].append(r["local"][b])
            cols["of"].append(r["face"][b])
            for k in ("nb", "nl", "nf"):
                cols[k].append(np.full(len(b), -1))
            cols["face_type"].append(np.full(len(b), ftype))
            cols["sub"].append(np.zeros(len(b), np.int64))
            on = np.full((len(b), 4), -1, np.int64)
            on[:, :nv] = r["nodes"][b]
            cols["onodes"].append(on)

        c = {k: np.concatenate(v) for k, v in cols.items()}
        F = len(c["ob"])
        nvert = np.where(c["face_type"] == 0, 4, 3)
        tags = np.full(F, -1, np.int32)
        oc = np.full((F, 4, 3), np.nan)
        ncn = np.full((F, 4, 3), np.nan)
        normal = np.zeros((F, 3))
        area = np.zeros(F)
        for nv in (3, 4):
            sel = np.flatnonzero(nvert == nv)
            if not len(sel):
                continue
            nodes = c["onodes"][sel, :nv]
            oc[sel] = self._ref_of(c["ob"][sel], c["ol"][sel], c["of"][sel], nodes)
            inner = sel[c["nb"][sel] >= 0]
            if len(inner):
                ncn[inner] = self._ref_of(c["nb"][inner], c["nl"][inner], c["nf"][inner],
                                          c["onodes"][inner, :nv])
            bnd = sel[c["nb"][sel] < 0]
            if len(bnd):
                tags[bnd] = self.mesh.tag_of(c["onodes"][bnd, :nv])
            v = self._phys_of(c["ob"][sel], c["ol"][sel], nodes)
            if nv == 3:
                av = 0.5 * np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
            else:
                av = 0.5 * (np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
                            + np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))
            center = v.mean(1)
            ocent = self._centers(c["ob"][sel], c["ol"][sel])
            flip = np.einsum("fd,fd->f", av, self._min_image(center - ocent)) < 0
            av[flip] *= -1
            a = np.linalg.norm(av, axis=1)
            n = av / a[:, None]
            has_n = c["nb"][sel] >= 0
            if np.any(has_n):
                ncent = self._centers(c["nb"][sel][has_n], c["nl"][sel][has_n])
                d = self._min_image(ncent - ocent[has_n])
                f2 = np.einsum("fd,fd->f", n[has_n], d) < 0
                idx = np.flatnonzero(has_n)[f2]
                n[idx] *= -1
            normal[sel] = n
            area[sel] = a
        return Faces(c["face_type"], c["sub"], c["ob"], c["ol"], c["of"], c["nb"], c["nl"],
                     c["nf"], oc, ncn, normal, area, tags)

    def _min_image(self, d):
        d = d.copy()
        for k in range(3):
            if self.period[k] > 0:
                d[:, k] -= self.period[k] * np.round(d[:, k] / self.period[k])
        return d

    _DECOMP = {"tetra": ((0, 1, 2, 3),),
               "wedge": ((0, 1, 2, 3), (1, 2, 4, 3), (2, 4, 5, 3)),
               "hexahedron": ((0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6),
                              (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6))}

    def _centers(self, block_ids, locals_):
        """Volume-weighted tetra-decomposition centres, hybrid/mesh.py:222-238."""
        out = np.zeros((len(block_ids), 3))
        for bid, b in enumerate(self.blocks):
            sel = np.flatnonzero(block_ids == bid)
            if not len(sel):
                continue
            X = b.coords[locals_[sel]]
            tot = np.zeros(len(sel))
            acc = np.zeros((len(sel), 3))
            for tet in self._DECOMP[b.element_type]:
                v = X[:, list(tet)]
                vol = np.abs(np.einsum("ed,ed->e", v[:, 1] - v[:, 0],
                                       np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6
                tot += vol
                acc += vol[:, None] * v.mean(1)
            out[sel] = acc / tot[:, None]
        return out
