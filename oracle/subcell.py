"""Subcell finite-volume recompute for MOOD, restated from hybrid/subcell.py.

Two indexings of the boundary subfaces are supported:

* ``"reference"`` (default, parity with the reference): hybrid/subcell.py:
  774-776, 805-807 and 842-848 index the (E, 64, V) subcell arrays with the
  *boundary-subface* number (0..95 for hexahedra, 0..79 for prisms, 0..63
  for tetrahedra) instead of the subcell that owns the subface.  Under JAX
  semantics the gathers clamp the index to 63 and the scatter-adds drop
  indices >= 64; both are reproduced here.
* ``"owner"``: the subcell that owns each boundary subface (the evident
  intent; conservative to round-off).

Sub-faces (a quad face cut into two triangles, which the reference mesh
rejects) pair each of the triangle side's 16 boundary subfaces with the quad
side's boundary subface containing its centroid; normal and area come from
the triangle side (extension, parity unpinned beyond conservation).

The TVD and WENO reconstructions are pinned by closed forms on the inner
subcells of a hexahedron (exact Green-Gauss gradients for linear and
quadratic data) and by the TVD trace bound on random data
(tests/test_oracle_mood.py).
"""

import numpy as np

from .elements import basis as lagrange_basis, basis_gradients, volume_quadrature
from .mesh import LOCAL_FACES


def p1(element_type, x):
    return lagrange_basis(element_type, 1, x)


def _triangle_lattice(n):
    """hybrid/subcell.py:52-72."""
    verts = np.asarray([(i / n, j / n) for i in range(n + 1) for j in range(n + 1 - i)])
    look = {tuple(np.rint(v * n).astype(int)): k for k, v in enumerate(verts)}
    tris = []
    for i in range(n):
        for j in range(n - i):
            tris.append([look[(i, j)], look[(i + 1, j)], look[(i, j + 1)]])
            if i + j <= n - 2:
                tris.append([look[(i + 1, j)], look[(i + 1, j + 1)], look[(i, j + 1)]])
    return verts, np.asarray(tris, np.int32)


def _tetra_subdivision():
    """hybrid/subcell.py:75-110: two red refinements (64 tetrahedra)."""
    verts = [np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0]),
             np.array([0.0, 0, 1])]
    cells = [(0, 1, 2, 3)]
    for _ in range(2):
        look = {tuple(p): k for k, p in enumerate(verts)}

        def mid(a, b):
            pt = 0.5 * (verts[a] + verts[b])
            key = tuple(pt)
            if key not in look:
                look[key] = len(verts)
                verts.append(pt)
            return look[key]

        ref = []
        for a, b, c, d in cells:
            ab, ac, ad = mid(a, b), mid(a, c), mid(a, d)
            bc, bd, cd = mid(b, c), mid(b, d), mid(c, d)
            ref += [(a, ab, ac, ad), (ab, b, bc, bd), (ac, bc, c, cd), (ad, bd, cd, d),
                    (ab, ac, ad, bd), (ab, bc, ac, bd), (ac, ad, bd, cd), (ac, bd, bc, cd)]
        cells = ref
    return np.asarray(verts), np.asarray(cells, np.int32)


def _wedge_subdivision(n):
    """hybrid/subcell.py:113-132."""
    tv, tris = _triangle_lattice(n)
    verts = np.asarray([(x, y, l / n) for l in range(n + 1) for x, y in tv])
    m = len(tv)
    cells = [[l * m + a, l * m + b, l * m + c, (l + 1) * m + a, (l + 1) * m + b, (l + 1) * m + c]
             for l in range(n) for a, b, c in tris]
    return verts, np.asarray(cells, np.int32)


def _hexahedron_subdivision(n):
    """hybrid/subcell.py:135-164."""
    verts = np.asarray([(i / n, j / n, k / n) for k in range(n + 1) for j in range(n + 1)
                        for i in range(n + 1)])

    def idx(i, j, k):
        return k * (n + 1) * (n + 1) + j * (n + 1) + i

    cells = [[idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
              idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1),
              idx(i, j + 1, k + 1)]
             for k in range(n) for j in range(n) for i in range(n)]
    return verts, np.asarray(cells, np.int32)


def _parent_face(element_type, x, tol=1e-10):
    """hybrid/subcell.py:167-180."""
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    tests = {"tetra": (Z, Y, X + Y + Z - 1.0, X),
             "wedge": (Z, Z - 1.0, Y, X + Y - 1.0, X),
             "hexahedron": (Z, Z - 1.0, Y, X - 1.0, Y - 1.0, X)}[element_type]
    for f, v in enumerate(tests):
        if np.all(np.abs(v) <= tol):
            return f
    return -1


class ReferenceSubcells:
    """hybrid/subcell.py:19-34, 183-233."""

    def __init__(self, element_type, n=4):
        if n != 4:
            raise ValueError("division_count must be 4")
        self.element_type = element_type
        if element_type == "tetra":
            self.vertices, self.cells = _tetra_subdivision()
        elif element_type == "wedge":
            self.vertices, self.cells = _wedge_subdivision(n)
        else:
            self.vertices, self.cells = _hexahedron_subdivision(n)
        pending = {}
        io, inb, ifaces = [], [], []
        for c, conn in enumerate(self.cells):
            for lf in LOCAL_FACES[element_type]:
                face = np.asarray(conn[list(lf)], np.int32)
                key = tuple(sorted(int(k) for k in face))
                if key in pending:
                    oc, of = pending.pop(key)
                    io.append(oc)
                    inb.append(c)
                    ifaces.append(of)
                else:
                    pending[key] = (c, face)
        bo, bp, bfaces = [], [], []
        for c, face in pending.values():
            pf = _parent_face(element_type, self.vertices[face])
            if pf < 0:
                raise ValueError("subcell boundary face off the parent boundary")
            bo.append(c)
            bp.append(pf)
            bfaces.append(face)
        self.internal_owner = np.asarray(io, np.int32)
        self.internal_neighbor = np.asarray(inb, np.int32)
        self.internal_faces = tuple(ifaces)
        self.boundary_owner = np.asarray(bo, np.int32)
        self.boundary_parent_face = np.asarray(bp, np.int32)
        self.boundary_faces = tuple(bfaces)

    @property
    def count(self):
        return len(self.cells)


def _mapped_quadrature(rs):
    """hybrid/subcell.py:236-274."""
    t = rs.element_type
    pts, w = volume_quadrature(t, 2)
    cv = rs.vertices[rs.cells]
    mapped = np.einsum("qi,sir->sqr", p1(t, pts), cv)
    cols = (1, 3, 4) if t == "hexahedron" else (1, 2, 3)
    M = np.stack([cv[:, k] - cv[:, 0] for k in cols], -1)
    return mapped, w[None, :] * np.abs(np.linalg.det(M))[:, None]


def _area_vectors(xyz_faces):
    v = xyz_faces
    a = 0.5 * np.cross(v[..., 1, :] - v[..., 0, :], v[..., 2, :] - v[..., 0, :])
    if v.shape[-2] == 4:
        a = a + 0.5 * np.cross(v[..., 2, :] - v[..., 0, :], v[..., 3, :] - v[..., 0, :])
    return a


class SubcellProjection:
    """hybrid/subcell.py:277-396 (linear geometry)."""

    def __init__(self, operator, block_id, rs):
        block = operator.mesh.element_blocks[block_id]
        t = block.element_type
        p = operator.degree
        X = operator.mesh.points[block.connectivity]
        pv = np.einsum("vi,eid->evd", p1(t, rs.vertices), X)
        pc = pv[:, rs.cells].mean(2)
        mpts, rw = _mapped_quadrature(rs)
        flat = mpts.reshape(-1, 3)
        phi = lagrange_basis(t, p, flat).reshape(rs.count, mpts.shape[1], -1)
        gg = basis_gradients(t, 1, flat).reshape(rs.count, mpts.shape[1], -1, 3)
        jac = np.einsum("eir,sqij->esqrj", X, gg)
        pw = np.abs(np.linalg.det(jac)) * rw[None]
        vol = pw.sum(-1)
        P = np.einsum("esq,sqi->esi", pw, phi) / vol[..., None]
        gram = np.einsum("esi,es,esj->eij", P, vol, P)
        R = np.einsum("eij,esj,es->eis", np.linalg.inv(gram), P, vol)
        self.projection, self.reconstruction, self.volumes = P, R, vol
        self.physical_centers = pc
        self.reference_centers = rs.vertices[rs.cells].mean(1)

        def metrics(faces, owners, nbrs=None):
            av = np.stack([_area_vectors(pv[:, np.asarray(f)]) for f in faces], 1)
            cen = np.stack([pv[:, np.asarray(f)].mean(1) for f in faces], 1)
            oc = pc[:, owners]
            d = cen - oc if nbrs is None else pc[:, nbrs] - oc
            s = np.where(np.sum(av * d, -1) >= 0.0, 1.0, -1.0)
            av = av * s[..., None]
            ar = np.linalg.norm(av, axis=-1)
            return av / ar[..., None], ar, cen

        self.internal_normals, self.internal_areas, _ = metrics(
            rs.internal_faces, rs.internal_owner, rs.internal_neighbor)
        self.boundary_normals, self.boundary_areas, self.boundary_physical_centers = metrics(
            rs.boundary_faces, rs.boundary_owner)
        self.boundary_reference_centers = np.asarray(
            [rs.vertices[f].mean(0) for f in rs.boundary_faces])


def dg_to_subcells(U, proj):
    """hybrid/subcell.py:399-403."""
    return np.einsum("esi,eiv->esv", proj.projection, U)


def subcells_to_dg(avg, proj, bound_limit=True):
    """hybrid/subcell.py:406-439: LSQ left inverse then a Zhang-Shu bound to
    the subcell min/max preserving the element mean."""
    raw = np.einsum("eis,esv->eiv", proj.reconstruction, avg)
    if not bound_limit:
        return raw
    vols = proj.volumes
    ev = vols.sum(1, keepdims=True)
    mean = (avg * vols[..., None]).sum(1) / np.maximum(ev, 1e-14)
    smin, smax = avg.min(1), avg.max(1)
    dmin, dmax = raw.min(1), raw.max(1)
    eps = 1e-12
    with np.errstate(invalid="ignore", divide="ignore"):
        tmax = np.where(dmax > mean + eps, (smax - mean) / (dmax - mean + eps), 1.0)
        tmin = np.where(dmin < mean - eps, (mean - smin) / (mean - dmin + eps), 1.0)
    th = np.clip(np.minimum(tmax, tmin), 0.0, 1.0)
    return mean[:, None, :] + th[:, None, :] * (raw - mean[:, None, :])


def _admissible_traces(U):
    """The clamps of hybrid/subcell.py:648-663, 727-740, 823-837."""
    U = U.copy()
    U[..., 0] = np.maximum(U[..., 0], 1e-14)
    ke = 0.5 * np.sum(U[..., 1:4] ** 2, -1) / U[..., 0]
    U[..., 4] = np.maximum(U[..., 4], ke + 1e-12)
    if U.shape[-1] > 5:
        U[..., 5:] = np.maximum(U[..., 5:], 0.0)
    return U


class SubcellFV:
    """hybrid/subcell.py:442-950."""

    def __init__(self, operator, division_count=4, reconstruction="weno", indexing="reference"):
        self.operator = operator
        self.solver = operator.solver
        self.mesh = operator.mesh
        self.reconstruction = reconstruction
        self.indexing = indexing
        self.references, self.projections = {}, {}
        for bid, b in enumerate(self.mesh.element_blocks):
            rs = ReferenceSubcells(b.element_type, division_count)
            self.references[b.element_type] = rs
            self.projections[b.element_type] = SubcellProjection(operator, bid, rs)
        self.face_groups = self._face_groups()

    def _boundary_ids(self, t, lf):
        return np.flatnonzero(self.references[t].boundary_parent_face == lf)

    def _face_groups(self):
        """hybrid/subcell.py:473-533 plus the sub-face pairing.  Each group
        lists, per face, S subface pairs: (owner boundary-subface id,
        neighbour boundary-subface id, owner ref point, neighbour ref point,
        normal, area)."""
        op = self.operator
        groups = []
        for ftype, ob, nb, idx in op.face_groups:
            fb = self.mesh.face_blocks[ftype]
            to = op._type(ob)
            opj = self.projections[to]
            rec = dict(ftype=ftype, ob=ob, nb=nb, idx=idx, osub=[], nsub=[], opt=[], npt=[],
                       nrm=[], area=[])
            for f in idx:
                ol, of = int(fb.owner_local[f]), int(fb.owner_face[f])
                oids = self._boundary_ids(to, of)
                if nb < 0:
                    rec["osub"].append(oids)
                    rec["opt"].append(opj.boundary_reference_centers[oids])
                    rec["nrm"].append(opj.boundary_normals[ol, oids])
                    rec["area"].append(opj.boundary_areas[ol, oids])
                    continue
                tn = op._type(nb)
                npj = self.projections[tn]
                nl, nf = int(fb.neighbor_local[f]), int(fb.neighbor_face[f])
                nids = self._boundary_ids(tn, nf)
                oxyz = opj.boundary_physical_centers[ol, oids]
                nxyz = npj.boundary_physical_centers[nl, nids]
                shift = self._period_shift(oxyz, nxyz)
                sub = len(LOCAL_FACES[to][of]) == 4 and ftype == "triangle"
                if not sub:
                    d = np.linalg.norm(oxyz[:, None] - (nxyz[None] + shift), axis=-1)
                    perm = np.argmin(d, 1)
                    scale = max(float(np.max(np.linalg.norm(oxyz, axis=1))), 1.0)
                    if len(np.unique(perm)) != len(oids) or \
                            d[np.arange(len(oids)), perm].max() > 2e-6 * scale:
                        raise ValueError(f"Nonconforming subfaces at {ftype} face {f}")
                    rec["osub"].append(oids)
                    rec["nsub"].append(nids[perm])
                    rec["opt"].append(opj.boundary_reference_centers[oids])
                    rec["npt"].append(npj.boundary_reference_centers[nids[perm]])
                    rec["nrm"].append(opj.boundary_normals[ol, oids])
                    rec["area"].append(opj.boundary_areas[ol, oids])
                else:
                    # neighbour sub-triangles inside this triangle half of the quad
                    fm = op.geometry.face_metrics[ftype]
                    tri_ref = fb.neighbor_corners[f]
                    inside = self._inside_triangle(npj.boundary_reference_centers[nids], tri_ref)
                    nsel = nids[inside]
                    cen_phys = npj.boundary_physical_centers[nl, nsel] + shift
                    owner_pts = self._owner_points(ob, ol, cen_phys)
                    osel = self._containing_subface(to, of, oids, owner_pts)
                    rec["osub"].append(osel)
                    rec["nsub"].append(nsel)
                    rec["opt"].append(owner_pts)
                    rec["npt"].append(npj.boundary_reference_centers[nsel])
                    rec["nrm"].append(-npj.boundary_normals[nl, nsel])
                    rec["area"].append(npj.boundary_areas[nl, nsel])
            for k in ("osub", "nsub", "opt", "npt", "nrm", "area"):
                rec[k] = np.asarray(rec[k]) if rec[k] else None
            groups.append(rec)
        return groups

    def _period_shift(self, oxyz, nxyz):
        per = self.mesh.period
        if per is None:
            return np.zeros(3)
        d = oxyz.mean(0) - nxyz.mean(0)
        s = np.zeros(3)
        for k in range(3):
            if per[k] > 0:
                s[k] = per[k] * np.round(d[k] / per[k])
        return s

    @staticmethod
    def _inside_triangle(pts, tri):
        """Reference points lying inside the (neighbour-side) triangle face."""
        a, b, c = tri[0], tri[1], tri[2]
        n = np.cross(b - a, c - a)
        out = []
        for x in pts:
            ok = True
            for p0, p1 in ((a, b), (b, c), (c, a)):
                if np.dot(np.cross(p1 - p0, x - p0), n) < -1e-12:
                    ok = False
            out.append(ok)
        return np.asarray(out)

    def _owner_points(self, ob, ol, xyz):
        """Owner reference coordinates of physical points (affine cells)."""
        block = self.mesh.element_blocks[ob]
        X = self.mesh.points[block.connectivity[ol]]
        t = block.element_type
        cols = (1, 3, 4) if t == "hexahedron" else (1, 2, 3)
        J = np.stack([X[k] - X[0] for k in cols], 1)
        return np.linalg.solve(J, (xyz - X[0]).T).T

    def _containing_subface(self, t, lf, ids, pts):
        rs = self.references[t]
        out = []
        for x in pts:
            best, bd = None, np.inf
            for j in ids:
                v = rs.vertices[rs.boundary_faces[j]]
                d = np.linalg.norm(v.mean(0) - x)
                if d < bd:
                    best, bd = j, d
            out.append(best)
        return np.asarray(out)

    # -- indexing of subcell arrays by boundary subface ids --------------------------
    def _gather_ids(self, t, ids):
        if self.indexing == "reference":
            return np.minimum(ids, self.references[t].count - 1)
        return self.references[t].boundary_owner[ids]

    def _scatter(self, target, cells, t, ids, values):
        if self.indexing == "reference":
            keep = ids < self.references[t].count
            c = np.broadcast_to(cells[:, None], ids.shape)
            np.add.at(target, (c[keep], ids[keep]), values[keep])
        else:
            sc = self.references[t].boundary_owner[ids]
            np.add.at(target, (np.broadcast_to(cells[:, None], ids.shape), sc), values)

    def project(self, states):
        return {k: dg_to_subcells(U, self.projections[k]) for k, U in states.items()}

    def _reconstruct(self, t, U):
        """hybrid/subcell.py:541-665."""
        rs = self.references[t]
        o, nb = rs.internal_owner, rs.internal_neighbor
        if self.reconstruction == "constant":
            return _admissible_traces(U[:, o]), _admissible_traces(U[:, nb])
        pj = self.projections[t]
        nrm, ar, vol, cen = pj.internal_normals, pj.internal_areas, pj.volumes, pj.physical_centers
        dU = U[:, nb] - U[:, o]
        dc = cen[:, nb] - cen[:, o]
        duo = np.zeros_like(dU)
        dun = np.zeros_like(dU)
        for d in range(3):
            fd = 0.5 * nrm[:, :, d] * ar
            contrib = fd[:, :, None] * dU
            g = np.zeros((U.shape[0], rs.count, U.shape[-1]))
            E = np.arange(U.shape[0])[:, None]
            np.add.at(g, (E, o[None, :]), contrib)
            np.add.at(g, (E, nb[None, :]), contrib)
            g = g / np.maximum(vol[:, :, None], 1e-14)
            duo = duo + g[:, o] * dc[:, :, d, None]
            dun = dun - g[:, nb] * dc[:, :, d, None]
        if self.reconstruction == "tvd":
            eps = 1e-12
            with np.errstate(invalid="ignore", divide="ignore"):
                ro = dU / (duo + np.where(duo >= 0, eps, -eps))
                rn = (-dU) / (dun + np.where(dun >= 0, eps, -eps))
                po = np.clip(np.where(ro > 0, (ro ** 2 + 2 * ro) / (ro ** 2 + ro + 2), 0), 0, 1)
                pn = np.clip(np.where(rn > 0, (rn ** 2 + 2 * rn) / (rn ** 2 + rn + 2), 0), 0, 1)
            UL = U[:, o] + 0.5 * po * duo
            UR = U[:, nb] + 0.5 * pn * dun
        else:
            so, sn = duo, -dun
            Ph = 0.5 * (U[:, o] + U[:, nb]) + 0.125 * (so - sn)
            Po = U[:, o] + 0.5 * so
            Pn = U[:, nb] - 0.5 * sn
            e = 1e-12
            sec = dU
            bh = (sn - so) ** 2 + 0.1 * sec ** 2
            bo = (sec - so) ** 2 + 0.1 * sec ** 2
            bn = (sec - sn) ** 2 + 0.1 * sec ** 2
            bc = sec ** 2
            tau = np.abs(bc - bh)
            a0 = 0.70 * (1 + (tau / (bh + e)) ** 2)
            a1 = 0.28 * (1 + (tau / (bo + e)) ** 2)
            a2 = 0.02 * (1 + (tau / (bc + e)) ** 2)
            sL = a0 + a1 + a2
            UL = (a0 / sL) * Ph + (a1 / sL) * Po + (a2 / sL) * U[:, o]
            b1 = 0.28 * (1 + (tau / (bn + e)) ** 2)
            sR = a0 + b1 + a2
            UR = (a0 / sR) * Ph + (b1 / sR) * Pn + (a2 / sR) * U[:, nb]
        return _admissible_traces(UL), _admissible_traces(UR)

    def _internal_residual(self, sub):
        """hybrid/subcell.py:667-688."""
        out = {}
        for t, U in sub.items():
            rs = self.references[t]
            pj = self.projections[t]
            UL, UR = self._reconstruct(t, U)
            flux = self.solver.numerical_flux(UL, UR, pj.internal_normals) * pj.internal_areas[..., None]
            r = np.zeros_like(U)
            E = np.arange(U.shape[0])[:, None]
            np.add.at(r, (E, rs.internal_owner[None, :]), -flux)
            np.add.at(r, (E, rs.internal_neighbor[None, :]), flux)
            out[t] = r
        return out

    def _dg_group_flux(self, states, g, boundary_state):
        """hybrid/subcell.py:690-746: DG face flux without the anchor, with
        the trace clamps."""
        op = self.operator
        fb = self.mesh.face_blocks[g["ftype"]]
        fm = op.geometry.face_metrics[g["ftype"]]
        idx = g["idx"]
        to = op._type(g["ob"])
        oc = fb.owner_local[idx]
        UL = np.einsum("fqi,fiv->fqv", op.geometry.reference_elements[to].basis(
            fm.owner_reference_points[idx]), states[to][oc])
        n = fm.normals[idx]
        if g["nb"] >= 0:
            tn = op._type(g["nb"])
            UR = np.einsum("fqi,fiv->fqv", op.geometry.reference_elements[tn].basis(
                fm.neighbor_reference_points[idx]), states[tn][fb.neighbor_local[idx]])
        else:
            UR = UL if boundary_state is None else boundary_state(UL, n, fb.boundary_tags[idx])
        UL, UR = _admissible_traces(UL), _admissible_traces(UR)
        return np.sum(self.solver.numerical_flux(UL, UR, n) * fm.integration_weights[idx][..., None], 1)

    def _external_fluxes(self, states, sub, troubled, residual, boundary_state):
        """hybrid/subcell.py:748-853."""
        op = self.operator
        new, old = [], []
        for g in self.face_groups:
            fb = self.mesh.face_blocks[g["ftype"]]
            idx = g["idx"]
            to = op._type(g["ob"])
            oc = fb.owner_local[idx]
            osub = g["osub"]
            odg = np.einsum("fki,fiv->fkv", op.geometry.reference_elements[to].basis(g["opt"]),
                            states[to][oc])
            ofv = sub[to][oc[:, None], self._gather_ids(to, osub)]
            oflag = troubled[to][oc]
            UL = np.where(oflag[:, None, None], ofv, odg)
            n, ar = g["nrm"], g["area"]
            if g["nb"] >= 0:
                tn = op._type(g["nb"])
                nc = fb.neighbor_local[idx]
                nsub = g["nsub"]
                ndg = np.einsum("fki,fiv->fkv", op.geometry.reference_elements[tn].basis(g["npt"]),
                                states[tn][nc])
                nfv = sub[tn][nc[:, None], self._gather_ids(tn, nsub)]
                UR = np.where(troubled[tn][nc][:, None, None], nfv, ndg)
            else:
                UR = UL if boundary_state is None else boundary_state(UL, n, fb.boundary_tags[idx])
            UL, UR = _admissible_traces(UL), _admissible_traces(UR)
            flux = self.solver.numerical_flux(UL, UR, n) * ar[..., None]
            self._scatter(residual[to], oc, to, osub, -flux)
            if g["nb"] >= 0:
                self._scatter(residual[tn], nc, tn, nsub, flux)
            new.append(flux.sum(1))
            old.append(self._dg_group_flux(states, g, boundary_state))
        return new, old

    def recompute(self, states, dg_candidate, dt, troubled, boundary_state=None,
                  additional_rhs=None):
        """hybrid/subcell.py:855-950."""
        op = self.operator
        sub = self.project(states)
        res = self._internal_residual(sub)
        new, old = self._external_fluxes(states, sub, troubled, res, boundary_state)

        def dts(t):
            if isinstance(dt, dict):
                v = np.asarray(dt[t])
                return v if v.ndim == 3 else v[..., None]
            return dt

        fv = {t: sub[t] + dts(t) * res[t] / self.projections[t].volumes[..., None] for t in sub}
        if additional_rhs is not None:
            fv = {t: v + dts(t) * dg_to_subcells(additional_rhs[t], self.projections[t])
                  for t, v in fv.items()}
        rec = {t: subcells_to_dg(v, self.projections[t]) for t, v in fv.items()}
        corr = {t: v.copy() for t, v in dg_candidate.items()}
        for g, nf, of in zip(self.face_groups, new, old):
            fb = self.mesh.face_blocks[g["ftype"]]
            idx = g["idx"]
            to = op._type(g["ob"])
            oc = fb.owner_local[idx]
            oflag = troubled[to][oc]
            if g["nb"] >= 0:
                tn = op._type(g["nb"])
                nc = fb.neighbor_local[idx]
                nflag = troubled[tn][nc]
            else:
                nflag = np.zeros_like(oflag)
            act = oflag | nflag
            diff = np.where(act[:, None], nf - of, 0.0)

            def fdt(t, cells):
                if isinstance(dt, dict):
                    return np.asarray(dt[t])[cells].reshape(len(cells), 1)
                return dt

            ovol = self.projections[to].volumes[oc].sum(1)
            oco = np.where((act & ~oflag)[:, None], -fdt(to, oc) * diff / ovol[:, None], 0.0)
            np.add.at(corr[to], oc, np.broadcast_to(oco[:, None, :], (len(oc),) + corr[to].shape[1:]))
            if g["nb"] >= 0:
                nvol = self.projections[tn].volumes[nc].sum(1)
                nco = np.where((act & ~nflag)[:, None], fdt(tn, nc) * diff / nvol[:, None], 0.0)
                np.add.at(corr[tn], nc,
                          np.broadcast_to(nco[:, None, :], (len(nc),) + corr[tn].shape[1:]))
        return {t: np.where(troubled[t][:, None, None], rec[t], corr[t]) for t in states}
