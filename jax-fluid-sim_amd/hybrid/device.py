"""Device-resident operator tables and the native context.

``DeviceMesh`` turns a ``Topology`` into the tables of include/fluid_gfx950.h:

* per element block: reference operators (basis and gradients at volume
  points in the layouts the element kernel reads, reference inverse mass)
  and per-element affine metrics;
* per face group (face type, owner block, neighbour block, sub-face flag;
  the grouping of hybrid/spatial.py:20-39): owner/neighbour element ids,
  normal and area, and de-duplicated trace tables for each distinct set of
  reference points (hybrid/geometry.py:238-256 computes these per face);
* a CSR from each element to its face-side records in the face-output
  buffer, so the element kernel gathers its surface terms without atomics.

Faces inside a group are ordered by owner element so neighbouring work items
touch neighbouring cells.
"""

import ctypes

import numpy as np
import torch

from . import _native as nat
from .elements import LOCAL_FACES, TYPE_ID, basis, basis_gradients, corners, face_quadrature, \
    reference_tables
from .topology import SLIVER_VOLUME, Topology


def _face_gauss(p):
    from .elements import _gauss01
    return _gauss01({2: 3}.get(p, 4))[0]


def hex_face_factors(p):
    """1-D factors at the quad-face Gauss points for the sum-factorised BR2
    face gradients of hexahedral sides: [B1 (Qf1, N1), D1 (Qf1, N1), l(0),
    l'(0), l(1), l'(1)], flattened; None below p = 2."""
    if p < 2:
        return None
    from .elements import _dline, _line
    g = _face_gauss(p)
    ends = [f(p, np.array([x]))[0] for x in (0.0, 1.0) for f in (_line, _dline)]
    return np.concatenate([_line(p, g).ravel(), _dline(p, g).ravel(), *ends])


def hex_face_map(element_type, p, pts):
    """For a trace table of a hexahedral side on a quad face (p >= 2): the
    face points are the Gauss grid on one face plane, so return [normal axis,
    plane 0/1, tangential axes t1 < t2, each point's Gauss index along t1,
    along t2] (int), else None."""
    if element_type != "hexahedron" or p < 2:
        return None
    g = _face_gauss(p)
    if len(pts) != len(g) ** 2:
        return None
    for dn in range(3):
        col = pts[:, dn]
        if np.all(np.abs(col - col[0]) < 1e-12) and min(abs(col[0]), abs(col[0] - 1)) < 1e-12:
            break
    else:
        return None
    t1, t2 = [d for d in range(3) if d != dn]
    idx = []
    for t in (t1, t2):
        d = np.abs(pts[:, t][:, None] - g[None, :])
        k = d.argmin(1)
        if d[np.arange(len(pts)), k].max() > 1e-12:
            return None
        idx.append(k)
    # the grid must be complete: every (a, b) pair exactly once
    if len(set(zip(idx[0].tolist(), idx[1].tolist()))) != len(pts):
        return None
    return np.concatenate([[dn, int(round(pts[0, dn])), t1, t2], idx[0], idx[1]])


def hex_factors(element_type, p, R):
    """1-D factors [B1 (Q1, N1), D1 (Q1, N1), M1^-1 (N1, N1)] of the
    hexahedral volume operators at p >= 2, flattened, or None.  The nodes and
    the Gauss points are tensor products with x fastest (hybrid/elements.py),
    so B = B1 (x) B1 (x) B1, d/dxi_x = B1 (x) B1 (x) D1 (cyclic) and
    M^-1 = M1^-1 (x) M1^-1 (x) M1^-1; checked here against the dense tables."""
    if element_type != "hexahedron" or p < 2:
        return None
    from .elements import _dline, _gauss01, _line
    g, w = _gauss01(5 if p == 3 else 3)
    B1, D1 = _line(p, g), _dline(p, g)
    M1inv = np.linalg.inv(B1.T @ (w[:, None] * B1))
    kron3 = lambda z, y, x: np.kron(z, np.kron(y, x))
    checks = [(kron3(B1, B1, B1), R["B"]), (kron3(B1, B1, D1), R["D"][..., 0]),
              (kron3(B1, D1, B1), R["D"][..., 1]), (kron3(D1, B1, B1), R["D"][..., 2])]
    for approx, dense in checks:
        if np.abs(approx - dense).max() > 1e-12 * max(1.0, np.abs(dense).max()):
            raise RuntimeError("hexahedral tables are not tensor products in the expected order")
    Mk = kron3(M1inv, M1inv, M1inv)
    if np.abs(Mk - R["Minv"]).max() > 1e-9 * np.abs(R["Minv"]).max():
        raise RuntimeError("hexahedral inverse mass is not the Kronecker product of its factors")
    return np.concatenate([B1.ravel(), D1.ravel(), M1inv.ravel()])

FACE_TYPE_NAME = ("quad", "triangle")


def n_face_nodes(face_type, sub, p):
    return (p + 1) ** 2 if (face_type == 0 or sub) else (p + 1) * (p + 2) // 2


def _support(element_type, local_face, p):
    """Nodes lying on the plane of a local face of the reference element."""
    c = corners(element_type)[list(LOCAL_FACES[element_type][local_face])]
    nrm = np.cross(c[1] - c[0], c[2] - c[0])
    x = reference_tables(element_type, p)["nodes"]
    return np.flatnonzero(np.abs((x - c[0]) @ nrm) < 1e-9)


class DeviceMesh:
    def __init__(self, topology: Topology, degree, nvars, params, dtype=torch.float64,
                 device="cuda"):
        self.topo = topology
        self.p = int(degree)
        self.nvars = int(nvars)
        self.dtype = dtype
        self.device = torch.device(device)
        self.keep = []          # every tensor the context points to
        self._build_blocks()
        self._build_groups()
        self._build_csr()
        self._create(params)

    # -- helpers -----------------------------------------------------------
    def _t(self, a, dtype=None):
        t = torch.as_tensor(np.ascontiguousarray(a),
                            dtype=dtype or self.dtype).to(self.device).contiguous()
        self.keep.append(t)
        return t

    # -- element blocks ------------------------------------------------------
    def _build_blocks(self):
        self.block_tables = []
        for b in self.topo.blocks:
            R = reference_tables(b.element_type, self.p)
            vol = np.abs(b.det) * R["weights"].sum()
            scale = np.minimum(1.0, vol / SLIVER_VOLUME) / np.abs(b.det)
            t = dict(
                type=b.element_type, count=b.count, N=R["N"], Q=R["Q"],
                Bt=self._t(R["B"].T), Dt=self._t(R["D"].transpose(1, 2, 0)),
                D=self._t(R["D"].transpose(0, 2, 1)), wq=self._t(R["weights"]),
                minv=self._t(R["Minv"]), minvT=self._t(R["Minv"].T),
                jinv=self._t(b.jinv.reshape(-1, 9)),
                scale=self._t(scale), detw=self._t(np.abs(b.det)), volume=vol,
                # WALE filter width V^(1/3) / (p + 1) (hybrid/navier_stokes.py:17-25)
                width=self._t(vol ** (1.0 / 3.0) / (self.p + 1)))
            sf = hex_factors(b.element_type, self.p, R)
            t["sf1d"] = None if sf is None else self._t(sf)
            self.block_tables.append(t)

    # -- face groups -------------------------------------------------------
    def _tables(self, element_type, lfaces, points):
        """De-duplicate per-face reference points into trace tables."""
        p = self.p
        F, Qf, _ = points.shape
        key = np.round(points.reshape(F, -1), 11)
        uniq, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
        R = reference_tables(element_type, p)
        blobs, sups, metas = [], [], []
        for u, f0 in enumerate(first):
            pts = points[f0]
            sup = _support(element_type, int(lfaces[f0]), p)
            phi = basis(element_type, p, pts)
            dphi = basis_gradients(element_type, p, pts)
            lift = (phi @ R["Minv"])[:, sup]
            blobs.append(np.concatenate([phi[:, sup].ravel(), lift.ravel(), dphi.ravel()]))
            sups.append(sup)
            metas.append(hex_face_map(element_type, p, pts))
        meta = None if any(m is None for m in metas) else np.stack(metas).astype(np.int32)
        self._last_meta = meta
        return np.stack(blobs), np.stack(sups).astype(np.int32), inv.ravel().astype(np.int32)

    def _build_groups(self):
        f = self.topo.faces
        p = self.p
        W = self.nvars + 12
        keys = np.stack([f.face_type, f.owner_block, f.neighbor_block, f.sub], 1)
        ukeys = np.unique(keys, axis=0)
        self.groups = []
        offset = 0
        support_chunks = []
        sup_base = 0
        for ft, ob, nb, sub in ukeys:
            sel = np.flatnonzero((keys == (ft, ob, nb, sub)).all(1))
            sel = sel[np.argsort(f.owner_local[sel], kind="stable")]
            F = len(sel)
            shapes, nw = face_quadrature(FACE_TYPE_NAME[ft], p)
            nv = shapes.shape[1]
            to = self.topo.blocks[ob].element_type
            opts = np.einsum("qc,fcr->fqr", shapes, f.owner_corners[sel, :nv])
            otab, osup, oidx = self._tables(to, f.owner_face[sel], opts)
            omap = self._last_meta
            FO = n_face_nodes(ft, sub, p)
            assert osup.shape[1] == FO, (to, ft, sub, osup.shape)
            g = dict(face_type=int(ft), owner_block=int(ob), neighbor_block=int(nb),
                     sub=int(sub), count=F, sel=sel, FO=FO, FN=0)
            g["owner_elem"] = self._t(f.owner_local[sel], torch.int32)
            g["geom"] = self._t(np.concatenate([f.normal[sel], f.area[sel, None]], 1))
            g["owner_table"] = self._t(oidx, torch.int32)
            g["owner_tables"] = self._t(otab)
            g["owner_sf"] = None if omap is None else self._t(omap, torch.int32)
            g["owner_sup_host"] = osup
            g["owner_sup_base"] = sup_base
            support_chunks.append(osup.ravel())
            sup_base += osup.size
            g["face_weights"] = self._t(nw)
            g["out_owner"] = offset
            offset += F * FO * W
            if nb >= 0:
                tn = self.topo.blocks[nb].element_type
                npts = np.einsum("qc,fcr->fqr", shapes, f.neighbor_corners[sel, :nv])
                ntab, nsup, nidx = self._tables(tn, f.neighbor_face[sel], npts)
                g["neighbor_sf"] = (None if self._last_meta is None
                                    else self._t(self._last_meta, torch.int32))
                FN = n_face_nodes(ft, 0, p)
                assert nsup.shape[1] == FN, (tn, ft, nsup.shape)
                g["FN"] = FN
                g["neighbor_elem"] = self._t(f.neighbor_local[sel], torch.int32)
                g["neighbor_table"] = self._t(nidx, torch.int32)
                g["neighbor_tables"] = self._t(ntab)
                g["neighbor_sup_host"] = nsup
                g["neighbor_sup_base"] = sup_base
                support_chunks.append(nsup.ravel())
                sup_base += nsup.size
                g["out_neighbor"] = offset
                offset += F * FN * W
                g["bflags"] = None
            else:
                g["tags"] = f.tag[sel]
                g["bflags"] = self._t(np.zeros(F), torch.int32)
                g["out_neighbor"] = 0
            self.groups.append(g)
        self.support = self._t(np.concatenate(support_chunks) if support_chunks
                               else np.zeros(1), torch.int32)
        for g in self.groups:
            g["owner_sup"] = self.support[g["owner_sup_base"]:]
            if g["FN"]:
                g["neighbor_sup"] = self.support[g["neighbor_sup_base"]:]
        self.face_out = torch.empty(max(offset, 1), dtype=self.dtype, device=self.device)
        self.keep.append(self.face_out)
        ff = hex_face_factors(p)
        self.face_sf1d = None if ff is None else self._t(ff)

    def _build_csr(self):
        W = self.nvars + 12
        recs = [[] for _ in self.topo.blocks]   # (element, offset, sup offset, nf)
        for g in self.groups:
            F = g["count"]
            j = np.arange(F)
            FO = g["FO"]
            sel = g["sel"]
            f = self.topo.faces
            oidx = g["owner_table"].cpu().numpy()
            recs[g["owner_block"]].append(np.stack([
                f.owner_local[sel], g["out_owner"] + j * FO * W,
                g["owner_sup_base"] + oidx * FO, np.full(F, FO)], 1))
            if g["FN"]:
                FN = g["FN"]
                nidx = g["neighbor_table"].cpu().numpy()
                recs[g["neighbor_block"]].append(np.stack([
                    f.neighbor_local[sel], g["out_neighbor"] + j * FN * W,
                    g["neighbor_sup_base"] + nidx * FN, np.full(F, FN)], 1))
        for bid, bt in enumerate(self.block_tables):
            r = np.concatenate(recs[bid]) if recs[bid] else np.zeros((0, 4), np.int64)
            r = r[np.argsort(r[:, 0], kind="stable")]
            counts = np.bincount(r[:, 0], minlength=bt["count"])
            ptr = np.concatenate([[0], np.cumsum(counts)])
            bt["rec_ptr"] = self._t(ptr, torch.int64)
            bt["rec_off"] = self._t(r[:, 1], torch.int64)
            bt["rec_sup"] = self._t(r[:, 2], torch.int32)
            bt["rec_nf"] = self._t(r[:, 3], torch.int32)

    def _create(self, params):
        blocks = (nat.BlockDesc * len(self.block_tables))()
        for i, t in enumerate(self.block_tables):
            d = blocks[i]
            d.element_type = TYPE_ID[t["type"]]
            d.count = t["count"]
            for k in ("Bt", "Dt", "D", "wq", "minv", "minvT", "jinv", "scale", "detw", "rec_ptr",
                      "rec_off", "rec_sup", "rec_nf", "width"):
                setattr(d, k, t[k].data_ptr())
            d.sf1d = 0 if t["sf1d"] is None else t["sf1d"].data_ptr()
        groups = (nat.GroupDesc * max(len(self.groups), 1))()
        for i, g in enumerate(self.groups):
            d = groups[i]
            d.face_type, d.owner_block, d.neighbor_block, d.sub = (
                g["face_type"], g["owner_block"], g["neighbor_block"], g["sub"])
            d.count = g["count"]
            for k in ("owner_elem", "geom", "owner_table", "owner_tables", "owner_sup",
                      "face_weights", "bflags", "neighbor_elem", "neighbor_table",
                      "neighbor_tables", "neighbor_sup", "owner_sf", "neighbor_sf"):
                v = g.get(k)
                setattr(d, k, 0 if v is None else v.data_ptr())
            d.face_sf1d = 0 if self.face_sf1d is None else self.face_sf1d.data_ptr()
            d.out_owner = g["out_owner"]
            d.out_neighbor = g["out_neighbor"]
        self._params = params
        ctx = ctypes.c_void_p()
        nat.check(nat.lib().fsg_create(ctypes.byref(params), len(self.block_tables), blocks,
                                       len(self.groups), groups, nat.ptr(self.support),
                                       nat.ptr(self.face_out), ctypes.byref(ctx)))
        self.ctx = ctx

    def set_boundary_flags(self, flags_of_tag):
        """Fill each boundary group's per-face flags from a tag -> flags map."""
        for g in self.groups:
            if g["neighbor_block"] < 0:
                fl = np.array([flags_of_tag(int(t)) for t in g["tags"]], np.int32)
                g["bflags"].copy_(torch.as_tensor(fl))

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                nat.lib().fsg_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    def set_hex_sum_factorisation(self, on=None):
        """Hexahedral volume operators at p >= 2 by their 1-D factors (True
        or None: the default) or dense (False); round-off apart, the same
        operator."""
        nat.check(nat.lib().fsg_set_option(self.ctx, 2, -1 if on is None else int(bool(on))))

    def set_element_path(self, mfma=None):
        """Element operators as batched matrix-core products (True) or one
        workgroup per element on the vector units (False, or None: the
        default); the discrete operator is the same."""
        nat.check(nat.lib().fsg_set_option(self.ctx, 1, -1 if mfma is None else int(bool(mfma))))
