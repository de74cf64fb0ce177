"""MOOD limiter facades on the GPU (hybrid/limiter.py:11-343 and
hybrid/subcell.py:442-950 of the reference).

``HybridMOODSubcellLimiter`` keeps the reference's methods: ``troubled``,
``recompute`` and ``positivity`` / ``__call__``.  ``HybridSSPRK3`` drives
its fused MOOD stage (``fsg_mood_stage``).  Everything runs in the
libfluid_gfx950.so kernels; this module only builds the device tables.
"""

import ctypes

import numpy as np
import torch

from . import _native as nat
from .elements import TYPES, TYPE_ID, basis, face_quadrature, nodes, reference_tables
from .subcells import NSUB, pair_subfaces, subcell_tables

RECONSTRUCTIONS = {"constant": 0, "tvd": 1, "weno": 2}
MAX_FACES = 12        # six local faces, each possibly cut in two sub-faces


def p1_projector(element_type, p):
    """hybrid/limiter.py:157-163: P1 values at the nodes times their
    pseudo-inverse."""
    phi = basis(element_type, 1, nodes(element_type, p))
    return phi @ np.linalg.pinv(phi)


class MoodTables:
    """Device tables of the MOOD kernels for one operator (affine elements)."""

    def __init__(self, operator, indexing="reference"):
        self.operator = operator
        dm = operator.device_mesh
        self.dm = dm
        topo = dm.topo
        p = dm.p
        self.indexing = indexing
        self.keep = []
        types_present = sorted({b.element_type for b in topo.blocks}, key=TYPE_ID.get)
        # Face sample points of each type: every face quadrature point the
        # mesh uses on that type (hybrid/limiter.py:28-67).
        pts = {t: [] for t in types_present}
        f = topo.faces
        for g in dm.groups:
            shapes, _ = face_quadrature(("quad", "triangle")[g["face_type"]], p)
            nv = shapes.shape[1]
            sel = g["sel"]
            sides = [(g["owner_block"], f.owner_corners)]
            if g["neighbor_block"] >= 0:
                sides.append((g["neighbor_block"], f.neighbor_corners))
            for blk, corners in sides:
                # the points depend only on the face's reference corners
                c = np.unique(np.round(corners[sel, :nv].reshape(len(sel), -1), 12), axis=0)
                pts[topo.blocks[blk].element_type].append(
                    np.einsum("qc,fcr->fqr", shapes, c.reshape(-1, nv, 3)).reshape(-1, 3))
        self.types = (nat.MoodType * 3)()
        for t in types_present:
            R = reference_tables(t, p)
            st = subcell_tables(t, p)
            face_pts = (np.unique(np.round(np.concatenate(pts[t]), 12), axis=0)
                        if pts[t] else np.zeros((0, 3)))
            samples = np.concatenate([R["B"], basis(t, p, face_pts)]) if len(face_pts) else R["B"]
            iptr, ilist = st.internal_csr()
            tptr, tlist, gather = st.boundary_targets(indexing)
            d = self.types[TYPE_ID[t]]
            d.N, d.Q, d.NS = R["N"], R["Q"], samples.shape[0]
            d.NI, d.NB = len(st.internal_owner), len(st.boundary_owner)
            d.p1proj = self._r(p1_projector(t, p).T)   # column j holds row j: coalesced
            d.samples = self._r(samples.T)        # (N, NS): coalesced over points
            d.wq = self._r(R["weights"])
            d.P = self._r(st.P)
            d.PT = self._r(st.P.T)
            d.RT = self._r(st.R.T)
            d.subvol = self._r(st.volume)
            d.int_o = self._i(st.internal_owner)
            d.int_n = self._i(st.internal_neighbor)
            d.int_area = self._r(st.internal_area)
            d.int_dc = self._r(st.internal_dc)
            d.int_ptr = self._i(iptr)
            d.int_list = self._i(ilist)
            d.bnd_face = self._i(st.boundary_parent_face)
            d.bnd_gather = self._i(gather)
            d.tgt_ptr = self._i(tptr)
            d.tgt_list = self._i(tlist)
        # Per-element face entries.
        nblk = len(topo.blocks)
        entries = [np.full((b.count, MAX_FACES, 2), -1, np.int32) for b in topo.blocks]
        fill = [np.zeros(b.count, np.int64) for b in topo.blocks]

        def add(block, elems, code, idx):
            k = fill[block][elems]
            if np.any(k >= MAX_FACES):
                raise ValueError("more than 12 face entries on one element")
            entries[block][elems, k, 0] = code
            entries[block][elems, k, 1] = idx
            fill[block][elems] += 1

        self.groups = (nat.MoodGroup * max(len(dm.groups), 1))()
        for gi, g in enumerate(dm.groups):
            sel = g["sel"]
            F = len(sel)
            idx = np.arange(F)
            ob, nb = g["owner_block"], g["neighbor_block"]
            oe = f.owner_local[sel]
            # an element can own several faces of one group: fill in rounds,
            # each round touching every element at most once
            for side, blk, elems, lfs in ((0, ob, oe, f.owner_face[sel]),) + (
                    ((1, nb, f.neighbor_local[sel], f.neighbor_face[sel]),) if nb >= 0 else ()):
                code = gi | (lfs.astype(np.int64) << 16) | (side << 24)
                order = np.argsort(elems, kind="stable")
                e_sorted = elems[order]
                # split repeated elements into rounds
                rank = np.zeros(F, np.int64)
                if F:
                    same = np.concatenate([[False], e_sorted[1:] == e_sorted[:-1]])
                    run = np.cumsum(~same) - 1
                    start = np.flatnonzero(~same)
                    rank = np.arange(F) - start[run]
                for r in range(int(rank.max()) + 1 if F else 0):
                    m = order[rank == r]
                    add(blk, elems[m], code[m].astype(np.int32), idx[m].astype(np.int32))
            self._pairs(gi, g, topo, p)
        self.blocks = (nat.MoodBlock * nblk)()
        self.flags, self.sensor = [], []
        for b in range(nblk):
            d = self.blocks[b]
            d.max_faces = MAX_FACES
            d.elem_faces = self._i(entries[b])
            fl = torch.zeros(topo.blocks[b].count, dtype=torch.int32, device=dm.device)
            sv = torch.zeros(topo.blocks[b].count, dtype=dm.dtype, device=dm.device)
            self.keep += [fl, sv]
            self.flags.append(fl)
            self.sensor.append(sv)
            d.flags = nat.ptr(fl)
            d.sensor = nat.ptr(sv)
        total = sum(b.count for b in topo.blocks)
        self.counter = torch.zeros(1, dtype=torch.int32, device=dm.device)
        self.list = torch.zeros(max(total, 1), dtype=torch.int32, device=dm.device)
        self.keep += [self.counter, self.list]

    def _r(self, a):
        t = torch.as_tensor(np.ascontiguousarray(a), dtype=self.dm.dtype).to(self.dm.device)
        self.keep.append(t)
        return nat.ptr(t)

    def _i(self, a):
        t = torch.as_tensor(np.ascontiguousarray(a, np.int32)).to(self.dm.device)
        self.keep.append(t)
        return ctypes.c_void_p(t.data_ptr())

    def _pairs(self, gi, g, topo, p):
        """Subface-pair tables of one face group, de-duplicated by the faces'
        reference orientation."""
        f = topo.faces
        sel = g["sel"]
        ob, nb = g["owner_block"], g["neighbor_block"]
        to = topo.blocks[ob].element_type
        tn = topo.blocks[nb].element_type if nb >= 0 else None
        nv = 4 if g["face_type"] == 0 else 3          # triangles leave corner 4 unset
        keys = np.concatenate([f.owner_face[sel, None].astype(float),
                               np.round(f.owner_corners[sel, :nv].reshape(len(sel), -1), 9)], 1)
        if nb >= 0:
            keys = np.concatenate([keys, f.neighbor_face[sel, None].astype(float),
                                   np.round(f.neighbor_corners[sel, :nv].reshape(len(sel), -1),
                                            9)], 1)
        uniq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
        ids, geom, bo, bn = [], [], [], []
        K = None
        for j in first:
            k = sel[j]
            pr = pair_subfaces(to, int(f.owner_face[k]), f.owner_corners[k],
                               tn, int(f.neighbor_face[k]) if tn else None,
                               f.neighbor_corners[k] if tn else None, p, bool(f.sub[k]))
            K = len(pr["osub"])
            nsub = pr["nsub"] if tn else np.full(K, -1)
            ids.append(np.stack([pr["osub"], nsub], 1))
            geom.append(np.concatenate([np.full((K, 1), float(pr["src"])), pr["area"]], 1))
            bo.append(basis(to, p, pr["opt"]))
            if tn:
                bn.append(basis(tn, p, pr["npt"]))
        d = self.groups[gi]
        d.K = K if K is not None else 0
        d.pair_table = self._i(inv.ravel())
        d.pair_ids = self._i(np.stack(ids) if ids else np.zeros((1, 1, 2)))
        d.pair_geom = self._r(np.stack(geom) if geom else np.zeros((1, 1, 4)))
        d.basis_o = self._r(np.stack(bo) if bo else np.zeros(1))
        d.basis_n = self._r(np.stack(bn)) if bn else ctypes.c_void_p(0)


class _MoodContext:
    def __init__(self, operator, threshold, density_floor, pressure_floor, relative_floor,
                 reconstruction, indexing):
        if reconstruction not in RECONSTRUCTIONS:
            raise ValueError(f"unknown reconstruction {reconstruction!r}")
        self.tables = MoodTables(operator, indexing)
        p = nat.MoodParams()
        p.threshold = float(threshold)
        p.density_floor = float(density_floor)
        p.pressure_floor = float(pressure_floor)
        p.relative_floor = float(relative_floor)
        p.reconstruction = RECONSTRUCTIONS[reconstruction]
        p.sensor_on = int(operator.solver.polynomial_degree >= 2)
        self.params = p
        ctx = ctypes.c_void_p()
        t = self.tables
        nat.check(nat.lib().fsg_mood_create(operator.device_mesh.ctx, ctypes.byref(p), t.types,
                                            t.blocks, t.groups, nat.ptr(t.counter),
                                            nat.ptr(t.list), ctypes.byref(ctx)))
        self.ctx = ctx

    def __del__(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            nat.lib().fsg_mood_destroy(self.ctx)
            self.ctx = None


class HybridPositivityLimiter:
    """hybrid/limiter.py:11-147 (affine elements)."""

    def __init__(self, operator, density_floor=1.0e-10, pressure_floor=1.0e-10,
                 relative_floor=0.05):
        self.operator = operator
        self.solver = operator.solver
        self.density_floor = float(density_floor)
        self.pressure_floor = float(pressure_floor)
        self.relative_floor = float(relative_floor)
        self._ctx = None

    def _context(self, threshold=5.0e-2, reconstruction="weno", indexing="reference"):
        if self._ctx is None:
            self._ctx = _MoodContext(self.operator, threshold, self.density_floor,
                                     self.pressure_floor, self.relative_floor, reconstruction,
                                     indexing)
        return self._ctx

    def __call__(self, states):
        c = self._context()
        op = self.operator
        out = {t: torch.empty_like(states[t]) for t in op.block_types}
        nat.check(nat.lib().fsg_mood_positivity(c.ctx, nat.ptr_array(op._inputs(states)),
                                                nat.ptr_array([out[t] for t in op.block_types]),
                                                nat.stream_handle(self.solver.device)))
        return out


class HybridModalShockLimiter:
    """hybrid/limiter.py:186-264: the conservative P2-to-P1 (or P0) fallback
    on sensor-flagged elements, then positivity; one device pass
    (fsg_mood_modal)."""

    def __init__(self, operator, threshold=5.0e-2, positivity=None, p0_fallback_factor=100.0,
                 p0_relative_floor=0.2):
        self.operator = operator
        self.solver = operator.solver
        self.threshold = float(threshold)
        self.positivity = HybridPositivityLimiter(operator) if positivity is None else positivity
        self.p0_fallback_factor = float(p0_fallback_factor)
        self.p0_relative_floor = float(p0_relative_floor)
        self._context_obj = None

    def _ctx(self):
        if self._context_obj is None:
            pos = self.positivity
            self._context_obj = _MoodContext(self.operator, self.threshold, pos.density_floor,
                                             pos.pressure_floor, pos.relative_floor, "weno",
                                             "reference")
        return self._context_obj

    def apply(self, states, out, diag=None):
        """Limit ``states`` into ``out`` (may be the same dict); with ``diag``
        (a device fsg_stage_diag) accumulate the admissibility of the
        result."""
        op = self.operator
        nat.check(nat.lib().fsg_mood_modal(
            self._ctx().ctx, nat.ptr_array(op._inputs(states)),
            nat.ptr_array([out[t] for t in op.block_types]), self.p0_fallback_factor,
            self.p0_relative_floor, diag, nat.stream_handle(self.solver.device)))
        return out

    def __call__(self, states):
        out = {t: torch.empty_like(states[t]) for t in self.operator.block_types}
        return self.apply(states, out)


class HybridTroubledCellSensor:
    """hybrid/limiter.py:150-185: returns (sensor values, sensor > threshold)."""

    def __init__(self, operator, threshold=5.0e-2, _mood=None):
        self.operator = operator
        self.solver = operator.solver
        self.threshold = float(threshold)
        self._mood = _mood

    def __call__(self, states):
        mood = self._mood or HybridMOODSubcellLimiter(self.operator, self.threshold)
        mood._detect(states)
        op = self.operator
        vals = {t: mood._ctx().tables.sensor[b].clone() for b, t in enumerate(op.block_types)}
        return vals, {t: v > self.threshold for t, v in vals.items()}


class HybridMOODSubcellLimiter:
    """hybrid/limiter.py:267-343 with the subcell recompute of
    hybrid/subcell.py:442-950.  ``indexing`` selects the boundary-subface
    indexing of the subcell arrays: "reference" reproduces the reference
    (subface ids used as subcell indices, clamped on gathers and dropped on
    scatters at 64), "owner" uses the owning subcell (conservative)."""

    def __init__(self, operator, threshold=5.0e-2, positivity=None, reconstruction="weno",
                 indexing="reference"):
        self.operator = operator
        self.solver = operator.solver
        self.threshold = float(threshold)
        self.positivity = HybridPositivityLimiter(operator) if positivity is None else positivity
        self.reconstruction = reconstruction
        self.indexing = indexing
        self.sensor = HybridTroubledCellSensor(operator, threshold, _mood=self)
        self._context_obj = None

    def _ctx(self):
        if self._context_obj is None:
            pos = self.positivity
            self._context_obj = _MoodContext(self.operator, self.threshold, pos.density_floor,
                                             pos.pressure_floor, pos.relative_floor,
                                             self.reconstruction, self.indexing)
            # the positivity facade shares this context (same floors)
            pos._ctx = self._context_obj
        return self._context_obj

    def _detect(self, states):
        op = self.operator
        nat.check(nat.lib().fsg_mood_troubled(self._ctx().ctx, nat.ptr_array(op._inputs(states)),
                                              nat.stream_handle(self.solver.device)))

    def troubled(self, candidate):
        """hybrid/limiter.py:285-300."""
        self._detect(candidate)
        t = self._ctx().tables
        return {name: t.flags[b].bool().clone() for b, name in enumerate(self.operator.block_types)}

    def troubled_count(self):
        """Troubled elements found by the last detection (synchronises)."""
        n = ctypes.c_int64()
        nat.check(nat.lib().fsg_mood_count(self._ctx().ctx, ctypes.byref(n),
                                           nat.stream_handle(self.solver.device)))
        return n.value

    def recompute(self, base, dg_candidate, dt, boundary_state=None, additional_rhs=None,
                  detection_state=None, troubled_mask=None):
        """hybrid/limiter.py:302-338: subcell recompute of the troubled
        elements from ``base`` and the mean correction of their neighbours,
        then positivity."""
        op = self.operator
        c = self._ctx()
        types = op.block_types
        out = {t: torch.empty_like(dg_candidate[t]) for t in types}
        use_flags = 0
        if troubled_mask is not None:
            for b, t in enumerate(types):
                c.tables.flags[b].copy_(troubled_mask[t].to(torch.int32))
            use_flags = 1
        dt_elem = None
        if isinstance(dt, dict):
            dts = [dt[t].reshape(-1).to(self.solver.dtype).contiguous() for t in types]
            dt_elem = nat.ptr_array(dts)
            dt = 0.0
        visc = nat.ptr_array(op._inputs(additional_rhs)) if additional_rhs is not None else None
        det = nat.ptr_array(op._inputs(detection_state)) if detection_state is not None else None
        bnd = op._boundary(boundary_state)
        nat.check(nat.lib().fsg_mood_recompute(
            c.ctx, nat.ptr_array(op._inputs(base)), nat.ptr_array(op._inputs(dg_candidate)),
            det, visc, float(dt), dt_elem, use_flags, ctypes.byref(bnd),
            nat.ptr_array([out[t] for t in types]), nat.stream_handle(self.solver.device)))
        return out

    def stage(self, base, curr, inv, visc, work, nxt, dt, alpha, beta, boundary_state, diag):
        """The MOOD branch of HybridSSPRK3.rk_stage after the residuals."""
        op = self.operator
        types = op.block_types
        dt_elem = None
        if isinstance(dt, dict):
            dts = [dt[t].reshape(-1).to(self.solver.dtype).contiguous() for t in types]
            dt_elem = nat.ptr_array(dts)
            dt = 0.0
        bnd = op._boundary(boundary_state)
        nat.check(nat.lib().fsg_mood_stage(
            self._ctx().ctx, nat.ptr_array(op._inputs(base)), nat.ptr_array(op._inputs(curr)),
            nat.ptr_array([inv[t] for t in types]),
            nat.ptr_array([visc[t] for t in types]) if visc is not None else None,
            nat.ptr_array([work[t] for t in types]), nat.ptr_array([nxt[t] for t in types]),
            float(dt), dt_elem, float(alpha), float(beta), ctypes.byref(bnd), diag,
            nat.stream_handle(self.solver.device)))

    def stage_fused(self, base, curr, total, visc, work, nxt, dt, alpha, beta, boundary_state,
                    diag):
        """The MOOD branch from one residual pass: ``total`` = rhs(curr), or
        None when ``work`` already holds the Euler candidate curr + dt rhs;
        the viscous rhs is evaluated into ``visc`` (or skipped when None) on
        the troubled elements only, after detection."""
        op = self.operator
        types = op.block_types
        dt_elem = None
        if isinstance(dt, dict):
            dts = [dt[t].reshape(-1).to(self.solver.dtype).contiguous() for t in types]
            dt_elem = nat.ptr_array(dts)
            dt = 0.0
        bnd = op._boundary(boundary_state)
        nat.check(nat.lib().fsg_mood_stage_fused(
            self._ctx().ctx, nat.ptr_array(op._inputs(base)), nat.ptr_array(op._inputs(curr)),
            None if total is None else nat.ptr_array([total[t] for t in types]),
            nat.ptr_array([visc[t] for t in types]) if visc is not None else None,
            nat.ptr_array([work[t] for t in types]), nat.ptr_array([nxt[t] for t in types]),
            float(dt), dt_elem, float(alpha), float(beta), ctypes.byref(bnd), diag,
            nat.stream_handle(self.solver.device)))

    def __call__(self, states):
        return self.positivity(states)
