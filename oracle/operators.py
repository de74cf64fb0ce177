"""DG Euler and BR2 Navier-Stokes operators, restated from hybrid/spatial.py
and hybrid/navier_stokes.py (numpy FP64), with the optional WALE eddy
viscosity of hybrid/solver.py:317-406 evaluated from the resolved polynomial
gradient (the lifts excluded) as hybrid/navier_stokes.py:220-290 does."""

import numpy as np

from .geometry import HybridGeometry


def polynomial_br2_penalty(degree):
    """hybrid/navier_stokes.py:350-352."""
    return (degree + 1) * (degree + 2) / 2


def _scatter_add(target, cells, values):
    np.add.at(target, cells, values)


class EulerOperator:
    def __init__(self, mesh, solver, degree):
        self.mesh = mesh
        self.solver = solver
        self.degree = degree
        self.geometry = HybridGeometry(mesh, degree)
        self.face_groups = self._group_faces()

    def _group_faces(self):
        """hybrid/spatial.py:20-39."""
        groups = []
        for ftype, fb in self.mesh.face_blocks.items():
            keys = sorted(set(zip(fb.owner_block.tolist(), fb.neighbor_block.tolist())))
            for ob, nb in keys:
                idx = np.flatnonzero((fb.owner_block == ob) & (fb.neighbor_block == nb))
                groups.append((ftype, ob, nb, idx))
        return groups

    def _type(self, bid):
        return self.mesh.element_blocks[bid].element_type

    @staticmethod
    def cell_anchor(states):
        """hybrid/spatial.py:41-45: node 0, not a mean."""
        return {k: U[:, 0, :] for k, U in states.items()}

    def _volume_terms(self, states, anchors):
        """hybrid/spatial.py:47-72."""
        out = {}
        s = self.solver
        for bid, block in enumerate(self.mesh.element_blocks):
            t = block.element_type
            ref = self.geometry.reference_elements[t]
            m = self.geometry.element_metrics[bid]
            U = states[t]
            a = anchors[t]
            Uq = a[:, None, :] + np.einsum("qi,eiv->eqv", ref.basis(ref.volume_points),
                                           U - a[:, None, :])
            F = s.physical_flux_tensor(Uq)
            Fa = s.physical_flux_tensor(a)
            out[t] = np.einsum("eqid,eqdv,eq->eiv", m.physical_basis_gradients,
                               F - Fa[:, None], m.integration_weights)
        return out

    def residual(self, states, boundary_state=None):
        """hybrid/spatial.py:137-234."""
        s = self.solver
        anchors = self.cell_anchor(states)
        res = self._volume_terms(states, anchors)
        for ftype, ob, nb, idx in self.face_groups:
            fb = self.mesh.face_blocks[ftype]
            fm = self.geometry.face_metrics[ftype]
            to = self._type(ob)
            oref = self.geometry.reference_elements[to]
            ocells = fb.owner_local[idx]
            obasis = oref.basis(fm.owner_reference_points[idx])
            oa = anchors[to][ocells]
            Uo = oa[:, None, :] + np.einsum("fqi,fiv->fqv", obasis,
                                            states[to][ocells] - oa[:, None, :])
            n = fm.normals[idx]
            if nb >= 0:
                tn = self._type(nb)
                nref = self.geometry.reference_elements[tn]
                ncells = fb.neighbor_local[idx]
                nbasis = nref.basis(fm.neighbor_reference_points[idx])
                na = anchors[tn][ncells]
                Un = na[:, None, :] + np.einsum("fqi,fiv->fqv", nbasis,
                                                states[tn][ncells] - na[:, None, :])
            else:
                Un = Uo if boundary_state is None else boundary_state(Uo, n, fb.boundary_tags[idx])
            Fs = s.numerical_flux(Uo, Un, n)
            w = fm.integration_weights[idx][..., None]
            base_o = s.normal_physical_flux(oa[:, None, :], n)
            _scatter_add(res[to], ocells, -np.einsum("fqi,fqv->fiv", obasis, (Fs - base_o) * w))
            if nb >= 0:
                base_n = s.normal_physical_flux(na[:, None, :], n)
                _scatter_add(res[tn], ncells, np.einsum("fqi,fqv->fiv", nbasis, (Fs - base_n) * w))
        return res

    def apply_inverse_mass(self, residuals):
        """hybrid/spatial.py:236-249."""
        out = {}
        for bid, block in enumerate(self.mesh.element_blocks):
            t = block.element_type
            out[t] = np.einsum("eij,ejv->eiv",
                               self.geometry.element_metrics[bid].inverse_mass_matrix,
                               residuals[t])
        return out

    def rhs(self, states, boundary_state=None):
        return self.apply_inverse_mass(self.residual(states, boundary_state))


class NavierStokesOperator(EulerOperator):
    """hybrid/navier_stokes.py:9-347."""

    def __init__(self, mesh, solver, degree, br2_stabilization=None):
        super().__init__(mesh, solver, degree)
        self.br2_stabilization = float(polynomial_br2_penalty(degree)
                                       if br2_stabilization is None else br2_stabilization)
        # hybrid/navier_stokes.py:17-25: WALE filter width V^(1/3) / (p + 1)
        self.filter_widths = {
            b.element_type: self.geometry.element_metrics[bid].integration_weights.sum(1)
            ** (1.0 / 3.0) / (degree + 1)
            for bid, b in enumerate(mesh.element_blocks)}

    def _trace_data(self, states, group, boundary_state):
        """hybrid/navier_stokes.py:27-110."""
        s = self.solver
        ftype, ob, nb, idx = group
        fb = self.mesh.face_blocks[ftype]
        fm = self.geometry.face_metrics[ftype]
        to = self._type(ob)
        ocells = fb.owner_local[idx]
        obasis = self.geometry.reference_elements[to].basis(fm.owner_reference_points[idx])
        Uo_nodes = states[to][ocells]
        Uo = np.einsum("fqi,fiv->fqv", obasis, Uo_nodes)
        vo_nodes = s.viscous_variables(Uo_nodes)
        vo = np.einsum("fqi,fim->fqm", obasis, vo_nodes)
        d = dict(ftype=ftype, idx=idx, fb=fb, fm=fm, ob=ob, to=to, ocells=ocells,
                 opoints=fm.owner_reference_points[idx], obasis=obasis, Uo=Uo,
                 vo_nodes=vo_nodes, nb=nb)
        if nb >= 0:
            tn = self._type(nb)
            ncells = fb.neighbor_local[idx]
            nbasis = self.geometry.reference_elements[tn].basis(fm.neighbor_reference_points[idx])
            Un_nodes = states[tn][ncells]
            Un = np.einsum("fqi,fiv->fqv", nbasis, Un_nodes)
            vn_nodes = s.viscous_variables(Un_nodes)
            vn = np.einsum("fqi,fim->fqm", nbasis, vn_nodes)
            d.update(tn=tn, ncells=ncells, nbasis=nbasis, Un=Un, vn_nodes=vn_nodes)
        else:
            n = fm.normals[idx]
            tags = fb.boundary_tags[idx]
            Un = Uo if boundary_state is None else boundary_state(Uo, n, tags)
            if boundary_state is None or not hasattr(boundary_state, "viscous_ghost"):
                vn = s.viscous_variables(Un)
            else:
                vn = boundary_state.viscous_ghost(Uo, n, tags)
            d.update(Un=Un)
        d["jump"] = vn - vo
        return d

    def _physical_basis_gradients(self, bid, cells, points):
        """hybrid/navier_stokes.py:112-136."""
        t = self._type(bid)
        g = self.geometry.geometry_reference_elements[t]
        X = self.mesh.points[self.mesh.element_blocks[bid].connectivity[cells]]
        jac = np.einsum("fid,fqir->fqdr", X, g.basis_gradients(points))
        inv = np.linalg.inv(jac)
        sg = self.geometry.reference_elements[t].basis_gradients(points)
        return np.einsum("fqir,fqrd->fqid", sg, inv)

    def br2_gradients(self, states, boundary_state):
        """hybrid/navier_stokes.py:138-208."""
        s = self.solver
        lifts = {t: np.zeros(U.shape[:2] + (4, 3)) for t, U in states.items()}
        face_data = []
        for group in self.face_groups:
            d = self._trace_data(states, group, boundary_state)
            w = d["fm"].integration_weights[d["idx"]]
            n = d["fm"].normals[d["idx"]]
            orhs = 0.5 * np.einsum("fq,fqm,fqd,fqi->fimd", w, d["jump"], n, d["obasis"])
            ominv = self.geometry.element_metrics[d["ob"]].inverse_mass_matrix[d["ocells"]]
            d["olift"] = np.einsum("fij,fjmd->fimd", ominv, orhs)
            _scatter_add(lifts[d["to"]], d["ocells"], d["olift"])
            if d["nb"] >= 0:
                nrhs = 0.5 * np.einsum("fq,fqm,fqd,fqi->fimd", w, d["jump"], n, d["nbasis"])
                nminv = self.geometry.element_metrics[d["nb"]].inverse_mass_matrix[d["ncells"]]
                d["nlift"] = np.einsum("fij,fjmd->fimd", nminv, nrhs)
                _scatter_add(lifts[d["tn"]], d["ncells"], d["nlift"])
            face_data.append(d)
        raw, grads = {}, {}
        for bid, block in enumerate(self.mesh.element_blocks):
            t = block.element_type
            ref = self.geometry.reference_elements[t]
            m = self.geometry.element_metrics[bid]
            v = s.viscous_variables(states[t])
            raw[t] = np.einsum("eim,eqid->eqmd", v, m.physical_basis_gradients)
            grads[t] = raw[t] + np.einsum("qi,eimd->eqmd", ref.basis(ref.volume_points), lifts[t])
        return grads, raw, face_data

    def viscous_residual(self, states, boundary_state=None):
        """hybrid/navier_stokes.py:210-324."""
        s = self.solver
        grads, raw, face_data = self.br2_gradients(states, boundary_state)
        res = {}
        for bid, block in enumerate(self.mesh.element_blocks):
            t = block.element_type
            ref = self.geometry.reference_elements[t]
            m = self.geometry.element_metrics[bid]
            Uq = np.einsum("qi,eiv->eqv", ref.basis(ref.volume_points), states[t])
            F = s.viscous_physical_flux(Uq, grads[t], self.filter_widths[t][:, None],
                                        turbulent_gradient=raw[t])
            res[t] = -np.einsum("eqid,eqdv,eq->eiv", m.physical_basis_gradients, F,
                                m.integration_weights)
        eta = self.br2_stabilization
        for d in face_data:
            fm, idx = d["fm"], d["idx"]
            n = fm.normals[idx]
            w = fm.integration_weights[idx]
            og = self._physical_basis_gradients(d["ob"], d["ocells"], d["opoints"])
            raw_o = np.einsum("fim,fqid->fqmd", d["vo_nodes"], og)
            loc_o = np.einsum("fqi,fimd->fqmd", d["obasis"], d["olift"])
            Fo = s.viscous_physical_flux(d["Uo"], raw_o + eta * loc_o,
                                         self.filter_widths[d["to"]][d["ocells"]][:, None],
                                         turbulent_gradient=raw_o)
            if d["nb"] >= 0:
                ng = self._physical_basis_gradients(d["nb"], d["ncells"],
                                                    fm.neighbor_reference_points[idx])
                raw_n = np.einsum("fim,fqid->fqmd", d["vn_nodes"], ng)
                loc_n = np.einsum("fqi,fimd->fqmd", d["nbasis"], d["nlift"])
                Fn = s.viscous_physical_flux(d["Un"], raw_n + eta * loc_n,
                                             self.filter_widths[d["tn"]][d["ncells"]][:, None],
                                             turbulent_gradient=raw_n)
                Ff = 0.5 * (Fo + Fn)
            else:
                Ff = Fo
            nf = np.einsum("fqdv,fqd->fqv", Ff, n)
            if d["nb"] < 0 and boundary_state is not None and hasattr(boundary_state, "wall_tags"):
                wall = boundary_state._matches(d["fb"].boundary_tags[idx], boundary_state.wall_tags)
                nf[..., 4] = np.where(wall[:, None], 0.0, nf[..., 4])
            wf = nf * w[..., None]
            _scatter_add(res[d["to"]], d["ocells"], np.einsum("fqi,fqv->fiv", d["obasis"], wf))
            if d["nb"] >= 0:
                _scatter_add(res[d["tn"]], d["ncells"], -np.einsum("fqi,fqv->fiv", d["nbasis"], wf))
        return res

    def viscous_rhs(self, states, boundary_state=None):
        return self.apply_inverse_mass(self.viscous_residual(states, boundary_state))

    def inviscid_rhs(self, states, boundary_state=None):
        return EulerOperator.rhs(self, states, boundary_state)

    def rhs(self, states, boundary_state=None):
        """hybrid/navier_stokes.py:338-343."""
        e = self.inviscid_rhs(states, boundary_state)
        if self.solver.dynamic_viscosity == 0.0 and not self.solver.use_wale:
            return e
        v = self.viscous_rhs(states, boundary_state)
        return {k: e[k] + v[k] for k in e}
