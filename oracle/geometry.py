"""Isoparametric element and face metrics, restated from hybrid/geometry.py.

Linear (P1) geometry only; curved P2 geometry (hybrid/geometry.py:103-109) is
not part of the synthetic box meshes.
"""

from dataclasses import dataclass

import numpy as np

from .elements import face_quadrature, make_reference_element


@dataclass(frozen=True)
class ElementMetrics:
    """hybrid/geometry.py:27-36."""
    physical_points: np.ndarray
    jacobian: np.ndarray
    inverse_jacobian: np.ndarray
    determinant: np.ndarray
    integration_weights: np.ndarray
    physical_basis_gradients: np.ndarray
    mass_matrix: np.ndarray
    inverse_mass_matrix: np.ndarray


@dataclass(frozen=True)
class FaceMetrics:
    """hybrid/geometry.py:39-45."""
    physical_points: np.ndarray
    normals: np.ndarray
    integration_weights: np.ndarray
    owner_reference_points: np.ndarray
    neighbor_reference_points: np.ndarray


SLIVER_VOLUME = 1.0e-4  # hybrid/geometry.py:190


class HybridGeometry:
    def __init__(self, mesh, degree):
        self.mesh = mesh
        self.polynomial_degree = degree
        types = {b.element_type for b in mesh.element_blocks}
        self.reference_elements = {t: make_reference_element(t, degree) for t in types}
        self.geometry_reference_elements = {t: make_reference_element(t, 1) for t in types}
        self.element_metrics = self._element_metrics()
        self.solution_node_coordinates = self._solution_nodes()
        self.face_metrics = self._face_metrics()

    def _coords(self, block):
        return self.mesh.points[block.connectivity]

    def _solution_nodes(self):
        """hybrid/geometry.py:115-137."""
        out = {}
        for block in self.mesh.element_blocks:
            ref = self.reference_elements[block.element_type]
            g = self.geometry_reference_elements[block.element_type]
            out[block.element_type] = np.einsum("qi,eid->eqd", g.basis(ref.nodes),
                                                self._coords(block))
        return out

    def _element_metrics(self):
        """hybrid/geometry.py:139-203 including the sliver scaling of M^-1."""
        out = {}
        for bid, block in enumerate(self.mesh.element_blocks):
            ref = self.reference_elements[block.element_type]
            g = self.geometry_reference_elements[block.element_type]
            X = self._coords(block)
            qp = ref.volume_points
            phys = np.einsum("qi,eid->eqd", g.basis(qp), X)
            jac = np.einsum("eid,qir->eqdr", X, g.basis_gradients(qp))
            det = np.linalg.det(jac)
            inv = np.linalg.inv(jac)
            pgrad = np.einsum("qir,eqrd->eqid", ref.basis_gradients(qp), inv)
            w = ref.volume_weights[None, :] * np.abs(det)
            phi = ref.basis(qp)
            mass = np.einsum("eq,qi,qj->eij", w, phi, phi)
            minv = np.linalg.inv(mass)
            vols = w.sum(1)
            minv = np.minimum(1.0, vols / SLIVER_VOLUME)[:, None, None] * minv
            out[bid] = ElementMetrics(phys, jac, inv, det, w, pgrad, mass, minv)
        return out

    def _face_metrics(self):
        """hybrid/geometry.py:226-362.  Owner and neighbour reference corners come
        from the mesh's canonical-id lookup (hybrid/geometry.py:205-224)."""
        out = {}
        for ftype, fb in self.mesh.face_blocks.items():
            shapes, nw = face_quadrature(ftype, self.polynomial_degree)
            own_pts = np.einsum("qc,fcr->fqr", shapes, fb.owner_corners)
            nbr_pts = np.einsum("qc,fcr->fqr", shapes, fb.neighbor_corners)
            if ftype == "triangle":
                tu = np.broadcast_to((fb.owner_corners[:, 1] - fb.owner_corners[:, 0])[:, None],
                                     own_pts.shape)
                tv = np.broadcast_to((fb.owner_corners[:, 2] - fb.owner_corners[:, 0])[:, None],
                                     own_pts.shape)
            else:
                u = shapes[:, 1] + shapes[:, 2]
                v = shapes[:, 2] + shapes[:, 3]
                c0, c1, c2, c3 = (fb.owner_corners[:, k][:, None] for k in range(4))
                tu = (-(1 - v)[None, :, None] * c0 + (1 - v)[None, :, None] * c1
                      + v[None, :, None] * c2 - v[None, :, None] * c3)
                tv = (-(1 - u)[None, :, None] * c0 - u[None, :, None] * c1
                      + u[None, :, None] * c2 + (1 - u)[None, :, None] * c3)
            F, Q = own_pts.shape[:2]
            phys = np.zeros((F, Q, 3))
            normals = np.zeros((F, Q, 3))
            weights = np.zeros((F, Q))
            measure = 0.5 if ftype == "triangle" else 1.0
            for ob in np.unique(fb.owner_block):
                idx = np.flatnonzero(fb.owner_block == ob)
                block = self.mesh.element_blocks[int(ob)]
                g = self.geometry_reference_elements[block.element_type]
                X = self.mesh.points[block.connectivity[fb.owner_local[idx]]]
                pts = own_pts[idx]
                mapped = np.einsum("fqi,fid->fqd", g.basis(pts), X)
                jac = np.einsum("fid,fqir->fqdr", X, g.basis_gradients(pts))
                t_u = np.einsum("fqdr,fqr->fqd", jac, tu[idx])
                t_v = np.einsum("fqdr,fqr->fqd", jac, tv[idx])
                av = np.cross(t_u, t_v)
                orient = np.where(np.sum(av * fb.normals[idx][:, None, :], -1) >= 0, 1.0, -1.0)
                av = av * orient[..., None]
                m = np.linalg.norm(av, axis=-1)
                phys[idx] = mapped
                normals[idx] = av / m[..., None]
                weights[idx] = nw[None, :] * measure * m
            out[ftype] = FaceMetrics(phys, normals, weights, own_pts, nbr_pts)
        return out
