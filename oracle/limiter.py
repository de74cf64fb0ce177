"""Positivity limiter, troubled-cell sensor and MOOD driver, restated from
hybrid/limiter.py (numpy FP64)."""

import numpy as np

from .elements import basis as lagrange_basis
from .subcell import SubcellFV


def p1_basis(element_type, points):
    """hybrid/elements.py:131-159."""
    return lagrange_basis(element_type, 1, points)


class PositivityLimiter:
    """hybrid/limiter.py:11-147."""

    def __init__(self, operator, density_floor=1.0e-10, pressure_floor=1.0e-10,
                 relative_floor=0.05):
        self.operator = operator
        self.solver = operator.solver
        self.density_floor = float(density_floor)
        self.pressure_floor = float(pressure_floor)
        self.relative_floor = float(relative_floor)
        self.face_sample_basis = self._face_sample_basis()

    def _face_sample_basis(self):
        """hybrid/limiter.py:28-67: distinct face points (rounded to 12
        decimals) of each element type."""
        op = self.operator
        pts = {b.element_type: [] for b in op.mesh.element_blocks}
        for ftype, ob, nb, idx in op.face_groups:
            fm = op.geometry.face_metrics[ftype]
            pts[op._type(ob)].append(fm.owner_reference_points[idx].reshape(-1, 3))
            if nb >= 0:
                pts[op._type(nb)].append(fm.neighbor_reference_points[idx].reshape(-1, 3))
        out = {}
        for t, chunks in pts.items():
            ref = op.geometry.reference_elements[t]
            if chunks:
                u = np.unique(np.round(np.concatenate(chunks), 12), axis=0)
                out[t] = ref.basis(u)
            else:
                out[t] = np.zeros((0, len(ref.nodes)))
        return out

    def mean(self, block_id, U):
        """hybrid/limiter.py:69-80."""
        t = self.operator._type(block_id)
        ref = self.operator.geometry.reference_elements[t]
        w = self.operator.geometry.element_metrics[block_id].integration_weights
        Uq = np.einsum("qi,eiv->eqv", ref.basis(ref.volume_points), U)
        return np.einsum("eqv,eq->ev", Uq, w) / w.sum(1)[:, None]

    def sample(self, element_type, U):
        """hybrid/limiter.py:82-92: nodes, volume points, face points."""
        ref = self.operator.geometry.reference_elements[element_type]
        Uq = np.einsum("qi,eiv->eqv", ref.basis(ref.volume_points), U)
        Uf = np.einsum("qi,eiv->eqv", self.face_sample_basis[element_type], U)
        return np.concatenate([U, Uq, Uf], 1)

    def _pressure_ok(self, mean, dev, theta, floor):
        """hybrid/limiter.py:94-98."""
        s = mean[:, None, :] + theta[:, None, None] * dev
        p = self.solver.conserved_to_primitive(s)[..., 4]
        return p.min(1) >= floor

    def limit_block(self, block_id, U):
        """hybrid/limiter.py:100-139: closed-form density theta, then a
        30-step bisection on the pressure theta."""
        t = self.operator._type(block_id)
        mean = self.mean(block_id, U)
        Wm = self.solver.conserved_to_primitive(mean)
        rho_floor = np.maximum(self.density_floor, self.relative_floor * Wm[:, 0])
        p_floor = np.maximum(self.pressure_floor, self.relative_floor * Wm[:, 4])
        smp = self.sample(t, U)
        dev = smp - mean[:, None, :]
        rmin = smp[..., 0].min(1)
        den = mean[:, 0] - rmin
        with np.errstate(invalid="ignore", divide="ignore"):
            th = np.where(rmin < rho_floor, (mean[:, 0] - rho_floor) / np.maximum(den, 1e-30), 1.0)
        upper = np.clip(th, 0.0, 1.0)
        lower = np.zeros_like(upper)
        hi = upper.copy()
        for _ in range(30):
            mid = 0.5 * (lower + hi)
            ok = self._pressure_ok(mean, dev, mid, p_floor)
            lower, hi = np.where(ok, mid, lower), np.where(ok, hi, mid)
        ok_upper = self._pressure_ok(mean, dev, upper, p_floor)
        theta = np.where(ok_upper, upper, lower)
        return mean[:, None, :] + theta[:, None, None] * (U - mean[:, None, :])

    def __call__(self, states):
        return {b.element_type: self.limit_block(bid, states[b.element_type])
                for bid, b in enumerate(self.operator.mesh.element_blocks)}


class TroubledCellSensor:
    """hybrid/limiter.py:150-185: Persson-like ratio of the non-P1 energy of
    density and pressure at the nodes."""

    def __init__(self, operator, threshold=5.0e-2):
        self.operator = operator
        self.solver = operator.solver
        self.threshold = float(threshold)
        self.p1_projectors = {}
        if operator.degree >= 2:
            for t, ref in operator.geometry.reference_elements.items():
                phi = p1_basis(t, ref.nodes)
                self.p1_projectors[t] = phi @ np.linalg.pinv(phi)

    def block_sensor(self, element_type, U):
        if self.operator.degree < 2:
            return np.zeros(U.shape[0])
        W = self.solver.conserved_to_primitive(U)
        m = W[..., [0, 4]]
        low = np.einsum("ij,ejm->eim", self.p1_projectors[element_type], m)
        high = ((m - low) ** 2).sum(1)
        scale = (m ** 2).sum(1)
        return (high / np.maximum(scale, 1e-12)).max(-1)

    def __call__(self, states):
        sensor = {k: self.block_sensor(k, U) for k, U in states.items()}
        return sensor, {k: v > self.threshold for k, v in sensor.items()}


class ModalShockLimiter:
    """hybrid/limiter.py:186-264: on sensor-flagged elements (degree >= 2)
    the P1 projection of every variable with the element mean restored, or
    the mean itself where that projection is non-finite at the samples or
    its density / pressure there drops below p0_fallback_factor x the floors
    or p0_relative_floor x the mean's; then the positivity limiter."""

    def __init__(self, operator, threshold=5.0e-2, positivity=None, p0_fallback_factor=100.0,
                 p0_relative_floor=0.2):
        self.operator = operator
        self.sensor = TroubledCellSensor(operator, threshold)
        self.positivity = PositivityLimiter(operator) if positivity is None else positivity
        self.p0_fallback_factor = float(p0_fallback_factor)
        self.p0_relative_floor = float(p0_relative_floor)

    def __call__(self, states):
        if self.operator.degree < 2:
            return self.positivity(states)
        _, troubled = self.sensor(states)
        pos = self.positivity
        s = self.operator.solver
        out = {}
        for bid, b in enumerate(self.operator.mesh.element_blocks):
            t = b.element_type
            U = states[t]
            low = np.einsum("ij,ejv->eiv", self.sensor.p1_projectors[t], U)
            old = pos.mean(bid, U)
            low = low + (old - pos.mean(bid, low))[:, None, :]
            smp = pos.sample(t, low)
            with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
                W = s.conserved_to_primitive(smp)
                dmin, pmin = W[..., 0].min(1), W[..., 4].min(1)
                Wm = s.conserved_to_primitive(old)
            md = np.maximum(Wm[..., 0], pos.density_floor)
            mp = np.maximum(Wm[..., 4], pos.pressure_floor)
            finite = np.all(np.isfinite(smp), axis=(1, 2))
            severe = (~finite | (dmin < self.p0_fallback_factor * pos.density_floor)
                      | (pmin < self.p0_fallback_factor * pos.pressure_floor)
                      | (dmin < self.p0_relative_floor * md)
                      | (pmin < self.p0_relative_floor * mp))
            fb = np.where(severe[:, None, None], old[:, None, :], low)
            out[t] = np.where(troubled[t][:, None, None], fb, U)
        return pos(out)


class MOODSubcellLimiter:
    """hybrid/limiter.py:267-343."""

    def __init__(self, operator, threshold=5.0e-2, positivity=None, reconstruction="weno"):
        self.operator = operator
        self.sensor = TroubledCellSensor(operator, threshold)
        self.positivity = PositivityLimiter(operator) if positivity is None else positivity
        self.subcell = SubcellFV(operator, reconstruction=reconstruction)

    def troubled(self, candidate):
        """hybrid/limiter.py:285-300: sensor flag or PAD failure at the samples
        and the mean."""
        _, flags = self.sensor(candidate)
        out = {}
        for bid, b in enumerate(self.operator.mesh.element_blocks):
            t = b.element_type
            smp = self.positivity.sample(t, candidate[t])
            mean = self.positivity.mean(bid, candidate[t])
            smp = np.concatenate([smp, mean[:, None, :]], 1)
            with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
                W = self.solver_c2p(smp)
                invalid = (~np.all(np.isfinite(W), axis=(1, 2))
                           | (W[..., 0].min(1) < self.positivity.density_floor)
                           | (W[..., 4].min(1) < self.positivity.pressure_floor))
            out[t] = flags[t] | invalid
        return out

    def solver_c2p(self, U):
        return self.operator.solver.conserved_to_primitive(U)

    def recompute(self, base, dg_candidate, dt, boundary_state=None, additional_rhs=None,
                  detection_state=None, troubled_mask=None):
        """hybrid/limiter.py:302-338."""
        troubled = (troubled_mask if troubled_mask is not None else
                    self.troubled(dg_candidate if detection_state is None else detection_state))
        if any(np.any(v) for v in troubled.values()):
            corrected = self.subcell.recompute(base, dg_candidate, dt, troubled,
                                               boundary_state, additional_rhs)
        else:
            corrected = dg_candidate
        return self.positivity(corrected)

    def __call__(self, states):
        return self.positivity(states)
