"""SSP-RK3 stage/step and the CFL estimate, restated from
hybrid/time_integration.py (numpy FP64)."""

import numpy as np

RK3_WEIGHTS = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))  # :456-461


class SSPRK3:
    def __init__(self, operator, cfl=0.15, limiter=None, density_floor=1.0e-7,
                 pressure_floor=1.0e-7):
        self.operator = operator
        self.solver = operator.solver
        self.cfl = float(cfl)
        self.limiter = limiter
        self.density_floor = float(density_floor)
        self.pressure_floor = float(pressure_floor)
        self.characteristic_lengths = self._characteristic_lengths()

    def _characteristic_lengths(self):
        """hybrid/time_integration.py:49-91: max(2V/sum(A), V^(1/3)/(4(p+1)))."""
        mesh = self.operator.mesh
        geo = self.operator.geometry
        sums = {b.element_type: np.zeros(len(b.connectivity)) for b in mesh.element_blocks}
        for ftype, fb in mesh.face_blocks.items():
            areas = geo.face_metrics[ftype].integration_weights.sum(1)
            for bid, b in enumerate(mesh.element_blocks):
                o = np.flatnonzero(fb.owner_block == bid)
                np.add.at(sums[b.element_type], fb.owner_local[o], areas[o])
                nbi = np.flatnonzero(fb.neighbor_block == bid)
                np.add.at(sums[b.element_type], fb.neighbor_local[nbi], areas[nbi])
        out = {}
        p = self.operator.degree
        for bid, b in enumerate(mesh.element_blocks):
            vol = geo.element_metrics[bid].integration_weights.sum(1)
            out[b.element_type] = np.maximum(2.0 * vol / sums[b.element_type],
                                             0.25 * vol ** (1.0 / 3.0) / (p + 1))
        return out

    def stable_dt(self, states):
        """hybrid/time_integration.py:93-169 (no WALE)."""
        s = self.solver
        p = self.operator.degree
        k = 2 * p + 1
        rates = []
        for b in self.operator.mesh.element_blocks:
            t = b.element_type
            W = s.conserved_to_primitive(states[t])
            rho = np.maximum(W[..., 0], 1e-14)
            pr = np.maximum(W[..., 4], 1e-14)
            speed = np.linalg.norm(W[..., 1:4], axis=-1) + np.sqrt(s.gamma * pr / rho)
            L = self.characteristic_lengths[t]
            rate = k * speed.max(1) / L
            if s.dynamic_viscosity > 0.0:
                rmin = np.maximum(rho.mean(1), 0.1)
                diff = np.maximum((4.0 / 3.0) * s.dynamic_viscosity / rmin,
                                  s.thermal_conductivity / (rmin * s.specific_heat_cp))
                pen = float(getattr(self.operator, "br2_stabilization", 1.0))
                rate = rate + 2.0 * pen * k ** 2 * diff / np.maximum(L, 0.01) ** 2
            rate = np.nan_to_num(rate, nan=1e7, posinf=1e7)
            rates.append(rate.max())
        return float(np.clip(self.cfl / max(max(rates), 1e-12), 1e-7, 1e-3))

    def state_admissibility(self, states):
        """hybrid/time_integration.py:291-312 (nodes only)."""
        finite = True
        rmins, pmins = [], []
        for U in states.values():
            finite = finite and bool(np.all(np.isfinite(U)))
            rho = U[..., 0]
            rmins.append(rho.min())
            ke = 0.5 * np.sum(U[..., 1:4] ** 2, -1) / np.maximum(rho, 1e-14)
            pmins.append(((self.solver.gamma - 1.0) * (U[..., 4] - ke)).min())
        rmin, pmin = min(rmins), min(pmins)
        ok = finite and rmin >= self.density_floor and pmin >= self.pressure_floor
        return bool(ok), float(rmin), float(pmin)

    def rk_stage(self, base, curr, dt, alpha, beta, boundary_state=None):
        """hybrid/time_integration.py:404-439 (limiter=None branch when no
        limiter; MOOD branch delegated to ``limiter``)."""
        mood = self.limiter is not None and hasattr(self.limiter, "recompute")
        if mood:
            inv = self.operator.inviscid_rhs(curr, boundary_state)
            visc = self.operator.viscous_rhs(curr, boundary_state)
            total = {k: inv[k] + visc[k] for k in inv}
        else:
            total = self.operator.rhs(curr, boundary_state)
            visc = None
        dts = dt if isinstance(dt, dict) else {k: dt for k in curr}
        euler = {k: curr[k] + dts[k] * total[k] for k in curr}
        unlim = {k: alpha * base[k] + beta * euler[k] for k in curr}
        if mood:
            recomputed = self.limiter.recompute(curr, euler, dt, boundary_state, visc,
                                                detection_state=unlim)
            blended = {k: alpha * base[k] + beta * recomputed[k] for k in curr}
            result = self.limiter.positivity(blended)
        else:
            result = unlim if self.limiter is None else self.limiter(unlim)
        ok, rmin, pmin = self.state_admissibility(result)
        return result, ok, rmin, pmin

    def checked_step(self, states, dt, boundary_state=None):
        """hybrid/time_integration.py:441-583 (stage sequence and the
        valid/min reductions of the non-MOOD branch; the MOOD branch returns
        the last stage's diagnostics, :544-554)."""
        mood = self.limiter is not None and hasattr(self.limiter, "recompute")
        s1, v1, r1, p1 = self.rk_stage(states, states, dt, *RK3_WEIGHTS[0], boundary_state)
        s2, v2, r2, p2 = self.rk_stage(states, s1, dt, *RK3_WEIGHTS[1], boundary_state)
        s3, v3, r3, p3 = self.rk_stage(states, s2, dt, *RK3_WEIGHTS[2], boundary_state)
        if mood:
            return s3, v3, r3, p3
        return s3, v1 and v2 and v3, min(r1, r2, r3), min(p1, p2, p3)
