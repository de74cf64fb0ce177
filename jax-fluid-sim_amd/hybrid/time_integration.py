"""HybridSSPRK3 facade (hybrid/time_integration.py:8-671 of the reference).

``rk_stage`` and ``checked_step`` keep the reference signatures and return
``(states, valid, rho_min, p_min)``; the stage is one fused residual + RK
epilogue launch per block and the diagnostics stay on the device as 0-d
tensors until the caller reads them.
"""

import ctypes

import numpy as np
import torch

from . import _native as nat

RK3_WEIGHTS = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))  # :456-461


class StageDiagnostics:
    """Device-side (rho_min, p_min, nonfinite) of include/fluid_gfx950.h."""

    def __init__(self, device):
        self.buf = torch.empty(3, dtype=torch.float64, device=device)

    def reset(self, stream):
        nat.check(nat.lib().fsg_diag_reset(nat.ptr(self.buf), stream))

    def values(self, density_floor, pressure_floor):
        rho_min = self.buf[0].clone()
        p_min = self.buf[1].clone()
        nonfinite = self.buf[2:3].view(torch.int32)[0]
        valid = (nonfinite == 0) & (rho_min >= density_floor) & (p_min >= pressure_floor)
        return valid, rho_min, p_min


class HybridSSPRK3:
    def __init__(self, operator, cfl=0.15, limiter=None, density_floor=1.0e-7,
                 pressure_floor=1.0e-7, maximum_rejections=16):
        if limiter is not None and not hasattr(limiter, "recompute"):
            raise NotImplementedError("only limiter=None or a MOOD limiter is supported")
        self.operator = operator
        self.solver = operator.solver
        self.cfl = float(cfl)
        self.limiter = limiter
        self.density_floor = float(density_floor)
        self.pressure_floor = float(pressure_floor)
        self.maximum_rejections = int(maximum_rejections)
        self._diag = StageDiagnostics(self.solver.device)
        self._buffers = None
        # MOOD stages from one residual pass (inviscid and viscous together,
        # the viscous rhs re-evaluated on troubled elements only); False
        # evaluates inviscid_rhs and viscous_rhs separately as the reference
        # does (the two differ by the rounding of inv + visc)
        self.fused_mood = True
        self.characteristic_lengths = self._characteristic_lengths()

    def _characteristic_lengths(self):
        """hybrid/time_integration.py:49-91: max(2V / sum of face areas,
        V^(1/3) / (4 (p+1))) per element."""
        topo = self.solver.mesh
        f = topo.faces
        p = self.solver.polynomial_degree
        out = {}
        for bid, b in enumerate(topo.blocks):
            area = np.zeros(b.count)
            np.add.at(area, f.owner_local[f.owner_block == bid], f.area[f.owner_block == bid])
            m = f.neighbor_block == bid
            np.add.at(area, f.neighbor_local[m], f.area[m])
            vol = self.operator.device_mesh.block_tables[bid]["volume"]
            L = np.maximum(2.0 * vol / area, 0.25 * vol ** (1.0 / 3.0) / (p + 1))
            out[b.element_type] = torch.as_tensor(L, dtype=self.solver.dtype,
                                                  device=self.solver.device)
        return out

    def stable_dt(self, states, boundary_state=None):
        """hybrid/time_integration.py:93-169 (no WALE), reduced on the device;
        returns a 0-d tensor clipped to [1e-7, 1e-3]."""
        s = self.solver
        k = 2 * s.polynomial_degree + 1
        rates = []
        for t in self.operator.block_types:
            U = states[t]
            rho = U[..., 0].clamp_min(1e-14)
            m = U[..., 1:4]
            p = ((s.gamma - 1.0) * (U[..., 4] - 0.5 * (m * m).sum(-1) / rho)).clamp_min(1e-14)
            speed = torch.linalg.vector_norm(m / rho[..., None], dim=-1) + \
                torch.sqrt(s.gamma * p / rho)
            L = self.characteristic_lengths[t]
            rate = k * speed.amax(1) / L
            if s.dynamic_viscosity > 0.0:
                rmin = rho.mean(1).clamp_min(0.1)
                diff = torch.maximum((4.0 / 3.0) * s.dynamic_viscosity / rmin,
                                     s.thermal_conductivity / (rmin * s.specific_heat_cp))
                rate = rate + 2.0 * self.operator.br2_stabilization * k ** 2 * diff / \
                    L.clamp_min(0.01) ** 2
            rates.append(torch.nan_to_num(rate, nan=1e7, posinf=1e7).amax())
        rate = torch.stack(rates).amax()
        return torch.clamp(self.cfl / rate.clamp_min(1e-12), 1e-7, 1e-3)

    def _stage(self, base, curr, nxt, dt, alpha, beta, boundary_state, reset):
        op = self.operator
        stream = nat.stream_handle(self.solver.device)
        if reset:
            self._diag.reset(stream)
        b = op._boundary(boundary_state)
        Ub, Uc = op._inputs(base), op._inputs(curr)
        Un = [nxt[t] for t in op.block_types]
        dt_elem = None
        if isinstance(dt, dict):
            dts = [dt[t].reshape(-1).to(self.solver.dtype).contiguous() for t in op.block_types]
            dt_elem = nat.ptr_array(dts)
            dt = 0.0
        nat.check(nat.lib().fsg_rk_stage(op.device_mesh.ctx, nat.ptr_array(Ub), nat.ptr_array(Uc),
                                         nat.ptr_array(Un), float(dt), dt_elem, float(alpha),
                                         float(beta), op.rhs_terms(), ctypes.byref(b),
                                         nat.ptr(self._diag.buf), stream))

    def _mood_stage(self, base, curr, nxt, dt, alpha, beta, boundary_state):
        """The MOOD branch of hybrid/time_integration.py:404-439: inviscid and
        viscous rhs, then the fused detection / subcell recompute / positivity
        / blend / positivity / admissibility sequence."""
        op = self.operator
        stream = nat.stream_handle(self.solver.device)
        self._diag.reset(stream)
        viscous = hasattr(op, "inviscid_rhs") and self.solver.dynamic_viscosity > 0.0
        if viscous and self.fused_mood:
            total = op.rhs(curr, boundary_state)
            visc = {t: torch.empty_like(curr[t]) for t in op.block_types}
            work = {t: torch.empty_like(curr[t]) for t in op.block_types}
            self.limiter.stage_fused(base, curr, total, visc, work, nxt, dt, alpha, beta,
                                     boundary_state, nat.ptr(self._diag.buf))
            return
        if hasattr(op, "inviscid_rhs"):
            inv = op.inviscid_rhs(curr, boundary_state)
            # a zero viscosity gives an exactly zero viscous rhs: skip it
            visc = (op.viscous_rhs(curr, boundary_state) if self.solver.dynamic_viscosity > 0.0
                    else None)
        else:
            inv, visc = op.rhs(curr, boundary_state), None
        work = {t: torch.empty_like(curr[t]) for t in op.block_types}
        self.limiter.stage(base, curr, inv, visc, work, nxt, dt, alpha, beta, boundary_state,
                           nat.ptr(self._diag.buf))

    def rk_stage(self, base_states, current_states, dt, base_weight, current_weight,
                 boundary_state=None):
        """hybrid/time_integration.py:404-439."""
        nxt = {t: torch.empty_like(current_states[t]) for t in self.operator.block_types}
        if self.limiter is not None:
            self._mood_stage(base_states, current_states, nxt, dt, base_weight, current_weight,
                             boundary_state)
        else:
            self._stage(base_states, current_states, nxt, dt, base_weight, current_weight,
                        boundary_state, True)
        return (nxt, *self._diag.values(self.density_floor, self.pressure_floor))

    def checked_step(self, states, dt, boundary_state=None):
        """hybrid/time_integration.py:441-583.  Without a limiter: three fused
        stages, valid = all stages valid, minima over the stages.  With MOOD
        (:463-554): three MOOD stages, diagnostics of the last one."""
        types = self.operator.block_types
        if self.limiter is not None:
            s = states
            for a, b in RK3_WEIGHTS:
                nxt = {t: torch.empty_like(states[t]) for t in types}
                self._mood_stage(states, s, nxt, dt, a, b, boundary_state)
                s = nxt
            return (s, *self._diag.values(self.density_floor, self.pressure_floor))
        if self._buffers is None or any(self._buffers[0][t].shape != states[t].shape
                                        for t in types):
            self._buffers = [{t: torch.empty_like(states[t]) for t in types} for _ in range(2)]
        s1, s2 = self._buffers
        out = {t: torch.empty_like(states[t]) for t in types}
        (a1, b1), (a2, b2), (a3, b3) = RK3_WEIGHTS
        self._stage(states, states, s1, dt, a1, b1, boundary_state, True)
        self._stage(states, s1, s2, dt, a2, b2, boundary_state, False)
        self._stage(states, s2, out, dt, a3, b3, boundary_state, False)
        return (out, *self._diag.values(self.density_floor, self.pressure_floor))

    def step(self, states, dt, boundary_state=None):
        """hybrid/time_integration.py:660-671."""
        return self.checked_step(states, dt, boundary_state)[0]
