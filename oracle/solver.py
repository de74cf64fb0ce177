"""Pointwise compressible physics, restated from hybrid/solver.py (numpy, FP64)."""

import numpy as np


class CompressibleSolver:
    """hybrid/solver.py:12-98 without the mesh: gamma, R, mu, Pr, species count."""

    def __init__(self, gamma=1.4, gas_constant=1.0, dynamic_viscosity=0.0,
                 prandtl=0.72, num_species=0, inviscid_flux="hllc", use_wale=False,
                 wale_constant=0.5, turbulent_prandtl=0.9):
        self.gamma = float(gamma)
        self.gas_constant = float(gas_constant)
        self.dynamic_viscosity = float(dynamic_viscosity)
        self.prandtl = float(prandtl)
        self.num_species = int(num_species)
        self.inviscid_flux = inviscid_flux
        # hybrid/solver.py:30-47
        self.use_wale = bool(use_wale)
        self.wale_constant = float(wale_constant)
        self.turbulent_prandtl = float(turbulent_prandtl)
        # hybrid/solver.py:48-53
        self.specific_heat_cp = self.gamma * self.gas_constant / (self.gamma - 1.0)
        self.thermal_conductivity = (self.dynamic_viscosity * self.specific_heat_cp
                                     / self.prandtl)

    @property
    def nvars(self):
        return 5 + self.num_species

    def primitive_to_conserved(self, W):
        """hybrid/solver.py:63-81."""
        rho = W[..., 0]
        vel = W[..., 1:4]
        E = W[..., 4] / (self.gamma - 1.0) + 0.5 * rho * np.sum(vel ** 2, -1)
        parts = [rho[..., None], rho[..., None] * vel, E[..., None]]
        if self.num_species:
            parts.append(rho[..., None] * W[..., 5:])
        return np.concatenate(parts, -1)

    def conserved_to_primitive(self, U):
        """hybrid/solver.py:83-98: rho floored at 1e-14 only for division."""
        rho = U[..., 0]
        m = U[..., 1:4]
        safe = np.maximum(rho, 1.0e-14)
        vel = m / safe[..., None]
        ke = 0.5 * np.sum(m ** 2, -1) / safe
        p = (self.gamma - 1.0) * (U[..., 4] - ke)
        parts = [rho[..., None], vel, p[..., None]]
        if self.num_species:
            parts.append(U[..., 5:] / safe[..., None])
        return np.concatenate(parts, -1)

    def normal_physical_flux(self, U, n):
        """hybrid/solver.py:100-122."""
        W = self.conserved_to_primitive(U)
        rho, vel, p = W[..., 0], W[..., 1:4], W[..., 4]
        un = np.sum(vel * n, -1)
        parts = [(rho * un)[..., None],
                 rho[..., None] * vel * un[..., None] + p[..., None] * n,
                 (un * (U[..., 4] + p))[..., None]]
        if self.num_species:
            parts.append(U[..., 5:] * un[..., None])
        return np.concatenate(parts, -1)

    def physical_flux_tensor(self, U):
        """hybrid/solver.py:124-130: (..., 3, V)."""
        eye = np.eye(3)
        return np.stack([self.normal_physical_flux(U, eye[d]) for d in range(3)], -2)

    def rusanov_flux(self, UL, UR, n):
        """hybrid/solver.py:132-152."""
        WL = self.conserved_to_primitive(UL)
        WR = self.conserved_to_primitive(UR)
        unl = np.sum(WL[..., 1:4] * n, -1)
        unr = np.sum(WR[..., 1:4] * n, -1)
        rl = np.maximum(WL[..., 0], 1.0e-14)
        rr = np.maximum(WR[..., 0], 1.0e-14)
        pl = np.maximum(WL[..., 4], 0.0)
        pr = np.maximum(WR[..., 4], 0.0)
        cl = np.sqrt(self.gamma * pl / rl)
        cr = np.sqrt(self.gamma * pr / rr)
        smax = np.maximum(np.abs(unl) + cl, np.abs(unr) + cr)
        return (0.5 * (self.normal_physical_flux(UL, n) + self.normal_physical_flux(UR, n))
                - 0.5 * smax[..., None] * (UR - UL))

    @staticmethod
    def _safe_signed(x, tol=1.0e-14):
        """hybrid/solver.py:154-161."""
        return np.where(np.abs(x) > tol, x, np.where(x >= 0.0, tol, -tol))

    def hllc_wave_speeds(self, UL, UR, n):
        """hybrid/solver.py:163-207."""
        WL = self.conserved_to_primitive(UL)
        WR = self.conserved_to_primitive(UR)
        rl, rr = WL[..., 0], WR[..., 0]
        pl, pr = WL[..., 4], WR[..., 4]
        unl = np.sum(WL[..., 1:4] * n, -1)
        unr = np.sum(WR[..., 1:4] * n, -1)
        with np.errstate(invalid="ignore", divide="ignore"):
            cl = np.sqrt(np.maximum(self.gamma * pl / np.maximum(rl, 1e-14), 0.0))
            cr = np.sqrt(np.maximum(self.gamma * pr / np.maximum(rr, 1e-14), 0.0))
        sl = np.minimum(unl - cl, unr - cr)
        sr = np.maximum(unl + cl, unr + cr)
        den = rl * (sl - unl) - rr * (sr - unr)
        sm = (pr - pl + rl * unl * (sl - unl) - rr * unr * (sr - unr)) / self._safe_signed(den)
        return sl, sm, sr

    def hllc_star_state(self, U, W, s, sm, n):
        """hybrid/solver.py:209-243."""
        rho, vel, p = W[..., 0], W[..., 1:4], W[..., 4]
        un = np.sum(vel * n, -1)
        den = self._safe_signed(s - sm)
        comp = (s - un) / den
        rs = rho * comp
        vs = vel + (sm - un)[..., None] * n
        ps = p + rho * (s - un) * (sm - un)
        Es = ((s - un) * U[..., 4] - p * un + ps * sm) / den
        parts = [rs[..., None], rs[..., None] * vs, Es[..., None]]
        if self.num_species:
            parts.append(U[..., 5:] * comp[..., None])
        return np.concatenate(parts, -1)

    def hllc_flux(self, UL, UR, n):
        """hybrid/solver.py:245-300 including the Rusanov fallback and the
        bitwise identical-state shortcut."""
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            WL = self.conserved_to_primitive(UL)
            WR = self.conserved_to_primitive(UR)
            FL = self.normal_physical_flux(UL, n)
            FR = self.normal_physical_flux(UR, n)
            sl, sm, sr = self.hllc_wave_speeds(UL, UR, n)
            SL = self.hllc_star_state(UL, WL, sl, sm, n)
            SR = self.hllc_star_state(UR, WR, sr, sm, n)
            FSL = FL + sl[..., None] * (SL - UL)
            FSR = FR + sr[..., None] * (SR - UR)
            h = np.where((sl >= 0)[..., None], FL,
                         np.where((sm >= 0)[..., None], FSL,
                                  np.where((sr > 0)[..., None], FSR, FR)))
            valid_in = (WL[..., 0] > 0) & (WR[..., 0] > 0) & (WL[..., 4] > 0) & (WR[..., 4] > 0)
            valid = valid_in & np.all(np.isfinite(h), -1)
            robust = np.where(valid[..., None], h, self.rusanov_flux(UL, UR, n))
            same = np.all(UL == UR, -1)
            return np.where(same[..., None], FL, robust)

    def numerical_flux(self, UL, UR, n):
        """hybrid/solver.py:302-305."""
        if self.inviscid_flux == "hllc":
            return self.hllc_flux(UL, UR, n)
        return self.rusanov_flux(UL, UR, n)

    def viscous_variables(self, U):
        """hybrid/solver.py:307-315: [u, v, w, T = p / (rho R)]."""
        W = self.conserved_to_primitive(U)
        T = W[..., 4] / (np.maximum(W[..., 0], 1.0e-14) * self.gas_constant)
        return np.concatenate([W[..., 1:4], T[..., None]], -1)

    def wale_dynamic_viscosity(self, U, grad, width):
        """hybrid/solver.py:317-355: WALE eddy viscosity rho nu_t from the
        velocity gradient grad[..., 0:3, :], with the Ducros sensor and the
        clips of the reference."""
        g = grad[..., 0:3, :]
        S = 0.5 * (g + np.swapaxes(g, -1, -2))
        g2 = np.einsum("...ik,...kj->...ij", g, g)
        sym2 = 0.5 * (g2 + np.swapaxes(g2, -1, -2))
        tr2 = np.trace(g2, axis1=-2, axis2=-1)
        Sd = sym2 - tr2[..., None, None] * np.eye(3) / 3.0
        snorm = np.sum(S ** 2, axis=(-2, -1))
        div = g[..., 0, 0] + g[..., 1, 1] + g[..., 2, 2]
        cx = g[..., 2, 1] - g[..., 1, 2]
        cy = g[..., 0, 2] - g[..., 2, 0]
        cz = g[..., 1, 0] - g[..., 0, 1]
        vort = cx ** 2 + cy ** 2 + cz ** 2
        ducros = vort / (vort + div ** 2 + 1.0e-12)
        sdnorm = np.sum(Sd ** 2, axis=(-2, -1))
        ss = np.clip(snorm, 0.0, 1.0e8)
        st = np.clip(sdnorm, 0.0, 1.0e10)
        with np.errstate(over="ignore", invalid="ignore"):
            nu = ((self.wale_constant * width) ** 2 * st ** 1.5
                  / (ss ** 2.5 + st ** 1.25 + 1.0e-30)) * ducros
        nu = np.nan_to_num(np.clip(nu, 0.0, 50.0), nan=0.0, posinf=50.0)
        return np.maximum(U[..., 0], 0.0) * nu

    def viscous_physical_flux(self, U, grad, width=None, turbulent_gradient=None):
        """hybrid/solver.py:357-406: grad is (..., 4, 3); with WALE and a
        filter width the eddy viscosity (from turbulent_gradient, the
        resolved polynomial gradient, when given) adds to mu and, over the
        turbulent Prandtl number, to kappa."""
        W = self.conserved_to_primitive(U)
        vel = W[..., 1:4]
        mu = self.dynamic_viscosity
        kappa = self.thermal_conductivity
        if self.use_wale and width is not None:
            mt = self.wale_dynamic_viscosity(
                U, grad if turbulent_gradient is None else turbulent_gradient, width)
            mu = mu + mt[..., None, None]
            kappa = kappa + mt * self.specific_heat_cp / self.turbulent_prandtl
        gu = grad[..., 0:3, :]
        gT = grad[..., 3, :]
        div = np.trace(gu, axis1=-2, axis2=-1)
        tau = mu * (gu + np.swapaxes(gu, -1, -2)
                    - (2.0 / 3.0) * div[..., None, None] * np.eye(3))
        mom = np.swapaxes(tau, -1, -2)
        work = np.einsum("...i,...id->...d", vel, tau)
        flux = np.zeros(U.shape[:-1] + (3, U.shape[-1]))
        flux[..., :, 1:4] = mom
        flux[..., :, 4] = work + np.asarray(kappa)[..., None] * gT
        return flux


class FarfieldWallBoundary:
    """hybrid/boundary.py:6-102 with tag sets given directly."""

    def __init__(self, solver, farfield_primitive, inlet_tags=(), outlet_tags=(),
                 wall_tags=(), no_slip_tags=(), slip_tags=(), wall_temperature=None):
        self.solver = solver
        self.farfield_primitive = np.asarray(farfield_primitive, float)
        self.inlet_tags = np.asarray(inlet_tags, int)
        self.outlet_tags = np.asarray(outlet_tags, int)
        self.wall_tags = np.asarray(wall_tags, int)
        self.no_slip_tags = np.asarray(no_slip_tags, int)
        self.slip_tags = np.asarray(slip_tags, int)
        self.wall_temperature = wall_temperature

    @staticmethod
    def _matches(tags, choices):
        """hybrid/boundary.py:48-52."""
        if choices.size == 0:
            return np.zeros(tags.shape, bool)
        return np.any(tags[..., None] == choices, -1)

    def __call__(self, U_in, n, tags):
        """hybrid/boundary.py:54-72."""
        s = self.solver
        W = s.conserved_to_primitive(U_in)
        ghost = W
        wall = self._matches(tags, self.wall_tags)[:, None]
        vel = W[..., 1:4]
        un = np.sum(vel * n, -1)
        Ww = W.copy()
        Ww[..., 1:4] = vel - 2.0 * un[..., None] * n
        ghost = np.where(wall[..., None], Ww, ghost)
        inlet = self._matches(tags, self.inlet_tags)[:, None]
        ghost = np.where(inlet[..., None], np.broadcast_to(self.farfield_primitive, W.shape), ghost)
        return s.primitive_to_conserved(ghost)

    def viscous_ghost(self, U_in, n, tags):
        """hybrid/boundary.py:74-102."""
        s = self.solver
        v = s.viscous_variables(U_in)
        ghost = v
        ns = self._matches(tags, self.no_slip_tags)[:, None]
        vn = v.copy()
        vn[..., 0:3] = -v[..., 0:3]
        if self.wall_temperature is not None:
            vn[..., 3] = 2.0 * self.wall_temperature - v[..., 3]
        ghost = np.where(ns[..., None], vn, ghost)
        slip = self._matches(tags, self.slip_tags)[:, None]
        vel = v[..., 0:3]
        un = np.sum(vel * n, -1)
        vs = v.copy()
        vs[..., 0:3] = vel - 2.0 * un[..., None] * n
        ghost = np.where(slip[..., None], vs, ghost)
        inlet = self._matches(tags, self.inlet_tags)[:, None]
        fU = s.primitive_to_conserved(np.broadcast_to(self.farfield_primitive, U_in.shape))
        ghost = np.where(inlet[..., None], s.viscous_variables(fU), ghost)
        return ghost
