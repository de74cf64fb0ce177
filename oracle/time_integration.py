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
        """hybrid/time_integration.py:93-169."""
        rates = [np.nan_to_num(r, nan=1e7, posinf=1e7).max()
                 for r in self._element_rates(states).values()]
        return float(np.clip(self.cfl / max(max(rates), 1e-12), 1e-7, 1e-3))

    def internal_connections(self):
        """hybrid/time_integration.py:27-47: (owner type, owner cells, neighbour
        type, neighbour cells) per face type (mesh order) and owner x
        neighbour block pair (ascending)."""
        mesh = self.operator.mesh
        out = []
        for ftype, fb in mesh.face_blocks.items():
            inner = np.flatnonzero(fb.neighbor_block >= 0)
            if not len(inner):
                continue
            ob, nb = fb.owner_block[inner], fb.neighbor_block[inner]
            for o in np.unique(ob):
                for n in np.unique(nb):
                    sel = inner[(ob == o) & (nb == n)]
                    if len(sel):
                        out.append((mesh.element_blocks[o].element_type, fb.owner_local[sel],
                                    mesh.element_blocks[n].element_type, fb.neighbor_local[sel]))
        return out

    def _element_rates(self, states):
        """Per-element stability rate of hybrid/time_integration.py:93-169,
        before the NaN replacement."""
        s = self.solver
        k = 2 * self.operator.degree + 1
        out = {}
        for b in self.operator.mesh.element_blocks:
            t = b.element_type
            W = s.conserved_to_primitive(states[t])
            rho = np.maximum(W[..., 0], 1e-14)
            pr = np.maximum(W[..., 4], 1e-14)
            speed = np.linalg.norm(W[..., 1:4], axis=-1) + np.sqrt(s.gamma * pr / rho)
            L = self.characteristic_lengths[t]
            rate = k * speed.max(1) / L
            if s.dynamic_viscosity > 0.0 or getattr(s, "use_wale", False):
                rmin = np.maximum(rho.mean(1), 0.1)
                mu = np.full(rmin.shape, s.dynamic_viscosity)
                kappa = np.full(rmin.shape, s.thermal_conductivity)
                if getattr(s, "use_wale", False):
                    # :119-148: the WALE viscosity of the resolved gradient,
                    # maximum over the volume points
                    bid = [b.element_type for b in self.operator.mesh.element_blocks].index(t)
                    geo = self.operator.geometry
                    ref = geo.reference_elements[t]
                    m = geo.element_metrics[bid]
                    v = s.viscous_variables(states[t])
                    grad = np.einsum("eim,eqid->eqmd", v, m.physical_basis_gradients)
                    Uq = np.einsum("qi,eiv->eqv", ref.basis(ref.volume_points), states[t])
                    width = m.integration_weights.sum(1) ** (1.0 / 3.0) / (self.operator.degree + 1)
                    turb = s.wale_dynamic_viscosity(Uq, grad, width[:, None]).max(1)
                    mu = mu + turb
                    kappa = kappa + turb * s.specific_heat_cp / s.turbulent_prandtl
                diff = np.maximum((4.0 / 3.0) * mu / rmin, kappa / (rmin * s.specific_heat_cp))
                pen = float(getattr(self.operator, "br2_stabilization", 1.0))
                rate = rate + 2.0 * pen * k ** 2 * diff / np.maximum(L, 0.01) ** 2
            out[t] = rate
        return out

    def local_dt(self, states, cfl=None):
        """hybrid/time_integration.py:171-262: per-element cfl / rate, two
        rounds of neighbour smoothing (ratio 1.3, each connection group reading
        both sides before updating them, groups in order), then the sliver
        factor min(1, V / 1e-4).  Returns {type: (E, 1, 1)}."""
        c = self.cfl if cfl is None else float(cfl)
        with np.errstate(invalid="ignore", divide="ignore"):
            dts = {t: c / np.maximum(r, 1e-12) for t, r in self._element_rates(states).items()}
        alpha = 1.3
        conns = self.internal_connections()
        for _ in range(2):
            for ot, oc, nt, nc in conns:
                d_own = dts[ot][oc].copy()
                d_nbr = dts[nt][nc].copy()
                np.minimum.at(dts[ot], oc, alpha * d_nbr)
                np.minimum.at(dts[nt], nc, alpha * d_own)
        geo = self.operator.geometry
        for bid, b in enumerate(self.operator.mesh.element_blocks):
            vol = geo.element_metrics[bid].integration_weights.sum(1)
            dts[b.element_type] = dts[b.element_type] * np.minimum(1.0, vol / 1.0e-4)
        return {t: v[:, None, None] for t, v in dts.items()}

    def conserved_totals(self, states):
        """sum over elements of int_K U dV per variable (FP64)."""
        geo = self.operator.geometry
        tot = 0.0
        for bid, b in enumerate(self.operator.mesh.element_blocks):
            ref = geo.reference_elements[b.element_type]
            w = geo.element_metrics[bid].integration_weights
            tot = tot + np.einsum("qi,eiv,eq->v", ref.basis(ref.volume_points),
                                  states[b.element_type], w)
        return tot

    def advance(self, states, target_time, boundary_state=None, coupler=None, lts=False,
                cfl_interval=1, maximum_steps=10 ** 9, maximum_rejections=16):
        """The step loop of run_hybrid_sphere.py:265-346 without soft start,
        output or GUI: dt from stable_dt (or local_dt) every cfl_interval
        steps, clipped to the target time, halved on a rejected step and
        grown by 1.25 (up to the CFL value) on an accepted one; the chemistry
        coupler after each accepted step.  Returns (states, time, steps,
        rejections)."""
        t, step, rej, consecutive, retry = 0.0, 0, 0, 0, 1.0
        cached = None
        while t < target_time and step < maximum_steps:
            if cached is None or step % cfl_interval == 0:
                cached = self.local_dt(states) if lts else self.stable_dt(states)
            if lts:
                dtd = {k: retry * v for k, v in cached.items()}
                dt = float(min(float(v.min()) for v in dtd.values()))
                adv = dtd
            else:
                dt = min(retry * cached, target_time - t)
                adv = dt
            if not np.isfinite(dt) or dt <= 0.0:
                raise FloatingPointError(f"invalid timestep at step {step}: {dt}")
            cand, ok, rmin, pmin = self.checked_step(states, adv, boundary_state)
            if ok:
                if coupler is not None:
                    reacted = {k: coupler.react_block(v, adv[k] if lts else dt)
                               for k, v in cand.items()}
                    finite = all(np.all(np.isfinite(v)) for v in reacted.values())
                    cand = reacted if finite else cand
                states = cand
                t += dt
                step += 1
                consecutive = 0
                retry = min(1.0, 1.25 * retry)
                if target_time - t <= 32.0 * np.finfo(np.float32).eps * max(abs(target_time),
                                                                             1e-12):
                    t = target_time
            else:
                rej += 1
                consecutive += 1
                retry *= 0.5
                if consecutive >= maximum_rejections:
                    raise FloatingPointError("Unable to find an admissible SSP-RK3 step")
        return states, t, step, rej

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
