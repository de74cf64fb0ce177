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
        if limiter is not None and not (hasattr(limiter, "recompute")
                                        or hasattr(limiter, "apply")):
            raise NotImplementedError("limiter must be None, a MOOD limiter or the modal "
                                      "shock limiter")
        self.operator = operator
        self.solver = operator.solver
        self.cfl = float(cfl)
        self.limiter = limiter
        self.density_floor = float(density_floor)
        self.pressure_floor = float(pressure_floor)
        self.maximum_rejections = int(maximum_rejections)
        self._diag = StageDiagnostics(self.solver.device)
        self._scratch_diag = StageDiagnostics(self.solver.device)
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
        self._lengths64 = {}
        for bid, b in enumerate(topo.blocks):
            area = np.zeros(b.count)
            np.add.at(area, f.owner_local[f.owner_block == bid], f.area[f.owner_block == bid])
            m = f.neighbor_block == bid
            np.add.at(area, f.neighbor_local[m], f.area[m])
            vol = self.operator.device_mesh.block_tables[bid]["volume"]
            L = np.maximum(2.0 * vol / area, 0.25 * vol ** (1.0 / 3.0) / (p + 1))
            self._lengths64[b.element_type] = L
            out[b.element_type] = torch.as_tensor(L, dtype=self.solver.dtype,
                                                  device=self.solver.device)
        return out

    # -- time-step control on the device (csrc/fsg_dt.hip) ---------------------
    def _dt_context(self):
        """The native dt context: per-block characteristic lengths, sliver
        factors, |det J| and basis integrals (FP64), the interior-face groups
        of the local-dt smoothing in the order of
        hybrid/time_integration.py:27-47, and its FP64 work buffer."""
        if getattr(self, "_dtc", None) is not None:
            return self._dtc
        from .elements import reference_tables
        op = self.operator
        topo = self.solver.mesh
        dev = self.solver.device
        keep = []

        def t64(a, dtype=torch.float64):
            x = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev).contiguous()
            keep.append(x)
            return x

        blocks = (nat.DtBlock * len(topo.blocks))()
        for b, (blk, bt) in enumerate(zip(topo.blocks, op.device_mesh.block_tables)):
            R = reference_tables(blk.element_type, self.solver.polynomial_degree)
            vol = bt["volume"]
            blocks[b].length = t64(self._lengths64[blk.element_type]).data_ptr()
            blocks[b].sliver = t64(np.minimum(1.0, vol / 1.0e-4)).data_ptr()
            blocks[b].detw = t64(np.abs(blk.det)).data_ptr()
            blocks[b].node_weights = t64(R["weights"] @ R["B"]).data_ptr()
        f = topo.faces
        conns = []
        for ft in (1, 0):                       # triangles, then quads (the mesh's order)
            inner = np.flatnonzero((f.face_type == ft) & (f.neighbor_block >= 0))
            if not len(inner):
                continue
            ob, nb = f.owner_block[inner], f.neighbor_block[inner]
            for o in np.unique(ob):
                for n in np.unique(nb):
                    sel = inner[(ob == o) & (nb == n)]
                    if len(sel):
                        conns.append((int(o), int(n), t64(f.owner_local[sel], torch.int32),
                                      t64(f.neighbor_local[sel], torch.int32)))
        carr = (nat.DtConn * max(len(conns), 1))()
        for k, (o, n, oc, nc) in enumerate(conns):
            carr[k].owner_block, carr[k].neighbor_block = o, n
            carr[k].count = oc.numel()
            carr[k].owner_cells, carr[k].neighbor_cells = oc.data_ptr(), nc.data_ptr()
        L = nat.lib()
        size = L.fsg_dt_work_size(op.device_mesh.ctx, len(conns), carr)
        work = torch.zeros(int(size), dtype=torch.float64, device=dev)
        keep.append(work)
        handle = ctypes.c_void_p()
        nat.check_dt(L.fsg_dt_create(op.device_mesh.ctx, blocks, len(conns), carr,
                                     work.data_ptr(), int(size), ctypes.byref(handle)))
        self._dtc = dict(handle=handle, keep=keep, conns=conns)
        return self._dtc

    def __del__(self):
        try:
            c = getattr(self, "_dtc", None)
            if c is not None:
                nat.lib().fsg_dt_destroy(c["handle"])
        except Exception:
            pass

    def stable_dt(self, states, boundary_state=None):
        """hybrid/time_integration.py:93-169 (with the WALE viscosity of the
        resolved gradient when the solver has use_wale): one wave per element
        forms its rate and an atomic maximum reduces them on the device;
        returns a 0-d FP64 tensor clipped to [1e-7, 1e-3] (no host
        synchronisation)."""
        c = self._dt_context()
        out = torch.empty((), dtype=torch.float64, device=self.solver.device)
        U = self.operator._inputs(states)
        nat.check_dt(nat.lib().fsg_stable_dt(c["handle"], nat.ptr_array(U), self.cfl,
                                             nat.ptr(out), None,
                                             nat.stream_handle(self.solver.device)))
        return out

    def local_dt(self, states, cfl=None, boundary_state=None):
        """hybrid/time_integration.py:171-262: element-local dt, neighbour
        smoothing (ratio 1.3, two rounds) and sliver scaling, on the device;
        {type: (E, 1, 1)} in the solver dtype."""
        c = self._dt_context()
        U = self.operator._inputs(states)
        out = {t: torch.empty(u.shape[0], dtype=self.solver.dtype, device=self.solver.device)
               for t, u in zip(self.operator.block_types, U)}
        nat.check_dt(nat.lib().fsg_local_dt(
            c["handle"], nat.ptr_array(U), float(self.cfl if cfl is None else cfl),
            nat.ptr_array([out[t] for t in self.operator.block_types]), None, None,
            nat.stream_handle(self.solver.device)))
        return {t: v[:, None, None] for t, v in out.items()}

    def conserved_totals(self, states):
        """sum over elements of int_K U dV per variable, accumulated in FP64
        on the device in a fixed order; (nvars,) FP64 tensor."""
        c = self._dt_context()
        out = torch.empty(self.solver.layout.number_of_variables, dtype=torch.float64,
                          device=self.solver.device)
        nat.check_dt(nat.lib().fsg_conserved_totals(
            c["handle"], nat.ptr_array(self.operator._inputs(states)), nat.ptr(out),
            nat.stream_handle(self.solver.device)))
        return out

    def advance(self, states, target_time, boundary_state=None, coupler=None, lts=False,
                cfl_interval=1, maximum_steps=10 ** 9, maximum_rejections=None, sync_every=16):
        """The step loop of run_hybrid_sphere.py:265-346 (without soft start,
        output or GUI) on the device: per attempt the CFL dt (refreshed every
        ``cfl_interval`` accepted steps) or the local dt, scaled by the retry
        factor and clipped to the target time, three SSP-RK3 stages, the
        chemistry ``coupler`` on the candidate, and the accept / reject
        decision from the stage diagnostics; the host reads the loop state
        once per ``sync_every`` attempts.  With ``lts`` the local dt is
        recomputed at every attempt (cfl_interval must be 1).  Returns
        (states, time, steps, rejections); raises FloatingPointError after
        ``maximum_rejections`` consecutive rejections, leaving the states at
        the last accepted step."""
        if lts and cfl_interval != 1:
            raise ValueError("the device LTS loop refreshes the local dt at every attempt")
        c = self._dt_context()
        L = nat.lib()
        dev = self.solver.device
        stream = nat.stream_handle(dev)
        types = self.operator.block_types
        U = {t: states[t].clone() for t in types}
        Uarr = nat.ptr_array([U[t] for t in types])
        dte = {t: torch.zeros(U[t].shape[0], dtype=self.solver.dtype, device=dev) for t in types}
        dt64 = ({t: dte[t] for t in types} if self.solver.dtype == torch.float64 else
                {t: torch.zeros(U[t].shape[0], dtype=torch.float64, device=dev) for t in types})
        dte_arr = nat.ptr_array([dte[t] for t in types])
        dt64_arr = nat.ptr_array([dt64[t] for t in types])
        st = nat.StepState()
        st.retry, st.target_time, st.min_dt = 1.0, float(target_time), float("inf")
        st.maximum_steps = int(maximum_steps)
        st.maximum_rejections = int(self.maximum_rejections if maximum_rejections is None
                                    else maximum_rejections)
        st.refresh_interval = int(cfl_interval)
        buf = torch.frombuffer(bytearray(bytes(st)), dtype=torch.uint8).to(dev)
        sp = ctypes.c_void_p(buf.data_ptr())
        dt_dict = {t: dte[t] for t in types}
        estimate = None
        while True:
            n = sync_every if estimate is None else max(1, min(sync_every, estimate))
            for _ in range(n):
                if not lts:
                    nat.check_dt(L.fsg_stable_dt(c["handle"], Uarr, self.cfl, None, sp, stream))
                nat.check_dt(L.fsg_step_begin(c["handle"], sp, dte_arr, dt64_arr, int(bool(lts)),
                                              Uarr, self.cfl, stream))
                cand = self.checked_step(U, dt_dict, boundary_state)[0]
                if coupler is not None:
                    cand = {t: coupler.react_block(cand[t], dt64[t]) for t in types}
                nat.check_dt(L.fsg_step_commit(
                    c["handle"], sp, nat.ptr(self._diag.buf), self.density_floor,
                    self.pressure_floor, nat.ptr_array([cand[t] for t in types]), Uarr, stream))
            host = nat.StepState.from_buffer_copy(bytes(buf.cpu().numpy().tobytes()))
            if host.status != 0:
                raise FloatingPointError(
                    "Unable to find an admissible SSP-RK3 step: "
                    f"t={host.time:.8g}, rejected={host.rejections}, "
                    f"rho_min={host.last_rho_min:.6g}, p_min={host.last_p_min:.6g}. "
                    "The state remains at the last accepted step.")
            if host.time >= host.target_time or host.step >= host.maximum_steps:
                break
            # attempts still needed at the present dt (plus one for a rejection)
            dt_now = max(host.retry * host.cfl_dt if not lts else host.dt, 1e-300)
            estimate = int(min(host.maximum_steps - host.step,
                               np.ceil((host.target_time - host.time) / dt_now))) + 1
        self.last_loop = host
        return U, host.time, int(host.step), int(host.rejections)

    def _stage(self, base, curr, nxt, dt, alpha, beta, boundary_state, reset):
        """One fused residual + RK-epilogue stage into nxt; with a
        limiter-after-stage (the modal shock limiter), the unlimited stage
        goes to a temporary and the limiter writes nxt and the diagnostics
        (hybrid/time_integration.py:432-439)."""
        op = self.operator
        stream = nat.stream_handle(self.solver.device)
        if reset:
            self._diag.reset(stream)
        if self.limiter is not None and hasattr(self.limiter, "apply"):
            tmp = {t: torch.empty_like(nxt[t]) for t in op.block_types}
            self._scratch_diag.reset(stream)
            self._stage_kernel(base, curr, tmp, dt, alpha, beta, boundary_state,
                               self._scratch_diag)
            self.limiter.apply(tmp, nxt, nat.ptr(self._diag.buf))
            return
        self._stage_kernel(base, curr, nxt, dt, alpha, beta, boundary_state, self._diag)

    def _stage_kernel(self, base, curr, nxt, dt, alpha, beta, boundary_state, diag):
        op = self.operator
        stream = nat.stream_handle(self.solver.device)
        b = op._boundary(boundary_state)
        Ub = None if base is None else op._inputs(base)
        Uc = op._inputs(curr)
        Un = [nxt[t] for t in op.block_types]
        dt_elem = None
        if isinstance(dt, dict):
            dts = [dt[t].reshape(-1).to(self.solver.dtype).contiguous() for t in op.block_types]
            dt_elem = nat.ptr_array(dts)
            dt = 0.0
        nat.check(nat.lib().fsg_rk_stage(op.device_mesh.ctx,
                                         None if base is None else nat.ptr_array(Ub),
                                         nat.ptr_array(Uc), nat.ptr_array(Un), float(dt), dt_elem,
                                         float(alpha), float(beta), op.rhs_terms(),
                                         ctypes.byref(b),
                                         None if diag is None else nat.ptr(diag.buf), stream))

    def _mood_stage(self, base, curr, nxt, dt, alpha, beta, boundary_state):
        """The MOOD branch of hybrid/time_integration.py:404-439: inviscid and
        viscous rhs, then the fused detection / subcell recompute / positivity
        / blend / positivity / admissibility sequence."""
        op = self.operator
        stream = nat.stream_handle(self.solver.device)
        self._diag.reset(stream)
        has_visc = self.solver.dynamic_viscosity > 0.0 or self.solver.use_wale
        viscous = hasattr(op, "inviscid_rhs") and has_visc
        if viscous and self.fused_mood:
            # the residual's epilogue writes the Euler candidate curr + dt rhs
            # (no base term, no diagnostics) straight into work, which the
            # detection then reads in place of rhs and curr
            visc = {t: torch.empty_like(curr[t]) for t in op.block_types}
            work = {t: torch.empty_like(curr[t]) for t in op.block_types}
            self._stage_kernel(None, curr, work, dt, 0.0, 1.0, boundary_state, None)
            self.limiter.stage_fused(base, curr, None, visc, work, nxt, dt, alpha, beta,
                                     boundary_state, nat.ptr(self._diag.buf))
            return
        if hasattr(op, "inviscid_rhs"):
            inv = op.inviscid_rhs(curr, boundary_state)
            # a zero viscosity without WALE gives an exactly zero viscous rhs: skip it
            visc = op.viscous_rhs(curr, boundary_state) if has_visc else None
        else:
            inv, visc = op.rhs(curr, boundary_state), None
        work = {t: torch.empty_like(curr[t]) for t in op.block_types}
        self.limiter.stage(base, curr, inv, visc, work, nxt, dt, alpha, beta, boundary_state,
                           nat.ptr(self._diag.buf))

    def _mood(self):
        return self.limiter is not None and hasattr(self.limiter, "recompute")

    def rk_stage(self, base_states, current_states, dt, base_weight, current_weight,
                 boundary_state=None):
        """hybrid/time_integration.py:404-439."""
        nxt = {t: torch.empty_like(current_states[t]) for t in self.operator.block_types}
        if self._mood():
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
        if self._mood():
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
